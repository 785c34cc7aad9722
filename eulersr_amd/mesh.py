"""Unstructured triangle mesh: connectivity and derived geometry.

Field names and conventions follow the reference ``Mesh`` pytree
(euler/jax_fvm/src/mesh.py:28-83):

* ``tris[i]`` = (v0, v1, v2), counter-clockwise; local edge k joins
  (v_k, v_{k+1 mod 3}).
* ``neighbors[i, k]`` = triangle across edge k, -1 on the boundary
  (mesh.py:113, meshing/mesh_utils.py:176).
* ``faces[f]`` = the two vertex ids of unique edge f; ``face_markers[f]``:
  0 interior, 1 periodic, 2 wall, 3 inlet, 4 outlet, 5 subsonic inlet
  (mesh.py:47-52, meshing/mesh_utils.py:16).
* ``face_connectivity[i, k]`` = id of the face on local edge k
  (mesh.py:242-260).
* ``normals[i, k]`` = (dy, -dx)/|d| for d = p_{k+1} - p_k, outward for a CCW
  triangle (mesh.py:226-230).
* ``barycenter``, ``area``, ``surface`` as in mesh.py:223-240.

Geometry is computed in float64 on the host; the FP32 solver casts it.
Storage is a plain ``.npz`` (no pickle): the reference pickles the whole object
(mesh.py:327-359), which this format replaces with the same field names.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

ARRAY_FIELDS = (
    "points", "tris", "neighbors", "faces", "face_markers", "barycenter", "area",
    "surface", "normals", "face_connectivity", "face_connectivity_opposite",
    "inlet_supersonic", "inlet_subsonic",
)


def edge_keys(a: np.ndarray, b: np.ndarray, n_points: int) -> np.ndarray:
    lo = np.minimum(a, b).astype(np.int64)
    hi = np.maximum(a, b).astype(np.int64)
    return lo * np.int64(n_points) + hi


def rcm_seed_order(neighbors) -> np.ndarray:
    """Order in which reverse Cuthill-McKee seeds the components of the cell
    graph: ``numpy.argsort`` of the node degrees (distinct neighbours, no self
    loops), the call scipy.sparse.csgraph.reverse_cuthill_mckee makes. Among
    cells of equal degree the default sort's order is numpy's choice, so
    passing this order to the library's ordering reproduces scipy's
    permutation on the machine that runs it."""
    nb = np.asarray(neighbors, dtype=np.int64)
    own = np.arange(nb.shape[0], dtype=np.int64)[:, None]
    ok = (nb >= 0) & (nb != own)
    srt = np.sort(np.where(ok, nb, -1), axis=1)
    distinct = (srt >= 0) & np.concatenate([np.ones((nb.shape[0], 1), bool), srt[:, 1:] != srt[:, :-1]], axis=1)
    deg = distinct.sum(axis=1).astype(np.int32)
    return np.argsort(deg).astype(np.int32)


def faces_neighbors_from_tris(tris: np.ndarray, n_points: int):
    """Unique edges in first-appearance order (triangle-major, local edge
    minor), neighbours across each local edge and a boundary mask.

    Same conventions as meshing/mesh_utils.py:268-288 (edge i = (v_i, v_i+1),
    faces stored as the sorted vertex pair, -1 for a boundary neighbour).
    """
    tris = np.asarray(tris, dtype=np.int64)
    n = tris.shape[0]
    a = tris.reshape(-1)
    b = np.roll(tris, -1, axis=1).reshape(-1)
    keys = edge_keys(a, b, n_points)
    uniq, first_idx, inverse, counts = np.unique(keys, return_index=True,
                                                 return_inverse=True, return_counts=True)
    if counts.max(initial=0) > 2:
        raise ValueError("non-manifold mesh: an edge is shared by more than two triangles")
    # renumber faces by first appearance
    order = np.argsort(first_idx, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    face_of_slot = rank[inverse]
    n_faces = uniq.size
    faces = np.empty((n_faces, 2), dtype=np.int32)
    faces[:, 0] = (uniq[order] // n_points).astype(np.int32)
    faces[:, 1] = (uniq[order] % n_points).astype(np.int32)
    boundary = counts[order] == 1
    # neighbours: for each face with two owner slots, link them
    slots = np.arange(3 * n)
    srt = np.argsort(face_of_slot, kind="stable")
    fs = face_of_slot[srt]
    pair = np.nonzero(fs[1:] == fs[:-1])[0]
    s0 = slots[srt][pair]
    s1 = slots[srt][pair + 1]
    neighbors = -np.ones(3 * n, dtype=np.int32)
    neighbors[s0] = (s1 // 3).astype(np.int32)
    neighbors[s1] = (s0 // 3).astype(np.int32)
    return faces, neighbors.reshape(n, 3), boundary, face_of_slot.reshape(n, 3).astype(np.int32)


class Mesh:
    """Container with the reference's field names (mesh.py:28-83)."""

    def __init__(self, **kwargs):
        self.metadata = {}
        for k, v in kwargs.items():
            setattr(self, k, v)

    # -- construction -----------------------------------------------------
    @classmethod
    def from_points(cls, points, tris, neighbors, faces, face_markers, metadata=None):
        """Derive the geometry as ``mesh_generator_from_points`` does
        (mesh.py:131-150)."""
        m = cls()
        m.points = np.asarray(points, dtype=np.float64)
        m.tris = np.asarray(tris, dtype=np.int32)
        m.neighbors = np.asarray(neighbors, dtype=np.int32)
        m.faces = np.asarray(faces, dtype=np.int32)
        m.face_markers = np.asarray(face_markers, dtype=np.int32)
        m.compute_geometry()
        m.face_connectivity = m._face_connectivity()
        m.face_connectivity_opposite = np.arange(m.faces.shape[0], dtype=np.int32)
        m._set_periodic()
        m.inlet_supersonic = np.array([1.0, 1.0, 0.0, 1.0])
        m.inlet_subsonic = np.array([1.0, 1.0, 0.0, 1.0])
        if metadata:
            m.metadata.update(metadata)
        return m

    def compute_geometry(self):
        p = self.points[self.tris]                      # (N,3,2)
        # barycenter: mean of the 3 vertices (mesh.py:223-224)
        self.barycenter = (p[:, 0] + p[:, 1] + p[:, 2]) / 3.0
        # normals (mesh.py:226-230)
        d = np.roll(p, -1, axis=1) - p
        nrm = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1])
        self.normals = np.stack([d[..., 1], -d[..., 0]], axis=-1) / nrm[..., None]
        # area (mesh.py:232-237)
        a = 0.5 * ((p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0])
                   - (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]))
        self.area = np.abs(a)
        # surface (mesh.py:239-240)
        e = self.points[self.faces[:, 1]] - self.points[self.faces[:, 0]]
        self.surface = np.sqrt(e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1])

    def _face_connectivity(self):
        """face id of each local edge (mesh.py:242-260), vectorised."""
        npts = self.points.shape[0]
        fk = edge_keys(self.faces[:, 0], self.faces[:, 1], npts)
        order = np.argsort(fk, kind="stable")
        t = self.tris.astype(np.int64)
        tk = edge_keys(t, np.roll(t, -1, axis=1), npts)
        pos = np.searchsorted(fk[order], tk)
        pos = np.minimum(pos, fk.size - 1)
        if not np.all(fk[order][pos] == tk):
            raise KeyError("a triangle edge has no matching face")
        return order[pos].astype(np.int32)

    def _set_periodic(self):
        """Periodic pairing of marker-1 faces (mesh.py:266-321).

        Without marker-1 faces the opposite map is the identity
        (mesh.py:279-281)."""
        bids = np.nonzero(self.face_markers == 1)[0]
        if bids.size == 0:
            return
        tol = 1e-10
        pts = self.points
        x_min, x_max = pts[:, 0].min(), pts[:, 0].max()
        y_min, y_max = pts[:, 1].min(), pts[:, 1].max()
        dx, dy = x_max - x_min, y_max - y_min
        fp = pts[self.faces[bids]]                      # (B,2,2)
        diff = fp[:, 1] - fp[:, 0]
        is_left = (np.abs(diff[:, 0]) < tol) & (np.abs(fp[:, 0, 0] - x_min) < tol)
        is_right = (np.abs(diff[:, 0]) < tol) & (np.abs(fp[:, 0, 0] - x_max) < tol)
        is_bottom = (np.abs(diff[:, 1]) < tol) & (np.abs(fp[:, 0, 1] - y_min) < tol)
        is_top = (np.abs(diff[:, 1]) < tol) & (np.abs(fp[:, 0, 1] - y_max) < tol)
        shifts = (is_left[:, None] * np.array([dx, 0.0]) + is_right[:, None] * np.array([-dx, 0.0])
                  + is_bottom[:, None] * np.array([0.0, dy]) + is_top[:, None] * np.array([0.0, -dy]))
        opp = fp + shifts[:, None, :]
        d0 = np.linalg.norm(fp[:, None, :, :] - opp[None, :, :, :], axis=-1)
        d1 = np.linalg.norm(fp[:, None, :, :] - np.roll(opp[None, :, :, :], 1, axis=-2), axis=-1)
        match = (np.sum(d0 < tol, axis=-1) == 2) | (np.sum(d1 < tol, axis=-1) == 2)
        id_opp = np.nonzero(match)[1]
        fc = self.face_connectivity
        for i, fid in enumerate(bids):
            ti, ki = np.nonzero(fc == fid)
            tj, _ = np.nonzero(fc == bids[id_opp[i]])
            self.neighbors[ti, ki] = tj
        self.face_connectivity_opposite = np.arange(self.faces.shape[0], dtype=np.int32)
        self.face_connectivity_opposite[bids] = bids[id_opp]

    # -- accessors --------------------------------------------------------
    @property
    def n_cells(self) -> int:
        return int(self.tris.shape[0])

    def set_metadata(self, **kwargs):
        self.metadata.update(kwargs)
        return self

    def statistics(self) -> dict:
        fm = np.asarray(self.face_markers)
        return {
            "n_points": int(self.points.shape[0]), "n_faces": int(self.faces.shape[0]),
            "n_cells": self.n_cells,
            "markers": {int(m): int(np.sum(fm == m)) for m in np.unique(fm)},
            "h_avg": float(np.sqrt(4 * float(np.mean(self.area)) / np.sqrt(3))),
        }

    # -- I/O ----------------------------------------------------------------
    def save(self, path):
        path = Path(path)
        if path.suffix != ".npz":
            path = path.with_suffix(".npz")
        arrays = {k: np.asarray(getattr(self, k)) for k in ARRAY_FIELDS if hasattr(self, k)}
        arrays["metadata_json"] = np.frombuffer(json.dumps(self.metadata).encode(), dtype=np.uint8)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(path, **arrays)
        return path

    @classmethod
    def load(cls, path):
        path = Path(path)
        if path.suffix != ".npz":
            path = path.with_suffix(".npz")
        m = cls()
        with np.load(path, allow_pickle=False) as z:
            for k in z.files:
                if k == "metadata_json":
                    m.metadata = json.loads(bytes(z[k]).decode())
                else:
                    setattr(m, k, z[k])
        return m


def load_mesh(mesh_path) -> Mesh:
    """Counterpart of euler/config.py:63-79 (``ValueError`` when no path)."""
    if mesh_path is None:
        raise ValueError("mesh_path must be given (--mesh-path or the default config)")
    return Mesh.load(mesh_path)
