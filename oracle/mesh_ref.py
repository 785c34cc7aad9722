"""CPU oracle (TEST INFRASTRUCTURE ONLY): loop-based restatement of the
reference mesh conventions, used to pin the product's vectorised mesh code
bit-exactly.

* edge map / faces / neighbours: meshing/mesh_utils.py:268-288
* face_connectivity via a sorted-pair dict: jax_fvm/src/mesh.py:242-260
* barycenter, normals, area, surface: jax_fvm/src/mesh.py:223-240
"""
from __future__ import annotations

from collections import defaultdict

import numpy as np


def faces_neighbors_from_tris(tris):
    edge_map = defaultdict(list)
    for t, tri in enumerate(tris):
        for i in range(3):
            a, b = int(tri[i]), int(tri[(i + 1) % 3])
            edge_map[(a, b) if a < b else (b, a)].append((t, i))
    n = len(tris)
    neighbors = -np.ones((n, 3), dtype=np.int32)
    faces = np.empty((len(edge_map), 2), dtype=np.int32)
    boundary = np.zeros(len(edge_map), dtype=bool)
    for fidx, (key, owners) in enumerate(edge_map.items()):
        faces[fidx] = key
        if len(owners) == 2:
            (t0, i0), (t1, i1) = owners
            neighbors[t0, i0] = t1
            neighbors[t1, i1] = t0
        else:
            boundary[fidx] = True
    return faces, neighbors, boundary


def face_connectivity(faces, tris):
    lookup = {tuple(sorted(f.tolist())): idx for idx, f in enumerate(faces)}
    fc = np.empty((len(tris), 3), dtype=np.int32)
    for t, tri in enumerate(tris):
        fc[t, 0] = lookup[tuple(sorted((int(tri[0]), int(tri[1]))))]
        fc[t, 1] = lookup[tuple(sorted((int(tri[1]), int(tri[2]))))]
        fc[t, 2] = lookup[tuple(sorted((int(tri[2]), int(tri[0]))))]
    return fc


def geometry(points, tris, faces):
    """Per-cell loop; each expression in the reference's order."""
    n = len(tris)
    bary = np.empty((n, 2))
    normals = np.empty((n, 3, 2))
    area = np.empty(n)
    for t in range(n):
        p = points[tris[t]]
        bary[t] = (p[0] + p[1] + p[2]) / 3.0
        for k in range(3):
            d = p[(k + 1) % 3] - p[k]
            nrm = np.sqrt(d[0] * d[0] + d[1] * d[1])
            normals[t, k] = (d[1] / nrm, -d[0] / nrm)
        area[t] = abs(0.5 * ((p[1, 1] - p[0, 1]) * (p[2, 0] - p[0, 0])
                             - (p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1])))
    surface = np.empty(len(faces))
    for f in range(len(faces)):
        e = points[faces[f, 1]] - points[faces[f, 0]]
        surface[f] = np.sqrt(e[0] * e[0] + e[1] * e[1])
    return bary, normals, area, surface
