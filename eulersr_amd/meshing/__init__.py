"""Mesh generation for the solver's cases.

Geometries and defaults follow euler/generate_mesh.py:18-79 (reference box
[0,4]x[0,4], chord 1, body centre (1.5, 2.0); diamond half-angle 5 degrees)
and meshing/airfoils.py:14-28 (NACA 4-digit table), extended with NACA4412
and NACA0015 (same thickness/camber formulas, airfoils.py:31-59).
"""
from __future__ import annotations

from math import radians, tan
from pathlib import Path

import numpy as np

from ..mesh import Mesh, edge_keys, faces_neighbors_from_tris
from . import shapes
from .cdt import Domain, refine

NACA = {
    "naca0012": dict(m=0.0, p=0.0, t=0.12),
    "naca2412": dict(m=0.02, p=0.4, t=0.12),
    "naca4412": dict(m=0.04, p=0.4, t=0.12),
    "naca0015": dict(m=0.0, p=0.0, t=0.15),
}
GEOMETRIES = ("diamond",) + tuple(NACA)


def _assemble(P, tris, segs, smark, metadata) -> Mesh:
    faces, neighbors, boundary, _ = faces_neighbors_from_tris(tris, len(P))
    fkey = edge_keys(faces[:, 0], faces[:, 1], len(P))
    skey = edge_keys(segs[:, 0], segs[:, 1], len(P))
    order = np.argsort(skey)
    pos = np.searchsorted(skey[order], fkey)
    pos = np.minimum(pos, len(skey) - 1)
    on_seg = skey[order][pos] == fkey
    if not np.array_equal(on_seg, boundary):
        raise RuntimeError("boundary faces and boundary segments disagree")
    markers = np.zeros(len(faces), dtype=np.int32)
    markers[on_seg] = smark[order][pos[on_seg]]
    return Mesh.from_points(P, tris, neighbors, faces, markers, metadata)


def _mesh_with_body(outer_pts, outer_ms, body_pts, body_ms, domain, size_many, verbose):
    n_o = len(outer_pts)
    pts = np.concatenate([outer_pts, body_pts])
    segs = [(i, (i + 1) % n_o) for i in range(n_o)]
    segs += [(n_o + i, n_o + (i + 1) % len(body_pts)) for i in range(len(body_pts))]
    # segment i of a loop carries the marker of its first point
    marks = list(outer_ms) + list(body_ms)
    return refine(pts, np.array(segs), np.array(marks), domain, size_many, verbose=verbose)


def build_diamond(h=0.05, Lx=4.0, Ly=4.0, chord=1.0, height=None, cx=1.5, cy=2.0,
                  x0=0.0, y0=0.0, case="diamond", verbose=False) -> Mesh:
    """Diamond airfoil (meshing/diamond.py:82-118); default height tan(5 deg)
    as in euler/generate_mesh.py:28-31."""
    if height is None:
        height = chord * tan(radians(5.0))
    size = shapes.diamond_size_fn(cx, cy, chord, height, h)
    shapes.diamond_size_many(size, cx, cy, chord, height, h)
    verts = shapes.diamond_vertices(cx, cy, chord, height)
    if verts[0][0] < x0 or verts[2][0] > x0 + Lx or verts[3][1] < y0 or verts[1][1] > y0 + Ly:
        raise ValueError("the diamond must lie inside the domain")
    outer, oms = shapes.outer_box(Lx, Ly, size, h, x0, y0)
    body, bms = shapes.sample_polygon(verts, h)
    dom = Domain(x0, y0, Lx, Ly, [np.array(verts)])
    P, tris, segs, smark = _mesh_with_body(outer, oms, body, bms, dom, size.many, verbose)
    meta = dict(case=case, h=h, domain={"Lx": Lx, "Ly": Ly, "x0": x0, "y0": y0},
                obstacle_length=chord, chord=chord, height=height,
                center={"cx": cx, "cy": cy}, symmetric=False, wall_markers=[2], force_marker=2)
    return _assemble(P, tris, segs, smark, meta)


def build_naca(name="naca0012", h=0.05, Lx=4.0, Ly=4.0, chord=1.0, cx=1.5, cy=2.0,
               x0=0.0, y0=0.0, verbose=False) -> Mesh:
    """NACA 4-digit airfoil with a closed trailing edge (airfoils.py:137-145,
    190-204; mesh_utils.py:204-234, non-symmetric path)."""
    spec = NACA[name]
    x_le = cx - chord / 2.0
    if x_le < x0 or cx + chord / 2.0 > x0 + Lx:
        raise ValueError("the airfoil must lie inside the domain")
    dense = shapes.naca_dense_contour(spec["m"], spec["p"], spec["t"], 2000, chord, x_le, cy)
    t_max = spec["t"] * chord
    size = shapes.airfoil_size_fn(dense, cx, chord, t_max, h)
    outer, oms = shapes.outer_box(Lx, Ly, size, h, x0, y0)
    body, bms = shapes.airfoil_boundary(dense, 2000, chord, x_le, cy, h)
    dom = Domain(x0, y0, Lx, Ly, [body])
    P, tris, segs, smark = _mesh_with_body(outer, oms, body, bms, dom, size.many, verbose)
    meta = dict(case=name, h=h, domain={"Lx": Lx, "Ly": Ly, "x0": x0, "y0": y0},
                obstacle_length=chord, chord=chord, thickness=spec["t"], camber=spec["m"],
                camber_pos=spec["p"], height=t_max, center={"cx": cx, "cy": cy},
                wall_markers=[2], force_marker=2)
    return _assemble(P, tris, segs, smark, meta)


def build_mesh(geometry: str, h: float, **kw) -> Mesh:
    if geometry == "diamond":
        return build_diamond(h=h, **kw)
    if geometry in NACA:
        return build_naca(geometry, h=h, **kw)
    raise ValueError(f"unknown geometry '{geometry}', expected one of {GEOMETRIES}")


def cached_mesh(geometry: str, h: float, cache_dir=None, **kw) -> Mesh:
    """Build once and keep the ``.npz`` under ``cache_dir`` (default:
    ``<repo>/data/meshes``)."""
    if cache_dir is None:
        cache_dir = Path(__file__).resolve().parents[2] / "data" / "meshes"
    tag = "".join(f"_{k}{v}" for k, v in sorted(kw.items()))
    path = Path(cache_dir) / f"{geometry}_h{h}{tag}.npz"
    if path.exists():
        return Mesh.load(path)
    m = build_mesh(geometry, h, **kw)
    m.save(path)
    return m
