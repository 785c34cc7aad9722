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


BUMP, INLET, OUTLET, CHANNEL_WALL = 2, 3, 4, 6     # bump.py:12 marker codes


def build_bump(h=0.05, Lx=3.0, Ly=1.0, thickness=0.04, chord=1.0, center=1.5, verbose=False) -> Mesh:
    """Channel with a circular-arc bump on its lower wall (the bump case of
    meshing/bump.py:139-210): inlet (3) on the left, outlet (4) on the right,
    flat walls (6) below and above, the bump (2) between center -/+ chord/2.
    Metadata case='bump' (no sponge, helper.py:266-267), wall markers [2, 6],
    forces on the bump. Size law (ours): h within one chord of the bump,
    growing linearly to 3h beyond."""
    if thickness <= 0.0 or chord <= 0.0:
        raise ValueError("the bump needs a positive thickness and chord")
    if center - chord / 2.0 < 0.0 or center + chord / 2.0 > Lx:
        raise ValueError("the bump must lie inside [0, Lx]")
    R = ((chord / 2.0) ** 2 + thickness ** 2) / (2.0 * thickness)
    yc = thickness - R
    xl, xr = center - chord / 2.0, center + chord / 2.0

    def size_many(pts):
        pts = np.atleast_2d(pts)
        dx = np.maximum(np.maximum(xl - pts[:, 0], pts[:, 0] - xr), 0.0)
        d = np.hypot(dx, pts[:, 1])
        return h * np.clip(1.0 + 2.0 * (d - 0.2 * chord) / chord, 1.0, 3.0)

    def seg(p0, p1):
        p0, p1 = np.asarray(p0, float), np.asarray(p1, float)
        n = max(1, int(np.ceil(np.linalg.norm(p1 - p0) / float(size_many((p0 + p1) / 2)[0]))))
        return p0 + (p1 - p0) * (np.arange(n) / n)[:, None]

    n_arc = max(3, int(np.ceil(chord / h)))
    xs = np.linspace(xl, xr, n_arc, endpoint=False)
    arc = np.stack([xs, yc + np.sqrt(np.clip(R * R - (xs - center) ** 2, 0.0, None))], axis=1)
    loops = [(seg((0.0, 0.0), (xl, 0.0)), CHANNEL_WALL), (arc, BUMP), (seg((xr, 0.0), (Lx, 0.0)), CHANNEL_WALL),
             (seg((Lx, 0.0), (Lx, Ly)), OUTLET), (seg((Lx, Ly), (0.0, Ly)), CHANNEL_WALL),
             (seg((0.0, Ly), (0.0, 0.0)), INLET)]
    pts = np.concatenate([p for p, _ in loops])
    marks = np.concatenate([np.full(len(p), m, dtype=np.int32) for p, m in loops])
    n = len(pts)
    segs = np.array([(i, (i + 1) % n) for i in range(n)])
    # the region under the arc is a body touching the lower wall
    body = np.concatenate([arc, [[xr, 0.0], [xr, -1.0], [xl, -1.0], [xl, 0.0]]])
    dom = Domain(0.0, 0.0, Lx, Ly, [body])
    P, tris, segs, smark = refine(pts, segs, marks, dom, size_many, verbose=verbose)
    meta = dict(case="bump", h=h, domain={"Lx": Lx, "Ly": Ly}, center=center, thickness=thickness, chord=chord,
                obstacle_length=chord, bump_marker=BUMP, force_marker=BUMP, wall_markers=[BUMP, CHANNEL_WALL],
                inlet_marker=INLET, outlet_marker=OUTLET)
    return _assemble(P, tris, segs, smark, meta)


def build_periodic_channel(nx=40, ny=20, Lx=2.0, Ly=1.0, jitter=0.25, seed=0) -> Mesh:
    """Channel periodic in y (marker 1 on the lower and upper sides, paired by
    Mesh._set_periodic as mesh.py:266-321 pairs them), inlet (3) on the left,
    outlet (4) on the right: a grid of nx x ny quads, each split along a
    diagonal whose direction alternates, interior points moved by a seeded
    jitter of up to `jitter` cells so the triangles are irregular."""
    rng = np.random.default_rng(seed)
    xs, ys = np.linspace(0.0, Lx, nx + 1), np.linspace(0.0, Ly, ny + 1)
    X, Y = np.meshgrid(xs, ys)
    P = np.stack([X.ravel(), Y.ravel()], axis=1)
    inner = (X.ravel() > 0) & (X.ravel() < Lx) & (Y.ravel() > 0) & (Y.ravel() < Ly)
    P[inner] += rng.uniform(-jitter, jitter, size=(int(inner.sum()), 2)) * np.array([Lx / nx, Ly / ny])
    tris = []
    for j in range(ny):
        for i in range(nx):
            a, b = j * (nx + 1) + i, j * (nx + 1) + i + 1
            c, d = (j + 1) * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i
            tris += [[a, b, c], [a, c, d]] if (i + j) % 2 == 0 else [[a, b, d], [b, c, d]]
    tris = np.array(tris, dtype=np.int64)
    faces, nb, bnd, _ = faces_neighbors_from_tris(tris, len(P))
    fp = P[faces]
    markers = np.zeros(len(faces), dtype=np.int32)
    tol = 1e-12
    on_x0 = bnd & (np.abs(fp[:, :, 0]) < tol).all(axis=1)
    on_x1 = bnd & (np.abs(fp[:, :, 0] - Lx) < tol).all(axis=1)
    markers[bnd] = 1
    markers[on_x0] = INLET
    markers[on_x1] = OUTLET
    meta = dict(case="periodic_channel", domain={"Lx": Lx, "Ly": Ly}, wall_markers=[2], force_marker=2,
                obstacle_length=1.0)
    return Mesh.from_points(P, tris, nb, faces, markers, meta)


def build_mesh(geometry: str, h: float, **kw) -> Mesh:
    if geometry == "diamond":
        return build_diamond(h=h, **kw)
    if geometry == "bump":
        return build_bump(h=h, **kw)
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
