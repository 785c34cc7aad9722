"""Post-processing of solved fields (host side, numpy) and the diamond
mesh-convergence campaign.

* Field diagnostics of euler/jax_fvm/src/helper.py: Mach number
  (:303-310), specific and total entropy and the entropy created between two
  states (:48-55, 312-320), and the wall profile -- Mach number and pressure
  of the wall cells under the lower wall faces, ordered along the wall by
  arc length (:418-506).
* Mesh convergence (euler/postproc/mesh_convergence_diamond.py): CD, CL and
  the wall Mach profile of a few (Mach, alpha) points on a chain of meshes,
  the relative L2 distance of each wall profile to the finest one (:96-110)
  and Richardson's observed order with Roache's grid convergence index on
  sliding triplets of a constant-ratio chain (:143-166). The reference reads
  the fields from its dataset bundles; here the chain is meshed and every
  (mesh, point) is solved in one batched fixed-time GPU solve.
"""
from __future__ import annotations

import math

import numpy as np

GAMMA = 1.4


# -- field diagnostics (helper.py) --------------------------------------------

def primitives(W, gamma=GAMMA, M=1.0):
    """getPrimitive (helper.py:37-46) in float64."""
    W = np.asarray(W, dtype=np.float64)
    rho = W[..., 0]
    u = W[..., 1] / rho
    v = W[..., 2] / rho
    P = (W[..., 3] - 0.5 * rho * (u * u + v * v)) * M * M * (gamma - 1.0)
    return np.stack([rho, u, v, P], axis=-1)


def get_mach_number(prim, gamma=GAMMA):
    """helper.py:303-310: |u| / sqrt(gamma P / rho)."""
    prim = np.asarray(prim, dtype=np.float64)
    c = np.sqrt(gamma * prim[..., 3] / prim[..., 0])
    return np.sqrt(prim[..., 1] ** 2 + prim[..., 2] ** 2) / c


def get_specific_entropy(W, gamma=GAMMA):
    """helper.py:48-55: s = ln(P / rho^gamma)."""
    p = primitives(W, gamma)
    return np.log(p[..., 3] / p[..., 0] ** gamma)


def get_total_entropy(W, mesh, gamma=GAMMA):
    """helper.py:312-315: sum of rho s area / (gamma - 1) over the cells."""
    W = np.asarray(W, dtype=np.float64)
    s = get_specific_entropy(W, gamma)
    return float(np.sum(W[..., 0] * s * np.asarray(mesh.area, dtype=np.float64) / (gamma - 1.0), axis=-1))


def get_entropy_creation(W_initial, W_final, mesh, gamma=GAMMA):
    """helper.py:317-320: S(W_final) - S(W_initial)."""
    return get_total_entropy(W_final, mesh, gamma) - get_total_entropy(W_initial, mesh, gamma)


def face_owner_data(face_connectivity, face_ids):
    """helper.py:426-436: for each face id, the first (cell, local face) of
    face_connectivity (row-major order) holding it; faces held by no cell
    are dropped."""
    fc = np.asarray(face_connectivity)
    flat = fc.ravel()
    order = np.argsort(flat, kind="stable")           # row-major first occurrence first
    sorted_ids = flat[order]
    face_ids = np.asarray(face_ids, dtype=np.int64)
    pos = np.searchsorted(sorted_ids, face_ids)
    ok = (pos < sorted_ids.size) & (sorted_ids[np.minimum(pos, sorted_ids.size - 1)] == face_ids)
    first = order[pos[ok]]
    return (first // fc.shape[1]).astype(np.int32), (first % fc.shape[1]).astype(np.int32)


def lower_wall_face_ids(mesh):
    """helper.py:439-461: wall faces (metadata wall_markers, default [2])
    whose midpoint lies in the lower half of the domain (all wall faces if
    none does), ordered by midpoint x, then y."""
    md = mesh.metadata if isinstance(getattr(mesh, "metadata", None), dict) else {}
    wm = md.get("wall_markers", [2])
    try:
        wm = tuple(int(m) for m in wm)
    except TypeError:
        wm = (int(wm),)
    points = np.asarray(mesh.points, dtype=np.float64)
    faces = np.asarray(mesh.faces)
    wall = np.isin(np.asarray(mesh.face_markers), wm)
    if not np.any(wall):
        return np.empty(0, dtype=np.int32)
    mid = points[faces].mean(axis=1)
    y_thr = 0.5 * float(points[:, 1].min() + points[:, 1].max())
    ids = np.nonzero(wall & (mid[:, 1] <= y_thr))[0]
    if ids.size == 0:
        ids = np.nonzero(wall)[0]
    sel = mid[ids]
    return ids[np.lexsort((sel[:, 1], sel[:, 0]))].astype(np.int32)


def get_wall_profile(W, mesh, gamma=GAMMA, M=1.0):
    """helper.py:464-506: arc length s, midpoint x, y, Mach number and
    pressure of the cells owning the lower wall faces, ordered along x (the
    sound speed floored at sqrt(1e-14))."""
    face_ids = lower_wall_face_ids(mesh)
    empty = np.empty(0, dtype=float)
    if face_ids.size == 0:
        return dict(s=empty, x=empty, y=empty, mach=empty, pressure=empty, face_ids=face_ids)
    cells, _ = face_owner_data(mesh.face_connectivity, face_ids)
    points = np.asarray(mesh.points, dtype=np.float64)
    mid = points[np.asarray(mesh.faces)[face_ids]].mean(axis=1)
    order = np.lexsort((mid[:, 1], mid[:, 0]))
    face_ids, cells, mid = face_ids[order], cells[order], mid[order]
    p = primitives(W, gamma, M)[cells]
    c = np.sqrt(np.maximum(gamma * p[:, 3] / p[:, 0], 1e-14))
    mach = np.sqrt(p[:, 1] ** 2 + p[:, 2] ** 2) / c
    seg = np.linalg.norm(np.diff(mid, axis=0), axis=1) if mid.shape[0] > 1 else empty
    s = np.concatenate(([0.0], np.cumsum(seg)))
    return dict(s=s, x=mid[:, 0], y=mid[:, 1], mach=mach, pressure=p[:, 3], face_ids=face_ids)


# -- mesh convergence (mesh_convergence_diamond.py) ---------------------------

def relative_profile_error(ref, cur, n_samples=400):
    """mesh_convergence_diamond.py:96-110: ||M_cur - M_ref|| / ||M_ref|| of
    the wall Mach profiles interpolated on n_samples points of their common
    arc-length range; None when undefined."""
    if ref is None or cur is None or ref["s"].size < 2 or cur["s"].size < 2:
        return None
    lo = max(ref["s"].min(), cur["s"].min())
    hi = min(ref["s"].max(), cur["s"].max())
    if hi <= lo:
        return None
    grid = np.linspace(lo, hi, n_samples)
    r = np.interp(grid, ref["s"], ref["mach"])
    c = np.interp(grid, cur["s"], cur["mach"])
    den = np.linalg.norm(r)
    if den <= 1e-16:
        return None
    return float(np.linalg.norm(c - r) / den)


def richardson_gci(values_by_h, chain, safety=1.25):
    """mesh_convergence_diamond.py:143-166: for each sliding triplet
    (coarse h1, mid h2, fine h3) of a constant-ratio chain r = h1/h2, the
    observed order p = ln((f1-f2)/(f2-f3)) / ln r, the Richardson
    extrapolation f3 + (f3-f2)/(r^p - 1) and GCI_fine = safety |(f3-f2)/f3|
    / (r^p - 1) in percent; p = None for oscillatory or stalled triplets."""
    r = chain[0] / chain[1]
    out = []
    for i in range(len(chain) - 2):
        h1, h2, h3 = chain[i], chain[i + 1], chain[i + 2]
        f1, f2, f3 = values_by_h.get(h1), values_by_h.get(h2), values_by_h.get(h3)
        if f1 is None or f2 is None or f3 is None:
            continue
        num, den = f1 - f2, f2 - f3
        if abs(den) < 1e-14 or (num / den) <= 0:
            out.append(dict(h_coarse=h1, h_mid=h2, h_fine=h3, p_obs=None, extrap=None, gci_fine_pct=None))
            continue
        p = math.log(num / den) / math.log(r)
        extrap = f3 + (f3 - f2) / (r ** p - 1.0)
        eps = abs((f3 - f2) / f3) if f3 != 0 else float("nan")
        out.append(dict(h_coarse=h1, h_mid=h2, h_fine=h3, p_obs=float(p), extrap=float(extrap),
                        gci_fine_pct=float(safety * eps / (r ** p - 1.0) * 100.0)))
    return out


# (label, Mach, alpha) of the reference's campaign (mesh_convergence_diamond.py:54-62)
TEST_POINTS = (("P1_transonic", 0.90, 0.0), ("P2_marginal_supersonic", 1.05, 0.0),
               ("P3_clean_supersonic", 1.50, 0.0), ("P4_strong_supersonic", 2.50, 0.0),
               ("P5_incidence", 2.00, 3.0), ("P6_mach2_symmetric", 2.00, 0.0),
               ("P7_strong_incidence", 2.50, 5.0))
H_GCI = (0.4, 0.2, 0.1, 0.05, 0.025, 0.0125)   # the r = 2 chain of the reference (:50-51)


def mesh_convergence(points=TEST_POINTS, chain=H_GCI, geometry="diamond", gamma=GAMMA):
    """Solves every (mesh of `chain`, point) with the dataset protocol
    (fixed time tf, euler/main.py defaults; pipeline.fixed_time), all points
    of a mesh in one batch, and returns per point: the rows (h, cells, CD,
    CL, wall-profile error to the finest mesh) and the Richardson / GCI
    triplets of CD and CL."""
    from . import pipeline as P
    from .meshing import build_mesh
    fields = {}
    for h in chain:
        mesh = build_mesh(geometry, h)
        specs, inits = P.make_cases(mesh, geometry, [(m, a) for _, m, a in points])
        reps, states = P.fixed_time(specs, inits)
        for (label, m, a), r, W in zip(points, reps, states):
            Wn = W.cpu().numpy()
            fields[(label, h)] = dict(h=h, n_cells=int(mesh.n_cells), cd=float(r["cd"]), cl=float(r["cl"]),
                                      wall=get_wall_profile(Wn, mesh, gamma))
    out = {}
    finest = min(chain)
    for label, m, a in points:
        ref = fields[(label, finest)]["wall"]
        rows = []
        for h in chain:
            f = fields[(label, h)]
            rows.append(dict(h=h, n_cells=f["n_cells"], cd=f["cd"], cl=f["cl"],
                             wall_profile_rel_err=relative_profile_error(ref, f["wall"])))
        out[label] = dict(mach=m, aoa=a, rows=rows,
                          gci_cd=richardson_gci({r["h"]: r["cd"] for r in rows}, list(chain)),
                          gci_cl=richardson_gci({r["h"]: r["cl"] for r in rows}, list(chain)))
    return out
