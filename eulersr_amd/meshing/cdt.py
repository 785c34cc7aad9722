"""Conforming Delaunay refinement mesher (our own; the reference uses
meshpy/Triangle, meshing/mesh_utils.py:182-198, which is not available here).

Algorithm (Ruppert-style, batched):

1. Delaunay-triangulate the current point set (Qhull via scipy).
2. Split every boundary segment that is missing from the triangulation or
   encroached by the apex of an adjacent in-domain triangle (apex inside the
   segment's diametral circle) at its midpoint.
3. Otherwise, for in-domain triangles that are too large for the local size
   law (area > sqrt(3)/4 h^2, the reference's refinement test,
   mesh_utils.py:147-151) or whose smallest angle is below ``min_angle``,
   propose their circumcentres; a circumcentre that encroaches a segment
   splits that segment instead. Proposals are thinned greedily (largest
   triangles first) so that no two accepted points are closer than half the
   circumradius.
4. Repeat until nothing changes.

The final triangles are oriented counter-clockwise and the body interior is
removed; boundary faces inherit the marker of the segment they lie on.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import Delaunay, cKDTree


def points_in_polygon(pts: np.ndarray, poly: np.ndarray) -> np.ndarray:
    """Even-odd rule, vectorised over points."""
    x, y = pts[:, 0], pts[:, 1]
    inside = np.zeros(len(pts), dtype=bool)
    n = len(poly)
    for i in range(n):
        x1, y1 = poly[i]
        x2, y2 = poly[(i + 1) % n]
        cond = (y1 > y) != (y2 > y)
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = x1 + (y - y1) * (x2 - x1) / (y2 - y1)
        inside ^= cond & (x < xint)
    return inside


def _circumcenters(p):
    ax, ay = p[:, 0, 0], p[:, 0, 1]
    bx, by = p[:, 1, 0], p[:, 1, 1]
    cx, cy = p[:, 2, 0], p[:, 2, 1]
    d = 2.0 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
    a2, b2, c2 = ax * ax + ay * ay, bx * bx + by * by, cx * cx + cy * cy
    with np.errstate(divide="ignore", invalid="ignore"):
        ux = (a2 * (by - cy) + b2 * (cy - ay) + c2 * (ay - by)) / d
        uy = (a2 * (cx - bx) + b2 * (ax - cx) + c2 * (bx - ax)) / d
    return np.stack([ux, uy], axis=1)


class Domain:
    """Axis-aligned outer box minus polygonal bodies."""

    def __init__(self, x0, y0, Lx, Ly, bodies):
        self.x0, self.y0, self.x1, self.y1 = x0, y0, x0 + Lx, y0 + Ly
        self.bodies = [np.asarray(b, dtype=np.float64) for b in bodies]

    def contains(self, pts, strict_box=True):
        eps = 1e-12 * max(self.x1 - self.x0, self.y1 - self.y0)
        ok = ((pts[:, 0] > self.x0 + eps) & (pts[:, 0] < self.x1 - eps)
              & (pts[:, 1] > self.y0 + eps) & (pts[:, 1] < self.y1 - eps))
        for b in self.bodies:
            ok &= ~points_in_polygon(pts, b)
        return ok

    def outside_bodies(self, pts):
        ok = np.ones(len(pts), dtype=bool)
        for b in self.bodies:
            ok &= ~points_in_polygon(pts, b)
        return ok


def refine(points, segments, seg_markers, domain: Domain, size_many, min_angle=30.0,
           max_iter=400, min_seg_frac=0.05, verbose=False):
    pts = [np.asarray(points, dtype=np.float64)]
    P = pts[0]
    segs = np.asarray(segments, dtype=np.int64)
    smark = np.asarray(seg_markers, dtype=np.int32)
    # shortest segment allowed (avoid endless splitting at sharp corners)
    seg_len0 = np.linalg.norm(P[segs[:, 1]] - P[segs[:, 0]], axis=1)
    min_seg = min_seg_frac * float(seg_len0.min())
    bound = 1.0 / (2.0 * np.sin(np.radians(min_angle)))

    for it in range(max_iter):
        n_pts = len(P)
        tri = Delaunay(P)
        simp = tri.simplices.astype(np.int64)
        cen = P[simp].mean(axis=1)
        in_dom = domain.outside_bodies(cen)

        # -- segment conformity -------------------------------------------
        # edge -> (triangle, apex) for in-domain triangles
        e_a = simp.reshape(-1)
        e_b = np.roll(simp, -1, axis=1).reshape(-1)
        apex = np.roll(simp, -2, axis=1).reshape(-1)
        tid = np.repeat(np.arange(len(simp)), 3)
        key = np.minimum(e_a, e_b) * n_pts + np.maximum(e_a, e_b)
        all_keys = np.unique(key)
        skey = np.minimum(segs[:, 0], segs[:, 1]) * n_pts + np.maximum(segs[:, 0], segs[:, 1])
        present = np.isin(skey, all_keys)
        # encroachment by in-domain apexes
        dom_slot = in_dom[tid]
        k_dom, apex_dom = key[dom_slot], apex[dom_slot]
        order = np.argsort(k_dom, kind="stable")
        k_sorted = k_dom[order]
        encroached = np.zeros(len(segs), dtype=bool)
        mid = 0.5 * (P[segs[:, 0]] + P[segs[:, 1]])
        r = 0.5 * np.linalg.norm(P[segs[:, 1]] - P[segs[:, 0]], axis=1)
        lo = np.searchsorted(k_sorted, skey, side="left")
        hi = np.searchsorted(k_sorted, skey, side="right")
        for rep in range(2):
            idx = lo + rep
            has = idx < hi
            ii = np.nonzero(has)[0]
            ap = apex_dom[order[idx[ii]]]
            d = np.linalg.norm(P[ap] - mid[ii], axis=1)
            encroached[ii] |= d < r[ii] * (1.0 - 1e-9)
        split = (~present | encroached) & (2 * r > 2 * min_seg)
        if np.any(split):
            P, segs, smark = _split_segments(P, segs, smark, np.nonzero(split)[0])
            if verbose:
                print(f"  it {it}: split {int(split.sum())} segments")
            continue
        if not np.all(present):
            raise RuntimeError("mesher could not recover all boundary segments")

        # -- triangle quality / size ----------------------------------------
        ts = simp[in_dom]
        tp = P[ts]
        e0 = np.linalg.norm(tp[:, 1] - tp[:, 0], axis=1)
        e1 = np.linalg.norm(tp[:, 2] - tp[:, 1], axis=1)
        e2 = np.linalg.norm(tp[:, 0] - tp[:, 2], axis=1)
        area = 0.5 * np.abs((tp[:, 1, 0] - tp[:, 0, 0]) * (tp[:, 2, 1] - tp[:, 0, 1])
                            - (tp[:, 1, 1] - tp[:, 0, 1]) * (tp[:, 2, 0] - tp[:, 0, 0]))
        lmin = np.minimum(np.minimum(e0, e1), e2)
        R = e0 * e1 * e2 / (4.0 * np.maximum(area, 1e-300))
        hloc = size_many(tp.mean(axis=1))
        too_big = area > np.sqrt(3.0) / 4.0 * hloc ** 2
        bad_shape = (R / lmin > bound) & (R > 0.15 * hloc)
        bad = too_big | bad_shape
        if not np.any(bad):
            break
        cc = _circumcenters(tp[bad])
        Rb = R[bad]
        ok = np.isfinite(cc).all(axis=1)
        cc, Rb = cc[ok], Rb[ok]
        # order: largest circumradius first; thin
        o = np.argsort(-Rb, kind="stable")
        cc, Rb = cc[o], Rb[o]
        ctree = cKDTree(cc)
        near = ctree.query_ball_point(cc, 0.5 * Rb)
        keep = np.zeros(len(cc), dtype=bool)
        supp = np.zeros(len(cc), dtype=bool)
        for i in range(len(cc)):
            if supp[i]:
                continue
            keep[i] = True
            for j in near[i]:
                if j > i:
                    supp[j] = True
        cc = cc[keep]
        # circumcentres outside the domain or encroaching a segment split it
        inside = domain.contains(cc)
        stree = cKDTree(mid)
        rmax = float(r.max())
        cand_segs = stree.query_ball_point(cc, rmax)
        enc_seg = -np.ones(len(cc), dtype=np.int64)
        for i, lst in enumerate(cand_segs):
            if not lst:
                continue
            lst = np.asarray(lst)
            d = np.linalg.norm(mid[lst] - cc[i], axis=1)
            hit = lst[d < r[lst]]
            if hit.size:
                enc_seg[i] = hit[np.argmin(d[d < r[lst]])]
        to_split = np.unique(enc_seg[enc_seg >= 0])
        to_split = to_split[2 * r[to_split] > 2 * min_seg]
        add = cc[inside & (enc_seg < 0)]
        if verbose:
            print(f"  it {it}: {len(P)} pts, bad {int(bad.sum())}, add {len(add)}, "
                  f"seg splits {len(to_split)}")
        if len(add) == 0 and len(to_split) == 0:
            break
        if len(add):
            P = np.concatenate([P, add])
        if len(to_split):
            P, segs, smark = _split_segments(P, segs, smark, to_split)
    return finalize(P, segs, smark, domain)


def _split_segments(P, segs, smark, which):
    which = np.asarray(which)
    a, b = segs[which, 0], segs[which, 1]
    newp = 0.5 * (P[a] + P[b])
    ids = np.arange(len(P), len(P) + len(which))
    P = np.concatenate([P, newp])
    keep = np.ones(len(segs), dtype=bool)
    keep[which] = False
    new_segs = np.concatenate([segs[keep], np.stack([a, ids], 1), np.stack([ids, b], 1)])
    new_mark = np.concatenate([smark[keep], smark[which], smark[which]])
    return P, new_segs, new_mark


def finalize(P, segs, smark, domain: Domain):
    tri = Delaunay(P)
    simp = tri.simplices.astype(np.int64)
    cen = P[simp].mean(axis=1)
    simp = simp[domain.outside_bodies(cen)]
    # drop unused points (inside bodies there are none by construction; the
    # body vertices are used) and renumber
    used = np.zeros(len(P), dtype=bool)
    used[simp.reshape(-1)] = True
    remap = -np.ones(len(P), dtype=np.int64)
    remap[used] = np.arange(int(used.sum()))
    P = P[used]
    simp = remap[simp]
    segs = remap[segs]
    if np.any(segs < 0):
        raise RuntimeError("boundary vertex not used by any triangle")
    # counter-clockwise orientation
    p = P[simp]
    cross = ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
             - (p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0]))
    cw = cross < 0
    simp[cw] = simp[cw][:, [0, 2, 1]]
    return P, simp.astype(np.int32), segs.astype(np.int32), smark.astype(np.int32)
