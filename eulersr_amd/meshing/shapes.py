"""Body contours, boundary sampling and mesh-size functions.

The steps follow the reference meshing modules so that meshes get the same
grading (the triangulator itself is our own, see ``cdt.py``):

* ``MeshSizeParams`` and the airfoil size law: meshing/mesh_utils.py:19-25, 114-141
* boundary sampling along a segment: mesh_utils.py:28-48
* curvature-driven airfoil contour sampling: mesh_utils.py:59-111
* outer box with markers (bottom/right/top = outlet, left = inlet):
  mesh_utils.py:153-170
* diamond contour and size law: meshing/diamond.py:20-60
* NACA 4-digit contour with a closed trailing edge: meshing/airfoils.py:31-59
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.spatial import cKDTree

WALL, INLET, OUTLET = 2, 3, 4


@dataclass(frozen=True)
class MeshSizeParams:
    growth_rate: float = 0.09
    left_growth_rate: float = 0.23
    obstacle_factor: float = 0.0
    max_size_factor: float = 6.0
    left_margin_factor: float = 5.0


DIAMOND_SIZE_PARAMS = MeshSizeParams(growth_rate=0.07, left_growth_rate=0.18)
AIRFOIL_SIZE_PARAMS = MeshSizeParams()


def smoothstep(x: float) -> float:
    x = min(max(x, 0.0), 1.0)
    return x * x * (3.0 - 2.0 * x)


def triangle_area_from_h(h):
    return np.sqrt(3.0) * np.asarray(h) ** 2 / 4.0


def sample_segment(p0, p1, size_fn, min_step) -> np.ndarray:
    """Points from p0 (included) towards p1 (excluded), step = local size."""
    p0 = np.asarray(p0, dtype=float)
    p1 = np.asarray(p1, dtype=float)
    delta = p1 - p0
    length = float(np.hypot(delta[0], delta[1]))
    if length == 0.0:
        return np.empty((0, 2))
    direction = delta / length
    out = [p0]
    s = 0.0
    while s < length:
        mid = p0 + direction * min(s + 0.5 * min_step, length)
        step = max(min_step, float(size_fn(mid)))
        nxt = min(length, s + step)
        if nxt >= length or (length - nxt) < 0.5 * min_step:
            break
        out.append(p0 + direction * nxt)
        s = nxt
    return np.asarray(out)


def outer_box(Lx, Ly, size_fn, h, x0=0.0, y0=0.0):
    corners = [(x0, y0), (x0 + Lx, y0), (x0 + Lx, y0 + Ly), (x0, y0 + Ly)]
    sides = [(corners[0], corners[1], OUTLET), (corners[1], corners[2], OUTLET),
             (corners[2], corners[3], OUTLET), (corners[3], corners[0], INLET)]
    pts, markers, last = [], [], None
    for a, b, m in sides:
        seg = sample_segment(a, b, size_fn, min_step=h)
        if last is not None and len(seg) and np.allclose(seg[0], last):
            seg = seg[1:]
        pts.append(seg)
        markers += [m] * len(seg)
        if len(seg):
            last = seg[-1]
    return np.concatenate(pts), markers


# -- diamond ------------------------------------------------------------------

def diamond_vertices(cx, cy, chord, height):
    hc, ht = chord / 2.0, height / 2.0
    return [(cx - hc, cy), (cx, cy + ht), (cx + hc, cy), (cx, cy - ht)]


def sample_polygon(verts, h):
    """Uniform points on each side, ceil(len/h) per side, endpoint excluded."""
    pts = []
    for i in range(len(verts)):
        p0 = np.asarray(verts[i], dtype=float)
        p1 = np.asarray(verts[(i + 1) % len(verts)], dtype=float)
        n = max(2, int(np.ceil(np.linalg.norm(p1 - p0) / h)))
        pts.append(np.linspace(p0, p1, n, endpoint=False))
    pts = np.concatenate(pts)
    return pts, [WALL] * len(pts)


def diamond_distance(pt, cx, cy, chord, height):
    px = (pt[0] - cx) / (chord / 2.0)
    py = (pt[1] - cy) / (height / 2.0)
    s = abs(px) + abs(py)
    if s <= 1.0:
        return 0.0
    near = (cx + np.sign(px) * (abs(px) / s) * (chord / 2.0),
            cy + np.sign(py) * (abs(py) / s) * (height / 2.0))
    return float(np.hypot(pt[0] - near[0], pt[1] - near[1]))


def diamond_size_fn(cx, cy, chord, height, h, sp: MeshSizeParams = DIAMOND_SIZE_PARAMS):
    x_le = cx - chord / 2.0
    growth_length = max(height, h)
    left_transition = max(sp.left_margin_factor * height, 4.0 * h)

    def size(pt):
        dist = diamond_distance(pt, cx, cy, chord, height)
        if dist <= sp.obstacle_factor * height:
            return float(h)
        blend = smoothstep((x_le - float(pt[0])) / left_transition)
        rate = sp.growth_rate + blend * (sp.left_growth_rate - sp.growth_rate)
        cap = sp.max_size_factor * (1.0 + 3.0 * blend)
        target = h * (1.0 + rate * dist / growth_length)
        return float(np.clip(target, h, cap * h))
    return size


# -- NACA 4-digit -------------------------------------------------------------

def naca_thickness(xc, t):
    xc = np.maximum(xc, 0.0)
    return 5 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2
                    + 0.2843 * xc ** 3 - 0.1036 * xc ** 4)


def naca_camber(xc, m, p):
    if m == 0.0 or p == 0.0:
        return np.zeros_like(xc), np.zeros_like(xc)
    yc = np.where(xc < p, m / p ** 2 * (2 * p * xc - xc ** 2),
                  m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * xc - xc ** 2))
    dyc = np.where(xc < p, 2 * m / p ** 2 * (p - xc), 2 * m / (1 - p) ** 2 * (p - xc))
    return yc, dyc


def naca_dense_contour(m, p, t, n_pts, chord, x0, y0):
    """Upper surface LE->TE then lower surface TE->LE, cosine spacing."""
    beta = np.linspace(0.0, np.pi, n_pts)
    xc = 0.5 * (1.0 - np.cos(beta))
    yt = naca_thickness(xc, t)
    yc, dyc = naca_camber(xc, m, p)
    th = np.arctan(dyc)
    upper = np.column_stack([(xc - yt * np.sin(th)) * chord + x0, (yc + yt * np.cos(th)) * chord + y0])
    lower = np.column_stack([(xc + yt * np.sin(th)) * chord + x0, (yc - yt * np.cos(th)) * chord + y0])[::-1]
    return np.concatenate([upper, lower[1:-1]])


def menger_curvature(pts):
    n = len(pts)
    k = np.zeros(n)
    a, b, c = pts[:-2], pts[1:-1], pts[2:]
    cross = np.abs((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))
    den = (np.linalg.norm(b - a, axis=1) * np.linalg.norm(c - b, axis=1) * np.linalg.norm(a - c, axis=1))
    k[1:-1] = np.where(den > 1e-20, 2.0 * cross / np.where(den > 1e-20, den, 1.0), 0.0)
    k[0], k[-1] = k[1], k[-2]
    return k


def _sample_half(half, h_body, h_min):
    seg = np.linalg.norm(np.diff(half, axis=0), axis=1)
    arc = np.concatenate([[0.0], np.cumsum(seg)])
    total = arc[-1]
    n_uni = max(int(total / (0.5 * h_min)) + 1, 8)
    s_uni = np.linspace(0.0, total, n_uni)
    xy = np.column_stack([np.interp(s_uni, arc, half[:, 0]), np.interp(s_uni, arc, half[:, 1])])
    kap = menger_curvature(xy)
    out, s = [], 0.0
    while s < total:
        out.append([float(np.interp(s, arc, half[:, 0])), float(np.interp(s, arc, half[:, 1]))])
        k = max(float(np.interp(s, s_uni, kap)), 1e-10)
        s += float(np.clip(h_body / k * 2, h_min, h_body))
    return np.array(out)


def airfoil_boundary(dense, n_half, chord, x0, y0, h):
    h_body, h_min = h * 2.0 / 3.0, h / 5.0
    upper = dense[:n_half]
    lower = np.concatenate([dense[:1], dense[n_half:][::-1], dense[n_half - 1:n_half]])
    up = _sample_half(upper, h_body, h_min)
    lo = _sample_half(lower, h_body, h_min)
    te = np.array([[x0 + chord, y0]])
    if len(up) > 1 and np.linalg.norm(up[-1] - te[0]) < 0.5 * h_body:
        up = up[:-1]
    if len(lo) > 1 and np.linalg.norm(lo[-1] - te[0]) < 0.5 * h_body:
        lo = lo[:-1]
    pts = np.concatenate([up, te, lo[::-1][:-1]])
    return pts, [WALL] * len(pts)


def airfoil_size_fn(dense, cx, chord, t_max, h, sp: MeshSizeParams = AIRFOIL_SIZE_PARAMS):
    tree = cKDTree(dense)
    kappa = menger_curvature(dense)
    x_le = cx - chord / 2.0
    growth_length = max(t_max, h)
    left_transition = max(sp.left_margin_factor * t_max, 4.0 * h)

    def size(pt):
        dist, idx = tree.query(np.asarray(pt, dtype=np.float64))
        dist = float(dist)
        if dist <= sp.obstacle_factor * t_max:
            return float(h)
        k = max(float(kappa[idx]), 1e-10)
        h_surf = float(np.clip(h / k, h / 5.0, h * 2.0 / 3.0))
        h_base = h_surf + min(1.0, dist / h) * (h - h_surf)
        blend = smoothstep((x_le - float(pt[0])) / left_transition)
        rate = sp.growth_rate + blend * (sp.left_growth_rate - sp.growth_rate)
        cap = sp.max_size_factor * (1.0 + 3.0 * blend)
        target = h_base * (1.0 + rate * dist / growth_length)
        return float(np.clip(target, h_base, cap * h))

    def size_many(pts):
        pts = np.asarray(pts, dtype=np.float64)
        dist, idx = tree.query(pts)
        k = np.maximum(kappa[idx], 1e-10)
        h_surf = np.clip(h / k, h / 5.0, h * 2.0 / 3.0)
        h_base = h_surf + np.minimum(1.0, dist / h) * (h - h_surf)
        x = np.clip((x_le - pts[:, 0]) / left_transition, 0.0, 1.0)
        blend = x * x * (3.0 - 2.0 * x)
        rate = sp.growth_rate + blend * (sp.left_growth_rate - sp.growth_rate)
        cap = sp.max_size_factor * (1.0 + 3.0 * blend)
        target = h_base * (1.0 + rate * dist / growth_length)
        out = np.clip(target, h_base, cap * h)
        return np.where(dist <= sp.obstacle_factor * t_max, h, out)

    size.many = size_many
    return size


def diamond_size_many(size_fn, cx, cy, chord, height, h, sp: MeshSizeParams = DIAMOND_SIZE_PARAMS):
    """Vectorised version of ``diamond_size_fn`` (same law)."""
    x_le = cx - chord / 2.0
    growth_length = max(height, h)
    left_transition = max(sp.left_margin_factor * height, 4.0 * h)

    def many(pts):
        pts = np.asarray(pts, dtype=np.float64)
        px = (pts[:, 0] - cx) / (chord / 2.0)
        py = (pts[:, 1] - cy) / (height / 2.0)
        s = np.abs(px) + np.abs(py)
        ss = np.where(s > 0, s, 1.0)
        nx = cx + np.sign(px) * (np.abs(px) / ss) * (chord / 2.0)
        ny = cy + np.sign(py) * (np.abs(py) / ss) * (height / 2.0)
        dist = np.where(s <= 1.0, 0.0, np.hypot(pts[:, 0] - nx, pts[:, 1] - ny))
        x = np.clip((x_le - pts[:, 0]) / left_transition, 0.0, 1.0)
        blend = x * x * (3.0 - 2.0 * x)
        rate = sp.growth_rate + blend * (sp.left_growth_rate - sp.growth_rate)
        cap = sp.max_size_factor * (1.0 + 3.0 * blend)
        target = h * (1.0 + rate * dist / growth_length)
        out = np.clip(target, h, cap * h)
        return np.where(dist <= sp.obstacle_factor * height, h, out)

    size_fn.many = many
    return size_fn
