"""Post-processing and validation: shock-expansion theory of the diamond
airfoil and its comparison with a solved field.

The reference validates the solver on the diamond against oblique-shock /
Prandtl-Meyer theory (euler/postproc/theory_validation_diamond.py:44-170,
345-372, 484): each flat panel turns the flow by a fixed angle, a shock for a
compression and an isentropic fan for an expansion, cascaded from upstream:
front panels deflect by eps -/+ alpha (upper/lower), the mid-chord ridges
expand by 2 eps. Measured quantities: CD/CL (theory pressures put on the same
wall cells and integrated by the same routine), the plateau pressure
coefficient and Mach number of each panel, and the leading-edge shock angle
fitted to the largest pressure gradients of the field (:293-320).

Here the weak-shock angle comes from the closed-form solution of the
theta-beta-M relation (a cubic in tan(beta); NACA Report 1135, eq. 150b)
instead of a root search, and the inverse Prandtl-Meyer function from a
bracketed Newton iteration.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

GAMMA = 1.4
PANELS = ("upper_front", "upper_rear", "lower_front", "lower_rear")


# -- gas dynamics ------------------------------------------------------------

def prandtl_meyer(M, g=GAMMA):
    """nu(M) in radians (M >= 1)."""
    M = max(float(M), 1.0)
    a = math.sqrt((g + 1.0) / (g - 1.0))
    s = math.sqrt(M * M - 1.0)
    return a * math.atan(s / a) - math.atan(s)


def inverse_prandtl_meyer(nu, g=GAMMA):
    """M such that nu(M) = nu (radians); Newton on [1, M_hi] with bisection
    safeguard. dnu/dM = sqrt(M^2-1) / (M (1 + (g-1)/2 M^2))."""
    if nu <= 0.0:
        return 1.0
    nu_max = 0.5 * math.pi * (math.sqrt((g + 1.0) / (g - 1.0)) - 1.0)
    if nu >= nu_max:
        raise ValueError(f"nu = {nu} rad is beyond the Prandtl-Meyer limit {nu_max}")
    lo, hi = 1.0, 2.0
    while prandtl_meyer(hi, g) < nu:
        lo, hi = hi, 2.0 * hi
    M = 0.5 * (lo + hi)
    for _ in range(100):
        f = prandtl_meyer(M, g) - nu
        if f > 0:
            hi = M
        else:
            lo = M
        d = math.sqrt(M * M - 1.0) / (M * (1.0 + 0.5 * (g - 1.0) * M * M))
        Mn = M - f / d if d > 0 else 0.5 * (lo + hi)
        if not (lo < Mn < hi):
            Mn = 0.5 * (lo + hi)
        if abs(Mn - M) < 1e-14 * M:
            return Mn
        M = Mn
    return M


def deflection_of(M, beta, g=GAMMA):
    """theta(beta) of the theta-beta-M relation (radians)."""
    s2 = math.sin(beta) ** 2
    num = M * M * s2 - 1.0
    den = M * M * (g + math.cos(2.0 * beta)) + 2.0
    return math.atan(2.0 * num / (math.tan(beta) * den))


def shock_angle(M, theta, g=GAMMA, weak=True):
    """Oblique shock angle beta (radians) for deflection theta > 0, or None if
    no attached shock exists (theta above the maximum deflection).

    tan(beta) = (M^2 - 1 + 2 l cos((4 pi d + acos(x)) / 3)) / (3 (1 + (g-1)/2 M^2) tan(theta))
    with l^2 = (M^2-1)^2 - 3 (1 + (g-1)/2 M^2)(1 + (g+1)/2 M^2) tan^2(theta),
    x = ((M^2-1)^3 - 9 (1 + (g-1)/2 M^2)(1 + (g-1)/2 M^2 + (g+1)/4 M^4) tan^2(theta)) / l^3,
    d = 1 for the weak root, 0 for the strong one."""
    if theta <= 0.0 or M <= 1.0:
        return None
    t = math.tan(theta)
    m2 = M * M
    a = 1.0 + 0.5 * (g - 1.0) * m2
    l2 = (m2 - 1.0) ** 2 - 3.0 * a * (1.0 + 0.5 * (g + 1.0) * m2) * t * t
    if l2 <= 0.0:
        return None
    lam = math.sqrt(l2)
    chi = ((m2 - 1.0) ** 3 - 9.0 * a * (a + 0.25 * (g + 1.0) * m2 * m2) * t * t) / lam ** 3
    if abs(chi) > 1.0:
        return None
    d = 1.0 if weak else 0.0
    tb = (m2 - 1.0 + 2.0 * lam * math.cos((4.0 * math.pi * d + math.acos(chi)) / 3.0)) / (3.0 * a * t)
    beta = math.atan(tb)
    # the closed form cancels for small deflections: polish with Newton steps
    # on theta(beta) = theta (central-difference slope), kept inside (mu, pi/2)
    mu = math.asin(1.0 / M)
    for _ in range(30):
        h = 1e-7
        f = deflection_of(M, beta, g) - theta
        df = (deflection_of(M, beta + h, g) - deflection_of(M, beta - h, g)) / (2.0 * h)
        if df == 0.0:
            break
        nb = beta - f / df
        if nb <= mu:
            nb = 0.5 * (beta + mu)
        elif nb >= 0.5 * math.pi:
            nb = 0.5 * (beta + 0.5 * math.pi)
        if abs(nb - beta) < 1e-15:
            beta = nb
            break
        beta = nb
    return beta


def turn(M, theta, g=GAMMA):
    """Turns a uniform supersonic stream by theta (radians, > 0 compression
    through an oblique shock, < 0 isentropic expansion). Returns (M after,
    p after / p before, attached, beta or None)."""
    if abs(theta) < 1e-14:
        return M, 1.0, True, None
    if theta > 0.0:
        beta = shock_angle(M, theta, g)
        if beta is None:
            return M, 1.0, False, None
        mn1 = M * math.sin(beta)
        pr = 1.0 + 2.0 * g / (g + 1.0) * (mn1 * mn1 - 1.0)
        mn2 = math.sqrt((1.0 + 0.5 * (g - 1.0) * mn1 * mn1) / (g * mn1 * mn1 - 0.5 * (g - 1.0)))
        return mn2 / math.sin(beta - theta), pr, True, beta
    M2 = inverse_prandtl_meyer(prandtl_meyer(M, g) - theta, g)
    e = g / (g - 1.0)
    pr = ((1.0 + 0.5 * (g - 1.0) * M * M) / (1.0 + 0.5 * (g - 1.0) * M2 * M2)) ** e
    return M2, pr, True, None


@dataclass
class Panel:
    name: str
    deflection: float        # radians, > 0 compression
    attached: bool
    M: float
    p_ratio: float           # p / p_inf
    Cp: float
    beta: float | None       # shock angle (radians), front panels with a shock


def diamond_panels(M_inf, alpha_deg, eps_deg, g=GAMMA):
    """Theory state of the four panels (upper/lower x front/rear)."""
    a, e = math.radians(alpha_deg), math.radians(eps_deg)
    q = 0.5 * g * M_inf * M_inf          # dynamic pressure / p_inf
    out = {}
    for side, th in (("upper", e - a), ("lower", e + a)):
        Mf, rf, okf, bf = turn(M_inf, th, g)
        Mr, rr, okr, _ = turn(Mf, -2.0 * e, g) if okf else (Mf, 1.0, False, None)
        out[f"{side}_front"] = Panel(f"{side}_front", th, okf, Mf, rf, (rf - 1.0) / q, bf)
        out[f"{side}_rear"] = Panel(f"{side}_rear", -2.0 * e, okr, Mr, rf * rr, (rf * rr - 1.0) / q, None)
    return out


# -- field measurements --------------------------------------------------------

def _wall_cells(mesh, marker=2):
    """Wall faces with the marker, their midpoints and owner cells."""
    faces = np.asarray(mesh.faces)
    fm = np.asarray(mesh.face_markers)
    pts = np.asarray(mesh.points, dtype=np.float64)
    ids = np.where(fm == marker)[0]
    mids = pts[faces[ids]].mean(axis=1)
    fc = np.asarray(mesh.face_connectivity)
    owner = np.full(len(fm), -1, np.int64)
    cells = np.repeat(np.arange(fc.shape[0]), 3)
    owner[fc.reshape(-1)] = cells
    return ids, mids, owner[ids]


def panel_of(x, y, cx, cy):
    return np.where(x < cx, np.where(y >= cy, 0, 2), np.where(y >= cy, 1, 3))


def theory_state(mesh, panels, p_inf=1.0, rho_inf=1.0, gamma=GAMMA):
    """Synthetic conserved state: freestream density at rest everywhere, the
    theory pressure of each panel on the cells that own its wall faces
    (theory_validation_diamond.py:345-372), for the same wall integration."""
    md = mesh.metadata
    cx, cy = md["center"]["cx"], md["center"]["cy"]
    _, mids, cells = _wall_cells(mesh)
    p = np.full(mesh.n_cells, p_inf, np.float64)
    which = panel_of(mids[:, 0], mids[:, 1], cx, cy)
    for k, name in enumerate(PANELS):
        p[cells[which == k]] = p_inf * panels[name].p_ratio
    W = np.zeros((mesh.n_cells, 4), np.float64)
    W[:, 0] = rho_inf
    W[:, 3] = p / (gamma - 1.0)
    return W


def panel_plateaus(mesh, prim, M_inf, p_inf=1.0, rho_inf=1.0, gamma=GAMMA, margin=0.15):
    """Mean wall Cp and Mach per panel, dropping `margin` of the faces at
    each end of the panel (:270-286)."""
    md = mesh.metadata
    cx, cy = md["center"]["cx"], md["center"]["cy"]
    _, mids, cells = _wall_cells(mesh)
    which = panel_of(mids[:, 0], mids[:, 1], cx, cy)
    q = 0.5 * rho_inf * (M_inf * math.sqrt(gamma * p_inf / rho_inf)) ** 2
    r, u, v, p = (prim[cells, i] for i in range(4))
    mach = np.sqrt(u * u + v * v) / np.sqrt(gamma * p / r)
    out = {}
    for k, name in enumerate(PANELS):
        sel = np.where(which == k)[0]
        sel = sel[np.argsort(mids[sel, 0])]
        n = len(sel)
        lo, hi = int(n * margin), int(n * (1 - margin))
        if hi - lo < 1:
            lo, hi = 0, n
        s = sel[lo:hi]
        out[name] = dict(Cp=float(np.mean((p[s] - p_inf) / q)), M=float(np.mean(mach[s])))
    return out


def lsq_gradient(mesh, f):
    """Least-squares gradient of a cell field over the face neighbours
    (weights 1/|dx|^2, boundary faces skipped)."""
    nb = np.asarray(mesh.neighbors)
    bc = np.asarray(mesh.barycenter, np.float64)
    A = np.zeros((mesh.n_cells, 2, 2))
    b = np.zeros((mesh.n_cells, 2))
    for k in range(3):
        j = nb[:, k]
        ok = j >= 0
        dx = np.where(ok[:, None], bc[np.where(ok, j, 0)] - bc, 0.0)
        w = np.where(ok, 1.0 / np.maximum((dx * dx).sum(1), 1e-300), 0.0)
        df = np.where(ok, f[np.where(ok, j, 0)] - f, 0.0)
        A += w[:, None, None] * dx[:, :, None] * dx[:, None, :]
        b += (w * df)[:, None] * dx
    det = A[:, 0, 0] * A[:, 1, 1] - A[:, 0, 1] * A[:, 1, 0]
    det = np.where(np.abs(det) > 1e-300, det, np.inf)
    gx = (A[:, 1, 1] * b[:, 0] - A[:, 0, 1] * b[:, 1]) / det
    gy = (A[:, 0, 0] * b[:, 1] - A[:, 1, 0] * b[:, 0]) / det
    return np.stack([gx, gy], 1)


def shock_angle_of_field(mesh, p, alpha_deg, side, window_x=(-0.35, 0.45), y_factor=3.0, pct=97.0):
    """Leading-edge shock angle (degrees, to the freestream direction): the
    cells with the top (100 - pct) % of |grad p| in a window above (below)
    the front panel, their principal direction (:293-320)."""
    md = mesh.metadata
    cx, cy, chord, height = md["center"]["cx"], md["center"]["cy"], md["chord"], md["height"]
    pos = np.asarray(mesh.barycenter, np.float64)
    g = np.linalg.norm(lsq_gradient(mesh, np.asarray(p, np.float64)), axis=1)
    x_le = cx - 0.5 * chord
    inx = (pos[:, 0] > x_le + window_x[0] * chord) & (pos[:, 0] < x_le + window_x[1] * chord)
    if side == "upper":
        iny = (pos[:, 1] > cy + 0.005) & (pos[:, 1] < cy + y_factor * height)
    else:
        iny = (pos[:, 1] < cy - 0.005) & (pos[:, 1] > cy - y_factor * height)
    win = inx & iny
    if win.sum() < 10:
        return None
    gw = g[win]
    sel = gw >= np.percentile(gw, pct)
    if sel.sum() < 5:
        return None
    pts = pos[win][sel]
    _, _, vt = np.linalg.svd(pts - pts.mean(0))
    a = math.radians(alpha_deg)
    c = abs(float(vt[0] @ np.array([math.cos(a), math.sin(a)])))
    return math.degrees(math.acos(min(1.0, c)))


def diamond_eps_deg(mesh):
    """Half-angle of the diamond: atan(thickness / chord)."""
    md = mesh.metadata
    return math.degrees(math.atan(md["height"] / md["chord"]))


def validate_diamond(mesh, W, M_inf, alpha_deg, p_inf=1.0, rho_inf=1.0, gamma=GAMMA):
    """Theory vs field on one solved diamond state W (conserved, float).
    CD/CL of both are integrated by the library's FP64 wall sums."""
    from . import helper
    eps = diamond_eps_deg(mesh)
    panels = diamond_panels(M_inf, alpha_deg, eps, gamma)
    if not all(p.attached for p in panels.values()):
        return dict(attached=False, panels=panels)
    U = M_inf * math.sqrt(gamma * p_inf / rho_inf)
    L = float(mesh.metadata["obstacle_length"])
    W = np.asarray(W, np.float32)
    Wt = theory_state(mesh, panels, p_inf, rho_inf, gamma).astype(np.float32)
    cd = float(helper.get_drag_coefficient(W, mesh, rho_inf=rho_inf, U_inf=U, L_ref=L))
    cl = float(helper.get_lift_coefficient(W, mesh, rho_inf=rho_inf, U_inf=U, L_ref=L))
    cd_t = float(helper.get_drag_coefficient(Wt, mesh, rho_inf=rho_inf, U_inf=U, L_ref=L))
    cl_t = float(helper.get_lift_coefficient(Wt, mesh, rho_inf=rho_inf, U_inf=U, L_ref=L))
    Wd = W.astype(np.float64)
    r = Wd[:, 0]
    u, v = Wd[:, 1] / r, Wd[:, 2] / r
    prim = np.stack([r, u, v, (Wd[:, 3] - 0.5 * r * (u * u + v * v)) * (gamma - 1.0)], 1)
    plate = panel_plateaus(mesh, prim, M_inf, p_inf, rho_inf, gamma)
    beta = {s: shock_angle_of_field(mesh, prim[:, 3], alpha_deg, s) for s in ("upper", "lower")}
    return dict(attached=True, eps_deg=eps, cd=cd, cd_theory=cd_t, cl=cl, cl_theory=cl_t, panels=panels,
                plateaus=plate, beta_field_deg=beta,
                beta_theory_deg={"upper": None if panels["upper_front"].beta is None
                                 else math.degrees(panels["upper_front"].beta),
                                 "lower": None if panels["lower_front"].beta is None
                                 else math.degrees(panels["lower_front"].beta)})
