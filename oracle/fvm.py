"""CPU oracle (TEST INFRASTRUCTURE ONLY): numpy restatement of the reference
explicit finite-volume Euler path.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module, and only as the checker. The product path never
calls it.

Every function mirrors the reference function named in its docstring, op for
op, in the array dtype given by the caller (float32 reproduces the reference's
``JAX_ENABLE_X64=false`` arithmetic, float64 the X64 runs). Python float
constants combine with arrays under numpy-2 weak promotion, which matches
JAX's weak typing. Integer powers are written as products (JAX lowers
``x**n`` to multiplies). The only deliberate deviations: the 2x2 solve is an
explicit LU with partial pivoting (XLA's LU rounding is not reproducible
here) and, as the project requires, the stationarity norms and the CD/CL wall
sums are accumulated in float64.

Pinning: see tests/test_oracle_*.py (closed forms, invariants, textbook
routines, brute force). A function with no such check is marked
"parity unpinned" in its docstring.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np

DEFAULT_INLET = (1.0, 1.0, 0.0, 1.0)


# ---------------------------------------------------------------------------
# mesh view
# ---------------------------------------------------------------------------

def omesh(mesh, dtype=np.float32):
    """Arrays of a mesh cast to the run dtype; integer arrays kept as int64."""
    dt = np.dtype(dtype)
    md = getattr(mesh, "metadata", None)
    om = SimpleNamespace(
        dtype=dt,
        points=np.asarray(mesh.points).astype(dt),
        tris=np.asarray(mesh.tris).astype(np.int64),
        neighbors=np.asarray(mesh.neighbors).astype(np.int64),
        faces=np.asarray(mesh.faces).astype(np.int64),
        face_markers=np.asarray(mesh.face_markers).astype(np.int64),
        face_connectivity=np.asarray(mesh.face_connectivity).astype(np.int64),
        face_connectivity_opposite=np.asarray(mesh.face_connectivity_opposite).astype(np.int64),
        barycenter=np.asarray(mesh.barycenter).astype(dt),
        area=np.asarray(mesh.area).astype(dt),
        surface=np.asarray(mesh.surface).astype(dt),
        normals=np.asarray(mesh.normals).astype(dt),
        inlet_subsonic=np.asarray(getattr(mesh, "inlet_subsonic", DEFAULT_INLET)).astype(dt),
        metadata=md if isinstance(md, dict) else {},
    )
    return om


def _c(x, like):
    return np.asarray(x, dtype=like.dtype)


def face_midpoints(om, face_ids):
    """mean of the two face vertices (euler_solver.py:49, helper.py:139)."""
    p = om.points[om.faces[face_ids]]          # (...,2,2)
    return (p[..., 0, :] + p[..., 1, :]) / 2


# ---------------------------------------------------------------------------
# state conversions (helper.py:25-127)
# ---------------------------------------------------------------------------

def getConserved(prim, gamma=1.4, M=1):
    """helper.py:25-35"""
    rho, u, v, P = prim[..., 0], prim[..., 1], prim[..., 2], prim[..., 3]
    E = P / ((gamma - 1) * M ** 2) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], axis=-1)


def getPrimitive(W, gamma=1.4, M=1):
    """helper.py:37-46"""
    rho, mx, my, E = W[..., 0], W[..., 1], W[..., 2], W[..., 3]
    u = mx / rho
    v = my / rho
    P = (E - 0.5 * rho * (u * u + v * v)) * M ** 2 * (gamma - 1)
    return np.stack([rho, u, v, P], axis=-1)


def get_specific_entropy(W, gamma=1.4):
    """helper.py:48-55"""
    rho = W[..., 0]
    u = W[..., 1] / rho
    v = W[..., 2] / rho
    P = (W[..., 3] - 0.5 * rho * (u * u + v * v)) * (gamma - 1)
    return np.log(P / rho ** gamma)


def getEntropyVariables(W, gamma=1.4):
    """helper.py:57-70"""
    rho = W[..., 0]
    u = W[..., 1] / rho
    v = W[..., 2] / rho
    P = (W[..., 3] - 0.5 * rho * (u * u + v * v)) * (gamma - 1)
    s = get_specific_entropy(W, gamma=gamma)
    V1 = (gamma - s) / (gamma - 1) - 0.5 * rho * (u * u + v * v) / P
    V2 = W[..., 1] / P
    V3 = W[..., 2] / P
    V4 = -rho / P
    return np.stack([V1, V2, V3, V4], axis=-1)


def getConserved_from_Entropy(eta, gamma=1.4):
    """helper.py:72-86"""
    e1, e2, e3, e4 = eta[..., 0], eta[..., 1], eta[..., 2], eta[..., 3]
    u = -e2 / e4
    v = -e3 / e4
    s = gamma - (gamma - 1) * (e1 - 0.5 * (u * u + v * v) * e4)
    rho = (-e4) ** (1 / (1 - gamma)) * np.exp(s / (1 - gamma))
    P = -rho / e4
    E = P / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], axis=-1)


def jvp_conserved_from_entropy(eta, deta, gamma=1.4):
    """Forward-mode derivative of ``getConserved_from_Entropy`` along
    ``deta`` (the ``jax.jvp`` call of euler_solver.py:234-238), written out by
    the chain rule over the same operation graph."""
    e1, e2, e3, e4 = (eta[..., i] for i in range(4))
    d1, d2, d3, d4 = (deta[..., i] for i in range(4))
    u = -e2 / e4
    du = -(d2 / e4 - e2 * d4 / (e4 * e4))
    v = -e3 / e4
    dv = -(d3 / e4 - e3 * d4 / (e4 * e4))
    q = 0.5 * (u * u + v * v)
    dq = u * du + v * dv
    s = gamma - (gamma - 1) * (e1 - q * e4)
    ds = -(gamma - 1) * (d1 - (dq * e4 + q * d4))
    k = 1 / (1 - gamma)
    a = (-e4) ** k
    da = k * (-e4) ** (k - 1) * (-d4)
    b = np.exp(s / (1 - gamma))
    db = b * ds / (1 - gamma)
    rho = a * b
    drho = da * b + a * db
    P = -rho / e4
    dP = -(drho / e4 - rho * d4 / (e4 * e4))
    kin = u * u + v * v
    dkin = 2 * u * du + 2 * v * dv
    dE = dP / (gamma - 1) + 0.5 * (drho * kin + rho * dkin)
    return np.stack([drho, drho * u + rho * du, drho * v + rho * dv, dE], axis=-1)


def get_IsmailRoe_variables(W, gamma=1.4):
    """helper.py:88-99"""
    rho = W[..., 0]
    u = W[..., 1] / rho
    v = W[..., 2] / rho
    P = (W[..., 3] - 0.5 * rho * (u * u + v * v)) * (gamma - 1)
    V = np.stack([np.ones_like(u), u, v, P], axis=-1)
    return np.sqrt(rho / P)[..., None] * V


def get_Roe_averaged_state(WL, WR, gamma=1.4):
    """helper.py:101-127"""
    def side(W):
        rho = W[..., 0]
        rho = np.where(rho <= 0, _c(1e-6, rho), rho)
        u = W[..., 1] / rho
        v = W[..., 2] / rho
        E = W[..., 3]
        P = (E - 0.5 * rho * (u * u + v * v)) * (gamma - 1)
        P = np.where(P <= 0, _c(1e-6, P), P)
        return rho, u, v, E, P
    rL, uL, vL, EL, PL = side(WL)
    rR, uR, vR, ER, PR = side(WR)
    sL, sR = np.sqrt(rL), np.sqrt(rR)
    rho = sL * sR
    u = (sL * uL + sR * uR) / (sL + sR)
    v = (sL * vL + sR * vR) / (sL + sR)
    HL = (EL + PL) / rL
    HR = (ER + PR) / rR
    H = (sL * HL + sR * HR) / (sL + sR)
    return np.stack([rho, u, v, H], axis=-1)


# ---------------------------------------------------------------------------
# boundary conditions (helper.py:168-256)
# ---------------------------------------------------------------------------

def _face_marker(om):
    return om.face_markers[om.face_connectivity]          # (N,3)


def BC_slipwall(WR, WL, om, bc_type=2):
    """helper.py:196-204 (default gamma=1.4, M=1 in both conversions)."""
    prim = getPrimitive(WL)
    n = om.normals
    vn = (prim[..., 1] - 0.0) * n[..., 0] + (prim[..., 2] - 0.0) * n[..., 1]
    vb = (prim[..., 1:3] - 0.0) - 2 * vn[..., None] * n
    pb = prim.copy()
    pb[..., 1:3] = vb + 0.0
    Wb = getConserved(pb)
    mask = (_face_marker(om) == bc_type)[..., None]
    return np.where(mask, Wb, WR)


def BC_inflow(W, om, value, bc_type=3):
    """helper.py:173-177"""
    mask = (_face_marker(om) == bc_type)[..., None]
    return np.where(mask, np.asarray(value, dtype=W.dtype), W)


def BC_outflow(WR, WL, om, bc_type=4):
    """helper.py:168-171"""
    mask = (_face_marker(om) == bc_type)[..., None]
    return np.where(mask, WL, WR)


def BC_subsonic_inlet(WR, WL, om, bc_type=5):
    """helper.py:179-194 (hardcoded gamma 1.4)."""
    prim = getPrimitive(WL)
    pb = prim.copy()
    pb[..., :3] = om.inlet_subsonic[:3]
    rho, u, v, P = pb[..., 0], pb[..., 1], pb[..., 2], pb[..., 3]
    E = P / (1.4 - 1) + 0.5 * rho * (u * u + v * v)
    Wb = np.stack([rho, rho * u, rho * v, E], axis=-1)
    mask = (_face_marker(om) == bc_type)[..., None]
    return np.where(mask, Wb, WR)


def wall_markers_of(om, kw):
    wm = kw.get("wall_markers", om.metadata.get("wall_markers", [2]))
    try:
        return tuple(int(m) for m in wm)
    except TypeError:
        return (int(wm),)


def BC_state(WR, WL, om, **kw):
    """helper.py:233-256: walls, then inflow(3), outflow(4), subsonic(5)."""
    value = kw.get("value", np.array(DEFAULT_INLET, dtype=WR.dtype))
    for m in wall_markers_of(om, kw):
        WR = BC_slipwall(WR, WL, om, bc_type=m)
    WR = BC_inflow(WR, om, value, bc_type=3)
    WR = BC_outflow(WR, WL, om, bc_type=4)
    WR = BC_subsonic_inlet(WR, WL, om, bc_type=5)
    return WR


# ---------------------------------------------------------------------------
# gradient, limiter, MUSCL (helper.py:136-153, euler_solver.py:36-122)
# ---------------------------------------------------------------------------

def solve2x2(A, b):
    """LU with partial pivoting of a 2x2 system, batched.
    A: (...,2,2), b: (...,2). Parity unpinned against XLA's LU rounding;
    pinned against numpy.linalg.solve to rounding in the tests."""
    a00, a01, a10, a11 = A[..., 0, 0], A[..., 0, 1], A[..., 1, 0], A[..., 1, 1]
    b0, b1 = b[..., 0], b[..., 1]
    swap = np.abs(a10) > np.abs(a00)
    p00 = np.where(swap, a10, a00)
    p01 = np.where(swap, a11, a01)
    p10 = np.where(swap, a00, a10)
    p11 = np.where(swap, a01, a11)
    c0 = np.where(swap, b1, b0)
    c1 = np.where(swap, b0, b1)
    l = p10 / p00
    u11 = p11 - l * p01
    y1 = c1 - l * c0
    x1 = y1 / u11
    x0 = (c0 - p01 * x1) / p00
    return np.stack([x0, x1], axis=-1)


def getgradientLSQ(WL, WR, om):
    """helper.py:136-153: weighted least-squares gradient, returns (N,4,2)."""
    dx = om.barycenter[om.neighbors] - om.barycenter[:, None, :]
    rep = face_midpoints(om, om.face_connectivity)
    rep = 2 * (rep - om.barycenter[:, None, :])
    dx = np.where((_face_marker(om) > 0)[..., None], rep, dx)
    dW = WR - WL
    nrm = np.sqrt(dx[..., 0] * dx[..., 0] + dx[..., 1] * dx[..., 1])
    w = 1 / (nrm * nrm)
    wdx = w[..., None] * dx
    wdW = w[..., None] * dW
    A = np.empty(dx.shape[:-2] + (2, 2), dtype=dx.dtype)
    for k in range(2):
        for l in range(2):
            A[..., k, l] = (wdx[:, 0, k] * dx[:, 0, l] + wdx[:, 1, k] * dx[:, 1, l]) + wdx[:, 2, k] * dx[:, 2, l]
    b = np.empty(dW.shape[:-2] + (dW.shape[-1], 2), dtype=dx.dtype)
    for v in range(dW.shape[-1]):
        for l in range(2):
            b[..., v, l] = (wdW[:, 0, v] * dx[:, 0, l] + wdW[:, 1, v] * dx[:, 1, l]) + wdW[:, 2, v] * dx[:, 2, l]
    Arep = np.repeat(A[:, None], b.shape[-2], axis=1)
    return solve2x2(Arep, b)


def venkatakrishnan(a, b):
    """euler_solver.py:36-39 with K=0 (omega = 0)."""
    return (a * a + 2 * a * b + 0.0) / (a * a + 2 * (b * b) + a * b + 0.0 + 1e-7)


def _delta(dx, grad):
    """einsum('ijl,ikl->ijk'): (N,3,2) x (N,4,2) -> (N,3,4)."""
    return dx[:, :, None, 0] * grad[:, None, :, 0] + dx[:, :, None, 1] * grad[:, None, :, 1]


def getlimiting(WL, WR, grad, om):
    """euler_solver.py:45-67"""
    cat = np.concatenate([WL, WR], axis=-2)
    Wm = np.min(cat, axis=-2)
    WM = np.max(cat, axis=-2)
    mid = face_midpoints(om, om.face_connectivity)
    dx = mid - om.barycenter[:, None, :]
    D = _delta(dx, grad)
    phi = np.ones_like(D)
    eps = _c(1e-12 if WL.dtype == np.float64 else 1e-7, WL)
    phi = np.where(D > eps, venkatakrishnan(WM[:, None, :] - WL, D), phi)
    phi = np.where(D < -eps, venkatakrishnan(Wm[:, None, :] - WL, D), phi)
    return np.min(phi, axis=-2)


def MUSCL(WL, WR, grad, om):
    """euler_solver.py:69-98"""
    phi = getlimiting(WL, WR, grad, om)
    mid = face_midpoints(om, om.face_connectivity)
    dx = mid - om.barycenter[:, None, :]
    D = _delta(dx, grad)
    mid_n = face_midpoints(om, om.face_connectivity_opposite[om.face_connectivity])
    dxn = mid_n - om.barycenter[om.neighbors]
    bnd = (_face_marker(om) > 1)
    dxn = np.where(bnd[..., None], -dx, dxn)
    gn = grad[om.neighbors]                               # (N,3,4,2)
    Dn = dxn[:, :, None, 0] * gn[..., 0] + dxn[:, :, None, 1] * gn[..., 1]
    phiL = np.where(bnd[..., None], _c(0.0, WL), np.repeat(phi[:, None, :], 3, axis=1))
    WLm = WL + phiL * D
    phiR = np.where(bnd[..., None], _c(0.0, WL), phi[om.neighbors])
    WRm = WR + phiR * Dn
    return WLm, WRm


RECON_CONSTANT = ("constant", "first_order", "none", "godunov")
RECON_MUSCL = ("muscl", "second_order", "2nd_order")
FLUXES = ("lf", "lax_friedrichs", "lax-friedrichs", "laxfriedrichs", "rusanov", "tadmor",
          "ausm", "hll", "hllc", "roe")


def reconstruction_mode(kw):
    return str(kw.get("reconstruction", kw.get("scheme", "muscl"))).lower()


def flux_type(kw):
    return str(kw.get("flux", kw.get("numerical_flux", "rusanov"))).lower()


def build_reconstruction(W, om, **kw):
    """euler_solver.py:106-122"""
    WL = np.repeat(W[:, None, :], 3, axis=1)
    WR = W[om.neighbors]
    WR = BC_state(WR, WL, om, **kw)
    mode = reconstruction_mode(kw)
    if mode in RECON_CONSTANT:
        return WL, WR
    if mode in RECON_MUSCL:
        g, M = kw.get("gamma", 1.4), kw.get("M", 1.0)
        PL = getPrimitive(WL, gamma=g, M=M)
        PR = getPrimitive(WR, gamma=g, M=M)
        grad = getgradientLSQ(PL, PR, om)
        PL, PR = MUSCL(PL, PR, grad, om)
        return getConserved(PL, gamma=g, M=M), getConserved(PR, gamma=g, M=M)
    raise ValueError(f"Unknown reconstruction scheme '{mode}'. Expected 'constant' or 'muscl'.")


# ---------------------------------------------------------------------------
# numerical fluxes (euler_solver.py:124-591)
# ---------------------------------------------------------------------------

def _prims(WL, WR, gamma, M):
    PL = getPrimitive(WL, gamma=gamma, M=M)
    PR = getPrimitive(WR, gamma=gamma, M=M)
    return (PL[..., 0], PL[..., 1], PL[..., 2], PL[..., 3],
            PR[..., 0], PR[..., 1], PR[..., 2], PR[..., 3])


def HLLC(WL, WR, n, gamma=1.4, M=1.0, **_):
    """euler_solver.py:273-374; returns central - F_HLLC."""
    rL, uL, vL, PL, rR, uR, vR, PR = _prims(WL, WR, gamma, M)
    nx, ny = n[..., 0], n[..., 1]
    unL = uL * nx + vL * ny
    unR = uR * nx + vR * ny
    tx, ty = -ny, nx
    utL = uL * tx + vL * ty
    utR = uR * tx + vR * ty
    aL = np.sqrt(np.abs(gamma * PL / rL)) / M
    aR = np.sqrt(np.abs(gamma * PR / rR)) / M
    SL = np.minimum(unL - aL, unR - aR)
    SR = np.maximum(unL + aL, unR + aR)
    SL = np.where(SL == SR, SL - 1e-6, SL)
    SR = np.where(SL == SR, SR + 1e-6, SR)
    EL, ER = WL[..., 3], WR[..., 3]
    FL = np.stack([rL * unL, rL * uL * unL + PL * nx / (M ** 2), rL * vL * unL + PL * ny / (M ** 2),
                   (EL + PL / (M ** 2)) * unL], axis=-1)
    FR = np.stack([rR * unR, rR * uR * unR + PR * nx / (M ** 2), rR * vR * unR + PR * ny / (M ** 2),
                   (ER + PR / (M ** 2)) * unR], axis=-1)
    Fc = 0.5 * (FL + FR)
    den = rL * (SL - unL) - rR * (SR - unR)
    den = np.where(den == 0, _c(1e-12, den), den)
    Ss = (PR - PL + rL * unL * (SL - unL) - rR * unR * (SR - unR)) / den
    PsL = PL + rL * (SL - unL) * (Ss - unL)
    PsR = PR + rR * (SR - unR) * (Ss - unR)
    Ps = 0.5 * (PsL + PsR)
    rsL = rL * (SL - unL) / (SL - Ss)
    usxL = Ss * nx + utL * tx
    usyL = Ss * ny + utL * ty
    EsL = ((SL - unL) * EL - PL * unL + Ps * Ss) / (SL - Ss)
    UsL = np.stack([rsL, rsL * usxL, rsL * usyL, EsL], axis=-1)
    rsR = rR * (SR - unR) / (SR - Ss)
    usxR = Ss * nx + utR * tx
    usyR = Ss * ny + utR * ty
    EsR = ((SR - unR) * ER - PR * unR + Ps * Ss) / (SR - Ss)
    UsR = np.stack([rsR, rsR * usxR, rsR * usyR, EsR], axis=-1)
    SLe, SRe, Sse = SL[..., None], SR[..., None], Ss[..., None]
    with np.errstate(all="ignore"):
        F = np.where(SLe >= 0, FL,
                     np.where((SLe <= 0) & (Sse >= 0), FL + SLe * (UsL - WL),
                              np.where((Sse <= 0) & (SRe >= 0), FR + SRe * (UsR - WR), FR)))
    return Fc - F


def Roe(WL, WR, n, gamma=1.4, **_):
    """euler_solver.py:145-199 (R^T, not R^-1, scaled by S)."""
    avg = get_Roe_averaged_state(WL, WR, gamma=gamma)
    r, u, v, H = avg[..., 0], avg[..., 1], avg[..., 2], avg[..., 3]
    a = np.sqrt(np.abs((gamma - 1) * (H - 0.5 * (u * u + v * v))))
    nx, ny = n[..., 0], n[..., 1]
    R = np.zeros(WL.shape + (4,), dtype=WL.dtype)
    R[..., 0, 0] = 1.0
    R[..., 1, 0] = u - a * nx
    R[..., 2, 0] = v - a * ny
    R[..., 3, 0] = H - a * (u * nx + v * ny)
    R[..., 0, 1] = 1.0
    R[..., 1, 1] = u
    R[..., 2, 1] = v
    R[..., 3, 1] = 0.5 * (u * u + v * v)
    R[..., 0, 2] = 0.0
    R[..., 1, 2] = -ny
    R[..., 2, 2] = nx
    R[..., 3, 2] = -u * ny + v * nx
    R[..., 0, 3] = 1.0
    R[..., 1, 3] = u + a * nx
    R[..., 2, 3] = v + a * ny
    R[..., 3, 3] = H + a * (u * nx + v * ny)
    un = u * nx + v * ny
    Lam = np.stack([un - a, un, un, un + a], axis=-1)
    dv = getEntropyVariables(WR, gamma=gamma) - getEntropyVariables(WL, gamma=gamma)
    S = np.stack([r / (2 * gamma), (gamma - 1) * r / gamma, r / gamma, r / (2 * gamma)], axis=-1)
    RH = ((R[..., 0, :] * dv[..., 0:1] + R[..., 1, :] * dv[..., 1:2]) + R[..., 2, :] * dv[..., 2:3]) \
        + R[..., 3, :] * dv[..., 3:4]
    DRH = np.abs(Lam * S) * RH
    RDRH = ((R[..., :, 0] * DRH[..., 0:1] + R[..., :, 1] * DRH[..., 1:2]) + R[..., :, 2] * DRH[..., 2:3]) \
        + R[..., :, 3] * DRH[..., 3:4]
    return 0.5 * RDRH


def Rusanov(WL, WR, n, gamma=1.4, M=1.0, entropy=False, **_):
    """euler_solver.py:201-244"""
    rL, uL, vL, PL, rR, uR, vR, PR = _prims(WL, WR, gamma, M)
    nx, ny = n[..., 0], n[..., 1]
    CL = np.sqrt(np.abs(gamma * PL / rL)) / M + np.abs(uL * nx + vL * ny)
    CR = np.sqrt(np.abs(gamma * PR / rR)) / M + np.abs(uR * nx + vR * ny)
    Cm = np.maximum(CR, CL)
    if entropy:
        eL = getEntropyVariables(WL, gamma=gamma)
        eR = getEntropyVariables(WR, gamma=gamma)
        ebar = 0.5 * (eL + eR)
        diss = jvp_conserved_from_entropy(ebar, eR - eL, gamma=gamma)
        return 0.5 * Cm[..., None] * diss
    return 0.5 * Cm[..., None] * (WR - WL)


def LaxFriedrichs(WL, WR, n, gamma=1.4, M=1.0, **_):
    """euler_solver.py:246-271 (global max wave speed)."""
    rL, uL, vL, PL, rR, uR, vR, PR = _prims(WL, WR, gamma, M)
    nx, ny = n[..., 0], n[..., 1]
    CL = np.sqrt(np.abs(gamma * PL / rL)) / M + np.abs(uL * nx + vL * ny)
    CR = np.sqrt(np.abs(gamma * PR / rR)) / M + np.abs(uR * nx + vR * ny)
    Cg = np.maximum(np.max(CL), np.max(CR))
    return 0.5 * Cg * (WR - WL)


def log_mean(aL, aR):
    """euler_solver.py:422-430"""
    xi = aL / aR
    f = (xi - 1.0) / (xi + 1.0)
    u = f * f
    with np.errstate(all="ignore"):
        F = np.where(u < 1e-3, 1.0 + u / 3.0 + (u * u) / 5.0 + (u * u * u) / 7.0,
                     np.log(xi) / (2.0 * f))
    return (aL + aR) / (2.0 * F)


def ausm_split_mach(Mn):
    """euler_solver.py:432-445"""
    absm = np.abs(Mn)
    sub = absm < 1.0
    beta = 1.0 / 8.0
    Mp_sub = 0.25 * ((Mn + 1.0) * (Mn + 1.0)) + beta * ((Mn * Mn - 1.0) * (Mn * Mn - 1.0))
    Mm_sub = -0.25 * ((Mn - 1.0) * (Mn - 1.0)) - beta * ((Mn * Mn - 1.0) * (Mn * Mn - 1.0))
    return np.where(sub, Mp_sub, 0.5 * (Mn + absm)), np.where(sub, Mm_sub, 0.5 * (Mn - absm))


def ausm_split_pressure(Mn):
    """euler_solver.py:447-460"""
    absm = np.abs(Mn)
    sub = absm < 1.0
    alpha = 3.0 / 16.0
    q = (Mn * Mn - 1.0) * (Mn * Mn - 1.0)
    Pp_sub = 0.25 * ((Mn + 1.0) * (Mn + 1.0)) * (2.0 - Mn) + alpha * Mn * q
    Pm_sub = 0.25 * ((Mn - 1.0) * (Mn - 1.0)) * (2.0 + Mn) - alpha * Mn * q
    return (np.where(sub, Pp_sub, 0.5 * (1.0 + np.sign(Mn))),
            np.where(sub, Pm_sub, 0.5 * (1.0 - np.sign(Mn))))


def AUSM(WL, WR, n, gamma=1.4, M=1.0, **_):
    """euler_solver.py:462-530; returns central - F_AUSM+."""
    rL, uL, vL, PL, rR, uR, vR, PR = _prims(WL, WR, gamma, M)
    nx, ny = n[..., 0], n[..., 1]
    unL = uL * nx + vL * ny
    unR = uR * nx + vR * ny
    aL = np.sqrt(np.abs(gamma * PL / rL)) / M
    aR = np.sqrt(np.abs(gamma * PR / rR)) / M
    af = 0.5 * (aL + aR)
    MnL = unL / (aL + 1e-12)
    MnR = unR / (aR + 1e-12)
    MpL, _ = ausm_split_mach(MnL)
    _, MmR = ausm_split_mach(MnR)
    PpL, _ = ausm_split_pressure(MnL)
    _, PmR = ausm_split_pressure(MnR)
    mL = af * MpL * rL
    mR = af * MmR * rR
    pf = PpL * PL + PmR * PR
    EL, ER = WL[..., 3], WR[..., 3]
    HL = (EL + PL / (M ** 2)) / rL
    HR = (ER + PR / (M ** 2)) / rR
    F = np.stack([mL + mR, mL * uL + mR * uR + pf * nx / (M ** 2),
                  mL * vL + mR * vR + pf * ny / (M ** 2), mL * HL + mR * HR], axis=-1)
    FcL = np.stack([rL * unL, rL * uL * unL + PL * nx / (M ** 2), rL * vL * unL + PL * ny / (M ** 2),
                    (EL + PL / (M ** 2)) * unL], axis=-1)
    FcR = np.stack([rR * unR, rR * uR * unR + PR * nx / (M ** 2), rR * vR * unR + PR * ny / (M ** 2),
                    (ER + PR / (M ** 2)) * unR], axis=-1)
    return 0.5 * (FcL + FcR) - F


def build_flux(ftype, WL, WR, n, **kw):
    """euler_solver.py:124-139"""
    kw = dict(kw)
    kw.pop("flux", None)
    kw.pop("numerical_flux", None)
    if ftype in ("lf", "lax_friedrichs", "lax-friedrichs", "laxfriedrichs"):
        return LaxFriedrichs(WL, WR, n, **kw)
    if ftype == "rusanov":
        return Rusanov(WL, WR, n, **kw)
    if ftype == "tadmor":
        kw["entropy"] = True
        return Rusanov(WL, WR, n, **kw)
    if ftype == "ausm":
        return AUSM(WL, WR, n, **kw)
    if ftype in ("hll", "hllc"):
        return HLLC(WL, WR, n, **kw)
    if ftype == "roe":
        return Roe(WL, WR, n, **kw)
    raise ValueError(f"Unknown flux type '{ftype}'. Expected one of: 'lf', 'rusanov', 'tadmor', "
                     f"'ausm', 'hll', 'hllc', 'roe'.")


def _solver_kw(kw):
    out = {k: v for k, v in kw.items() if k in ("gamma", "M", "entropy")}
    return out


def apply_face_surfaces(F, s):
    """euler_solver.py:141-143"""
    Fs = s[..., None] * F
    return (Fs[:, 0] + Fs[:, 1]) + Fs[:, 2]


def getFlux(WL, WR, n, s, **kw):
    """euler_solver.py:376-420"""
    g, M = kw.get("gamma", 1.4), kw.get("M", 1.0)
    ft = flux_type(kw)
    rL, uL, vL, PL, rR, uR, vR, PR = _prims(WL, WR, g, M)
    nx, ny = n[..., 0], n[..., 1]
    eL, eR = WL[..., 3], WR[..., 3]
    i2 = 1 / M ** 2
    frL = rL * (uL * nx + vL * ny)
    fuL = rL * uL * (uL * nx + vL * ny) + i2 * PL * nx
    fvL = rL * vL * (uL * nx + vL * ny) + i2 * PL * ny
    fEL = (eL + i2 * PL) * (uL * nx + vL * ny)
    frR = rR * (uR * nx + vR * ny)
    fuR = rR * uR * (uR * nx + vR * ny) + i2 * PR * nx
    fvR = rR * vR * (uR * nx + vR * ny) + i2 * PR * ny
    fER = (eR + i2 * PR) * (uR * nx + vR * ny)
    F = np.stack([0.5 * (frL + frR), 0.5 * (fuL + fuR), 0.5 * (fvL + fvR), 0.5 * (fEL + fER)], axis=-1)
    F = F - build_flux(ft, WL, WR, n, **_solver_kw(kw))
    return apply_face_surfaces(F, s)


def getFlux_Tadmor(WL, WR, n, s, **kw):
    """euler_solver.py:532-591 (Ismail-Roe flux minus the chosen dissipation)."""
    g = kw.get("gamma", 1.4)
    ft = flux_type(kw)

    def prim(W):
        rho = W[..., 0]
        u = W[..., 1] / rho
        v = W[..., 2] / rho
        P = (g - 1) * (W[..., 3] - 0.5 * rho * (u * u + v * v))
        return rho, u, v, P
    rL, uL, vL, PL = prim(WL)
    rR, uR, vR, PR = prim(WR)
    ZL = get_IsmailRoe_variables(WL, gamma=g)
    ZR = get_IsmailRoe_variables(WR, gamma=g)
    Zb = 0.5 * (ZL + ZR)
    nx, ny = n[..., 0], n[..., 1]
    Z1 = log_mean(ZL[..., 0], ZR[..., 0])
    Z4 = log_mean(ZL[..., 3], ZR[..., 3])
    rh = Zb[..., 0] * Z4
    uh = Zb[..., 1] / Zb[..., 0]
    vh = Zb[..., 2] / Zb[..., 0]
    P1 = Zb[..., 3] / Zb[..., 0]
    P2 = (g + 1) / (2 * g) * Z4 / Z1 + (g - 1) / (2 * g) * Zb[..., 3] / Zb[..., 0]
    ah = np.sqrt(g * P2 / rh)
    Hh = ah * ah / (g - 1) + (uh * uh + vh * vh) / 2
    F1 = rh * (uh * nx + vh * ny)
    F2 = rh * uh * (uh * nx + vh * ny) + P1 * nx
    F3 = rh * vh * (uh * nx + vh * ny) + P1 * ny
    F4 = rh * (uh * nx + vh * ny) * Hh
    F = np.stack([F1, F2, F3, F4], axis=-1)
    F = F - build_flux(ft, WL, WR, n, **_solver_kw(kw))
    return apply_face_surfaces(F, s)


# ---------------------------------------------------------------------------
# sponge, residual, time steppers (helper.py:262-295, euler_solver.py:602-667)
# ---------------------------------------------------------------------------

def get_sponge_source(W, om, value, gamma=1.4, M=1.0, width=None, strength=None):
    """helper.py:262-295"""
    if om.metadata.get("case") == "bump":
        return np.zeros_like(W)
    bary = om.barycenter
    y_min = np.min(om.points[:, 1])
    y_max = np.max(om.points[:, 1])
    Ly = y_max - y_min
    if width is None:
        width = 0.2 * Ly
    value = np.asarray(value, dtype=W.dtype)
    pinf = getPrimitive(value, gamma=gamma, M=M)
    U = np.sqrt(pinf[1] * pinf[1] + pinf[2] * pinf[2])
    a = np.sqrt(gamma * pinf[3] / pinf[0]) / M
    if strength is None:
        strength = 1.0 * (U + a) / width
    dist = np.minimum(bary[:, 1] - y_min, y_max - bary[:, 1])
    blend = np.clip((width - dist) / width, 0.0, 1.0).astype(W.dtype)
    sigma = strength * (blend * blend * blend)
    return sigma[:, None] * (np.broadcast_to(value, W.shape) - W)


def residual(W, mesh_or_om, **kw):
    """euler_solver.py:602-628: R = sum(Flux*len)/area - sponge."""
    om = mesh_or_om if isinstance(mesh_or_om, SimpleNamespace) else omesh(mesh_or_om, W.dtype)
    WL, WR = build_reconstruction(W, om, **kw)
    s = om.surface[om.face_connectivity]
    ft = flux_type(kw)
    if ft == "tadmor" or kw.get("entropy", False):
        F = getFlux_Tadmor(WL, WR, om.normals, s, **kw)
    else:
        F = getFlux(WL, WR, om.normals, s, **kw)
    value = kw.get("value", np.array(DEFAULT_INLET, dtype=W.dtype))
    sp = get_sponge_source(W, om, value, gamma=kw.get("gamma", 1.4), M=kw.get("M", 1.0),
                           width=kw.get("sponge_width", None), strength=kw.get("sponge_strength", None))
    return F / om.area[:, None] - sp


def time_step_Euler(W, om, dt, **kw):
    """euler_solver.py:630-633"""
    return W - dt * residual(W, om, **kw)


def time_step_RK2_SSP(W, om, dt, **kw):
    """euler_solver.py:635-643"""
    k1 = residual(W, om, **kw)
    W1 = W - dt * k1
    k2 = residual(W1, om, **kw)
    return 0.5 * (W + W1 - dt * k2)


def time_step_RK2(W, om, dt, **kw):
    """euler_solver.py:645-653"""
    F1 = residual(W, om, **kw)
    W1 = W - dt * F1
    F2 = residual(W1, om, **kw)
    Wn = W - dt * F2
    return 0.5 * (Wn + W1)


def time_step_RK4(W, om, dt, **kw):
    """euler_solver.py:655-667"""
    F1 = residual(W, om, **kw)
    W1 = W - dt / 2 * F1
    F2 = residual(W1, om, **kw)
    W2 = W - dt / 2 * F2
    F3 = residual(W2, om, **kw)
    W3 = W - dt * F3
    F4 = residual(W3, om, **kw)
    return W - dt / 6 * (F1 + 2 * F2 + 2 * F3 + F4)


SCHEMES = {"EE": time_step_Euler, "RK2": time_step_RK2, "RK4": time_step_RK4,
           "SRK2": time_step_RK2_SSP, "SSP_RK2": time_step_RK2_SSP}


def cell_dt(W, om, gamma=1.4, M=1.0):
    """area / sum_k (c + |u . n_k|) len_k per cell: helper.py:8-18 before the
    minimum over cells and the CFL factor."""
    p = getPrimitive(W, gamma=gamma, M=M)
    rho, u, v, P = p[:, 0], p[:, 1], p[:, 2], p[:, 3]
    c = np.sqrt(np.abs(gamma * P / rho)) / M
    un = u[:, None] * om.normals[..., 0] + v[:, None] * om.normals[..., 1]
    lam = c[:, None] + np.abs(un)
    t = lam * om.surface[om.face_connectivity]
    return om.area / ((t[:, 0] + t[:, 1]) + t[:, 2])


def get_dt(W, om, CFL=0.5, gamma=1.4, M=1.0):
    """helper.py:8-18"""
    return np.min(cell_dt(W, om, gamma=gamma, M=M)) * CFL


def local_dt(W, om, CFL=0.5, gamma=1.4, M=1.0):
    """Per-cell time steps of local time stepping, shaped (N, 1) to broadcast
    over the state: cell_dt * CFL from the step's start state. An
    acceleration mode of this project, not a reference function (the
    reference steps every cell with get_dt's global minimum); passed as `dt`
    to the steppers below it gives each cell its own step. Parity unpinned
    against the reference (it has none); pinned by its minimum, which is
    get_dt, and by the stepper reducing to the global step when every entry
    is the global dt."""
    return (cell_dt(W, om, gamma=gamma, M=M) * CFL)[:, None]


# ---------------------------------------------------------------------------
# initial state, forces, solve loop (main.py:28-230, helper.py:366-415)
# ---------------------------------------------------------------------------

AOA_CASES = ("diamond", "naca0012", "naca2412", "rae2822", "onerad", "oa209", "diamond_tx",
             "diamond_txy", "diamond_alpha10", "rae2822_txy", "diamond_sym", "naca23012",
             "diamond_big", "cylinder", "naca4412", "naca0015")


def initialize(n_cells, cfg, dtype=np.float32):
    """main.py:28-40"""
    dt = np.dtype(dtype)
    rho, p = cfg["rho_inf"], cfg["p_inf"]
    gamma, Mach = cfg["gamma"], cfg["Mach"]
    c = (gamma * p / rho) ** 0.5
    aoa = float(cfg.get("aoa", 0.0)) if str(cfg.get("case", "")).lower() in AOA_CASES else 0.0
    rad = np.asarray(aoa, dtype=dt) * np.asarray(math.pi / 180, dtype=dt)
    u = Mach * c * np.cos(rad)
    v = Mach * c * np.sin(rad)
    prim0 = np.array([rho, u, v, p], dtype=dt)
    W = getConserved(np.tile(prim0[None, :], (n_cells, 1)), gamma=gamma, M=1.0)
    inlet = getConserved(prim0[None], gamma=gamma, M=1.0)[0]
    return W, inlet


def wall_faces(om):
    """Owner cell and local slot of each force-marker face
    (helper.py:370-380, main.py:83-96): first cell containing the face."""
    marker = int(om.metadata.get("force_marker", 2))
    fids = np.nonzero(om.face_markers == marker)[0]
    cells, locs = [], []
    for fid in fids:
        hit = np.any(om.face_connectivity == fid, axis=1)
        c = int(np.argmax(hit))
        cells.append(c)
        locs.append(int(np.argmax(om.face_connectivity[c] == fid)))
    return fids, np.array(cells, dtype=np.int64), np.array(locs, dtype=np.int64)


def forces(W, om, rho_inf, U_inf, L_ref):
    """get_drag_coefficient / get_lift_coefficient (helper.py:366-415):
    pressure (default gamma 1.4) times body-axis normal times face length,
    summed in float64."""
    P = getPrimitive(W)[:, 3]
    fids, cells, locs = wall_faces(om)
    n = om.normals[cells, locs].astype(np.float64)
    ds = om.surface[fids].astype(np.float64)
    Pw = P[cells].astype(np.float64)
    q = 0.5 * rho_inf * U_inf ** 2
    cd = float(np.sum(Pw * n[:, 0] * ds) / (q * L_ref))
    cl = float(np.sum(Pw * n[:, 1] * ds) / (q * L_ref))
    return cd, cl


def rel_norm(Wn, Wr, block_steps):
    """main.py:150-152 with float64 norms."""
    d = (Wn.astype(np.float64) - Wr.astype(np.float64))
    return math.sqrt(float(np.sum(d * d))) / (max(1, block_steps) * (math.sqrt(float(np.sum(Wr.astype(np.float64) ** 2))) + 1e-16))


def run(W, om, inlet, cfg, max_steps=None):
    """The stationarity loop of main.py:116-230 (without export).

    Returns W, a dict with converged, stopping_step, last_rel, dt, N and the
    final one-step residual. ``max_steps`` truncates the scan (test use)."""
    kw = dict(gamma=cfg["gamma"], M=1.0, reconstruction=cfg["reconstruction"],
              flux=cfg["flux"], value=inlet, entropy=False)
    fn = SCHEMES[str(cfg["time_scheme"]).upper()]
    thr = float(cfg["stationarity_threshold"])
    K = int(cfg["stationarity_check_every"])
    pat = max(1, int(cfg.get("stationarity_patience", 1)))
    CFL = cfg["CFL"]
    if 0.6 < cfg["Mach"] < 1.1 and CFL > 0.45:
        CFL = 0.45
    dt = get_dt(W, om, CFL=CFL, gamma=cfg["gamma"], M=1.0)
    N = int(np.asarray(cfg["tf"], dtype=W.dtype) / dt) + 1
    if max_steps is not None:
        N_run = min(N, max_steps)
    else:
        N_run = N
    W_ref, ref_step = W, 0
    stop, stopping_step, streak, streak_start = False, N, 0, 0
    last_rel = float("inf")
    for step in range(N_run):
        nxt = step + 1
        Wn = W if stop else fn(W, om, dt, **kw)
        if nxt % K == 0 or nxt == N:
            rel = rel_norm(Wn, W_ref, nxt - ref_step)
            under = rel <= thr
            new_streak = streak + 1 if under else 0
            if under and streak == 0:
                streak_start = nxt
            confirmed = new_streak >= pat and not stop
            if confirmed:
                stopping_step = streak_start
            stop = stop or confirmed
            streak = new_streak
            W_ref, ref_step, last_rel = Wn, nxt, rel
        W = Wn
        if stop:
            break
    Wp = fn(W, om, dt, **kw)
    fin = math.sqrt(float(np.sum((Wp.astype(np.float64) - W.astype(np.float64)) ** 2))) / (
        math.sqrt(float(np.sum(W.astype(np.float64) ** 2))) + 1e-16)
    return W, dict(converged=stop, stopping_step=stopping_step, last_rel=last_rel, dt=dt,
                   N=N, final_residual=fin)
