"""CPU oracle (TEST INFRASTRUCTURE ONLY): the LR -> HR warm-start inputs by
brute force -- k nearest LR barycentres of each HR barycentre
(preprocessing/knn.py:20-40), normalised inverse-distance weights p = 2 with
eps 1e-12 (utils/metrics.py:9-14), the IDW prediction (eval/core.py:135-142)
and the wall features [wd/L, exp(-wd/(2 h_LR)), n_x, n_y] from 8 samples per
wall face (utils/aero.py:114-181).
"""
from __future__ import annotations

import numpy as np


def knn_brute(lr_pts, hr_pts, k=6):
    d2 = ((hr_pts[:, None, :] - lr_pts[None, :, :]) ** 2).sum(-1)
    idx = np.argsort(d2, axis=1, kind="stable")[:, :k]
    dist = np.sqrt(np.take_along_axis(d2, idx, axis=1))
    return idx, dist


def idw_weights(dist, eps=1e-12):
    w = 1.0 / (np.asarray(dist, dtype=np.float64) ** 2 + eps)
    return w / w.sum(axis=1, keepdims=True)


def idw_predict(lr_prim, idx, w):
    return (w[:, :, None] * lr_prim[idx]).sum(axis=1)


def wall_features(points, faces, face_markers, query, L_ref, h_lr):
    pts = np.asarray(points, dtype=np.float64)
    bnd = faces[face_markers > 0]
    mid = pts[bnd].mean(axis=1)
    lo, hi = pts.min(0), pts.max(0)
    tol = 1e-6 * float((hi - lo).max())
    on_box = ((np.abs(mid[:, 0] - lo[0]) < tol) | (np.abs(mid[:, 0] - hi[0]) < tol)
              | (np.abs(mid[:, 1] - lo[1]) < tol) | (np.abs(mid[:, 1] - hi[1]) < tol))
    wall = bnd[~on_box]
    a, b = pts[wall[:, 0]], pts[wall[:, 1]]
    tang = (b - a) / (np.linalg.norm(b - a, axis=1, keepdims=True) + 1e-12)
    nrm = np.stack([-tang[:, 1], tang[:, 0]], axis=1)
    t = np.linspace(0.0, 1.0, 8)[None, :, None]
    samples = (a[:, None, :] * (1 - t) + b[:, None, :] * t).reshape(-1, 2)
    d2 = ((query[:, None, :] - samples[None, :, :]) ** 2).sum(-1)
    j = np.argmin(d2, axis=1)
    wd = np.sqrt(d2[np.arange(len(query)), j])
    seg = j // 8
    n = nrm[seg]
    to_q = query - 0.5 * (a + b)[seg]
    s = np.sign((n * to_q).sum(axis=1, keepdims=True))
    n = n * np.where(s == 0, 1.0, s)
    return np.column_stack([wd / L_ref, np.exp(-wd / (2.0 * h_lr)), n])


def features(idx, rel, w, wall, lr_states, case_scal, mu, sig, gamma, in_pad):
    """Rows (case-major) of the network input, float32 before bfloat16
    rounding, and the IDW primitives; the column order is: per neighbour
    j = 0..k-1 the normalised primitives (4), offset (2), weight (1); then
    the wall features (4) and the case scalars (2); zero padding."""
    lr = np.asarray(lr_states, dtype=np.float32)
    n_cases, n_lr, _ = lr.shape
    n_hr, k = idx.shape
    g1 = np.float32(gamma - 1.0)
    r = lr[..., 0]
    u = lr[..., 1] / r
    v = lr[..., 2] / r
    p = (lr[..., 3] - np.float32(0.5) * r * (u * u + v * v)) * g1
    prim = np.stack([r, u, v, p], axis=-1).astype(np.float32)        # (n_cases, n_lr, 4)
    isig = (np.float32(1.0) / np.asarray(sig, np.float32)).astype(np.float32)
    mu = np.asarray(mu, np.float32)
    X = np.zeros((n_cases, n_hr, in_pad), np.float32)
    idw = np.zeros((n_cases, n_hr, 4), np.float64)
    for j in range(k):
        pj = prim[:, idx[:, j], :]                                     # (n_cases, n_hr, 4)
        X[:, :, 7 * j:7 * j + 4] = (pj - mu) * isig
        X[:, :, 7 * j + 4:7 * j + 6] = rel[None, :, j, :]
        X[:, :, 7 * j + 6] = w[None, :, j]
        idw += w[None, :, j, None].astype(np.float64) * pj
    X[:, :, 7 * k:7 * k + 4] = wall[None]
    X[:, :, 7 * k + 4:7 * k + 6] = np.asarray(case_scal, np.float32)[:, None, :]
    return X.reshape(n_cases * n_hr, in_pad), idw.reshape(n_cases * n_hr, 4)


def apply(idw, delta, sig, gamma, n_hr):
    """prim = IDW + sig * delta -> conserved (float64); per-case flag when a
    cell has rho <= 0, p <= 0 or a non-finite value."""
    p = np.asarray(idw, np.float64).copy()
    if delta is not None:
        p = p + np.asarray(sig, np.float64) * np.asarray(delta, np.float64)
    ok = (p[:, 0] > 0) & (p[:, 3] > 0) & np.isfinite(p).all(axis=1)
    bad = ~ok.reshape(-1, n_hr).all(axis=1)
    E = p[:, 3] / (gamma - 1.0) + 0.5 * p[:, 0] * (p[:, 1] ** 2 + p[:, 2] ** 2)
    W = np.stack([p[:, 0], p[:, 0] * p[:, 1], p[:, 0] * p[:, 2], E], axis=1)
    return W, bad
