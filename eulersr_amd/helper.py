"""Drop-in counterpart of the hot-path parts of euler/jax_fvm/src/helper.py.

``get_dt`` and the force coefficients run on the GPU. The state conversions
are cheap host helpers (float32 numpy, reference operation order), used for
initial states and exports; the solver kernels carry their own device copy.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .device import as_device_state, device_mesh, stream_ptr


def getConserved(prim, gamma=1.4, M=1):
    """helper.py:25-35 (host, dtype of the input)."""
    prim = np.asarray(prim)
    rho, u, v, P = prim[..., 0], prim[..., 1], prim[..., 2], prim[..., 3]
    E = P / ((gamma - 1) * M ** 2) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], axis=-1)


def getPrimitive(W, gamma=1.4, M=1):
    """helper.py:37-46 (host, dtype of the input)."""
    W = np.asarray(W)
    rho = W[..., 0]
    u = W[..., 1] / rho
    v = W[..., 2] / rho
    P = (W[..., 3] - 0.5 * rho * (u * u + v * v)) * M ** 2 * (gamma - 1)
    return np.stack([rho, u, v, P], axis=-1)


def get_mach_number(prim, gamma=1.4):
    """helper.py:303-310"""
    prim = np.asarray(prim)
    c = np.sqrt(gamma * prim[..., 3] / prim[..., 0])
    return np.sqrt(prim[..., 1] ** 2 + prim[..., 2] ** 2) / c


def get_dt(W, mesh, CFL=0.5, gamma=1.4, M=1.0):
    """helper.py:8-18 on the GPU: CFL * min_i area_i / sum_k (c_i + |u_i.n_k|) len_k."""
    dm = device_mesh(mesh)
    t, _ = as_device_state(W)
    out = C.c_float()
    L.check(L.lib().esr_get_dt(dm.handle, C.c_void_p(t.data_ptr()), C.c_float(float(CFL)),
                               C.c_double(float(gamma)), C.c_double(float(M)), C.byref(out),
                               stream_ptr()), "esr_get_dt")
    return np.float32(out.value)


def _forces(W, mesh, rho_inf, U_inf, L_ref):
    dm = device_mesh(mesh)
    t, _ = as_device_state(W)
    q = 0.5 * rho_inf * U_inf ** 2 * L_ref
    cd, cl = C.c_double(), C.c_double()
    L.check(L.lib().esr_forces(dm.handle, C.c_void_p(t.data_ptr()), C.c_double(q), C.byref(cd),
                               C.byref(cl), stream_ptr()), "esr_forces")
    return cd.value, cl.value


def get_drag_coefficient(W, mesh, rho_inf, U_inf, L_ref):
    """helper.py:366-390 (body-axis x force, FP64 wall sum)."""
    return _forces(W, mesh, rho_inf, U_inf, L_ref)[0]


def get_lift_coefficient(W, mesh, rho_inf, U_inf, L_ref):
    """helper.py:392-415 (body-axis y force, FP64 wall sum)."""
    return _forces(W, mesh, rho_inf, U_inf, L_ref)[1]


def get_force_coefficients(W, mesh, rho_inf, U_inf, L_ref):
    return _forces(W, mesh, rho_inf, U_inf, L_ref)


def relative_change(W_next, W, mesh):
    """||W_next - W||_F / (||W||_F + 1e-16) on the GPU, in float64: the
    one-step probe residual of main.py:228-230 and
    eval/_fvm_solve_worker.py:133-134."""
    dm = device_mesh(mesh)
    a, _ = as_device_state(W_next)
    b, _ = as_device_state(W)
    d, r = C.c_double(), C.c_double()
    L.check(L.lib().esr_norms(dm.handle, C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), C.byref(d),
                              C.byref(r), stream_ptr()), "esr_norms")
    return d.value / (r.value + 1e-16)


def get_total_entropy(W, mesh, gamma=1.4):
    """helper.py:312-315 (host diagnostic)."""
    W = np.asarray(W, dtype=np.float64)
    rho = W[:, 0]
    u, v = W[:, 1] / rho, W[:, 2] / rho
    P = (W[:, 3] - 0.5 * rho * (u * u + v * v)) * (gamma - 1)
    s = np.log(P / rho ** gamma)
    return float(np.sum(rho * s * np.asarray(mesh.area) / (gamma - 1)))


def get_entropy_creation(W_initial, W_final, mesh, gamma=1.4):
    """helper.py:317-320"""
    return get_total_entropy(W_final, mesh, gamma) - get_total_entropy(W_initial, mesh, gamma)
