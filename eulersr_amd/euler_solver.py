"""Drop-in counterpart of euler/jax_fvm/src/euler_solver.py.

Same functions, arguments and keyword arguments as the reference; every call
runs the gfx950 kernels of libeulersr.so. ``W`` may be a numpy array (copied
to the GPU and back) or a float32 CUDA tensor (used in place); the result has
the type of ``W``. Inputs are never modified (functional semantics of the
reference, euler_solver.py:602-667).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .device import as_device_state, back, device_mesh, stream_ptr

RECON = {"constant": L.RECON_CONSTANT, "first_order": L.RECON_CONSTANT, "none": L.RECON_CONSTANT,
         "godunov": L.RECON_CONSTANT, "muscl": L.RECON_MUSCL, "second_order": L.RECON_MUSCL,
         "2nd_order": L.RECON_MUSCL}
FLUX = {"lf": L.FLUX_LF, "lax_friedrichs": L.FLUX_LF, "lax-friedrichs": L.FLUX_LF,
        "laxfriedrichs": L.FLUX_LF, "rusanov": L.FLUX_RUSANOV, "tadmor": L.FLUX_TADMOR,
        "ausm": L.FLUX_AUSM, "hll": L.FLUX_HLLC, "hllc": L.FLUX_HLLC, "roe": L.FLUX_ROE}
SCHEME = {"EE": L.SCHEME_EE, "RK2": L.SCHEME_RK2, "RK4": L.SCHEME_RK4, "SRK2": L.SCHEME_SRK2,
          "SSP_RK2": L.SCHEME_SRK2}
DEFAULT_VALUE = (1.0, 1.0, 0.0, 1.0)


def get_reconstruction_mode(kwargs):
    """euler_solver.py:100-101"""
    return str(kwargs.get("reconstruction", kwargs.get("scheme", "muscl"))).lower()


def get_flux_type(kwargs):
    """euler_solver.py:103-104"""
    return str(kwargs.get("flux", kwargs.get("numerical_flux", "rusanov"))).lower()


def make_opts(kwargs) -> L.Opts:
    """Translate the reference keyword arguments into the C options struct,
    raising ValueError for unknown names as euler_solver.py:122 and :139 do."""
    rec = get_reconstruction_mode(kwargs)
    if rec not in RECON:
        raise ValueError(f"Unknown reconstruction scheme '{rec}'. Expected 'constant' or 'muscl'.")
    fl = get_flux_type(kwargs)
    if fl not in FLUX:
        raise ValueError(f"Unknown flux type '{fl}'. Expected one of: 'lf', 'rusanov', 'tadmor', "
                         f"'ausm', 'hll', 'hllc', 'roe'.")
    o = L.Opts()
    o.flux = FLUX[fl]
    o.recon = RECON[rec]
    o.entropy = 1 if bool(kwargs.get("entropy", False)) else 0
    o.gamma = float(kwargs.get("gamma", 1.4))
    o.M = float(kwargs.get("M", 1.0))
    val = kwargs.get("value", DEFAULT_VALUE)
    if hasattr(val, "detach"):
        val = val.detach().cpu().numpy()
    val = np.asarray(val, dtype=np.float32).reshape(4)
    for i in range(4):
        o.value[i] = float(val[i])
    w = kwargs.get("sponge_width", None)
    s = kwargs.get("sponge_strength", None)
    o.sponge_width = -1.0 if w is None else float(w)
    o.sponge_strength = -1.0 if s is None else float(s)
    return o


def residual(W, mesh, **kwargs):
    """euler_solver.py:602-628: sum_k(F_k len_k)/area - sponge."""
    opts = make_opts(kwargs)
    dm = device_mesh(mesh, kwargs)
    t, was_np = as_device_state(W)
    torch = __import__("torch")
    out = torch.empty_like(t)
    L.check(L.lib().esr_residual(dm.handle, C.byref(opts), C.c_void_p(t.data_ptr()),
                                 C.c_void_p(out.data_ptr()), stream_ptr()), "esr_residual")
    return back(out, was_np)


def _step(scheme, W, mesh, dt, kwargs):
    opts = make_opts(kwargs)
    dm = device_mesh(mesh, kwargs)
    t, was_np = as_device_state(W)
    torch = __import__("torch")
    out = torch.empty_like(t)
    L.check(L.lib().esr_step(dm.handle, C.byref(opts), scheme, C.c_float(float(dt)),
                             C.c_void_p(t.data_ptr()), C.c_void_p(out.data_ptr()), stream_ptr()), "esr_step")
    return back(out, was_np)


def time_step_Euler(W, mesh, dt, **kwargs):
    """euler_solver.py:630-633"""
    return _step(L.SCHEME_EE, W, mesh, dt, kwargs)


def time_step_RK2_SSP(W, mesh, dt, **kwargs):
    """euler_solver.py:635-643"""
    return _step(L.SCHEME_SRK2, W, mesh, dt, kwargs)


def time_step_RK2(W, mesh, dt, **kwargs):
    """euler_solver.py:645-653"""
    return _step(L.SCHEME_RK2, W, mesh, dt, kwargs)


def time_step_RK4(W, mesh, dt, **kwargs):
    """euler_solver.py:655-667"""
    return _step(L.SCHEME_RK4, W, mesh, dt, kwargs)


def time_step(scheme: str, W, mesh, dt, **kwargs):
    """Dispatch by the names of main.py:50-51 (KeyError for an unknown name)."""
    return _step(SCHEME[str(scheme).upper()], W, mesh, dt, kwargs)
