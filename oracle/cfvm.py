"""CPU oracle, C restatement (TEST INFRASTRUCTURE ONLY): ctypes binding of
oracle/libcfvm.so (oracle/cfvm.c, built by ``make -C oracle``).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may use it. It computes, in float32 and in parallel over cells (OpenMP),
the same operations as oracle/fvm.py for HLLC or Roe with constant or MUSCL
reconstruction: ``residual``, ``time_step_RK2_SSP`` (and ``srk2_steps``, n
steps without leaving C) and ``get_dt``; with HLLC the results are
bit-identical to oracle/fvm.py, with Roe they agree to a few ulps of the C
library's logf / powf against numpy's (tests/test_oracle_c.py). Other fluxes,
the entropy flag and float64 are not restated here.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from . import fvm as F

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcfvm.so")


class _Mesh(C.Structure):
    _fields_ = [("n", C.c_int64), ("npts", C.c_int64), ("points", C.c_void_p), ("bary", C.c_void_p),
                ("area", C.c_void_p), ("nb", C.c_void_p), ("fc", C.c_void_p), ("fmark", C.c_void_p),
                ("faces", C.c_void_p), ("fco", C.c_void_p), ("surface", C.c_void_p), ("normals", C.c_void_p),
                ("inlet_sub", C.c_void_p), ("walls", C.c_void_p), ("n_walls", C.c_int32), ("bump", C.c_int32)]


class _Opts(C.Structure):
    _fields_ = [("gamma", C.c_double), ("M", C.c_double), ("muscl", C.c_int32), ("value", C.c_float * 4),
                ("has_width", C.c_int32), ("has_strength", C.c_int32), ("sponge_width", C.c_double),
                ("sponge_strength", C.c_double), ("roe", C.c_int32)]


LIB_RCP = os.path.join(HERE, "libcfvm_rcp.so")
_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(rcp=False):
    """The restatement (rcp=False), or its build whose divisions go through
    rounded reciprocals, a * (1 / b) (rcp=True)."""
    path = LIB_RCP if rcp else LIB
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C oracle)")
        L = C.CDLL(path)
        L.cf_residual.argtypes = [C.c_void_p] * 6
        L.cf_srk2.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_int64] + [C.c_void_p] * 4
        L.cf_get_dt.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.cf_get_dt.restype = C.c_float
        L.cf_set_threads.argtypes = [C.c_int]
        L.cf_max_threads.restype = C.c_int
        _libs[path] = L
    return _libs[path]


def set_threads(n: int):
    lib().cf_set_threads(int(n))
    if os.path.exists(LIB_RCP):
        lib(True).cf_set_threads(int(n))


def max_threads() -> int:
    return int(lib().cf_max_threads())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class CMesh:
    """The arrays of an oracle mesh view (oracle.fvm.omesh, float32) held
    contiguous for the C code."""

    def __init__(self, om, wall_markers=None):
        if om.dtype != np.float32:
            raise TypeError("the C oracle is float32 only")
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
        i64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)  # noqa: E731
        self.om = om
        self.arrays = dict(points=f32(om.points), bary=f32(om.barycenter), area=f32(om.area),
                           nb=i64(om.neighbors), fc=i64(om.face_connectivity), fmark=i64(om.face_markers),
                           faces=i64(om.faces), fco=i64(om.face_connectivity_opposite), surface=f32(om.surface),
                           normals=f32(om.normals), inlet_sub=f32(om.inlet_subsonic),
                           walls=i64(F.wall_markers_of(om, {} if wall_markers is None else
                                                       {"wall_markers": wall_markers})))
        a = self.arrays
        self.n = int(a["area"].shape[0])
        self.s = _Mesh(self.n, a["points"].shape[0], *[_ptr(a[k]) for k in (
            "points", "bary", "area", "nb", "fc", "fmark", "faces", "fco", "surface", "normals", "inlet_sub",
            "walls")], a["walls"].shape[0], int(om.metadata.get("case") == "bump"))
        self.grad = np.empty((self.n, 8), np.float32)
        self.phi = np.empty((self.n, 4), np.float32)


def _opts(kw):
    ft = F.flux_type(kw)
    if ft not in ("hll", "hllc", "roe"):
        raise ValueError(f"the C oracle restates HLLC and Roe only, not '{ft}'")
    if kw.get("entropy", False):
        raise ValueError("the C oracle does not restate the entropy flux")
    mode = F.reconstruction_mode(kw)
    if mode not in F.RECON_MUSCL + F.RECON_CONSTANT:
        raise ValueError(f"Unknown reconstruction scheme '{mode}'")
    value = np.asarray(kw.get("value", F.DEFAULT_INLET), dtype=np.float32)
    w, s = kw.get("sponge_width"), kw.get("sponge_strength")
    return _Opts(float(kw.get("gamma", 1.4)), float(kw.get("M", 1.0)), int(mode in F.RECON_MUSCL),
                 (C.c_float * 4)(*value.tolist()), int(w is not None), int(s is not None),
                 0.0 if w is None else float(w), 0.0 if s is None else float(s), int(ft == "roe"))


def _cm(mesh_or_cm, kw):
    if isinstance(mesh_or_cm, CMesh):
        return mesh_or_cm
    return CMesh(mesh_or_cm, kw.get("wall_markers"))


def residual(W, cm, **kw):
    """euler_solver.py:602-628 (as oracle.fvm.residual)."""
    cm = _cm(cm, kw)
    W = np.ascontiguousarray(W, dtype=np.float32)
    R = np.empty_like(W)
    o = _opts(kw)
    lib().cf_residual(C.byref(cm.s), C.byref(o), _ptr(W), _ptr(R), _ptr(cm.grad), _ptr(cm.phi))
    return R


def srk2_steps(W, cm, dt, n_steps, rcp=False, **kw):
    """n_steps SSP-RK2 steps (euler_solver.py:635-643); returns the new state.
    rcp: on the build whose divisions go through rounded reciprocals."""
    cm = _cm(cm, kw)
    W = np.array(W, dtype=np.float32, order="C", copy=True)
    W1 = np.empty_like(W)
    k = np.empty_like(W)
    o = _opts(kw)
    lib(rcp).cf_srk2(C.byref(cm.s), C.byref(o), _ptr(W), np.float32(dt), int(n_steps), _ptr(W1), _ptr(k),
                  _ptr(cm.grad), _ptr(cm.phi))
    return W


def time_step_RK2_SSP(W, cm, dt, **kw):
    return srk2_steps(W, cm, dt, 1, **kw)


def get_dt(W, cm, CFL=0.5, gamma=1.4, M=1.0):
    """helper.py:8-18 (as oracle.fvm.get_dt)."""
    cm = _cm(cm, {})
    W = np.ascontiguousarray(W, dtype=np.float32)
    return np.float32(lib().cf_get_dt(C.byref(cm.s), _ptr(W), float(CFL), float(gamma), float(M)))
