"""ctypes binding of libeulersr.so (include/eulersr.h).

The library is loaded from the package directory only (built in-tree by
``python -m eulersr_amd.build``); there is no fallback implementation: if the
library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libeulersr.so"

ESR_OK, ESR_ERR_VALUE, ESR_ERR_KEY, ESR_ERR_DEVICE, ESR_ERR_MEMORY, ESR_ERR_STATE = range(6)
FLUX_LF, FLUX_RUSANOV, FLUX_TADMOR, FLUX_AUSM, FLUX_HLLC, FLUX_ROE = range(6)
RECON_CONSTANT, RECON_MUSCL = 0, 1
SCHEME_EE, SCHEME_RK2, SCHEME_SRK2, SCHEME_RK4 = range(4)

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_points", C.c_int32), ("n_cells", C.c_int32), ("n_faces", C.c_int32),
        ("points", f32p), ("tris", i32p), ("neighbors", i32p), ("faces", i32p),
        ("face_markers", i32p), ("face_connectivity", i32p), ("face_connectivity_opposite", i32p),
        ("barycenter", f32p), ("area", f32p), ("surface", f32p), ("normals", f32p),
        ("inlet_subsonic", f32p), ("n_wall_markers", C.c_int32), ("wall_markers", i32p),
        ("force_marker", C.c_int32), ("sponge", C.c_int32), ("renumber", C.c_int32),
        ("rcm_seed_order", i32p),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("flux", C.c_int32), ("recon", C.c_int32), ("entropy", C.c_int32),
        ("gamma", C.c_double), ("M", C.c_double), ("value", C.c_float * 4),
        ("sponge_width", C.c_float), ("sponge_strength", C.c_float),
    ]


class Case(C.Structure):
    _fields_ = [
        ("mesh", C.c_void_p), ("opts", Opts), ("cfl", C.c_float), ("tf", C.c_float),
        ("q_ref", C.c_double), ("track_engineering", C.c_int32),
        ("target_cd", C.c_double), ("target_cl", C.c_double), ("cd_tol_abs", C.c_double),
        ("cl_tol_abs", C.c_double), ("engineering_patience", C.c_int32), ("threshold", C.c_double),
        ("max_steps", C.c_int64), ("group", C.c_int32),
    ]


class SolveOpts(C.Structure):
    _fields_ = [
        ("scheme", C.c_int32), ("threshold", C.c_double), ("check_every", C.c_int32),
        ("patience", C.c_int32), ("max_steps", C.c_int64), ("local_dt", C.c_int32), ("no_graph", C.c_int32),
    ]


STAGE_AUTO, STAGE_CELL, STAGE_FACE = 0, 1, 2


class KnnMap(C.Structure):
    _fields_ = [
        ("n_hr", C.c_int32), ("n_lr", C.c_int32), ("k", C.c_int32), ("knn_idx", C.c_void_p),
        ("knn_rel", C.c_void_p), ("knn_w", C.c_void_p), ("wall", C.c_void_p),
    ]


class CaseReport(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("n_steps", C.c_int64), ("converged", C.c_int32),
        ("stopping_step", C.c_int64), ("steps_run", C.c_int64), ("last_rel", C.c_double),
        ("final_residual", C.c_double), ("cd", C.c_double), ("cl", C.c_double),
        ("engineering_converged", C.c_int32), ("engineering_stopping_step", C.c_int64),
        ("diverged", C.c_int32), ("recent_rel_max", C.c_double),
    ]


# symbol -> (restype, argtypes); every symbol declared in include/eulersr.h
SIGNATURES = {
    "esr_mesh_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]),
    "esr_mesh_destroy": (C.c_int, [C.c_void_p]),
    "esr_mesh_permutation": (C.c_int, [C.c_void_p, i32p]),
    "esr_mesh_info": (C.c_int, [C.c_void_p, i32p, i32p, C.POINTER(C.c_int64)]),
    "esr_rcm": (C.c_int, [C.c_int32, i32p, i32p, i32p]),
    "esr_residual": (C.c_int, [C.c_void_p, C.POINTER(Opts), C.c_void_p, C.c_void_p, C.c_void_p]),
    "esr_step": (C.c_int, [C.c_void_p, C.POINTER(Opts), C.c_int32, C.c_float, C.c_void_p, C.c_void_p,
                           C.c_void_p]),
    "esr_step_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Opts), C.c_int32, C.c_float, C.c_void_p, C.c_void_p,
                                   C.c_void_p]),
    "esr_get_dt": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_double, C.c_double, f32p, C.c_void_p]),
    "esr_forces": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_double),
                             C.POINTER(C.c_double), C.c_void_p]),
    "esr_norms": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double),
                            C.POINTER(C.c_double), C.c_void_p]),
    "esr_solver_create": (C.c_int, [C.c_int32, C.POINTER(Case), C.POINTER(SolveOpts),
                                    C.POINTER(C.c_void_p)]),
    "esr_solver_destroy": (C.c_int, [C.c_void_p]),
    "esr_solver_set_state": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "esr_solver_get_state": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "esr_solver_prepare": (C.c_int, [C.c_void_p, C.c_void_p]),
    "esr_solver_run": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.c_void_p]),
    "esr_solver_finalize": (C.c_int, [C.c_void_p, C.c_void_p]),
    "esr_solver_replace_case": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Case), C.c_void_p, C.c_void_p]),
    "esr_solver_reports": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(CaseReport)]),
    "esr_solver_compact": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_void_p]),
    "esr_solver_report": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(CaseReport)]),
    "esr_solver_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "esr_solver_time_kernels": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.c_void_p]),
    "esr_solver_time_kernel_runs": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.c_void_p]),
    "esr_check_device_math": (C.c_int, [C.POINTER(C.c_uint64), C.c_void_p]),
    "esr_set_stage_path": (C.c_int, [C.c_int32]),
    "esr_solver_set_launch": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    "esr_solver_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_void_p]),
    "esr_get_stage_path": (C.c_int32, []),
    "esr_mlp_create": (C.c_int, [C.c_int32, i32p, C.c_int64, C.POINTER(C.c_void_p)]),
    "esr_mlp_destroy": (C.c_int, [C.c_void_p]),
    "esr_mlp_param_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "esr_mlp_set_params": (C.c_int, [C.c_void_p, f32p]),
    "esr_mlp_get_params": (C.c_int, [C.c_void_p, f32p]),
    "esr_mlp_forward": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "esr_mlp_loss_grad": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.POINTER(C.c_double),
                                    f32p, C.c_void_p]),
    "esr_mlp_adamw": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32,
                                C.c_void_p]),
    "esr_warm_features": (C.c_int, [C.POINTER(KnnMap), C.c_int32, C.c_void_p, C.c_void_p, f32p, f32p,
                                    C.c_double, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "esr_warm_apply": (C.c_int, [C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, f32p, C.c_double, C.c_void_p,
                                 C.c_void_p, C.c_void_p]),
    "esr_last_error": (C.c_char_p, []),
    "esr_version": (C.c_int32, []),
}

_lib = None


def lib():
    """Load the in-tree library (raises if it has not been built).

    torch is imported first: its wheel bundles a HIP runtime with the same
    soname (libamdhip64.so.7) as the ROCm install the library links against.
    Loaded after torch, the library binds to torch's runtime, so the process
    has one runtime and device pointers and streams from torch are valid in
    the library. Loaded before torch, the process would end up with two
    runtimes, and the library's own would not see the GPU."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m eulersr_amd.build`")
        import torch  # noqa: F401  (one HIP runtime per process, see above)
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class ESRError(RuntimeError):
    pass


def check(status: int, what: str = ""):
    if status == ESR_OK:
        return
    msg = lib().esr_last_error()
    msg = msg.decode() if msg else ""
    text = f"{what}: {msg}" if what else msg
    if status == ESR_ERR_VALUE:
        raise ValueError(text)
    if status == ESR_ERR_KEY:
        raise KeyError(text)
    if status == ESR_ERR_MEMORY:
        raise MemoryError(text)
    raise ESRError(text)
