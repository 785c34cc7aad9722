"""Batched solve loop on the GPU (esr_solver_* of libeulersr.so).

One ``BatchSolver`` advances any number of (mesh, Mach, alpha, flux) cases
together: cases on the same mesh with the same kernel configuration share a
group whose state is stored case-innermost, so a wave reads one geometry
record for up to 64 cases. Each case keeps the reference loop semantics
(euler/main.py:116-230): its own global dt from its initial state, its own
N = int(tf/dt)+1, its own stationarity streak, frozen once stopped.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .device import as_device_state, device_mesh, require_gpu, stream_ptr
from .euler_solver import SCHEME, make_opts


@dataclass
class CaseSpec:
    mesh: object
    kw: dict                       # solver kwargs (flux, reconstruction, gamma, value, ...)
    cfl: float = 0.6               # after the transonic clamp
    tf: float = 4.0
    q_ref: float = 0.0             # 0.5 rho U^2 L (0: no forces)
    engineering: tuple | None = None  # (target_cd, target_cl, cd_tol_abs, cl_tol_abs, patience)
    threshold: float = 0.0         # > 0 overrides the solver's stationarity threshold
    tag: dict = field(default_factory=dict)
    max_steps: int = 0             # > 0: per-case step cap (N = min(N, max_steps))
    group: int = 0                 # cases of different groups never share a launch group


def transonic_cfl(mach: float, cfl: float) -> float:
    """main.py:117-120"""
    return 0.45 if (0.6 < mach < 1.1 and cfl > 0.45) else cfl


def set_stage_path(path: str):
    """"cell" (default), "face" or "auto": the stage kernels of solvers and
    per-mesh engines built afterwards."""
    code = {"auto": L.STAGE_AUTO, "cell": L.STAGE_CELL, "face": L.STAGE_FACE}[path]
    L.check(L.lib().esr_set_stage_path(code), "esr_set_stage_path")


def get_stage_path() -> str:
    return {L.STAGE_CELL: "cell", L.STAGE_FACE: "face"}[L.lib().esr_get_stage_path()]


class _LazyReports:
    """Indexable reports of a solver: entry j becomes its dict on first read."""
    __slots__ = ("_arr", "_conv", "_cache")

    def __init__(self, arr, conv):
        self._arr, self._conv, self._cache = arr, conv, {}

    def __len__(self):
        return len(self._arr)

    def __getitem__(self, j):
        d = self._cache.get(j)
        if d is None:
            if not 0 <= j < len(self._arr):
                raise IndexError(j)
            d = self._cache[j] = self._conv(self._arr[j])
        return d


class BatchSolver:
    def __init__(self, cases, scheme="SRK2", threshold=1e-6, check_every=5, patience=1, max_steps=0,
                 graph=True, autotune=False, launch=None, local_dt=False):
        require_gpu()
        self.autotune = bool(autotune)
        self.launch = tuple(launch) if launch else None   # (path, block, xcd_order), applied at prepare
        self.tuning = None
        self.specs = list(cases)
        n = len(self.specs)
        arr = (L.Case * n)()
        self._meshes = []
        for i, cs in enumerate(self.specs):
            dm = device_mesh(cs.mesh, cs.kw)
            self._meshes.append(dm)
            self._fill_case(arr[i], cs, dm)
        so = L.SolveOpts()
        so.scheme = SCHEME[str(scheme).upper()]
        so.threshold = float(threshold)
        so.check_every = max(1, int(check_every))
        so.patience = max(1, int(patience))
        so.max_steps = int(max_steps)
        # local time stepping (acceleration mode; the reference steps every
        # cell with one global dt): each cell advances with its own CFL dt
        so.local_dt = 1 if local_dt else 0
        so.no_graph = 0 if graph else 1
        h = C.c_void_p()
        L.check(L.lib().esr_solver_create(n, arr, C.byref(so), C.byref(h)), "esr_solver_create")
        self.handle = h
        self.n = n
        self.steps = 0

    @staticmethod
    def _fill_case(c, cs, dm):
        c.mesh = dm.handle
        c.opts = make_opts(cs.kw)
        c.cfl = float(cs.cfl)
        c.tf = float(cs.tf)
        c.q_ref = float(cs.q_ref)
        c.threshold = float(cs.threshold)
        c.max_steps = int(cs.max_steps)
        c.group = int(cs.group)
        if cs.engineering is not None:
            tcd, tcl, cdt, clt, pat = cs.engineering
            c.track_engineering = 1
            c.target_cd, c.target_cl = float(tcd), float(tcl)
            c.cd_tol_abs, c.cl_tol_abs = float(cdt), float(clt)
            c.engineering_patience = int(pat)

    def replace_case(self, i, spec, W):
        """Continuous batching: case `spec` (same mesh and kernel
        configuration as slot i) enters slot i with initial state W; its
        steps count from the current solver step."""
        import torch
        t, _ = as_device_state(W)
        dm = device_mesh(spec.mesh, spec.kw)
        c = L.Case()
        self._fill_case(c, spec, dm)
        L.check(L.lib().esr_solver_replace_case(self.handle, int(i), C.byref(c), C.c_void_p(t.data_ptr()),
                                                stream_ptr()), "esr_solver_replace_case")
        torch.cuda.current_stream().synchronize()
        self.specs[i] = spec
        self._meshes[i] = dm

    def compact(self):
        """Drops groups whose cases are all done from the launches; returns
        the number of groups still launched."""
        n = C.c_int32()
        L.check(L.lib().esr_solver_compact(self.handle, C.byref(n), stream_ptr()), "esr_solver_compact")
        return n.value

    def set_state(self, i, W):
        t, _ = as_device_state(W)
        if t.shape != (self.specs[i].mesh.tris.shape[0], 4):
            raise ValueError(f"state of case {i} has shape {tuple(t.shape)}")
        L.check(L.lib().esr_solver_set_state(self.handle, i, C.c_void_p(t.data_ptr()), stream_ptr()),
                "esr_solver_set_state")
        # keep the source alive until the copy has run
        import torch
        torch.cuda.current_stream().synchronize()

    def get_state(self, i):
        import torch
        n = int(self.specs[i].mesh.tris.shape[0])
        out = torch.empty((n, 4), dtype=torch.float32, device="cuda")
        L.check(L.lib().esr_solver_get_state(self.handle, i, C.c_void_p(out.data_ptr()), stream_ptr()),
                "esr_solver_get_state")
        return out

    def prepare(self):
        L.check(L.lib().esr_solver_prepare(self.handle, stream_ptr()), "esr_solver_prepare")
        if self.launch:
            self.set_launch(*self.launch)
            self.tuning = dict(path=self.launch[0], block=int(self.launch[1]), xcd_order=bool(self.launch[2]),
                               us_per_stage=None)
        elif self.autotune:
            self.tune()

    def set_launch(self, path="cell", block=256, xcd_order=True):
        """Pins the stage path ("cell" / "face"), block size (64/128/256) and
        block order (XCD-contiguous runs or the hardware's round robin)."""
        code = {"cell": L.STAGE_CELL, "face": L.STAGE_FACE}[path]
        L.check(L.lib().esr_solver_set_launch(self.handle, code, int(block), int(bool(xcd_order)), stream_ptr()),
                "esr_solver_set_launch")

    def tune(self, reps=10):
        """Autotuner: fastest stage path and block size for this solver's
        launch shape (cached per shape by the library)."""
        p, b, x, us = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
        L.check(L.lib().esr_solver_autotune(self.handle, int(reps), C.byref(p), C.byref(b), C.byref(x),
                                            C.byref(us), stream_ptr()), "esr_solver_autotune")
        self.tuning = dict(path={L.STAGE_CELL: "cell", L.STAGE_FACE: "face"}[p.value], block=b.value,
                           xcd_order=bool(x.value), us_per_stage=us.value)
        return self.tuning

    def run(self, until_step=0):
        s = C.c_int64()
        L.check(L.lib().esr_solver_run(self.handle, int(until_step), C.byref(s), stream_ptr()), "esr_solver_run")
        self.steps = s.value
        return self.steps

    def finalize(self):
        L.check(L.lib().esr_solver_finalize(self.handle, stream_ptr()), "esr_solver_finalize")

    def report(self, i) -> dict:
        r = L.CaseReport()
        L.check(L.lib().esr_solver_report(self.handle, i, C.byref(r)), "esr_solver_report")
        return self._report_dict(r)

    def reports(self) -> list:
        """Reports of every case (one device copy)."""
        arr = (L.CaseReport * self.n)()
        L.check(L.lib().esr_solver_reports(self.handle, self.n, arr), "esr_solver_reports")
        return [self._report_dict(r) for r in arr]

    def reports_lazy(self):
        """Reports of every case from one device copy, each converted to its
        dict when first indexed: a continuous-batching loop reads only the
        slots in use, out of every slot of every sub-group."""
        arr = (L.CaseReport * self.n)()
        L.check(L.lib().esr_solver_reports(self.handle, self.n, arr), "esr_solver_reports")
        return _LazyReports(arr, self._report_dict)

    @staticmethod
    def _report_dict(r) -> dict:
        return dict(dt=np.float32(r.dt), N=int(r.n_steps), converged=bool(r.converged),
                    stopping_step=int(r.stopping_step), steps_run=int(r.steps_run),
                    last_rel=float(r.last_rel), final_residual=float(r.final_residual),
                    cd=float(r.cd), cl=float(r.cl),
                    engineering_converged=bool(r.engineering_converged),
                    engineering_stopping_step=int(r.engineering_stopping_step),
                    diverged=bool(r.diverged), recent_rel_max=float(r.recent_rel_max))

    def info(self):
        """(cell work-items, group cells, face work-items, group faces) of one stage."""
        v = [C.c_int64() for _ in range(4)]
        L.check(L.lib().esr_solver_info(self.handle, *[C.byref(x) for x in v]), "esr_solver_info")
        return tuple(x.value for x in v)

    def time_kernels(self, reps=20):
        """Average ms per launch set of the (reconstruction, face flux,
        gather + update, cell-centric flux + update) kernels, timed with HIP
        events between the kernels of real stage sequences on the current
        stream."""
        ms = (C.c_double * 4)()
        L.check(L.lib().esr_solver_time_kernels(self.handle, int(reps), ms, stream_ptr()), "esr_solver_time_kernels")
        return tuple(ms)

    def time_kernel_runs(self, reps=50):
        """Average ms per launch set of the same four kernels, each launched
        `reps` times back to back between two HIP events (the kernel's
        duration plus only the gap between queued launches)."""
        ms = (C.c_double * 4)()
        L.check(L.lib().esr_solver_time_kernel_runs(self.handle, int(reps), ms, stream_ptr()),
                "esr_solver_time_kernel_runs")
        return tuple(ms)

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            L.lib().esr_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
