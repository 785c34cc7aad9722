"""Solver driver, counterpart of euler/main.py (initialize, run, CLI).

    python -m eulersr_amd.main --case diamond --mach 2.5 --aoa 5 --flux HLLC \
        --reconstruction MUSCL --time-scheme SRK2 --mesh-path gen:diamond:0.1

The whole time loop runs on the GPU (esr_solver_*): global dt from the
initial state, N = int(tf/dt)+1 steps at most, a stationarity check every
``stationarity_check_every`` steps and at step N, patience, freeze on stop,
then a one-step probe residual (main.py:116-230). Residual norms and CD/CL
sums are float64; everything else float32.
"""
from __future__ import annotations

import json
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import helper
from .config import (AOA_CASES_LC, KNOWN_CASES_LC, build_config_from_cli, format_snapshot_name,
                     load_default_config, load_mesh, print_config, setup_dirs)
from .solve import BatchSolver, CaseSpec, transonic_cfl


def initialize(mesh, cfg):
    """main.py:28-40: freestream primitives -> conserved state (float32)."""
    rho, p = cfg["rho_inf"], cfg["p_inf"]
    gamma, Mach = cfg["gamma"], cfg["Mach"]
    c = (gamma * p / rho) ** 0.5
    aoa = float(cfg.get("aoa", 0.0)) if str(cfg.get("case", "")).lower() in AOA_CASES_LC else 0.0
    rad = np.float32(aoa) * np.float32(math.pi / 180)
    u = Mach * c * np.cos(rad)
    v = Mach * c * np.sin(rad)
    prim0 = np.array([rho, u, v, p], dtype=np.float32)
    n = int(np.asarray(mesh.area).shape[0])
    W = helper.getConserved(np.tile(prim0[None, :], (n, 1)), gamma=gamma, M=1.0)
    inlet = helper.getConserved(prim0[None], gamma=gamma, M=1.0)[0]
    return W.astype(np.float32), inlet.astype(np.float32)


def solver_kwargs(cfg, inlet):
    """main.py:54-55"""
    return dict(gamma=cfg["gamma"], M=1.0, reconstruction=cfg["reconstruction"], flux=cfg["flux"],
                value=np.asarray(inlet, dtype=np.float32), entropy=False)


def freestream_speed(cfg):
    return cfg["Mach"] * float(np.sqrt(cfg["gamma"] * cfg["p_inf"] / cfg["rho_inf"]))


def case_spec(mesh, inlet, cfg) -> CaseSpec:
    """One solver case from a run configuration (CFL clamp of main.py:117-120,
    engineering criterion of main.py:66-114)."""
    cfl = transonic_cfl(cfg["Mach"], cfg["CFL"])
    L_ref = mesh.metadata.get("obstacle_length")
    U = freestream_speed(cfg)
    q_ref = 0.5 * cfg["rho_inf"] * U ** 2 * L_ref if L_ref else 0.0
    eng = None
    tcd, tcl = cfg.get("engineering_target_cd"), cfg.get("engineering_target_cl")
    if tcd is not None and tcl is not None:
        if L_ref is None:
            raise ValueError("engineering_target_cd/cl given but the mesh metadata has no 'obstacle_length'")
        pat = max(1, int(cfg.get("engineering_patience", cfg.get("stationarity_patience", 1))))
        eng = (float(tcd), float(tcl), float(cfg.get("engineering_cd_tol", 0.001)) * max(abs(tcd), 1e-8),
               float(cfg.get("engineering_cl_tol", 0.001)) * max(abs(tcl), 1e-8), pat)
    return CaseSpec(mesh=mesh, kw=solver_kwargs(cfg, inlet), cfl=cfl, tf=float(cfg["tf"]), q_ref=q_ref,
                    engineering=eng)


def build_run_summary(cfg, mesh, cd, cl, delta_s, wall_time_s, **kw):
    """Field set of euler/utils.py:147-182."""
    return {
        "case": cfg["case"], "h": mesh.metadata.get("h"), "n_cells": int(np.asarray(mesh.tris).shape[0]),
        "cd": None if cd is None else float(cd), "cl": None if cl is None else float(cl),
        "deltaS": None if delta_s is None else float(delta_s),
        "stationarity_rel": kw.get("stationarity_rel"), "converged": kw.get("converged"),
        "stopping_step": kw.get("stopping_step"),
        "stationarity_threshold": float(cfg.get("stationarity_threshold")),
        "stationarity_check_every": int(cfg.get("stationarity_check_every")),
        "stationarity_patience": kw.get("stationarity_patience"),
        "wall_time_s": float(wall_time_s), "compile_time_s": kw.get("compile_time_s"),
        "solve_time_s": kw.get("solve_time_s"),
        "engineering_converged": kw.get("engineering_converged"),
        "engineering_stopping_step": kw.get("engineering_stopping_step"),
        "engineering_target_cd": kw.get("engineering_target_cd"),
        "engineering_target_cl": kw.get("engineering_target_cl"),
        "time_scheme": cfg["time_scheme"], "flux": cfg["flux"], "reconstruction": cfg["reconstruction"],
        "mach": cfg["Mach"], "mesh_path": cfg.get("mesh_path"), "aoa": float(cfg.get("aoa", 0.0)),
    }


def export_bundle(W, mesh, t, cfg, out_dirs):
    """Bundle fields of euler/utils.py:51-69 (gradient field omitted: it is
    a post-processing product, not part of the solve)."""
    gamma = cfg["gamma"]
    prims = helper.getPrimitive(np.asarray(W, dtype=np.float32), gamma=gamma, M=1.0)
    path = Path(out_dirs["res"]) / f"{format_snapshot_name(cfg, t)}.npz"
    np.savez_compressed(
        path, node_pos=np.asarray(mesh.barycenter, dtype=np.float32),
        node_area=np.asarray(mesh.area, dtype=np.float32), conservatives=np.asarray(W, dtype=np.float32),
        primitives=prims.astype(np.float32), mach=helper.get_mach_number(prims, gamma).astype(np.float32),
        time=np.asarray([t], dtype=np.float32), mach_in=np.asarray([cfg["Mach"]], dtype=np.float32),
        aoa_in=np.asarray([cfg.get("aoa", 0.0)], dtype=np.float32), case=np.asarray([str(cfg["case"])]),
        mesh_path=np.asarray([str(cfg.get("mesh_path", ""))]), n_cells=np.asarray([mesh.tris.shape[0]]),
        h=np.asarray([float(mesh.metadata.get("h", float("nan")))]), flux=np.asarray([str(cfg["flux"])]),
        reconstruction=np.asarray([str(cfg["reconstruction"])]),
        time_scheme=np.asarray([str(cfg["time_scheme"])]))
    return path


def run(W, mesh, inlet, cfg, out_dirs=None):
    """main.py:43-270. Returns the final state (float32 numpy); the run
    report is stored in cfg['_last_run']."""
    verbose = bool(cfg.get("verbose", True))
    W_initial = np.asarray(W, dtype=np.float32)
    t0 = time.time()
    spec = case_spec(mesh, inlet, cfg)
    solver = BatchSolver([spec], scheme=cfg["time_scheme"], threshold=float(cfg["stationarity_threshold"]),
                         check_every=int(cfg["stationarity_check_every"]),
                         patience=max(1, int(cfg.get("stationarity_patience", 1))),
                         max_steps=int(cfg.get("max_steps", 0) or 0))
    solver.set_state(0, W_initial)
    solver.prepare()
    compile_time_s = time.time() - t0
    if spec.cfl != cfg["CFL"] and verbose:
        print(f"transonic regime: CFL reduced to {spec.cfl}")
    t1 = time.time()
    solver.run()
    solve_time_s = time.time() - t1
    solver.finalize()
    rep = solver.report(0)
    Wf = solver.get_state(0).cpu().numpy()
    solver.close()
    wall = compile_time_s + solve_time_s
    final_time = rep["stopping_step"] * float(rep["dt"])
    if verbose:
        print(f"    dt={float(rep['dt']):.2e}, Nt={rep['N']}")
        print("-" * 78)
        print(f"Simulation terminée en {wall:.2f}s (converged={rep['converged']}, t={final_time:.4f}s)")
        print(f"Résidu final : {rep['final_residual']:.6e}")
    else:
        print(f"Simulation terminée en {wall:.2f}s (converged={rep['converged']}, t={final_time:.4f}s, "
              f"résidu={rep['final_residual']:.2e})")
    summary = None
    if str(cfg["case"]).lower() in KNOWN_CASES_LC and spec.q_ref > 0:
        dS = helper.get_entropy_creation(W_initial, Wf, mesh, gamma=cfg["gamma"])
        eng = spec.engineering
        summary = build_run_summary(
            cfg, mesh, rep["cd"], rep["cl"], dS, wall, stationarity_rel=rep["final_residual"],
            converged=rep["converged"], stopping_step=rep["stopping_step"], compile_time_s=compile_time_s,
            solve_time_s=solve_time_s, stationarity_patience=max(1, int(cfg.get("stationarity_patience", 1))),
            engineering_converged=rep["engineering_converged"] if eng else None,
            engineering_stopping_step=rep["engineering_stopping_step"] if eng else None,
            engineering_target_cd=eng[0] if eng else None, engineering_target_cl=eng[1] if eng else None)
        if verbose:
            print(f"C_D = {rep['cd']:.6f}, C_L = {rep['cl']:.6f}, dS = {dS:.6e}")
    exp = cfg.get("export", {})
    if out_dirs is not None:
        if exp.get("results"):
            np.save(Path(out_dirs["res"]) / f"{format_snapshot_name(cfg, final_time)}.npy", Wf)
        if exp.get("bundle", True):
            export_bundle(Wf, mesh, final_time, cfg, out_dirs)
        if exp.get("summary", True) and summary is not None:
            p = Path(out_dirs["res"]) / f"summary_{format_snapshot_name(cfg, final_time)}.json"
            p.write_text(json.dumps(summary, indent=2))
    cfg["_last_run"] = dict(report=rep, summary=summary, compile_time_s=compile_time_s,
                            solve_time_s=solve_time_s, steps=solver.steps)
    return Wf


def main(argv=None):
    cfg = build_config_from_cli(load_default_config(), argv)
    mesh = load_mesh(cfg)
    out_dirs = setup_dirs(cfg, mesh)
    print_config(cfg, mesh)
    W, inlet = initialize(mesh, cfg)
    run(W, mesh, inlet, cfg, out_dirs)
    return 0


if __name__ == "__main__":
    sys.exit(main())
