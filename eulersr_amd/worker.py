"""Subprocess worker of the warm-start harness, counterpart of
eval/_fvm_solve_worker.py:31-155: one solve, a one-step probe or the forces
of a given field, configured from the command line (same arguments), with
the reference's stdout protocol, which its parent parses by regex
(eval/live_fvm.py:25-28):

    FVM_SUMMARY_JSON={...}       the run summary (--emit-summary)
    FVM_PROBE_JSON={...}         one-step residual and dt of --init-field (--probe-only)
    FVM_FORCES_JSON={...}        CD/CL of --init-field (--forces-only)
    Simulation terminée en Xs    printed by every full run (main.run)

    python -m eulersr_amd.worker --case naca0012 --mesh-path gen:naca0012:0.025 \\
        --mach 0.8 --aoa 1.25 --tf 3 --stationarity-threshold 1e-5 \\
        --scratch-dir /tmp/w [--init-field prim.npy] [--probe-only | --forces-only] \\
        [--target-cd CD --target-cl CL] [--emit-summary]

--init-field is a (n_cells, 4) float array of primitives in the mesh's cell
order (loaded without pickles). Errors keep the reference's types:
ValueError for a field of the wrong shape, SystemExit for --probe-only /
--forces-only without a field or for forces on a mesh without an obstacle
length.
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

import numpy as np

from . import euler_solver, helper
from . import main as M
from .config import load_default_config, load_mesh
from .solve import transonic_cfl

SUMMARY_TAG, PROBE_TAG, FORCES_TAG = "FVM_SUMMARY_JSON=", "FVM_PROBE_JSON=", "FVM_FORCES_JSON="


def build_parser():
    """eval/_fvm_solve_worker.py:32-63"""
    ap = argparse.ArgumentParser(description="one FVM solve with the harness's stdout protocol")
    ap.add_argument("--case", required=True)
    ap.add_argument("--mesh-path", required=True)
    ap.add_argument("--mach", type=float, required=True)
    ap.add_argument("--aoa", type=float, required=True)
    ap.add_argument("--tf", type=float, required=True)
    ap.add_argument("--stationarity-threshold", type=float, required=True)
    ap.add_argument("--scratch-dir", required=True)
    ap.add_argument("--stationarity-check-every", type=int, default=None,
                    help="overrides stationarity_check_every (config default: 10000)")
    ap.add_argument("--stationarity-patience", type=int, default=1,
                    help="consecutive checks under the threshold before a stop")
    ap.add_argument("--target-cd", type=float, default=None, help="engineering criterion: target C_D")
    ap.add_argument("--target-cl", type=float, default=None, help="engineering criterion: target C_L")
    ap.add_argument("--cd-tol", type=float, default=0.001, help="relative tolerance on C_D (fraction of |target|)")
    ap.add_argument("--cl-tol", type=float, default=0.001, help="relative tolerance on C_L (fraction of |target|)")
    ap.add_argument("--engineering-patience", type=int, default=None,
                    help="patience of the engineering criterion (default: --stationarity-patience)")
    ap.add_argument("--init-field", default=None, help="(n_cells, 4) primitives to warm-start from (.npy)")
    ap.add_argument("--emit-summary", action="store_true", help="print the run summary as FVM_SUMMARY_JSON=")
    ap.add_argument("--probe-only", action="store_true",
                    help="with --init-field: the residual of one step only (FVM_PROBE_JSON=)")
    ap.add_argument("--forces-only", action="store_true",
                    help="with --init-field: C_D/C_L of that field, no evolution (FVM_FORCES_JSON=)")
    return ap


def configure(args):
    """The run configuration of eval/_fvm_solve_worker.py:67-90."""
    cfg = load_default_config()
    cfg.update(case=args.case, Mach=args.mach, aoa=args.aoa, mesh_path=args.mesh_path, tf=args.tf,
               stationarity_threshold=args.stationarity_threshold, verbose=False)
    if args.stationarity_check_every is not None:
        cfg["stationarity_check_every"] = args.stationarity_check_every
    cfg["stationarity_patience"] = max(1, int(args.stationarity_patience))
    if args.target_cd is not None and args.target_cl is not None:
        cfg["engineering_target_cd"] = float(args.target_cd)
        cfg["engineering_target_cl"] = float(args.target_cl)
        cfg["engineering_cd_tol"] = float(args.cd_tol)
        cfg["engineering_cl_tol"] = float(args.cl_tol)
        cfg["engineering_patience"] = (int(args.engineering_patience) if args.engineering_patience is not None
                                       else cfg["stationarity_patience"])
    # only the summary (in the scratch directory) is ever written
    cfg["export"] = dict(results=False, figures=[], graph=False, summary=bool(args.emit_summary), bundle=False,
                         n_snaps=1)
    return cfg


def initial_state(args, mesh, cfg):
    """Freestream, or the conserved form of --init-field
    (eval/_fvm_solve_worker.py:97-105)."""
    W, inlet = M.initialize(mesh, cfg)
    if args.init_field:
        prim = np.load(args.init_field)
        n = int(np.asarray(mesh.area).shape[0])
        if prim.shape != (n, 4):
            raise ValueError(f"--init-field has shape {prim.shape}, expected ({n}, 4) for this mesh.")
        W = helper.getConserved(np.asarray(prim, dtype=np.float32), gamma=cfg["gamma"], M=1.0).astype(np.float32)
    return W, inlet


def forces(W, mesh, cfg):
    """CD/CL of a field (eval/_fvm_solve_worker.py:107-118)."""
    rho_inf, p_inf, gamma = cfg["rho_inf"], cfg["p_inf"], cfg["gamma"]
    U_inf = cfg["Mach"] * float((gamma * p_inf / rho_inf) ** 0.5)
    L_ref = mesh.metadata.get("obstacle_length")
    if L_ref is None:
        raise SystemExit(f"the mesh metadata has no 'obstacle_length' (case='{cfg['case']}').")
    cd, cl = helper.get_force_coefficients(W, mesh, rho_inf, U_inf, L_ref)
    return dict(cd=float(cd), cl=float(cl))


def probe(W, mesh, inlet, cfg):
    """Residual of one step of the configured scheme from W, with the dt a
    run would take (transonic CFL clamp included): ||W1 - W|| / (||W|| +
    1e-16) (eval/_fvm_solve_worker.py:120-139)."""
    cfl = transonic_cfl(cfg["Mach"], cfg["CFL"])
    dt = helper.get_dt(W, mesh, CFL=cfl, gamma=cfg["gamma"], M=1.0)
    kw = dict(gamma=cfg["gamma"], M=1.0, reconstruction=cfg["reconstruction"], flux=cfg["flux"], value=inlet,
              entropy=False)
    W1 = euler_solver.time_step(str(cfg["time_scheme"]).upper(), W, mesh, dt, **kw)
    return dict(residual=float(helper.relative_change(W1, W, mesh)), dt=float(dt))


def main(argv=None) -> int:
    args = build_parser().parse_args(argv)
    cfg = configure(args)
    mesh = load_mesh(cfg)
    scratch = Path(args.scratch_dir)
    out_dirs = {"fig": scratch / "fig", "res": scratch / "res"}
    for d in out_dirs.values():
        d.mkdir(parents=True, exist_ok=True)
    W, inlet = initial_state(args, mesh, cfg)
    if args.forces_only:
        if not args.init_field:
            raise SystemExit("--forces-only needs --init-field.")
        print(FORCES_TAG + json.dumps(forces(W, mesh, cfg)), flush=True)
        return 0
    if args.probe_only:
        if not args.init_field:
            raise SystemExit("--probe-only needs --init-field.")
        print(PROBE_TAG + json.dumps(probe(W, mesh, inlet, cfg)), flush=True)
        return 0
    if args.emit_summary:
        # a summary left by an earlier run in the same scratch would be read back
        for stale in out_dirs["res"].glob("summary_*.json"):
            stale.unlink()
    M.run(W, mesh, inlet, cfg, out_dirs)
    if args.emit_summary:
        found = sorted(out_dirs["res"].glob("summary_*.json"))
        if not found:
            raise RuntimeError(f"--emit-summary was given but no summary was written (case='{args.case}').")
        print(SUMMARY_TAG + json.dumps(json.loads(found[-1].read_text())), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
