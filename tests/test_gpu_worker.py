"""The subprocess worker on the GPU, read through the stdout protocol its
parent parses (eval/live_fvm.py:25-28): --forces-only and --probe-only on a
warm-start field against the oracle, and a full run with --emit-summary that
prints the summary line and the timing line."""
import json
import re

import numpy as np
import pytest

from conftest import base_cfg, perturbed_state

pytestmark = pytest.mark.gpu

TIME_LINE = re.compile(r"Simulation termin.+? en ([\d.]+)s")
SUMMARY = re.compile(r"^FVM_SUMMARY_JSON=(.+)$", re.MULTILINE)
PROBE = re.compile(r"^FVM_PROBE_JSON=(.+)$", re.MULTILINE)
FORCES = re.compile(r"^FVM_FORCES_JSON=(.+)$", re.MULTILINE)
MACH, AOA = 2.0, 3.0


@pytest.fixture(scope="module")
def setup(tmp_path_factory):
    from eulersr_amd.meshing import build_mesh
    from oracle import fvm as F
    d = tmp_path_factory.mktemp("worker")
    mesh = build_mesh("diamond", 0.3)
    path = mesh.save(d / "diamond_h0.3.npz")
    cfg = base_cfg(Mach=MACH, aoa=AOA)
    W, inlet = perturbed_state(mesh, cfg, seed=9)
    prim = F.getPrimitive(W.astype(np.float64)).astype(np.float32)
    field = d / "prim.npy"
    np.save(field, prim)
    return mesh, path, field, d, prim, inlet


def _run(capsys, setup, *extra):
    from eulersr_amd import worker as WK
    mesh, path, field, d, _, _ = setup
    argv = ["--case", "diamond", "--mesh-path", str(path), "--mach", str(MACH), "--aoa", str(AOA),
            "--tf", "0.5", "--stationarity-threshold", "1e-5", "--scratch-dir", str(d / "scratch"), *extra]
    assert WK.main(argv) == 0
    return capsys.readouterr().out


def test_forces_only(capsys, setup):
    from oracle import fvm as F
    mesh, _, field, _, prim, _ = setup
    out = _run(capsys, setup, "--init-field", str(field), "--forces-only")
    got = json.loads(FORCES.search(out).group(1))
    W = F.getConserved(prim.astype(np.float32)).astype(np.float32)
    om = F.omesh(mesh, np.float32)
    U = MACH * np.sqrt(1.4)
    cd, cl = F.forces(W, om, 1.0, U, 1.0)
    # rounding bound: 1e-6 of the sum of |P n ds| / (q L) (the wall sum cancels)
    fids, cells, locs = F.wall_faces(om)
    P = F.getPrimitive(W.astype(np.float64))[cells, 3]
    q = 0.5 * U ** 2
    sx = np.sum(np.abs(P * om.normals[cells, locs, 0] * om.surface[fids])) / q
    sy = np.sum(np.abs(P * om.normals[cells, locs, 1] * om.surface[fids])) / q
    print(f"case: worker forces {got} oracle ({cd}, {cl}) scales ({sx}, {sy})", flush=True)
    assert abs(got["cd"] - cd) <= 1e-6 * sx and abs(got["cl"] - cl) <= 1e-6 * sy


def test_probe_only(capsys, setup):
    from oracle import fvm as F
    mesh, _, field, _, prim, inlet = setup
    out = _run(capsys, setup, "--init-field", str(field), "--probe-only")
    got = json.loads(PROBE.search(out).group(1))
    om = F.omesh(mesh, np.float32)
    W = F.getConserved(prim.astype(np.float32)).astype(np.float32)
    dt = F.get_dt(W, om, CFL=0.6)
    kw = dict(gamma=1.4, M=1.0, reconstruction="MUSCL", flux="HLLC", value=inlet, entropy=False)
    W1 = F.time_step_RK2_SSP(W, om, dt, **kw)
    res = float(np.linalg.norm((W1 - W).astype(np.float64)) / (np.linalg.norm(W.astype(np.float64)) + 1e-16))
    print(f"case: worker probe {got} oracle residual {res} dt {dt}", flush=True)
    assert abs(got["dt"] - float(dt)) <= 2e-7 * float(dt)
    assert abs(got["residual"] - res) <= 1e-4 * res


def test_full_run_with_summary(capsys, setup):
    out = _run(capsys, setup, "--stationarity-check-every", "5", "--stationarity-patience", "3",
               "--target-cd", "0.02", "--target-cl", "0.1", "--emit-summary")
    assert TIME_LINE.search(out), out
    summary = json.loads(SUMMARY.search(out).group(1))
    print(f"case: worker summary {summary}", flush=True)
    for key in ("cd", "cl", "converged", "stopping_step", "stationarity_rel", "wall_time_s",
                "engineering_stopping_step", "engineering_target_cd"):
        assert key in summary, key
    assert summary["engineering_target_cd"] == 0.02 and np.isfinite(summary["cd"])
