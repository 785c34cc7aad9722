"""The single-case GPU configuration at its full size: diamond and NACA0012 at
h = 0.025 (22k and 28k cells), M in {0.9, 1.1, 2.0}, alpha = 2 deg, HLLC +
MUSCL, SSP-RK2 at CFL 0.6 (0.45 in the transonic band), from freestream. The
six cases run as one GPU batch for 100 steps and are compared cell by cell
with the C restatement of the oracle (oracle/cfvm.c, float32, bit-identical
to oracle/fvm.py for HLLC) fed the same time step: every primitive within
1e-5 of max(|p|, freestream scale), the bar of the hot path; the time step
against the oracle's get_dt within 2e-7. The same six cases with Roe (the
other half of the headline campaign's cases) are held to the same bar."""
import numpy as np
import pytest

from conftest import base_cfg, mesh_of

pytestmark = pytest.mark.gpu

MACHS, AOA, STEPS = (0.9, 1.1, 2.0), 2.0, 100
_ORACLE = {}


def _cases(flux="HLLC"):
    from eulersr_amd import main as M
    specs, inits, cfgs = [], [], []
    for geom in ("diamond", "naca0012"):
        mesh = mesh_of(geom, 0.025)
        for mach in MACHS:
            cfg = base_cfg(Mach=mach, aoa=AOA, case=geom, flux=flux, reconstruction="MUSCL", CFL=0.6)
            W, inlet = M.initialize(mesh, cfg)
            specs.append(M.case_spec(mesh, inlet, cfg))
            inits.append(W)
            cfgs.append(cfg)
    return specs, inits, cfgs


def _oracle(i, sp, W0, dt, flux="HLLC"):
    """100 SSP-RK2 steps of case i on the C oracle (cached across launches)."""
    from oracle import cfvm as CF
    from oracle import fvm as F
    key = (i, float(dt), flux)
    if key not in _ORACLE:
        cm = CF.CMesh(F.omesh(sp.mesh, np.float32))
        kw = dict(gamma=1.4, M=1.0, reconstruction="MUSCL", flux=flux, value=sp.kw["value"], entropy=False)
        _ORACLE[key] = (CF.get_dt(W0, cm, CFL=sp.cfl), CF.srk2_steps(W0, cm, np.float32(dt), STEPS, **kw))
    return _ORACLE[key]


@pytest.mark.parametrize("flux,launch", [("HLLC", ("cell", 64, True)), ("HLLC", ("face", 64, True)),
                                         ("ROE", ("cell", 64, True)), ("ROE", ("face", 64, True))])
def test_c2_cases_100_steps_within_1e5(flux, launch):
    from eulersr_amd.solve import BatchSolver
    from oracle import fvm as F
    specs, inits, cfgs = _cases(flux)
    print(f"flux {flux} launch {launch}: {[sp.mesh.n_cells for sp in specs]} cells, cfl {[sp.cfl for sp in specs]}",
          flush=True)
    bs = BatchSolver(specs, scheme="SRK2", threshold=0.0, check_every=5, patience=1, max_steps=STEPS,
                     launch=launch)
    try:
        for i, W in enumerate(inits):
            bs.set_state(i, W)
        bs.prepare()
        bs.run()
        reps = bs.reports()
        states = [bs.get_state(i).cpu().numpy() for i in range(len(specs))]
    finally:
        bs.close()
    for i, (sp, W0, cfg) in enumerate(zip(specs, inits, cfgs)):
        dt = reps[i]["dt"]
        assert reps[i]["steps_run"] == STEPS
        dt_or, Wr = _oracle(i, sp, W0, dt, flux)
        assert abs(float(dt) - float(dt_or)) <= 2e-7 * float(dt_or), (float(dt), float(dt_or))
        pg = F.getPrimitive(states[i].astype(np.float64))
        pr = F.getPrimitive(Wr.astype(np.float64))
        U = np.hypot(sp.kw["value"][1], sp.kw["value"][2]) / sp.kw["value"][0]
        scale = np.array([1.0, U, U, 1.0])
        rel = np.abs(pg - pr) / np.maximum(np.abs(pr), scale)
        print(f"case {cfg['case']} M={cfg['Mach']}: dt {float(dt):.6e}, max rel per primitive "
              f"{np.array2string(rel.max(axis=0), precision=2)}", flush=True)
        assert np.all(np.isfinite(pg))
        assert rel.max() <= 1e-5


def test_largest_and_smallest_mesh_in_one_launch_within_1e5():
    """The extremes of the ragged configuration's mesh sizes in one launch:
    the diamond at h = 0.0125 (89k cells) and at h = 0.1 (1.6k cells),
    100 SSP-RK2 steps from freestream, every primitive of every cell within
    1e-5 of the C oracle given the same dt."""
    from eulersr_amd import main as M
    from eulersr_amd.solve import BatchSolver
    from oracle import cfvm as CF
    from oracle import fvm as F
    specs, inits = [], []
    for h, mach in ((0.0125, 2.0), (0.1, 1.5)):
        mesh = mesh_of("diamond", h)
        cfg = base_cfg(Mach=mach, aoa=AOA, case="diamond", flux="HLLC", reconstruction="MUSCL", CFL=0.6)
        W, inlet = M.initialize(mesh, cfg)
        specs.append(M.case_spec(mesh, inlet, cfg))
        inits.append(W)
    print(f"cells {[sp.mesh.n_cells for sp in specs]}", flush=True)
    bs = BatchSolver(specs, scheme="SRK2", threshold=0.0, check_every=5, patience=1, max_steps=STEPS)
    try:
        for i, W in enumerate(inits):
            bs.set_state(i, W)
        bs.prepare()
        bs.run()
        reps = bs.reports()
        states = [bs.get_state(i).cpu().numpy() for i in range(len(specs))]
    finally:
        bs.close()
    for i, (sp, W0) in enumerate(zip(specs, inits)):
        kw = dict(gamma=1.4, M=1.0, reconstruction="MUSCL", flux="HLLC", value=sp.kw["value"], entropy=False)
        Wr = CF.srk2_steps(W0, CF.CMesh(F.omesh(sp.mesh, np.float32)), np.float32(reps[i]["dt"]), STEPS, **kw)
        pg = F.getPrimitive(states[i].astype(np.float64))
        pr = F.getPrimitive(Wr.astype(np.float64))
        U = np.hypot(sp.kw["value"][1], sp.kw["value"][2]) / sp.kw["value"][0]
        rel = np.abs(pg - pr) / np.maximum(np.abs(pr), np.array([1.0, U, U, 1.0]))
        print(f"{sp.mesh.n_cells} cells: max rel per primitive {np.array2string(rel.max(axis=0), precision=2)}",
              flush=True)
        assert reps[i]["steps_run"] == STEPS and rel.max() <= 1e-5
