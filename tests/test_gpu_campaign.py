"""Continuous batching (esr_solver_replace_case): a case that enters a slot
mid-run behaves exactly as if it were solved alone -- same dt, stationarity
checks, stopping step and in-loop CD -- and a replacement of another kind is
refused."""
import copy

import numpy as np
import pytest

from conftest import base_cfg, mesh_of

pytestmark = pytest.mark.gpu

LAUNCH = ("cell", 128, True)


def _cases():
    from eulersr_amd import main as M
    m1 = mesh_of("diamond", 0.2)
    m2 = mesh_of("naca0012", 0.2)
    specs, inits = [], []
    for mesh, mach, aoa, case in [(m1, 2.0, 3.0, "diamond"), (m1, 1.5, -2.0, "diamond"), (m2, 0.8, 1.0, "naca0012"),
                                  (m1, 2.6, 0.0, "diamond"), (m2, 1.8, 4.0, "naca0012"), (m1, 1.2, 1.0, "diamond")]:
        cfg = base_cfg(Mach=mach, aoa=aoa, case=case, tf=1.0e4)
        W, inlet = M.initialize(mesh, cfg)
        sp = M.case_spec(mesh, inlet, cfg)
        sp.threshold = 2e-4
        sp.engineering = (0.05, 0.05, 1e-9, 1e-9, 3)   # in-loop CD/CL every check
        specs.append(sp)
        inits.append(W)
    return specs, inits


def test_replaced_cases_match_cases_run_alone():
    from eulersr_amd import pipeline as P
    from eulersr_amd.solve import BatchSolver
    specs, inits = _cases()
    old = (P.AUTOTUNE, P.LAUNCH)
    P.AUTOTUNE, P.LAUNCH = False, LAUNCH
    try:
        reps, t, restarted = P.campaign(specs, inits, slots=2, max_steps=3000, chunk=40)
        assert restarted == []
    finally:
        P.AUTOTUNE, P.LAUNCH = old
    for i, (sp, W) in enumerate(zip(specs, inits)):
        sp1 = copy.copy(sp)
        sp1.max_steps = 3000
        one = BatchSolver([sp1], threshold=1.0, check_every=5, patience=3, launch=LAUNCH)
        one.set_state(0, W)
        one.prepare()
        one.run()
        r = one.report(0)
        one.close()
        c = reps[i]
        print(f"case {i}: campaign steps {c['stopping_step']} conv {c['converged']} cd {c['cd']!r}; "
              f"alone steps {r['stopping_step']} conv {r['converged']} cd {r['cd']!r}", flush=True)
        assert c["converged"] == r["converged"]
        assert c["stopping_step"] == r["stopping_step"]
        assert c["N"] == r["N"] and c["dt"] == r["dt"]
        assert c["cd"] == r["cd"] and c["cl"] == r["cl"]


def test_replacement_of_another_kind_is_refused():
    from eulersr_amd.solve import BatchSolver
    specs, inits = _cases()
    bs = BatchSolver(specs[:2], threshold=1.0, check_every=5, patience=3, max_steps=20, launch=LAUNCH)
    for i in range(2):
        bs.set_state(i, inits[i])
    bs.prepare()
    bs.run(10)
    with pytest.raises(ValueError):
        bs.replace_case(0, specs[2], inits[2])        # NACA mesh into a diamond slot
    bs.replace_case(1, specs[3], inits[3])            # same mesh and kernels: accepted
    bs.run(20)
    assert bs.report(1)["steps_run"] == 10
    bs.close()


def test_hedged_campaign_matches_solo_runs_from_the_winning_start():
    """Warm starts (a perturbed freestream) with freestream fallbacks in 4
    slots (sub-groups of 2): once nothing waits, free slots run fallbacks
    beside the warm starts still in flight. Every case's result equals, bit
    for bit, a solo run from the start that produced it: its warm start, or
    freestream for a case that was restarted or whose hedge won."""
    from conftest import perturbed_state
    from eulersr_amd import pipeline as P
    from eulersr_amd.solve import BatchSolver
    specs, cold = _cases()
    warm = []
    for i, sp in enumerate(specs):
        cfg = base_cfg(Mach=float(np.hypot(*sp.kw["value"][1:3]) / sp.kw["value"][0] / np.sqrt(1.4)))
        W, _ = perturbed_state(sp.mesh, dict(cfg, aoa=0.0), amp=0.03, seed=i)
        warm.append(W)
    # a poor prediction for case 4 (NACA0012, M = 1.8): the subsonic
    # freestream of M = 0.9; its hedge from freestream should finish first
    from eulersr_amd import main as M
    warm[4], _ = M.initialize(specs[4].mesh, base_cfg(Mach=0.9, aoa=0.0, case="naca0012"))
    old = (P.AUTOTUNE, P.LAUNCH)
    P.AUTOTUNE, P.LAUNCH = False, LAUNCH
    try:
        reps, t, restarted = P.campaign(specs, warm, slots=4, max_steps=3000, chunk=40, fallback_inits=cold)
    finally:
        P.AUTOTUNE, P.LAUNCH = old
    print(f"restarted {restarted}", flush=True)
    assert 4 in restarted
    for i, sp in enumerate(specs):
        sp1 = copy.copy(sp)
        sp1.max_steps = 3000
        one = BatchSolver([sp1], threshold=1.0, check_every=5, patience=3, launch=LAUNCH)
        one.set_state(0, cold[i] if i in restarted else warm[i])
        one.prepare()
        one.run()
        r = one.report(0)
        one.close()
        c = reps[i]
        print(f"case {i}: campaign steps {c['stopping_step']} conv {c['converged']}; alone steps "
              f"{r['stopping_step']} conv {r['converged']}", flush=True)
        assert c["converged"] == r["converged"] and c["stopping_step"] == r["stopping_step"]
        assert c["cd"] == r["cd"] and c["cl"] == r["cl"]
