"""Ragged plain batch in the shape of the 256-case configuration: four
geometries (diamond, NACA0012, NACA2412, NACA4412) at several mesh sizes,
(M, alpha) drawn with default_rng(0) from M = 0.70 + 0.01 i (i = 0..230) and
alpha = -5 + 0.5 j (j = 0..20), the (geometry, h) pair of each case drawn by
the same generator, HLLC and Roe alternating. Every case runs in one solver
(one launch per stage kernel over every group, per-case stationarity masks,
graph-captured chunks) and must report and hold exactly what it reports and
holds when solved alone: final state, stopping step, convergence flag, CD
and CL bitwise; the FP64 stationarity and probe residuals to 1e-12 relative
(their per-range partial sums are laid out by the group's case count, so
their summation order differs from a solo run). Sizes are cut for
a test: h in {0.1, 0.07, 0.05} (1.3k-6.3k cells), 12 cases, at most 600
steps, every fourth case stopping at its third check; both stage paths."""
import numpy as np
import pytest

from conftest import base_cfg, mesh_of

pytestmark = pytest.mark.gpu

GEOMS = ("diamond", "naca0012", "naca2412", "naca4412")
HS = (0.1, 0.07, 0.05)
N_CASES = 12


def _cases():
    from eulersr_amd import main as M
    rng = np.random.default_rng(0)
    specs, inits, keys = [], [], []
    for i in range(N_CASES):
        mach = round(0.70 + 0.01 * int(rng.integers(0, 231)), 2)
        aoa = -5.0 + 0.5 * int(rng.integers(0, 21))
        geom = GEOMS[int(rng.integers(0, len(GEOMS)))]
        h = HS[int(rng.integers(0, len(HS)))]
        flux = "HLLC" if i % 2 == 0 else "ROE"
        mesh = mesh_of(geom, h)
        cfg = base_cfg(Mach=mach, aoa=aoa, case=geom, flux=flux)
        W, inlet = M.initialize(mesh, cfg)
        sp = M.case_spec(mesh, inlet, cfg)
        # every fourth case stops when its third check (step 15, patience 3)
        # confirms the streak that began at step 5 (main.py's stopping step
        # is the streak's start); the others run to the 600-step cap
        sp.threshold = 1.0 if i % 4 == 0 else 0.0
        specs.append(sp)
        inits.append(W)
        keys.append((geom, h, mach, aoa, flux, mesh.n_cells))
    return specs, inits, keys


def _solve(specs, inits, launch):
    from eulersr_amd.solve import BatchSolver
    bs = BatchSolver(specs, scheme="SRK2", threshold=0.0, check_every=5, patience=3, max_steps=600,
                     launch=launch)
    try:
        for i, W in enumerate(inits):
            bs.set_state(i, W)
        bs.prepare()
        bs.run()
        bs.finalize()
        states = [bs.get_state(i).cpu().numpy() for i in range(len(specs))]
        reps = bs.reports()
    finally:
        bs.close()
    return states, reps


@pytest.mark.parametrize("launch", [("cell", 64, True), ("face", 256, False)])
def test_ragged_batch_equals_cases_solved_alone(launch):
    specs, inits, keys = _cases()
    print(f"launch {launch}: {len(specs)} cases, cells {sorted(k[-1] for k in keys)}", flush=True)
    assert len({(k[0], k[1]) for k in keys}) >= 6          # ragged: many meshes in one launch
    states, reps = _solve(specs, inits, launch)
    stopped = sum(r["converged"] for r in reps)
    print(f"batched: {stopped} of {len(reps)} stopped early, steps {[r['stopping_step'] for r in reps]}",
          flush=True)
    assert stopped == N_CASES // 4                           # per-case masks are exercised
    assert all(r["stopping_step"] == 5 for r in reps if r["converged"])
    for i, (sp, W) in enumerate(zip(specs, inits)):
        s1, r1 = _solve([sp], [W], launch)
        print(f"case {i} {keys[i]}: batched vs alone", flush=True)
        assert np.array_equal(states[i].view(np.uint32), s1[0].view(np.uint32))
        for key in ("converged", "stopping_step", "steps_run", "N"):
            assert reps[i][key] == r1[0][key], key
        for key in ("cd", "cl"):          # one block per case, fixed order: batch-invariant
            assert reps[i][key] == r1[0][key] or (reps[i][key] != reps[i][key] and r1[0][key] != r1[0][key]), key
        # the FP64 norms' cell ranges depend on the group's case count, so
        # their summation order (not their inputs) differs from a solo run
        for key in ("last_rel", "final_residual"):
            a, b = reps[i][key], r1[0][key]
            assert (a != a and b != b) or abs(a - b) <= 1e-12 * abs(b), (key, a, b)
