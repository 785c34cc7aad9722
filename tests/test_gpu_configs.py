"""The batched configurations at their full sizes, checked through
size-independent properties: every case reported with finite results, and
a seeded sample of cases equal, bit for bit, to the same cases solved alone.

* Ragged plain-method batch: 256 cases over 20 meshes (diamond, NACA0012,
  NACA2412, NACA4412 at h = 0.1, 0.05, 0.025, 0.0175, 0.0125: 1.6k-113k
  cells) with (M, alpha) drawn by default_rng(0) from the reference grid,
  HLLC + MUSCL, SSP-RK2, all in one solver with per-case convergence masks
  (threshold 5e-5, patience 3, at most 4,000 steps).
* The 1,024-case campaign at h = 0.025 over five geometries, half HLLC and
  half Roe, learned warm start from an LR (h = 0.1) solve with the
  freestream fallback, continuous batching with 64 slots. To keep the test
  short, the reference fields that set each case's threshold stop after
  3,000 steps, the training pairs after 3,000 steps and the networks train
  300 steps, and every solver runs the launch the autotuner picks for the
  campaign (cell path, 64 threads, XCD order) so the solo runs take the same
  path; the campaign machinery is the bench's."""
import copy
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))

pytestmark = pytest.mark.gpu


def _solo(spec, init, kwargs):
    from eulersr_amd.solve import BatchSolver
    one = BatchSolver([spec], **kwargs)
    try:
        one.set_state(0, init)
        one.prepare()
        one.run()
        return one.report(0), one.get_state(0).cpu().numpy()
    finally:
        one.close()


def test_ragged_256_cases_20_meshes_one_solver():
    import bench
    from eulersr_amd import main as M
    from eulersr_amd.meshing import build_mesh
    from eulersr_amd.solve import BatchSolver
    from conftest import base_cfg
    grid = bench.mixed_grid(256)
    meshes = {key: build_mesh(*key) for key in dict.fromkeys((g, h) for g, h, _, _ in grid)}
    assert len(meshes) == 20
    specs, inits = [], []
    for g, h, m, a in grid:
        cfg = base_cfg(Mach=m, aoa=a, case=g, CFL=0.6, tf=1.0e4)
        W, inlet = M.initialize(meshes[(g, h)], cfg)
        specs.append(M.case_spec(meshes[(g, h)], inlet, cfg))
        inits.append(W)
    cells = [sp.mesh.n_cells for sp in specs]
    print(f"case: ragged 256 cases, {min(cells)}-{max(cells)} cells, {sum(cells)} in total", flush=True)
    kwargs = dict(scheme="SRK2", threshold=5e-5, check_every=5, patience=3, max_steps=4000)
    bs = BatchSolver(specs, autotune=True, **kwargs)
    try:
        for i, W in enumerate(inits):
            bs.set_state(i, W)
        bs.prepare()
        bs.run()
        reps = bs.reports()
        sample = sorted(np.random.default_rng(5).choice(256, 8, replace=False).tolist())
        early = [i for i, r in enumerate(reps) if r["converged"]]
        if early:
            sample = sorted(set(sample) | {early[0]})
        states = {i: bs.get_state(i).cpu().numpy() for i in sample}
        tuning = bs.tuning
    finally:
        bs.close()
    print(f"tuning {tuning}; converged before the cap: {len(early)}", flush=True)
    assert len(reps) == 256
    assert all(np.isfinite(r["last_rel"]) and r["steps_run"] > 0 for r in reps)
    assert 0 < len(early) < 256          # masks stop some cases and keep others running
    for i in sample:
        r, W = _solo(specs[i], inits[i], dict(kwargs, launch=(tuning["path"], tuning["block"], tuning["xcd_order"])))
        g, h, m, a = grid[i]
        print(f"case {i} ({g} h={h} M={m} a={a}, {cells[i]} cells): batch stop {reps[i]['stopping_step']} "
              f"conv {reps[i]['converged']}; alone stop {r['stopping_step']} conv {r['converged']}", flush=True)
        assert np.all(np.isfinite(W))
        assert r["converged"] == reps[i]["converged"] and r["stopping_step"] == reps[i]["stopping_step"]
        assert r["steps_run"] == reps[i]["steps_run"] and r["dt"] == reps[i]["dt"]
        assert np.array_equal(W.view(np.uint32), states[i].view(np.uint32))


LAUNCH = ("cell", 64, True)     # the launch the autotuner picks for the campaign's batch


def test_campaign_1024_cases_warm_start_matches_solo_runs(monkeypatch):
    import bench
    import torch
    from eulersr_amd import pipeline as P
    monkeypatch.setattr(P, "AUTOTUNE", False)
    monkeypatch.setattr(P, "LAUNCH", LAUNCH)
    meshes = bench.build_meshes(bench.GEOMETRIES, 0.025, 0.1)
    train_grid = [(m, a) for m in bench.TRAIN_MACHS for a in bench.TRAIN_AOAS]
    starter = P.WarmStarter()
    for g, (lm, hm) in meshes.items():
        pairs = P.training_pairs(g, lm, hm, train_grid, max_steps=3000)
        kmap = P.WS.KnnMap(lm, hm)
        net, _ = P.train(kmap, pairs, steps=300, batch=16384, seed=0)
        starter.add(g, lm, hm, net=net, kmap=kmap)
        del pairs
    grid = bench.ragged_grid(1024)
    cases, ref, _ = bench.make_workload(meshes, grid, cap_ref=3000)
    print(f"case: 1,024-case campaign, {sum(c.hr_spec.mesh.n_cells for c in cases)} HR cells", flush=True)
    run = starter.campaign(cases, mode="mlp", slots=64, max_steps=20000)
    reps = run["reports"]
    assert len(reps) == 1024 and all(r is not None for r in reps)
    assert all(np.isfinite(r["cd"]) and np.isfinite(r["cl"]) for r in reps)
    assert all(r["converged"] or r["steps_run"] >= r["N"] for r in reps)
    n_conv = sum(r["converged"] for r in reps)
    print(f"converged {n_conv}/1024, fallback {sum(run['fallback'])}, HR {run['hr_s']:.1f}s", flush=True)
    # the warm-start states again (LR solves and inference are deterministic)
    inits, fallback, _, _ = starter.initial_states(cases, "mlp")
    sample = sorted(np.random.default_rng(3).choice(1024, 8, replace=False).tolist())
    fb = [i for i, f in enumerate(run["fallback"]) if f]
    if fb:
        sample = sorted(set(sample) | {fb[0]})
    for i in sample:
        init = cases[i].hr_init if run["fallback"][i] else inits[i]
        sp = copy.copy(cases[i].hr_spec)
        sp.max_steps = 20000
        r, W = _solo(sp, init, dict(scheme="SRK2", threshold=1.0, check_every=P.CHECK_EVERY, patience=P.PATIENCE,
                                    launch=P.LAUNCH))
        c = reps[i]
        print(f"case {i} {grid[i]}: campaign stop {c['stopping_step']} cd {c['cd']!r} fallback "
              f"{run['fallback'][i]}; alone stop {r['stopping_step']} cd {r['cd']!r}", flush=True)
        assert np.all(np.isfinite(W))
        assert c["converged"] == r["converged"] and c["stopping_step"] == r["stopping_step"]
        assert c["cd"] == r["cd"] and c["cl"] == r["cl"]
    torch.cuda.empty_cache()
