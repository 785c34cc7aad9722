"""The campaign scheduler (pipeline.campaign) against a mock solver on the
CPU: every case is solved exactly once in a slot of its own kind, slots are
refilled in place, sub-groups keep their kind and launch-group tag, idle
sub-groups open for the kinds with waiting cases so the cases in flight stay
near the slot budget, stalled warm starts restart from the fallback state,
and runs end on multiples of the chunk."""
import copy
from dataclasses import dataclass, field

import pytest

from eulersr_amd import pipeline as P


@dataclass
class Spec:
    mesh: object
    kw: dict
    cost: int                       # steps to converge from `inits`
    max_steps: int = 0
    group: int = 0
    tag: dict = field(default_factory=dict)


class MockSolver:
    """Per-slot step counters; a case converges after its cost in steps (a
    placeholder with max_steps ends after max_steps)."""

    instances = []

    def __init__(self, specs, **kw):
        self.specs = [copy.copy(s) for s in specs]
        self.start = [0] * len(specs)
        self.init = [None] * len(specs)
        self.step = 0
        self.log = []
        self.compacts = 0
        MockSolver.instances.append(self)

    def set_state(self, j, W):
        self.init[j] = W

    def prepare(self):
        pass

    def tune(self):
        self.log.append(("tune", self.step))

    def _cost(self, j):
        sp = self.specs[j]
        c = sp.cost if self.init[j] != "fallback" else sp.tag.get("cold", sp.cost)
        if self.init[j] == "stuck":
            c = 10 ** 9
        return c

    def _report(self, j):
        sp = self.specs[j]
        ran = self.step - self.start[j]
        N = sp.max_steps if sp.max_steps else 10 ** 9
        conv = ran >= self._cost(j) and ran < N
        hist = 1e-3 if self.init[j] == "stuck" else 1e-3 / (1 + ran)
        return dict(converged=bool(ran >= self._cost(j)), diverged=False, steps_run=min(ran, N), N=N,
                    stopping_step=self._cost(j) if conv else N, last_rel=hist, kind=P._kind(sp), group=sp.group,
                    case=sp.tag.get("id"), cd=float(ran), cl=0.0)

    def run(self, until):
        assert until > self.step
        self.log.append(("run", self.step, until))
        self.step = until
        return self.step

    def reports(self):
        return [self._report(j) for j in range(len(self.specs))]

    def reports_lazy(self):
        return self.reports()

    def replace_case(self, j, spec, W):
        assert P._kind(spec) == P._kind(self.specs[j]), "a slot only takes cases of its own kind"
        g = self.specs[j].group
        self.specs[j] = copy.copy(spec)
        self.specs[j].group = g
        self.start[j] = self.step
        self.init[j] = W
        self.log.append(("replace", j, spec.tag.get("id"), self.step))

    def compact(self):
        self.compacts += 1

    def close(self):
        pass


@pytest.fixture
def mock(monkeypatch):
    import torch
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)
    MockSolver.instances = []
    monkeypatch.setattr(P, "BatchSolver", MockSolver)
    monkeypatch.setattr(P, "AUTOTUNE", True)
    monkeypatch.setattr(P, "LAUNCH", None)
    return MockSolver


def _cases(kinds, per_kind, costs):
    meshes = [object() for _ in range(kinds)]
    specs, inits = [], []
    for i in range(kinds * per_kind):
        k = i % kinds
        sp = Spec(meshes[k], dict(flux="HLLC", reconstruction="MUSCL"), cost=costs(i, k))
        sp.tag = {"id": i, "cold": 2 * costs(i, k)}
        specs.append(sp)
        inits.append(f"init{i}")
    return specs, inits


def test_every_case_solved_once_in_its_kind(mock):
    specs, inits = _cases(3, 20, lambda i, k: 300 + 700 * k + 37 * (i % 5))
    reps, t, restarted = P.campaign(specs, inits, slots=12, chunk=100)
    assert restarted == []
    assert all(r is not None and r["converged"] for r in reps)
    for i, r in enumerate(reps):
        assert r["case"] == i and r["kind"] == P._kind(specs[i])
        assert r["stopping_step"] == specs[i].cost
    bs = mock.instances[0]
    # sub-groups: S slots of one kind and one launch-group tag
    S = P.subgroup_size(specs, 12)
    assert S == 4
    tags = {}
    for j, sp in enumerate(bs.specs):
        tags.setdefault(sp.group, set()).add(P._kind(sp))
    assert all(len(v) == 1 for v in tags.values())
    assert all(sum(1 for sp in bs.specs if sp.group == g) == S for g in tags)
    # runs end on chunk multiples after the one-step placeholder run
    runs = [x for x in bs.log if x[0] == "run"]
    assert runs[0][1:] == (0, 1)
    assert all(u % 100 == 0 for _, _, u in runs[1:])
    assert ("tune", 1) in bs.log


def test_fast_kinds_hand_their_slots_to_slow_kinds(mock):
    # kind 0 converges in 100 steps, kind 1 in 2000: once kind 0 has no
    # waiting case, its share goes to kind 1 and the slots stay busy
    specs, inits = _cases(2, 40, lambda i, k: 100 if k == 0 else 2000)
    reps, t, _ = P.campaign(specs, inits, slots=8, chunk=100)
    assert all(r["converged"] for r in reps)
    bs = mock.instances[0]
    # kind 1 starts with half of the 8 slots and ends up using more of them
    slot_kind1 = {x[1] for x in bs.log if x[0] == "replace" and specs[x[2]].cost == 2000}
    assert len(slot_kind1) > 4
    # 40 kind-1 cases x 2000 steps over about 8 slots; the kind-0 work is small
    total_steps = [x for x in bs.log if x[0] == "run"][-1][2]
    assert total_steps <= 40 * 2000 / 8 + 2 * 2000


def test_stalled_warm_start_restarts_from_fallback(mock):
    specs, inits = _cases(1, 6, lambda i, k: 1500)
    inits[2] = "stuck"
    fb = ["fallback"] * len(specs)
    reps, t, restarted = P.campaign(specs, inits, slots=3, chunk=100, fallback_inits=fb, stall=(1000, 500, 0.97))
    assert restarted == [2]
    assert all(r["converged"] for r in reps)
    assert reps[2]["stopping_step"] == specs[2].tag["cold"]


def test_chunk_must_follow_the_check_period(mock):
    specs, inits = _cases(1, 2, lambda i, k: 100)
    with pytest.raises(ValueError):
        P.campaign(specs, inits, slots=2, chunk=42)


def test_hedge_finishes_a_stuck_warm_start_once_nothing_waits(mock):
    # case 5 (the last to start) never converges from its warm start and the
    # stall rule is out of reach: once the queue is empty a free slot runs its
    # fallback beside it, and the fallback's result is the case's
    specs, inits = _cases(1, 6, lambda i, k: 1500)
    inits[5] = "stuck"
    fb = ["fallback"] * len(specs)
    reps, t, restarted = P.campaign(specs, inits, slots=3, chunk=100, fallback_inits=fb,
                                    stall=(10 ** 8, 500, 0.97))
    assert restarted == [5]
    assert all(r is not None and r["converged"] for r in reps)
    assert reps[5]["stopping_step"] == specs[5].tag["cold"]
    bs = mock.instances[0]
    hedges = [x for x in bs.log if x[0] == "replace" and x[2] == 5]
    assert len(hedges) >= 2                      # its warm start, then its hedge (and the retirement)
    # the hedge started only after the last queued case had been placed
    last_refill = max(x[3] for x in bs.log if x[0] == "replace" and x[2] != 5)
    first_hedge = [x[3] for x in hedges][1]
    assert first_hedge >= last_refill


def test_warm_start_that_converges_first_keeps_its_result(mock):
    # a slow but converging warm start beats its hedge (cold cost 2x)
    specs, inits = _cases(1, 4, lambda i, k: 1000 if i != 3 else 1800)
    fb = ["fallback"] * len(specs)
    reps, t, restarted = P.campaign(specs, inits, slots=2, chunk=100, fallback_inits=fb)
    assert restarted == []
    assert reps[3]["converged"] and reps[3]["stopping_step"] == 1800


def test_no_hedges_without_fallback_or_when_off(mock):
    specs, inits = _cases(1, 4, lambda i, k: 1000 if i != 3 else 1800)
    for kw in (dict(), dict(fallback_inits=["fallback"] * 4, hedge=False)):
        MockSolver.instances = []
        reps, t, restarted = P.campaign(specs, inits, slots=2, chunk=100, **kw)
        assert restarted == [] and all(r["converged"] for r in reps)
        bs = MockSolver.instances[0]
        assert not [x for x in bs.log if x[0] == "replace" and x[2] == 3 and x[3] > 1000]


def test_stall_with_a_running_hedge_hands_the_case_to_the_hedge(mock):
    specs, inits = _cases(1, 4, lambda i, k: 1000)
    inits[3] = "stuck"
    fb = ["fallback"] * len(specs)
    reps, t, restarted = P.campaign(specs, inits, slots=2, chunk=100, fallback_inits=fb,
                                    stall=(1500, 500, 0.97))
    assert restarted == [3]
    assert reps[3]["converged"] and reps[3]["stopping_step"] == specs[3].tag["cold"]
    bs = mock.instances[0]
    # the stuck warm start was not restarted in its own slot beside the hedge:
    # its case id was placed twice (warm start, hedge) plus one retirement
    placed = [x for x in bs.log if x[0] == "replace" and x[2] == 3]
    assert len(placed) == 3


def test_hedge_goes_to_the_case_longest_in_its_slot(mock):
    # cases 0 and 1 finish together while case 3 (the last in the queue)
    # takes one of their slots: the other free slot hedges case 2, which has
    # run since the start, not case 3, which has just entered
    costs = {0: 2000, 1: 2000, 2: 9000, 3: 9000}
    specs, inits = _cases(1, 4, lambda i, k: costs[i])
    fb = ["fallback"] * len(specs)
    P.campaign(specs, inits, slots=3, chunk=100, fallback_inits=fb)
    bs = mock.instances[0]
    later = [x for x in bs.log if x[0] == "replace" and x[3] >= 2000]
    assert later[0][2] == 3                  # the refill
    assert later[1][2] == 2                  # the first hedge


def test_steps_total_counts_the_discarded_warm_attempt(mock):
    # case 2 stalls from its warm start and restarts from its fallback in its
    # slot: its total is the warm attempt plus the fallback run, its stopping
    # step only the latter
    specs, inits = _cases(1, 6, lambda i, k: 1500)
    inits[2] = "stuck"
    fb = ["fallback"] * len(specs)
    reps, t, restarted = P.campaign(specs, inits, slots=3, chunk=100, fallback_inits=fb, stall=(1000, 500, 0.97))
    assert restarted == [2]
    bs = mock.instances[0]
    placed = [x[3] for x in bs.log if x[0] == "replace" and x[2] == 2]
    assert len(placed) == 1                       # in flight from step 0, restarted once
    warm_steps = placed[0]
    assert reps[2]["steps_total"] == reps[2]["steps_run"] + warm_steps
    assert reps[2]["steps_total"] >= specs[2].tag["cold"] + 1000
    # the abandoned warm attempt is on record, with its last in-loop CD
    assert len(reps[2]["discarded"]) == 1 and reps[2]["discarded"][0]["steps_run"] == warm_steps
    assert all(r["discarded"] == [] for i, r in enumerate(reps) if i != 2)
    # cases 0 and 1 finish while cases still wait (no hedges yet); every
    # total covers at least the run that gave the result
    assert all(reps[i]["steps_total"] == reps[i]["steps_run"] for i in (0, 1))
    assert all(r["steps_total"] >= r["steps_run"] for r in reps)


def test_steps_total_counts_a_losing_hedge(mock):
    # a slow warm start that converges before its hedge: its total includes
    # the hedge's steps
    specs, inits = _cases(1, 4, lambda i, k: 1000 if i != 3 else 1800)
    fb = ["fallback"] * len(specs)
    reps, t, restarted = P.campaign(specs, inits, slots=2, chunk=100, fallback_inits=fb)
    assert restarted == [] and reps[3]["stopping_step"] == 1800
    bs = mock.instances[0]
    hedge = [x for x in bs.log if x[0] == "replace" and x[2] == 3 and x[3] > 0]
    assert hedge, "a hedge ran for case 3"
    assert reps[3]["steps_total"] > reps[3]["steps_run"]


def test_cases_already_at_their_fallback_are_neither_hedged_nor_restarted(mock):
    # case 3 starts from its fallback (a non-physical prediction): a hedge or
    # a restart would replay the same trajectory
    specs, inits = _cases(1, 4, lambda i, k: 1000 if i != 3 else 1800)
    inits[3] = "stuck"
    fb = ["fallback"] * len(specs)
    reps, t, restarted = P.campaign(specs, inits, slots=2, chunk=100, fallback_inits=fb, cold_ids=[3],
                                    stall=(1500, 500, 0.97), max_steps=4000)
    assert restarted == []
    bs = mock.instances[0]
    assert not [x for x in bs.log if x[0] == "replace" and x[2] == 3 and x[3] > 1000]
    assert not reps[3]["converged"] and reps[3]["steps_total"] == reps[3]["steps_run"]


def test_lazy_reports_convert_each_slot_once_on_demand():
    import ctypes as C
    from eulersr_amd.solve import _LazyReports
    calls = []
    arr = (C.c_int * 600)(*range(600))
    reps = _LazyReports(arr, lambda x: calls.append(x) or {"v": x})
    assert len(reps) == 600 and calls == []
    assert reps[17] == {"v": 17} and reps[17] is reps[17] and calls == [17]
    with pytest.raises(IndexError):
        reps[600]


def test_stall_watch_equals_the_stall_rule_on_random_histories():
    import random
    rng = random.Random(5)
    for trial in range(300):
        w, hist, steps = P.StallWatch(), [], 0
        after, window = rng.choice([(0, 300), (6000, 3000), (1000, 500)])
        for _ in range(rng.randint(0, 120)):
            steps += rng.choice([5, 100, 100, 200])
            # residuals falling, flat or noisy, with repeats
            r = rng.choice([1e-3 / (1 + steps), 1e-4, rng.uniform(1e-6, 1e-3), 3e-5])
            hist.append((steps, r))
            w.add(steps, r)
            assert w.stalled(steps, after, window, 0.97) == P.stalled(hist, steps, after, window, 0.97)


def test_steps_key_orders_cases_by_their_step_count():
    """On one mesh, fixed_time's packing key orders freestream starts as
    their N = int(tf / dt) + 1 does, with dt from the oracle's get_dt."""
    import numpy as np
    from conftest import base_cfg, mesh_of
    from eulersr_amd import main as M
    from oracle import fvm as F
    mesh = mesh_of("naca0012", 0.2)
    om = F.omesh(mesh, np.float64)
    keys, ns = [], []
    for mach, aoa in ((0.7, 0.0), (0.85, 3.0), (1.05, -2.0), (1.2, 5.0), (1.8, 1.0), (2.4, -4.0), (3.0, 2.0)):
        cfg = base_cfg(Mach=mach, aoa=aoa, case="naca0012", tf=1.0)
        W, inlet = M.initialize(mesh, cfg)
        sp = M.case_spec(mesh, inlet, cfg)
        dt = F.get_dt(np.asarray(W, np.float64), om, CFL=sp.cfl)
        keys.append(P._steps_key(sp))
        ns.append(int(sp.tf / dt) + 1)
    assert np.argsort(keys).tolist() == np.argsort(ns).tolist(), (keys, ns)
