"""Grid driver and warm-start harness on one GPU.

* ``fixed_time``: the dataset runs of slurm/run_euler_grid.py:25-104 -- the
  euler/main.py defaults, tf = 5 for the diamond and 3 otherwise
  (eval/live_fvm.py:21-22), no early stop -- for a whole (Mach, alpha) grid in
  one batched solver instead of one process per case.
* ``training_pairs`` / ``train``: LR/HR pairs generated by the solver itself
  and the warm-start network trained on them (warmstart.WarmNet).
* ``WarmStarter.run``: the protocol of eval/convergence_case.py:30-142 --
  for each case the residual threshold comes from the HR reference field
  (x 1.001, see ``with_targets``), CD/CL targets are the reference field's,
  stationarity is checked every 5 steps with patience 3
  (eval/live_fvm.py:37-47) -- started cold (freestream), from the IDW
  prediction, or from the network's prediction (LR solve -> features ->
  network -> HR state), with the freestream fallback for a prediction with
  rho <= 0 or p <= 0, a solve that diverges, or a solve that has not reached
  its threshold within the warm-start budget (15,000 steps: with the K = 0
  limiter a warm start occasionally settles on a limit cycle whose residual
  stays above the threshold set by the cold trajectory). All cases of a call,
  of any geometry, run in one batched solver.
"""
from __future__ import annotations

import copy
import math
import time
from dataclasses import dataclass

import numpy as np

from . import main as M
from . import warmstart as WS
from .solve import BatchSolver

REF_TF = {"diamond": 5.0}        # eval/live_fvm.py:21-22
REF_TF_DEFAULT = 3.0
CHECK_EVERY, PATIENCE = 5, 3     # eval/live_fvm.py:37-47
THRESHOLD_MARGIN = 1.001         # eval/convergence_case.py:61-66
ENGINEERING_TOL = 0.001          # eval/live_fvm.py:47
AUTOTUNE = True                  # solvers pick their stage path / block size (esr_solver_autotune)
WARM_BUDGET = 15000              # steps a warm start gets before its freestream fallback
LAUNCH = None                    # or a pinned (path, block, xcd_order) for every solver


def ref_tf(geometry: str) -> float:
    return REF_TF.get(geometry, REF_TF_DEFAULT)


def case_config(geometry, mach, aoa, **over):
    """euler/main.py defaults for one (Mach, alpha) case."""
    cfg = dict(rho_inf=1.0, p_inf=1.0, gamma=1.4, Mach=float(mach), aoa=float(aoa), case=geometry,
               time_scheme="SRK2", flux="HLLC", reconstruction="MUSCL", CFL=0.6, tf=ref_tf(geometry))
    cfg.update(over)
    return cfg


def make_cases(mesh, geometry, grid, **over):
    specs, inits = [], []
    for mach, aoa in grid:
        cfg = case_config(geometry, mach, aoa, **over)
        W, inlet = M.initialize(mesh, cfg)
        sp = M.case_spec(mesh, inlet, cfg)
        sp.tag = dict(geometry=geometry, mach=float(mach), aoa=float(aoa), h=mesh.metadata.get("h"))
        specs.append(sp)
        inits.append(W)
    return specs, inits


PACK_SIZE, PACK_CHUNK = 16, 250   # fixed_time(pack=True): cases per sub-group, steps between compactions


def _steps_key(sp):
    """tf x (|u| + c) / CFL of a case's inlet state: on one mesh a freestream
    start's dt is CFL x (a mesh quantity) / (|u| + c), so this orders its N."""
    kw = sp.kw
    g, m = float(kw.get("gamma", 1.4)), float(kw.get("M", 1.0))
    r, ru, rv, E = (float(x) for x in kw["value"])
    u, v = ru / r, rv / r
    P = (E - 0.5 * r * (u * u + v * v)) * m * m * (g - 1.0)
    c = math.sqrt(abs(g * P / r)) / m
    return sp.tf * (math.hypot(u, v) + c) / sp.cfl


def fixed_time(specs, inits, max_steps=0, pack=False):
    """Runs every case to its N = int(tf/dt)+1 (no early stop). Returns the
    reports (probe residual, CD, CL) and the final states (CUDA tensors).

    pack: the cases of each kind run in launch sub-groups of PACK_SIZE taken
    in the order of their N (`_steps_key`), and every PACK_CHUNK steps the
    sub-groups whose cases are all done drop out of the launches
    (esr_solver_compact); otherwise every case is computed (its stores
    skipped once done) until the last one ends. A case's state does not
    depend on its group; the FP64 residuals of the reports do at the 1e-12
    level (their partial sums are laid out by the group's case count), so
    solves whose residuals set thresholds keep pack off."""
    if pack:
        specs = [copy.copy(sp) for sp in specs]
        by_kind = {}
        for i, sp in enumerate(specs):
            by_kind.setdefault(_kind(sp), []).append(i)
        tag = 0
        for ids in by_kind.values():
            ids.sort(key=lambda i: _steps_key(specs[i]))
            for j in range(0, len(ids), PACK_SIZE):
                tag += 1
                for i in ids[j:j + PACK_SIZE]:
                    specs[i].group = tag
    bs = BatchSolver(specs, scheme="SRK2", threshold=0.0, check_every=CHECK_EVERY, patience=PATIENCE,
                     max_steps=max_steps, autotune=AUTOTUNE, launch=LAUNCH)
    try:
        for i, W in enumerate(inits):
            bs.set_state(i, W)
        bs.prepare()
        if pack:
            step = 0
            while True:
                nxt = bs.run(step + PACK_CHUNK)
                if bs.compact() == 0 or nxt == step:
                    break
                step = nxt
        else:
            bs.run()
        bs.finalize()
        reps = [bs.report(i) for i in range(len(specs))]
        states = [bs.get_state(i) for i in range(len(specs))]
    finally:
        bs.close()
    return reps, states


def converge(specs, inits, max_steps=0, sync=True):
    """Runs every case until its own stationarity criterion (spec.threshold,
    check every 5, patience 3) or N; returns (reports, states, solve seconds)."""
    import torch
    bs = BatchSolver(specs, scheme="SRK2", threshold=1.0, check_every=CHECK_EVERY, patience=PATIENCE,
                     max_steps=max_steps, autotune=AUTOTUNE, launch=LAUNCH)
    try:
        for i, W in enumerate(inits):
            bs.set_state(i, W)
        bs.prepare()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bs.run()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        bs.finalize()
        reps = [bs.report(i) for i in range(len(specs))]
        states = [bs.get_state(i) for i in range(len(specs))]
    finally:
        bs.close()
    return reps, states, t


def with_targets(specs, ref_reports, tf=1.0e4):
    """Copies of the specs carrying the reference protocol: engineering
    targets = reference CD/CL (0.1 %), tf large enough that stationarity, not
    N, ends the solve, and threshold = 1.001 x the reference field's residual
    (eval/convergence_case.py:61-66). The reference takes that residual from a
    one-step probe and expects the reference field to pass its own check at
    once; with the limiter's residual plateau (a limit cycle whose check
    residual wanders within a band) a one-step probe or a single check can
    sit at the bottom of that band, so the residual used here is the larger of
    the probe and the reference run's largest residual over its last 16
    checks (80 steps)."""
    out = []
    for sp, r in zip(specs, ref_reports):
        s = copy.copy(sp)
        s.tf = float(tf)
        base = r["final_residual"]
        rm = r.get("recent_rel_max", float("nan"))
        if rm == rm:
            base = max(base, rm)
        s.threshold = base * THRESHOLD_MARGIN
        s.engineering = (r["cd"], r["cl"], ENGINEERING_TOL * max(abs(r["cd"]), 1e-8),
                         ENGINEERING_TOL * max(abs(r["cl"]), 1e-8), PATIENCE)
        out.append(s)
    return out


def training_pairs(geometry, lr_mesh, hr_mesh, grid, max_steps=0):
    """Fixed-time LR and HR solves of every (Mach, alpha) in `grid` (one
    batched solver for both resolutions). Returns a dict of CUDA tensors
    lr (n, n_lr, 4), hr (n, n_hr, 4) and the grid."""
    import torch
    ls, li = make_cases(lr_mesh, geometry, grid)
    hs, hi = make_cases(hr_mesh, geometry, grid)
    reps, states = fixed_time(ls + hs, li + hi, max_steps=max_steps, pack=True)
    n = len(grid)
    return dict(lr=torch.stack(states[:n]), hr=torch.stack(states[n:]), lr_reports=reps[:n], hr_reports=reps[n:],
                machs=np.array([g[0] for g in grid]), aoas=np.array([g[1] for g in grid]))


def primitive_stats(hr_states, gamma=1.4):
    """Mean and standard deviation of the HR primitives over all cells of all
    cases (the role of the dataset stats mu/sig, preprocessing/dataset.py)."""
    P = WS.primitives(hr_states.reshape(-1, 4).cpu().numpy(), gamma)
    return P.mean(0).astype(np.float32), np.maximum(P.std(0), 1e-6).astype(np.float32)


def targets(kmap, pairs, mu, sig, gamma=1.4):
    """Network inputs and targets T = (prim_HR - IDW) / sig on the GPU."""
    import torch
    X, idw = WS.features(kmap, pairs["lr"], pairs["machs"], pairs["aoas"], mu, sig, gamma)
    P = WS.primitives(pairs["hr"].reshape(-1, 4).cpu().numpy(), gamma)
    T = ((P - idw.cpu().numpy().astype(np.float64)) / sig.astype(np.float64)).astype(np.float32)
    return X, idw, torch.from_numpy(T).to(X.device)


def train(kmap, pairs, steps=3000, batch=16384, lr=1e-3, weight_decay=1e-5, seed=0, gamma=1.4, log_every=0):
    """Trains a WarmNet on solver-generated pairs; returns (net, info)."""
    import torch
    mu, sig = primitive_stats(pairs["hr"], gamma)
    X, idw, T = targets(kmap, pairs, mu, sig, gamma)
    net = WS.WarmNet(in_dim=X.shape[1], max_rows=max(batch, 1 << 15), seed=seed)
    net.mu, net.sig = mu, sig
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = net.fit(X, T, steps=steps, batch=batch, lr=lr, weight_decay=weight_decay, seed=seed,
                   log_every=log_every)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    idw_mse = float(torch.mean(T.double() ** 2))
    return net, dict(train_s=t, rows=int(X.shape[0]), steps=steps, batch=batch, loss_first=hist[0],
                     loss_last=float(np.mean(hist[-20:])), idw_mse=idw_mse)


def _kind(sp):
    """The solver's grouping key of a case: cases of one kind can take each
    other's slot."""
    kw = sp.kw
    return (id(sp.mesh), str(kw.get("flux", kw.get("numerical_flux", ""))).lower(),
            str(kw.get("reconstruction", kw.get("scheme", ""))).lower(), bool(kw.get("entropy", False)),
            float(kw.get("gamma", 1.4)), float(kw.get("M", 1.0)), kw.get("sponge_width"))


STALL = (6000, 3000, 0.97)   # (steps before the test, window, required improvement factor)


def stalled(hist, steps, after, window, factor):
    """True when a solve past `after` steps has not lowered its best check
    residual below factor x the best one before the last `window` steps."""
    if steps < after:
        return False
    before = [r for s, r in hist if s <= steps - window]
    recent = [r for s, r in hist if s > steps - window]
    return bool(before) and bool(recent) and min(recent) > factor * min(before)


class StallWatch:
    """`stalled` for one solve whose check residuals arrive in step order,
    kept incrementally: the checks at or before steps - window are a prefix
    that only grows (its minimum kept as it grows), the later ones a window
    whose minimum a monotonic queue holds, so each check costs O(1)
    amortised instead of a pass over the whole history."""
    __slots__ = ("hist", "p", "before_min", "q")

    def __init__(self):
        from collections import deque
        self.hist = []                 # (steps, residual), steps non-decreasing
        self.p = 0                     # hist[:p] lie at or before the window
        self.before_min = math.inf
        self.q = deque()               # indices >= p, residuals increasing: q[0] is the window's minimum

    def add(self, steps, rel):
        h, q = self.hist, self.q
        h.append((steps, rel))
        while q and h[q[-1]][1] >= rel:
            q.pop()
        q.append(len(h) - 1)

    def stalled(self, steps, after, window, factor):
        if steps < after:
            return False
        h, q = self.hist, self.q
        cut = steps - window
        while self.p < len(h) and h[self.p][0] <= cut:
            self.before_min = min(self.before_min, h[self.p][1])
            if q and q[0] == self.p:
                q.popleft()
            self.p += 1
        return self.p > 0 and self.p < len(h) and h[q[0]][1] > factor * self.before_min


def _queues(specs):
    from collections import OrderedDict, deque
    queues = OrderedDict()
    for i, sp in enumerate(specs):
        queues.setdefault(_kind(sp), deque()).append(i)
    return queues


SUBGROUP_MAX = 16     # cases per launch sub-group of the campaign scheduler, at most
LANE_FACTOR = 1.25    # launched slots (live sub-groups x their size) <= LANE_FACTOR x slots


def subgroup_size(specs, slots):
    """Cases per launch sub-group of the campaign scheduler: the open slots
    split over the kinds (at least 1, at most SUBGROUP_MAX)."""
    return max(1, min(SUBGROUP_MAX, int(slots) // max(1, len(_queues(specs)))))


def initial_slots(specs, slots):
    """Per-kind queues of case ids and the cases in flight at the start of a
    campaign: `slots` // S sub-groups of S cases (subgroup_size) opened over
    the kinds in proportion to their case counts (at least one each), cases
    in input order."""
    queues = _queues(specs)
    n = len(specs)
    S = subgroup_size(specs, slots)
    n_open = max(1, int(slots) // S)
    share = {k: max(1, int(round(n_open * len(queues[k]) / n))) for k in queues}
    while sum(share.values()) > max(n_open, len(queues)) and any(v > 1 for v in share.values()):
        k = max(share, key=lambda q: share[q])
        share[k] -= 1
    first = []
    for k in queues:
        for _ in range(min(share[k] * S, len(queues[k]))):
            first.append(queues[k].popleft())
    return queues, first


def campaign(specs, inits, slots=128, max_steps=0, chunk=100, fallback_inits=None, stall=STALL, trace=None,
             hedge=True, cold_ids=()):
    """Continuous batching: solves every case of `specs` (initial states
    `inits`) to its own stationarity criterion with about `slots` cases in
    flight; whenever a case is done (converged, diverged or at its step cap
    `max_steps`), the next waiting case of the same kind (mesh and kernel
    configuration) takes its slot (esr_solver_replace_case).

    Slots live in launch sub-groups of S cases of one kind (subgroup_size;
    esr_case.group keeps them apart). Each kind holds enough sub-groups to
    take every open slot, but only `slots` // S of them start open; a
    sub-group whose cases are all done drops out of every launch
    (esr_solver_compact), and while room is left (cases in flight and
    launched slots, LANE_FACTOR) a whole idle sub-group of the kind with the
    most waiting cases is opened. Kinds whose cases converge fast therefore
    hand their share of the GPU to the slower kinds instead of leaving it
    idle.

    fallback_inits (warm starts): a case whose solve diverges, or stalls (its
    best check residual no longer falls, `stalled`) above its threshold, is
    restarted in its slot from fallback_inits[i] (freestream). With the K = 0
    limiter a warm start can settle on a limit cycle whose residual stays
    above the threshold set by the cold trajectory.

    hedge (warm starts): once no case waits in any queue, a slot that falls
    free in a sub-group still launched runs the fallback of a warm-started
    case of its kind in flight (the one that has run longest, not yet
    restarted or hedged), beside the warm start. The first of the two to
    converge is the case's result (a hedge that wins counts as a restart),
    and the other slot is retired; a warm start that stalls or diverges
    while its hedge runs is retired in favour of the hedge. The free lanes of
    the campaign's tail would otherwise idle in launches that run anyway.

    cold_ids: cases whose initial state already is their fallback (a
    non-physical prediction): they are neither restarted nor hedged, since
    either would replay the same trajectory.

    trace: a list that receives (seconds, step, cases in flight, cases
    waiting) after every chunk.

    Each report carries ``steps_total``: the steps the case took in every
    slot it ran in -- a warm attempt that was restarted or lost to its hedge,
    a hedge that lost, and the run that gave the result -- so step counts and
    cell-updates of warm starts include the work they discarded; and
    ``discarded``: the last in-loop CD, CL, check residual and steps of each
    warm attempt it gave up (stall, divergence, or a hedge that converged
    first).

    Returns (reports in input order, seconds of solving, restarted case ids)."""
    import math
    import torch
    if chunk % CHECK_EVERY:
        raise ValueError(f"chunk must be a multiple of the check period {CHECK_EVERY}")
    specs = [copy.copy(sp) for sp in specs]
    kinds = [_kind(sp) for sp in specs]
    if max_steps:
        for sp in specs:
            sp.max_steps = int(max_steps)
    n = len(specs)
    S = subgroup_size(specs, slots)
    n_open = max(1, int(slots) // S)
    queues, first = initial_slots(specs, slots)
    # sub-groups: (kind, slot ids); placeholders (one step, then done) fill
    # the slots of sub-groups that start idle
    start = {}
    for ci in first:
        start.setdefault(kinds[ci], []).append(ci)
    sub_kind, sub_slots, slot_case, slot_spec, slot_init = [], [], [], [], []
    tag = 0
    for k in list(queues.keys()):
        kind_cases = start.get(k, []) + list(queues[k])
        n_sub = min(math.ceil(len(kind_cases) / S), n_open)
        mine = start.get(k, [])
        proto = kind_cases[0]
        for j in range(n_sub):
            tag += 1
            ids = []
            for _ in range(S):
                sid = len(slot_case)
                if mine:
                    ci = mine.pop(0)
                    sp = copy.copy(specs[ci])
                    init = inits[ci]
                else:
                    ci = None
                    sp = copy.copy(specs[proto])
                    sp.max_steps = 1
                    init = inits[proto]
                sp.group = tag
                slot_case.append(ci)
                slot_spec.append(sp)
                slot_init.append(init)
                ids.append(sid)
            sub_kind.append(k)
            sub_slots.append(ids)
    bs = BatchSolver(slot_spec, scheme="SRK2", threshold=1.0, check_every=CHECK_EVERY, patience=PATIENCE,
                     autotune=False, launch=LAUNCH)
    results = [None] * n
    try:
        for j, init in enumerate(slot_init):
            bs.set_state(j, init)
        bs.prepare()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        # retire the placeholders (one step) and tune the launch of the
        # sub-groups in flight
        step = bs.run(1)
        bs.compact()
        if AUTOTUNE and LAUNCH is None:
            bs.tune()
        hist = {ci: StallWatch() for ci in first}
        placed = {ci: step for ci in first}    # step at which each case entered its slot
        restarted = []
        was_restarted = set()                  # the ids of `restarted`, for membership tests
        cold = set(cold_ids)
        spent = {}                             # case id -> steps of its discarded runs
        dropped = {}                           # case id -> last reports of its discarded warm attempts

        def discard(ci, r):
            dropped.setdefault(ci, []).append(dict(cd=r["cd"], cl=r["cl"], last_rel=r["last_rel"],
                                                   steps_run=r["steps_run"], diverged=r["diverged"]))
        changed = False
        slot_hedge = [None] * len(slot_case)   # case id whose fallback runs in the slot
        hedge_of = {}                          # case id -> slot of its hedge

        def retire(j, ci):
            # one step of a placeholder (then done) ends whatever ran in slot j
            sp = copy.copy(specs[ci])
            sp.max_steps = 1
            bs.replace_case(j, sp, fallback_inits[ci])

        def busy(j):
            # a case, or the hedge of one, runs in slot j
            return slot_case[j] is not None or slot_hedge[j] is not None

        while any(ci is not None for ci in slot_case) or any(queues.values()):
            # runs end on multiples of `chunk` (a multiple of the check
            # period), so replaced cases keep the check cadence of a solo run
            step = bs.run((step // chunk + 1) * chunk)
            reps = bs.reports_lazy()
            for j, ci in enumerate(slot_hedge):
                if ci is None:
                    continue
                r = reps[j]
                p = slot_case.index(ci)
                if r["converged"] and not reps[p]["converged"]:
                    # the hedge wins: the warm start's slot is retired
                    retire(p, ci)
                    slot_case[p] = None
                    spent[ci] = spent.get(ci, 0) + reps[p]["steps_run"]
                    discard(ci, reps[p])
                    results[ci] = dict(r, steps_total=spent[ci] + r["steps_run"], discarded=dropped.get(ci, []))
                    restarted.append(ci)
                    was_restarted.add(ci)
                    slot_hedge[j] = None
                    del hedge_of[ci]
                    changed = True
                elif r["converged"] or r["diverged"] or r["steps_run"] >= r["N"]:
                    # done without winning (a finished copy stops computing):
                    # the warm start converged in the same chunk and is kept,
                    # or the hedge failed and the warm start goes on
                    slot_hedge[j] = None
                    del hedge_of[ci]
                    spent[ci] = spent.get(ci, 0) + r["steps_run"]
            for j, ci in enumerate(slot_case):
                if ci is None:
                    continue
                r = reps[j]
                done = r["converged"] or r["diverged"] or r["steps_run"] >= r["N"]
                if fallback_inits is not None and ci not in was_restarted and ci not in cold and not r["converged"]:
                    if r["last_rel"] == r["last_rel"] and r["last_rel"] != float("inf"):
                        hist[ci].add(r["steps_run"], r["last_rel"])
                    if r["diverged"] or (not done and hist[ci].stalled(r["steps_run"], *stall)):
                        spent[ci] = spent.get(ci, 0) + r["steps_run"]
                        discard(ci, r)
                        if ci in hedge_of:
                            # its fallback already runs: that slot carries the case on
                            h = hedge_of.pop(ci)
                            retire(j, ci)
                            slot_case[j] = None
                            slot_hedge[h] = None
                            slot_case[h] = ci
                            changed = True
                        else:
                            bs.replace_case(j, specs[ci], fallback_inits[ci])
                        restarted.append(ci)
                        was_restarted.add(ci)
                        continue
                if not done:
                    continue
                if ci in hedge_of:
                    h = hedge_of.pop(ci)
                    retire(h, ci)
                    slot_hedge[h] = None
                    spent[ci] = spent.get(ci, 0) + reps[h]["steps_run"]
                results[ci] = dict(r, steps_total=spent.get(ci, 0) + r["steps_run"], discarded=dropped.get(ci, []))
                q = queues[kinds[ci]]
                if q:
                    nxt = q.popleft()
                    bs.replace_case(j, specs[nxt], inits[nxt])
                    slot_case[j] = nxt
                    hist[nxt] = StallWatch()
                    placed[nxt] = step
                else:
                    slot_case[j] = None
                    changed = True
            # open idle sub-groups of the kinds with the most waiting cases
            # (a live sub-group costs all of its slots in every launch, so the
            # launched slots are bounded as well as the cases in flight)
            # (one pass over the slots: cases in flight, sub-groups in use)
            active = 0
            sub_busy = []
            for sl in sub_slots:
                b = False
                for j in sl:
                    if slot_case[j] is not None:
                        active += 1
                        b = True
                    elif slot_hedge[j] is not None:
                        b = True
                sub_busy.append(b)
            lanes = S * sum(sub_busy)
            while active <= slots - S and lanes + S <= LANE_FACTOR * slots:
                waiting = [(len(queues[sub_kind[g]]), g) for g in range(len(sub_kind))
                           if queues[sub_kind[g]] and not sub_busy[g]]
                if not waiting:
                    break
                g = max(waiting, key=lambda x: (x[0], -x[1]))[1]
                q = queues[sub_kind[g]]
                for j in sub_slots[g]:
                    if not q:
                        break
                    nxt = q.popleft()
                    bs.replace_case(j, specs[nxt], inits[nxt])
                    slot_case[j] = nxt
                    hist[nxt] = StallWatch()
                    placed[nxt] = step
                    active += 1
                sub_busy[g] = True
                lanes += S
                changed = True
            # hedges: free slots of launched sub-groups, once nothing waits
            if hedge and fallback_inits is not None and not any(queues.values()):
                for g, sl in enumerate(sub_slots):
                    if not any(slot_case[j] is not None for j in sl):
                        continue
                    for j in sl:
                        if busy(j):
                            continue
                        cand = [(step - placed[ci], ci) for ci in slot_case
                                if ci is not None and ci not in was_restarted and ci not in hedge_of
                                and ci not in cold and kinds[ci] == sub_kind[g]]
                        if not cand:
                            break
                        ci = max(cand)[1]
                        bs.replace_case(j, specs[ci], fallback_inits[ci])
                        slot_hedge[j] = ci
                        hedge_of[ci] = j
            if changed:
                bs.compact()        # sub-groups with no case left stop launching
                changed = False
            if trace is not None:
                trace.append((time.perf_counter() - t0, step, active, sum(len(q) for q in queues.values())))
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
    finally:
        bs.close()
    return results, t, restarted


@dataclass
class WarmCase:
    """One (geometry, Mach, alpha) case of the warm-start harness: its LR
    fixed-time case, its HR case (threshold and targets set) and the HR
    freestream state (cold start and fallback)."""
    geometry: str
    mach: float
    aoa: float
    lr_spec: object
    lr_init: object
    hr_spec: object
    hr_init: object


def warm_cases(geometry, lr_mesh, hr_mesh, grid, ref_reports):
    """WarmCase list for `grid` on one mesh pair; ref_reports are the HR
    reference fields' reports (threshold and CD/CL targets)."""
    ls, li = make_cases(lr_mesh, geometry, grid)
    hs, hi = make_cases(hr_mesh, geometry, grid)
    hs = with_targets(hs, ref_reports)
    return [WarmCase(geometry, float(m), float(a), ls[i], li[i], hs[i], hi[i]) for i, (m, a) in enumerate(grid)]


class WarmStarter:
    """LR solve -> features -> network -> HR warm start -> HR solve, batched
    over cases of any number of geometries (one mesh pair and network each)."""

    def __init__(self):
        self.models = {}

    def add(self, geometry, lr_mesh, hr_mesh, net=None, kmap=None):
        self.models[geometry] = (kmap or WS.KnnMap(lr_mesh, hr_mesh), net)
        return self

    def predict(self, geometry, lr_states, grid, use_net=True, gamma=1.4):
        """HR conserved states (n, n_hr, 4) and per-case fallback flags."""
        import torch
        kmap, net = self.models[geometry]
        use = use_net and net is not None
        mu, sig = (net.mu, net.sig) if use else (np.zeros(4, np.float32), np.ones(4, np.float32))
        X, idw = WS.features(kmap, lr_states, [g[0] for g in grid], [g[1] for g in grid], mu, sig, gamma)
        Y = net.forward(X) if use else None
        W, bad = WS.assemble(kmap.n_hr, idw, Y, sig, gamma)
        torch.cuda.synchronize()
        return W.reshape(len(grid), kmap.n_hr, 4), bad.cpu().numpy().astype(bool)

    def initial_states(self, cases, mode="mlp", max_steps=0):
        """HR initial states of `cases` for a start mode: freestream ("cold"),
        or LR fixed-time solves (one batch) -> features -> IDW ("idw") or
        network ("mlp"), with the freestream fallback for non-physical
        predictions. Returns (states, fallback flags, LR seconds, inference
        seconds)."""
        import torch
        n = len(cases)
        inits = [c.hr_init for c in cases]
        fallback = [False] * n
        t_lr = t_inf = 0.0
        if mode != "cold":
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, lr_states = fixed_time([c.lr_spec for c in cases], [c.lr_init for c in cases], max_steps=max_steps,
                                      pack=True)
            torch.cuda.synchronize()
            t_lr = time.perf_counter() - t0
            t0 = time.perf_counter()
            for g in dict.fromkeys(c.geometry for c in cases):
                ids = [i for i, c in enumerate(cases) if c.geometry == g]
                W, bad = self.predict(g, torch.stack([lr_states[i] for i in ids]),
                                      [(cases[i].mach, cases[i].aoa) for i in ids], use_net=(mode == "mlp"))
                for j, i in enumerate(ids):
                    fallback[i] = bool(bad[j])
                    if not bad[j]:
                        inits[i] = W[j]
            t_inf = time.perf_counter() - t0
        return inits, fallback, t_lr, t_inf

    def campaign(self, cases, mode="mlp", slots=128, max_steps=0, lr_max_steps=0, trace=None):
        """The warm-start protocol over a campaign of cases with continuous
        batching (`campaign`): LR solves and predictions for every case, then
        the HR solves, `slots` at a time. A warm-started case that diverges
        is solved again from freestream. Returns a summary like `run`;
        `trace` receives the HR campaign's (seconds, step, cases in flight,
        cases waiting) after every chunk."""
        import torch
        torch.cuda.synchronize()
        t_all = time.perf_counter()
        inits, fallback, t_lr, t_inf = self.initial_states(cases, mode, lr_max_steps)
        fb = None if mode == "cold" else [c.hr_init for c in cases]
        reps, t, restarted = campaign([c.hr_spec for c in cases], inits, slots=slots, max_steps=max_steps,
                                      fallback_inits=fb, trace=trace,
                                      cold_ids=[i for i, f in enumerate(fallback) if f])
        for i in restarted:
            fallback[i] = True
        torch.cuda.synchronize()
        return dict(mode=mode, lr_s=t_lr, infer_s=t_inf, hr_s=t, total_s=time.perf_counter() - t_all,
                    fallback=fallback, retried=restarted, reports=reps)

    def run(self, cases, mode="mlp", max_steps=0, warm_budget=WARM_BUDGET):
        """One batched pass. mode: "cold" (freestream), "idw" or "mlp".
        A warm-started case falls back to freestream (a second batch of such
        cases, up to max_steps) when its prediction is non-physical, its solve
        diverges, or it has not reached its threshold within `warm_budget`
        steps (0: no budget). Returns per-case reports and the wall-time split
        (LR solve, inference, HR solve, total)."""
        import torch
        n = len(cases)
        out = dict(mode=mode, lr_s=0.0, infer_s=0.0, hr_s=0.0, fallback=[False] * n, retried=[])
        torch.cuda.synchronize()
        t_all = time.perf_counter()
        inits = [c.hr_init for c in cases]
        if mode != "cold":
            t0 = time.perf_counter()
            _, lr_states = fixed_time([c.lr_spec for c in cases], [c.lr_init for c in cases], max_steps=max_steps,
                                      pack=True)
            torch.cuda.synchronize()
            out["lr_s"] = time.perf_counter() - t0
            t0 = time.perf_counter()
            for g in dict.fromkeys(c.geometry for c in cases):
                ids = [i for i, c in enumerate(cases) if c.geometry == g]
                W, bad = self.predict(g, torch.stack([lr_states[i] for i in ids]),
                                      [(cases[i].mach, cases[i].aoa) for i in ids], use_net=(mode == "mlp"))
                for j, i in enumerate(ids):
                    out["fallback"][i] = bool(bad[j])
                    if not bad[j]:
                        inits[i] = W[j]
            out["infer_s"] = time.perf_counter() - t0
        warm_cap = max_steps
        if mode != "cold" and warm_budget:
            warm_cap = min(max_steps, warm_budget) if max_steps else warm_budget
        reps, states, t = converge([c.hr_spec for c in cases], inits, max_steps=warm_cap)
        out["hr_s"] = t
        # a warm start whose solve diverges, or that has not converged within
        # the budget, is re-run from freestream
        div = [i for i, r in enumerate(reps)
               if mode != "cold" and not out["fallback"][i] and (r["diverged"] or not r["converged"])]
        if div:
            r2, s2, t2 = converge([cases[i].hr_spec for i in div], [cases[i].hr_init for i in div],
                                  max_steps=max_steps)
            out["hr_s"] += t2
            for j, i in enumerate(div):
                reps[i], states[i] = r2[j], s2[j]
                out["fallback"][i] = True
            out["retried"] = div
        torch.cuda.synchronize()
        out["total_s"] = time.perf_counter() - t_all
        out["reports"] = reps
        out["states"] = states
        return out
