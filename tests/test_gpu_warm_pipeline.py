"""End-to-end learned warm start on a small diamond pair (LR h = 0.2, HR
h = 0.05): solver-generated LR/HR pairs, network training on the GPU, and
the warm-start protocol (threshold = HR probe residual x 1.001, patience 3)
for held-out (Mach, alpha) cases. The network's prediction must beat the
IDW baseline on the HR field, and warm-started solves must converge to the
same threshold with drag within 1e-4 of the cold-started solve's, plus the
amount by which the cold solve's own drag still moves after it stops."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TRAIN = [(m, a) for m in (0.8, 1.3, 1.7, 2.2, 2.7) for a in (-3.0, 0.0, 3.0)]
TEST = [(1.1, 2.0), (2.0, 2.0)]


@pytest.fixture(scope="module")
def setup():
    import torch
    from eulersr_amd import pipeline as P
    from eulersr_amd import warmstart as WS
    from eulersr_amd.device import require_gpu
    from eulersr_amd.meshing import build_mesh
    require_gpu()
    lr, hr = build_mesh("diamond", 0.2), build_mesh("diamond", 0.05)
    km = WS.KnnMap(lr, hr)
    print(f"pipeline: LR {lr.n_cells} cells, HR {hr.n_cells} cells; {len(TRAIN)} training cases", flush=True)
    pairs = P.training_pairs("diamond", lr, hr, TRAIN)
    print("pipeline: training pairs done", flush=True)
    net, info = P.train(km, pairs, steps=600, batch=8192, lr=1e-3, seed=0)
    print(f"pipeline: trained {info}", flush=True)
    test = P.training_pairs("diamond", lr, hr, TEST)
    torch.cuda.synchronize()
    return dict(lr=lr, hr=hr, km=km, net=net, info=info, test=test)


def _field_error(W, ref, sig):
    from eulersr_amd.warmstart import primitives
    a = primitives(W.reshape(-1, 4).cpu().numpy())
    b = primitives(ref.reshape(-1, 4).cpu().numpy())
    return float(np.sqrt(np.mean(((a - b) / sig) ** 2)))


def test_training_converges(setup):
    info = setup["info"]
    assert info["loss_last"] < 0.5 * info["idw_mse"], info


def test_network_beats_idw(setup):
    from eulersr_amd.pipeline import WarmStarter
    pipe = WarmStarter().add("diamond", setup["lr"], setup["hr"], net=setup["net"], kmap=setup["km"])
    t = setup["test"]
    W_idw, bad_i = pipe.predict("diamond", t["lr"], TEST, use_net=False)
    W_net, bad_n = pipe.predict("diamond", t["lr"], TEST, use_net=True)
    assert not bad_i.any() and not bad_n.any()
    sig = setup["net"].sig
    e_idw = _field_error(W_idw, t["hr"], sig)
    e_net = _field_error(W_net, t["hr"], sig)
    print(f"held-out field error (sig units): IDW {e_idw:.4f}  network {e_net:.4f}", flush=True)
    assert e_net < 0.8 * e_idw


def test_warm_start_protocol(setup):
    from eulersr_amd import pipeline as P
    lr, hr = setup["lr"], setup["hr"]
    cases = P.warm_cases("diamond", lr, hr, TEST, setup["test"]["hr_reports"])
    pipe = P.WarmStarter().add("diamond", lr, hr, net=setup["net"], kmap=setup["km"])
    res = {}
    for mode in ("cold", "idw", "mlp"):
        r = pipe.run(cases, mode=mode, max_steps=60000)
        res[mode] = r
        steps = [x["stopping_step"] for x in r["reports"]]
        print(f"{mode}: steps {steps} converged {[x['converged'] for x in r['reports']]} "
              f"cd {[round(x['cd'], 6) for x in r['reports']]} lr {r['lr_s']:.3f}s infer {r['infer_s']:.4f}s "
              f"hr {r['hr_s']:.3f}s fallback {r['fallback']}", flush=True)
        assert all(x["converged"] for x in r["reports"]), mode
        assert not any(r["fallback"])
    # |dCD| <= 1e-4 against the plain (cold) solve, widened by how far the
    # plain solve's own CD still moves after it stops: the residual criterion
    # stops on the K = 0 limiter's plateau, where CD can drift or cycle by
    # more than 1e-4 (the transonic M = 1.1 case here); a warm solve that
    # stops at the same residual elsewhere on that drift is not in error.
    import copy
    from eulersr_amd.solve import BatchSolver
    for i in range(len(TEST)):
        cold = res["cold"]["reports"][i]
        sp = copy.copy(cases[i].hr_spec)
        sp.threshold = 1e-30
        sp.max_steps = 2 * cold["stopping_step"]
        bs = BatchSolver([sp], scheme="SRK2", threshold=1.0, check_every=P.CHECK_EVERY, patience=P.PATIENCE)
        try:
            bs.set_state(0, cases[i].hr_init)
            bs.prepare()
            bs.run()
            later = bs.report(0)["cd"]
        finally:
            bs.close()
        drift = abs(later - cold["cd"])
        for mode in ("idw", "mlp"):
            w = res[mode]["reports"][i]
            print(f"case {TEST[i]} {mode}: |dCD| {abs(w['cd'] - cold['cd']):.2e}; plain CD drift from its stop "
                  f"to twice that {drift:.2e}", flush=True)
            assert abs(w["cd"] - cold["cd"]) <= 1e-4 + drift, (mode, i, w["cd"], cold["cd"], drift)
    # the supersonic case has settled: the bar holds as it stands
    for mode in ("idw", "mlp"):
        assert abs(res[mode]["reports"][1]["cd"] - res["cold"]["reports"][1]["cd"]) <= 1e-4
    s_cold = sum(x["stopping_step"] for x in res["cold"]["reports"])
    s_mlp = sum(x["stopping_step"] for x in res["mlp"]["reports"])
    assert s_mlp < s_cold
