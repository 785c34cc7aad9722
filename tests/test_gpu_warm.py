"""GPU parity of the warm-start kernels against the CPU oracles: the BF16
MFMA GEMM (forward of one- and multi-layer networks, odd sizes, split-K
weight gradients), the MSE gradient of the whole network, AdamW, the feature
kernel and the state assembly with its fallback flags."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev():
    import torch
    from eulersr_amd.device import require_gpu
    require_gpu()
    return torch


def _bf16_tensor(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to("cuda").to(torch.bfloat16)


@pytest.mark.parametrize("rows,widths", [(1, [48, 4]), (1000, [48, 70]), (333, [40, 256, 4]),
                                         (4099, [48, 256, 256, 256, 256, 4])])
def test_forward_matches_oracle(rows, widths):
    torch = _dev()
    from oracle import mlp as OM
    from eulersr_amd.warmstart import WarmNet
    print(f"forward rows={rows} widths={widths}", flush=True)
    rng = np.random.default_rng(rows)
    net = WarmNet(in_dim=widths[0], hidden=widths[1:-1], out_dim=widths[-1], max_rows=2048)
    p = (rng.normal(0, 1, net.n_params) / np.sqrt(widths[0])).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    X = rng.normal(0, 1, (rows, widths[0])).astype(np.float32)
    Y = net.forward(_bf16_tensor(torch, X)).cpu().numpy()
    Zs, _ = OM.forward(p, widths, X, emulate_bf16=True)
    ref = Zs[-1]
    scale = np.abs(ref).max() + 1e-30
    # fp32 accumulation vs float64; hidden layers may differ by one bf16 ulp
    tol = 1e-5 if len(widths) == 2 else 2e-2
    err = np.abs(Y - ref).max() / scale
    assert err <= tol, err
    if len(widths) > 2:
        assert np.abs(Y - ref).mean() / scale <= 1e-3
    net.close()


@pytest.mark.parametrize("rows,widths", [(333, [40, 256, 4]), (4099, [48, 256, 256, 256, 256, 4])])
def test_forward_is_its_layers_composed_exactly(rows, widths):
    """The multi-layer forward equals, bit for bit, a chain of one-layer
    networks with the same weights, ReLU and round-to-nearest-even bfloat16
    applied on the host between them; and every layer of the chain, fed the
    GPU's own activations, matches the float64 oracle within 1e-5. The
    tolerance above for whole networks covers only where the GPU's and the
    oracle's hidden activations round to neighbouring bfloat16 values."""
    torch = _dev()
    from oracle import mlp as OM
    from eulersr_amd.warmstart import WarmNet
    print(f"forward composition rows={rows} widths={widths}", flush=True)
    rng = np.random.default_rng(11 + rows)
    net = WarmNet(in_dim=widths[0], hidden=widths[1:-1], out_dim=widths[-1], max_rows=2048)
    p = (rng.normal(0, 1, net.n_params) / np.sqrt(widths[0])).astype(np.float32)
    net.set_params(p)
    X = OM.bf16(rng.normal(0, 1, (rows, widths[0])).astype(np.float32))
    Y = net.forward(_bf16_tensor(torch, X)).cpu().numpy()
    net.close()
    H = X
    for li, (W, b) in enumerate(OM.unpack(p, widths)):
        one = WarmNet(in_dim=widths[li], hidden=[], out_dim=widths[li + 1], max_rows=2048)
        one.set_params(np.concatenate([W.reshape(-1), b]))
        Z = one.forward(_bf16_tensor(torch, H)).cpu().numpy()
        one.close()
        ref = H.astype(np.float64) @ OM.bf16(W).astype(np.float64).T + b.astype(np.float64)
        err = np.abs(Z - ref).max() / (np.abs(ref).max() + 1e-30)
        print(f"layer {li}: max relative error against float64 {err:.2e}", flush=True)
        assert err <= 1e-5
        H = Z if li == len(widths) - 2 else OM.bf16(np.maximum(Z, 0.0))
    assert np.array_equal(Y.view(np.uint32), H.view(np.uint32))


@pytest.mark.parametrize("rows,widths", [(300, [48, 70, 4]), (3000, [48, 70, 4]), (5000, [48, 256, 256, 4])])
def test_loss_grad_matches_oracle(rows, widths):
    torch = _dev()
    from oracle import mlp as OM
    from eulersr_amd.warmstart import WarmNet
    print(f"loss_grad rows={rows} widths={widths}", flush=True)
    rng = np.random.default_rng(7 + rows)
    net = WarmNet(in_dim=widths[0], hidden=widths[1:-1], out_dim=widths[-1], max_rows=rows)
    p = (rng.normal(0, 1, net.n_params) / np.sqrt(64)).astype(np.float32)
    net.set_params(p)
    X = rng.normal(0, 1, (rows, widths[0])).astype(np.float32)
    T = rng.normal(0, 1, (rows, widths[-1])).astype(np.float32)
    loss, g = net.loss_grad(_bf16_tensor(torch, X), torch.from_numpy(T).cuda(), want_grads=True)
    lref, gref = OM.loss_grad(p, widths, X, T, emulate_bf16=True)
    assert abs(loss - lref) <= 1e-5 * abs(lref)
    # per layer: weight and bias gradients
    o = 0
    for i in range(len(widths) - 1):
        n = widths[i] * widths[i + 1] + widths[i + 1]
        a, b = g[o:o + n], gref[o:o + n]
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
        assert rel <= 5e-3, (i, rel)
        o += n
    net.close()


def test_adamw_matches_oracle():
    torch = _dev()
    from oracle import mlp as OM
    from eulersr_amd.warmstart import WarmNet
    widths = [48, 64, 4]
    rng = np.random.default_rng(11)
    net = WarmNet(in_dim=48, hidden=(64,), out_dim=4, max_rows=512)
    p0 = (rng.normal(0, 1, net.n_params) * 0.1).astype(np.float32)
    net.set_params(p0)
    X = _bf16_tensor(torch, rng.normal(0, 1, (512, 48)))
    T = torch.from_numpy(rng.normal(0, 1, (512, 4)).astype(np.float32)).cuda()
    p, m, v = p0.astype(np.float64), np.zeros(net.n_params), np.zeros(net.n_params)
    for t in range(1, 4):
        _, g = net.loss_grad(X, T, want_grads=True)
        net.adamw(1e-3, weight_decay=0.01)
        p, m, v = OM.adamw(p, g.astype(np.float64), m, v, t, 1e-3, wd=0.01, widths=widths)
        np.testing.assert_allclose(net.params(), p, rtol=1e-5, atol=1e-7)
    net.close()


def test_training_reduces_loss():
    torch = _dev()
    from eulersr_amd.warmstart import WarmNet
    rng = np.random.default_rng(12)
    X = rng.normal(0, 1, (20000, 48)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2], np.maximum(X[:, 3], 0), X[:, 4] ** 2 - 1,
                  np.tanh(X[:, 5] - X[:, 6])], 1).astype(np.float32)
    net = WarmNet(in_dim=48, max_rows=4096, seed=3)
    Xt, Tt = _bf16_tensor(torch, X), torch.from_numpy(T).cuda()
    l0 = net.loss_grad(Xt[:4096], Tt[:4096])
    hist = net.fit(Xt, Tt, steps=400, batch=2048, lr=1e-3, warmup=20, seed=0)
    l1 = net.loss_grad(Xt[:4096], Tt[:4096])
    print(f"train loss {l0:.4f} -> {l1:.4f}", flush=True)
    assert l1 < 0.1 * l0
    # deterministic: the same seed and data give bit-identical parameters
    net2 = WarmNet(in_dim=48, max_rows=4096, seed=3)
    net2.fit(Xt, Tt, steps=400, batch=2048, lr=1e-3, warmup=20, seed=0)
    assert np.array_equal(net.params(), net2.params())
    net.close()
    net2.close()


@pytest.fixture(scope="module")
def pair():
    from eulersr_amd.meshing import build_mesh
    return build_mesh("diamond", 0.2), build_mesh("diamond", 0.1)


def _lr_states(lr_mesh, n_cases, seed):
    rng = np.random.default_rng(seed)
    n = lr_mesh.n_cells
    prim = np.stack([rng.uniform(0.5, 2.0, (n_cases, n)), rng.normal(1.0, 0.5, (n_cases, n)),
                     rng.normal(0.0, 0.5, (n_cases, n)), rng.uniform(0.5, 2.0, (n_cases, n))], -1)
    E = prim[..., 3] / 0.4 + 0.5 * prim[..., 0] * (prim[..., 1] ** 2 + prim[..., 2] ** 2)
    return np.stack([prim[..., 0], prim[..., 0] * prim[..., 1], prim[..., 0] * prim[..., 2], E],
                    -1).astype(np.float32)


@pytest.mark.parametrize("k,in_pad", [(6, None), (6, 64), (4, None), (4, 40)])
def test_features_and_apply_match_oracle(pair, k, in_pad):
    torch = _dev()
    from oracle import mlp as OM
    from oracle import warm as OW
    from eulersr_amd import warmstart as WS
    print(f"features k={k} in_pad={in_pad}", flush=True)
    lr, hr = pair
    km = WS.KnnMap(lr, hr, k=k)
    width = in_pad or WS.n_features(k)
    n_cases = 3
    W = _lr_states(lr, n_cases, 0)
    machs, aoas = [0.8, 1.5, 2.5], [-2.0, 0.0, 4.5]
    mu = np.array([1.1, 0.9, 0.05, 1.2], np.float32)
    sig = np.array([0.4, 0.5, 0.3, 0.6], np.float32)
    X, idw = WS.features(km, torch.from_numpy(W).cuda(), machs, aoas, mu, sig, in_pad=in_pad)
    assert X.shape == (n_cases * km.n_hr, width)
    Xg = X.float().cpu().numpy()
    Xo, idw_o = OW.features(km.idx, km.rel, km.w, km.wall, W, WS.case_scalars(machs, aoas), mu, sig, 1.4, width)
    Xo16 = OM.bf16(Xo)
    exact = (Xg == Xo16)
    assert exact.mean() > 0.999, exact.mean()
    # the rest within one bf16 ulp (contracted vs separate float32 roundings)
    np.testing.assert_allclose(Xg, Xo16, rtol=2.0 ** -7, atol=1e-30)
    np.testing.assert_allclose(idw.cpu().numpy(), idw_o, rtol=2e-6, atol=1e-6)
    # assembly: IDW + sig * delta -> conserved
    rng = np.random.default_rng(1)
    delta = (rng.normal(0, 0.05, (n_cases * km.n_hr, 4))).astype(np.float32)
    Wa, bad = WS.assemble(km.n_hr, idw, torch.from_numpy(delta).cuda(), sig)
    Wo, bado = OW.apply(idw.cpu().numpy(), delta, sig, 1.4, km.n_hr)
    np.testing.assert_allclose(Wa.cpu().numpy(), Wo, rtol=2e-6, atol=1e-6)
    assert not bad.cpu().numpy().any() and not bado.any()
    # IDW only (no network)
    W0, _ = WS.assemble(km.n_hr, idw, None, sig)
    Wo0, _ = OW.apply(idw.cpu().numpy(), None, sig, 1.4, km.n_hr)
    np.testing.assert_allclose(W0.cpu().numpy(), Wo0, rtol=2e-6, atol=1e-6)
    # fallback flags: negative pressure in case 1, NaN in case 2
    d2 = delta.copy()
    d2[km.n_hr + 5, 3] = -1e3
    d2[2 * km.n_hr + 9, 1] = np.nan
    _, bad = WS.assemble(km.n_hr, idw, torch.from_numpy(d2).cuda(), sig)
    assert list(bad.cpu().numpy()) == [0, 1, 1]
