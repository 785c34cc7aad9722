"""CPU pins of the warm-start oracles (oracle/mlp.py, oracle/warm.py) and of
the host-side kNN / wall-feature code of eulersr_amd.warmstart.

The MLP oracle is pinned by central finite differences and by torch CPU
autograd (float64), its bfloat16 rounding by torch's own conversion, its
AdamW by torch.optim.AdamW; the kNN by brute force; IDW by its closed forms;
the wall features by geometric invariants of a convex body."""
import math

import numpy as np
import pytest
import torch

from oracle import mlp as OM
from oracle import warm as OW
from eulersr_amd import warmstart as WS


def _params(widths, rng, scale=0.5):
    n = sum(widths[i] * widths[i + 1] + widths[i + 1] for i in range(len(widths) - 1))
    return rng.normal(0, scale, n)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-30, 100), rng.normal(0, 1e30, 100),
                        [0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(OM.bf16(x), ref)
    # ties go to even: 1 + 2^-8 is half-way between 1 and 1 + 2^-7
    assert OM.bf16(np.float32(1.0 + 2.0 ** -8)) == np.float32(1.0)
    assert OM.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == np.float32(1.0 + 2.0 ** -6)


def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(1)
    widths = [5, 7, 6, 3]
    p = _params(widths, rng)
    X = rng.normal(0, 1, (11, 5))
    T = rng.normal(0, 1, (11, 3))
    loss, g = OM.loss_grad(p, widths, X, T, emulate_bf16=False)
    h = 1e-6
    for i in rng.choice(len(p), 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (OM.loss_grad(pp, widths, X, T, emulate_bf16=False)[0]
              - OM.loss_grad(pm, widths, X, T, emulate_bf16=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_gradient_matches_torch_autograd():
    rng = np.random.default_rng(2)
    widths = [8, 16, 16, 4]
    p = _params(widths, rng)
    X = rng.normal(0, 1, (32, 8))
    T = rng.normal(0, 1, (32, 4))
    loss, g = OM.loss_grad(p, widths, X, T, emulate_bf16=False)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    h = torch.tensor(X)
    o = 0
    for i in range(len(widths) - 1):
        ni, no = widths[i], widths[i + 1]
        W = pt[o:o + ni * no].view(no, ni)
        o += ni * no
        b = pt[o:o + no]
        o += no
        h = h @ W.T + b
        if i < len(widths) - 2:
            h = torch.relu(h)
    lt = torch.mean((h - torch.tensor(T)) ** 2)
    lt.backward()
    assert abs(lt.item() - loss) < 1e-12
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_mlp_forward_output_layer_zero_gives_zero():
    rng = np.random.default_rng(3)
    widths = [6, 10, 4]
    p = _params(widths, rng)
    p[-(10 * 4 + 4):] = 0.0
    Zs, _ = OM.forward(p, widths, rng.normal(0, 1, (5, 6)), emulate_bf16=True)
    assert np.all(Zs[-1] == 0.0)


def test_adamw_matches_torch():
    rng = np.random.default_rng(4)
    widths = [3, 5, 2]
    n = 3 * 5 + 5 + 5 * 2 + 2
    p0 = rng.normal(0, 1, n)
    grads = [rng.normal(0, 1, n) for _ in range(3)]
    lr, wd = 1e-2, 0.1
    # torch: decay on the weight matrices only, as two parameter groups
    wmask = np.zeros(n, bool)
    wmask[0:15] = True
    wmask[20:30] = True
    pw = torch.tensor(p0[wmask], requires_grad=True)
    pb = torch.tensor(p0[~wmask], requires_grad=True)
    opt = torch.optim.AdamW([{"params": [pw], "weight_decay": wd}, {"params": [pb], "weight_decay": 0.0}],
                            lr=lr, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(n), np.zeros(n)
    for t, g in enumerate(grads, start=1):
        pw.grad = torch.tensor(g[wmask])
        pb.grad = torch.tensor(g[~wmask])
        opt.step()
        p, m, v = OM.adamw(p, g, m, v, t, lr, wd=wd, widths=widths)
    ref = np.empty(n)
    ref[wmask] = pw.detach().numpy()
    ref[~wmask] = pb.detach().numpy()
    np.testing.assert_allclose(p, ref, rtol=1e-12, atol=1e-14)


def test_idw_closed_forms():
    d = np.array([[1.0, 1.0, 1.0, 1.0], [0.0, 1.0, 2.0, 3.0], [1.0, 2.0, 2.0, 4.0]])
    w = WS.idw_weights(d)
    np.testing.assert_allclose(w.sum(1), 1.0, rtol=1e-15)
    np.testing.assert_allclose(w[0], 0.25)
    assert w[1, 0] > 1 - 1e-11
    inv = 1 / np.array([1.0, 4.0, 4.0, 16.0])
    np.testing.assert_allclose(w[2], inv / inv.sum(), rtol=1e-11)
    np.testing.assert_allclose(OW.idw_weights(d), w, rtol=0, atol=0)
    # IDW reproduces a constant field exactly (up to rounding)
    prim = np.tile(np.array([1.2, 0.3, -0.1, 2.0]), (10, 1))
    idx = np.array([[0, 3, 5, 7], [1, 2, 8, 9], [4, 4, 6, 0]])
    np.testing.assert_allclose(OW.idw_predict(prim, idx, w), prim[:3], rtol=1e-14)


@pytest.fixture(scope="module")
def pair():
    from eulersr_amd.meshing import build_mesh
    return build_mesh("diamond", 0.2), build_mesh("diamond", 0.1)


def test_knn_matches_brute_force(pair):
    lr, hr = pair
    host = WS.knn_host(lr, hr, k=6)
    idx, dist = OW.knn_brute(np.asarray(lr.barycenter, np.float64), np.asarray(hr.barycenter, np.float64), 6)
    np.testing.assert_allclose(host["dist"], dist, rtol=1e-12, atol=1e-14)
    # indices agree wherever the distances are not tied
    untied = np.abs(np.diff(dist, axis=1)).min(axis=1) > 1e-12
    assert untied.mean() > 0.99
    assert np.array_equal(host["idx"][untied], idx[untied])
    rel = (np.asarray(lr.barycenter)[host["idx"]] - np.asarray(hr.barycenter)[:, None, :]) / 0.2
    np.testing.assert_allclose(host["rel"], rel, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(np.linalg.norm(host["rel"], axis=2) * 0.2, host["dist"], rtol=1e-5)


def test_wall_features_geometry(pair):
    lr, hr = pair
    q = np.asarray(hr.barycenter, np.float64)
    f = WS.wall_feature_array(hr, q, 0.2)
    o = OW.wall_features(hr.points, hr.faces, hr.face_markers, q, 1.0, 0.2)
    np.testing.assert_allclose(f[:, :2], o[:, :2], rtol=1e-6, atol=1e-7)
    # the normals agree wherever one wall face holds the nearest sample
    # (a corner sample is shared by two faces and the tie is the tree's)
    a, b, _ = WS.wall_segments(hr)
    t = np.linspace(0.0, 1.0, 8)[None, :, None]
    samp = (a[:, None, :] * (1 - t) + b[:, None, :] * t)
    ds = np.linalg.norm(q[:, None, None, :] - samp[None], axis=3).min(axis=2)   # (n_q, n_faces)
    unique = ((ds <= ds.min(1, keepdims=True) + 1e-12).sum(1) == 1)
    assert unique.mean() > 0.3
    np.testing.assert_allclose(f[unique, 2:], o[unique, 2:], rtol=1e-6, atol=1e-7)
    # unit normals, exp(-wd/(2h)) consistent with wd
    np.testing.assert_allclose(np.hypot(f[:, 2], f[:, 3]), 1.0, rtol=1e-6)
    np.testing.assert_allclose(f[:, 1], np.exp(-f[:, 0] * 1.0 / (2 * 0.2)), rtol=1e-5)
    # above and below the diamond (away from its two tips) the oriented
    # normal points away from the body: upwards above it, downwards below
    md = hr.metadata
    c = np.array([md["center"]["cx"], md["center"]["cy"]])
    mid = np.abs(q[:, 0] - c[0]) < 0.4
    assert mid.sum() > 50
    assert np.all(np.sign(f[mid, 3]) == np.sign(q[mid, 1] - c[1]))
    # the wall distance is at least the distance to the diamond (sampled
    # segments) and at most that plus half a sample spacing
    a, b, _ = WS.wall_segments(hr)
    t = np.clip((((q[:, None, :] - a[None]) * (b - a)[None]).sum(-1)) / ((b - a) ** 2).sum(-1)[None], 0, 1)
    proj = a[None] + t[..., None] * (b - a)[None]
    dseg = np.linalg.norm(q[:, None, :] - proj, axis=2).min(1)
    spacing = np.linalg.norm(b - a, axis=1).max() / 7
    assert np.all(f[:, 0] >= dseg - 1e-6)
    assert np.all(f[:, 0] <= np.hypot(dseg, spacing / 2) + 1e-6)


def test_wall_segments_exclude_box(pair):
    _, hr = pair
    a, b, n = WS.wall_segments(hr)
    wall = np.asarray(hr.face_markers) == 2
    assert len(a) == int(wall.sum())
    assert WS.obstacle_length(hr) == 1.0


def test_features_oracle_columns():
    rng = np.random.default_rng(5)
    n_lr, n_hr, k, n_cases = 9, 7, 3, 2
    prim = np.stack([rng.uniform(0.5, 2, (n_cases, n_lr)), rng.normal(0, 1, (n_cases, n_lr)),
                     rng.normal(0, 1, (n_cases, n_lr)), rng.uniform(0.5, 2, (n_cases, n_lr))], -1)
    g = 1.4
    E = prim[..., 3] / (g - 1) + 0.5 * prim[..., 0] * (prim[..., 1] ** 2 + prim[..., 2] ** 2)
    W = np.stack([prim[..., 0], prim[..., 0] * prim[..., 1], prim[..., 0] * prim[..., 2], E], -1)
    idx = rng.integers(0, n_lr, (n_hr, k)).astype(np.int32)
    w = WS.idw_weights(rng.uniform(0.1, 1, (n_hr, k))).astype(np.float32)
    rel = rng.normal(0, 1, (n_hr, k, 2)).astype(np.float32)
    wall = rng.normal(0, 1, (n_hr, 4)).astype(np.float32)
    scal = WS.case_scalars([1.85, 3.0], [0.0, 5.0])
    np.testing.assert_allclose(scal, [[0, 0], [1, 1]], atol=1e-7)
    mu, sig = np.zeros(4, np.float32), np.ones(4, np.float32)
    X, idw = OW.features(idx, rel, w, wall, W.astype(np.float32), scal, mu, sig, g, 48)
    assert X.shape == (n_cases * n_hr, 48)
    r = 1 * n_hr + 4                      # case 1, HR cell 4
    np.testing.assert_allclose(X[r, 0:4], prim[1, idx[4, 0]], rtol=1e-5)
    np.testing.assert_allclose(X[r, 7:11], prim[1, idx[4, 1]], rtol=1e-5)
    np.testing.assert_allclose(X[r, 4:6], rel[4, 0])
    assert X[r, 6] == w[4, 0]
    np.testing.assert_allclose(X[r, 21:25], wall[4])
    np.testing.assert_allclose(X[r, 25:27], scal[1])
    assert np.all(X[:, 27:] == 0)
    np.testing.assert_allclose(idw[r], (w[4][:, None] * prim[1, idx[4]]).sum(0), rtol=1e-5)
    # apply inverts the conversion and flags non-physical predictions
    Wb, bad = OW.apply(idw, None, sig, g, n_hr)
    assert not bad.any()
    d = np.zeros_like(idw)
    d[n_hr + 2, 3] = -10.0
    _, bad = OW.apply(idw, d, sig, g, n_hr)
    assert list(bad) == [False, True]


def test_stall_rule():
    from eulersr_amd.pipeline import stalled
    falling = [(s, 1e-4 * 0.9995 ** s) for s in range(100, 12000, 100)]
    flat = [(s, 2e-5 * (1.0 + 0.1 * ((s // 100) % 2))) for s in range(100, 12000, 100)]
    assert not stalled(falling, 11900, 6000, 3000, 0.97)
    assert stalled(flat, 11900, 6000, 3000, 0.97)
    assert not stalled(flat, 5000, 6000, 3000, 0.97)       # too early
    assert not stalled([], 11900, 6000, 3000, 0.97)
