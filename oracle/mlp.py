"""CPU oracle (TEST INFRASTRUCTURE ONLY): the warm-start multilayer perceptron
(ReLU hidden layers, linear output, mean-squared-error loss) and its
gradients, with optional bfloat16 rounding at the points where the GPU path
stores bfloat16 (inputs, weights, hidden activations, output and hidden
gradients fed to the GEMMs). Products are accumulated in float64.

Parameter layout: per layer W (out x in, row-major) then b (out), flattened.
Pinned by central finite differences in float64 (tests/test_oracle_warm.py);
the GPU kernels are checked against it (tests/test_gpu_warm.py).
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """float32 -> bfloat16 -> float32, round to nearest even."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(params, widths):
    out, o = [], 0
    for i in range(len(widths) - 1):
        n_in, n_out = widths[i], widths[i + 1]
        W = params[o:o + n_out * n_in].reshape(n_out, n_in)
        o += n_out * n_in
        b = params[o:o + n_out]
        o += n_out
        out.append((W, b))
    return out


def forward(params, widths, X, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=np.float32 if emulate_bf16 else np.float64), widths)
    h = r(X).astype(np.float64)
    Zs, Hs = [], [h]
    for li, (W, b) in enumerate(layers):
        z = h @ r(W).astype(np.float64).T + np.asarray(b, dtype=np.float64)
        Zs.append(z)
        if li < len(layers) - 1:
            z32 = z.astype(np.float32) if emulate_bf16 else z
            h = r(np.maximum(z32, 0)).astype(np.float64)
            Hs.append(h)
    return Zs, Hs


def loss_grad(params, widths, X, T, emulate_bf16=True):
    """Loss = mean((Y - T)^2); returns (loss, flat gradient)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=np.float32 if emulate_bf16 else np.float64), widths)
    Zs, Hs = forward(params, widths, X, emulate_bf16)
    Y = Zs[-1].astype(np.float32) if emulate_bf16 else Zs[-1]
    T = np.asarray(T, dtype=np.float64)
    n = Y.size
    loss = float(np.mean((np.asarray(Y, dtype=np.float64) - T) ** 2))
    dz = (2.0 / n) * (np.asarray(Y, dtype=np.float64) - T)
    if emulate_bf16:
        dz = dz.astype(np.float32)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        dz_g = r(dz).astype(np.float64)
        gW = dz_g.T @ Hs[li]
        gb = np.sum(np.asarray(dz, dtype=np.float64), axis=0)
        grads[li] = (gW, gb)
        if li > 0:
            dh = dz_g @ r(W).astype(np.float64)
            if emulate_bf16:
                dh = dh.astype(np.float32)
            dz = np.where(Zs[li - 1].astype(np.float32 if emulate_bf16 else np.float64) > 0, dh, 0.0)
    flat = np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])
    return loss, flat


def adamw(params, grads, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0, widths=None):
    """Decoupled weight decay on the weight matrices only."""
    p = np.asarray(params, dtype=np.float64).copy()
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    decay = np.zeros_like(p)
    if widths is not None:
        o = 0
        for i in range(len(widths) - 1):
            nw = widths[i] * widths[i + 1]
            decay[o:o + nw] = wd
            o += nw + widths[i + 1]
    p = p - lr * (mh / (np.sqrt(vh) + eps) + decay * p)
    return p, m, v
