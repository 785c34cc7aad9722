"""Learned warm start of HR solves from an LR solution, on the GPU.

The reference reconstructs an HR field from an LR solve with attention
networks over kNN/IDW features (preprocessing/knn.py:20-40,
utils/metrics.py:9-14, eval/core.py:135-142, utils/aero.py:114-181) and uses
the reconstruction as the initial state of the HR solver
(eval/convergence_case.py:30-142, eval/_fvm_solve_worker.py:97-155). Here:

* ``KnnMap``: k = 6 nearest LR barycentres of each HR barycentre (cKDTree, as
  preprocessing/knn.py:25-27), normalised inverse-distance weights
  w = 1/(d^2 + 1e-12) (utils/metrics.py:9-14), neighbour offsets divided by
  the LR mesh size, and the wall features [wd/L, exp(-wd/(2 h_LR)), n_x, n_y]
  (utils/aero.py:160-181). Built once per mesh pair on the host, resident on
  the GPU.
* ``features``: one HIP kernel assembles, per (case, HR cell), the BF16 input
  row of the network and the IDW prediction sum_j w_j prim_LR,j
  (eval/core.py:135-142) straight from the LR solver's conserved states.
* ``WarmNet``: an MLP (4 hidden layers of 256, ReLU) trained and evaluated by
  the library's BF16 MFMA GEMM kernels with FP32 accumulation and FP32
  master weights (AdamW). It predicts the HR primitives as a correction to
  the IDW prediction in units of the HR primitive spread: prim = IDW + sig*Y.
  The output layer starts at zero, so an untrained network is the IDW
  baseline.
* ``assemble``: one HIP kernel turns the prediction into conserved HR states
  (gamma from the case, M = 1, as eval/_fvm_solve_worker.py:104-105) and flags
  every case whose prediction has rho <= 0, p <= 0 or a non-finite value; such
  a case starts from freestream instead.

Conditioning of the case scalars follows preprocessing/dataset.py:12-14:
M -> (M - 1.85) / 1.15, alpha -> alpha / 5.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from .device import require_gpu, stream_ptr

K_NEIGHBOURS = 6
MACH_MID, MACH_SCALE = (0.7 + 3.0) / 2, (3.0 - 0.7) / 2    # preprocessing/dataset.py:12-13
AOA_SCALE = 5.0                                             # preprocessing/dataset.py:14
HIDDEN = (256, 256, 256, 256)
N_SAMPLES_PER_WALL_FACE = 8                                 # utils/aero.py:164


def n_features(k: int = K_NEIGHBOURS) -> int:
    """Input width: per neighbour 4 normalised primitives, 2 offsets, 1 weight;
    then 4 wall features and the 2 case scalars."""
    return 7 * k + 6


def idw_weights(dist: np.ndarray, eps: float = 1e-12) -> np.ndarray:
    """utils/metrics.py:9-14: w = 1/(d^2 + eps), normalised per row (float64)."""
    d = np.asarray(dist, dtype=np.float64)
    w = 1.0 / (d * d + eps)
    return w / w.sum(axis=1, keepdims=True)


def wall_segments(mesh):
    """Obstacle wall faces: boundary faces whose midpoint is not on the
    bounding box (utils/aero.py:114-138). Returns endpoints a, b and the
    left-hand unit normals (-t_y, t_x)."""
    pts = np.asarray(mesh.points, dtype=np.float64)
    faces = np.asarray(mesh.faces)
    fm = np.asarray(mesh.face_markers)
    bnd = faces[fm > 0]
    mid = pts[bnd].mean(axis=1)
    lo, hi = pts.min(0), pts.max(0)
    tol = 1e-6 * float((hi - lo).max())
    on_box = ((np.abs(mid[:, 0] - lo[0]) < tol) | (np.abs(mid[:, 0] - hi[0]) < tol)
              | (np.abs(mid[:, 1] - lo[1]) < tol) | (np.abs(mid[:, 1] - hi[1]) < tol))
    wall = bnd[~on_box]
    if len(wall) == 0:
        raise ValueError("the mesh has no obstacle wall face")
    a, b = pts[wall[:, 0]], pts[wall[:, 1]]
    t = b - a
    t = t / (np.linalg.norm(t, axis=1, keepdims=True) + 1e-12)
    return a, b, np.stack([-t[:, 1], t[:, 0]], axis=1)


def obstacle_length(mesh) -> float:
    """utils/aero.py:141-157: metadata obstacle_length / chord, else the
    diagonal of the wall's bounding box."""
    md = getattr(mesh, "metadata", None) or {}
    for key in ("obstacle_length", "chord"):
        v = md.get(key)
        if v is not None and float(v) > 0:
            return float(v)
    a, b, _ = wall_segments(mesh)
    wp = np.concatenate([a, b])
    return float(np.linalg.norm(wp.max(0) - wp.min(0)))


def wall_feature_array(mesh, query: np.ndarray, h_lr: float) -> np.ndarray:
    """[wd/L, exp(-wd/(2 h_LR)), n_x, n_y] per query point: distance to the
    nearest of 8 samples per wall face and that face's normal, oriented
    towards the query (utils/aero.py:160-181)."""
    from scipy.spatial import cKDTree
    a, b, nrm = wall_segments(mesh)
    t = np.linspace(0.0, 1.0, N_SAMPLES_PER_WALL_FACE)[None, :, None]
    samples = (a[:, None, :] * (1 - t) + b[:, None, :] * t).reshape(-1, 2)
    q = np.asarray(query, dtype=np.float64)
    wd, js = cKDTree(samples).query(q, workers=-1)
    seg = js // N_SAMPLES_PER_WALL_FACE
    n = nrm[seg]
    s = np.sign((n * (q - 0.5 * (a + b)[seg])).sum(axis=1, keepdims=True))
    n = n * np.where(s == 0, 1.0, s)
    return np.column_stack([wd / obstacle_length(mesh), np.exp(-wd / (2.0 * h_lr)), n]).astype(np.float32)


def mesh_size(mesh) -> float:
    h = (getattr(mesh, "metadata", None) or {}).get("h")
    if h is None:
        raise ValueError("mesh metadata has no 'h' (target element size)")
    return float(h)


def knn_host(lr_mesh, hr_mesh, k: int = K_NEIGHBOURS, h_lr: float | None = None) -> dict:
    """Host side of the LR -> HR map: neighbour indices (n_hr, k) int32,
    distances (float64), IDW weights, offsets / h_LR and wall features
    (float32)."""
    from scipy.spatial import cKDTree
    k = int(k)
    h_lr = float(h_lr) if h_lr is not None else mesh_size(lr_mesh)
    lr = np.asarray(lr_mesh.barycenter, dtype=np.float64)
    hr = np.asarray(hr_mesh.barycenter, dtype=np.float64)
    if k < 1 or k > len(lr):
        raise ValueError(f"k = {k} must be in [1, {len(lr)}] (LR cells)")
    dist, idx = cKDTree(lr).query(hr, k=k, workers=-1)
    dist, idx = dist.reshape(len(hr), k), idx.reshape(len(hr), k)
    return dict(k=k, h_lr=h_lr, n_hr=len(hr), n_lr=len(lr), idx=idx.astype(np.int32), dist=dist,
                w=idw_weights(dist).astype(np.float32),
                rel=((lr[idx] - hr[:, None, :]) / h_lr).astype(np.float32),
                wall=wall_feature_array(hr_mesh, hr, h_lr))


class KnnMap:
    """LR -> HR neighbour map of one mesh pair, resident on the GPU."""

    def __init__(self, lr_mesh, hr_mesh, k: int = K_NEIGHBOURS, h_lr: float | None = None):
        import torch
        require_gpu()
        self.lr_mesh, self.hr_mesh = lr_mesh, hr_mesh
        host = knn_host(lr_mesh, hr_mesh, k, h_lr)
        self.k, self.h_lr, self.n_hr, self.n_lr = host["k"], host["h_lr"], host["n_hr"], host["n_lr"]
        self.idx, self.dist, self.w, self.rel, self.wall = (host[x] for x in ("idx", "dist", "w", "rel", "wall"))
        dev = torch.device("cuda")
        self._t = dict(idx=torch.from_numpy(self.idx).to(dev), rel=torch.from_numpy(self.rel).to(dev),
                       w=torch.from_numpy(self.w).to(dev), wall=torch.from_numpy(self.wall).to(dev))
        self.c = L.KnnMap(n_hr=self.n_hr, n_lr=self.n_lr, k=self.k, knn_idx=self._t["idx"].data_ptr(),
                          knn_rel=self._t["rel"].data_ptr(), knn_w=self._t["w"].data_ptr(),
                          wall=self._t["wall"].data_ptr())


def case_scalars(machs, aoas) -> np.ndarray:
    m = np.asarray(machs, dtype=np.float64)
    a = np.asarray(aoas, dtype=np.float64)
    return np.stack([(m - MACH_MID) / MACH_SCALE, a / AOA_SCALE], axis=1).astype(np.float32)


def _f32p(a):
    return C.cast(C.c_void_p(a.ctypes.data), C.POINTER(C.c_float))


def features(kmap: KnnMap, lr_states, machs, aoas, mu, sig, gamma=1.4, in_pad=None):
    """Network inputs (rows x in_pad, bfloat16) and IDW primitives (rows x 4,
    float32) for n_cases LR states ((n_cases, n_lr, 4) conserved, float32 on
    the GPU). Rows are case-major in HR original cell order."""
    import torch
    lr = lr_states.contiguous()
    if lr.dtype != torch.float32 or not lr.is_cuda or lr.shape[-2:] != (kmap.n_lr, 4):
        raise ValueError(f"LR states must be float32 CUDA (n_cases, {kmap.n_lr}, 4)")
    n_cases = lr.numel() // (kmap.n_lr * 4)
    if len(machs) != n_cases or len(aoas) != n_cases:
        raise ValueError("one (Mach, alpha) per LR state is required")
    in_pad = int(in_pad or n_features(kmap.k))
    rows = n_cases * kmap.n_hr
    X = torch.empty((rows, in_pad), dtype=torch.bfloat16, device=lr.device)
    idw = torch.empty((rows, 4), dtype=torch.float32, device=lr.device)
    scal = torch.from_numpy(case_scalars(machs, aoas)).to(lr.device)
    mu = np.ascontiguousarray(mu, dtype=np.float32)
    sig = np.ascontiguousarray(sig, dtype=np.float32)
    L.check(L.lib().esr_warm_features(C.byref(kmap.c), n_cases, C.c_void_p(lr.data_ptr()),
                                      C.c_void_p(scal.data_ptr()), _f32p(mu), _f32p(sig), float(gamma), in_pad,
                                      C.c_void_p(X.data_ptr()), C.c_void_p(idw.data_ptr()), stream_ptr()),
            "esr_warm_features")
    return X, idw


def assemble(n_hr: int, idw, delta, sig, gamma=1.4):
    """Conserved HR states (rows x 4) = cons(IDW + sig * delta) and a per-case
    flag (1: negative density/pressure or non-finite -> freestream fallback)."""
    import torch
    rows = idw.shape[0]
    n_cases = rows // n_hr
    W = torch.empty((rows, 4), dtype=torch.float32, device=idw.device)
    bad = torch.zeros(n_cases, dtype=torch.int32, device=idw.device)
    sig = np.ascontiguousarray(sig, dtype=np.float32)
    dptr = C.c_void_p(delta.data_ptr()) if delta is not None else None
    L.check(L.lib().esr_warm_apply(rows, n_hr, C.c_void_p(idw.data_ptr()), dptr, _f32p(sig), float(gamma),
                                   C.c_void_p(W.data_ptr()), C.c_void_p(bad.data_ptr()), stream_ptr()),
            "esr_warm_apply")
    return W, bad


class WarmNet:
    """MLP over the warm-start features, trained and run by the library's
    BF16 MFMA kernels. ``mu``/``sig`` are the HR primitive statistics of the
    training set (inputs are normalised by them, outputs scaled by sig)."""

    def __init__(self, in_dim=None, hidden=HIDDEN, out_dim=4, max_rows=1 << 15, seed=0):
        require_gpu()
        self.widths = [int(in_dim or n_features()), *[int(h) for h in hidden], int(out_dim)]
        self.max_rows = int(max_rows)
        w = (C.c_int32 * len(self.widths))(*self.widths)
        h = C.c_void_p()
        L.check(L.lib().esr_mlp_create(len(self.widths) - 1, w, self.max_rows, C.byref(h)), "esr_mlp_create")
        self.handle = h
        n = C.c_int64()
        L.check(L.lib().esr_mlp_param_count(self.handle, C.byref(n)), "esr_mlp_param_count")
        self.n_params = n.value
        self.mu = np.zeros(4, np.float32)
        self.sig = np.ones(4, np.float32)
        self.t = 0
        self.set_params(self.init_params(seed))

    def layer_slices(self):
        out, o = [], 0
        for i in range(len(self.widths) - 1):
            ni, no = self.widths[i], self.widths[i + 1]
            out.append((slice(o, o + no * ni), slice(o + no * ni, o + no * ni + no)))
            o += no * ni + no
        return out

    def init_params(self, seed=0):
        """He-uniform weights on the hidden layers, zero biases, and a zero
        output layer (the untrained network predicts no correction)."""
        rng = np.random.default_rng(seed)
        p = np.zeros(self.n_params, np.float32)
        sl = self.layer_slices()
        for i, (ws, _) in enumerate(sl[:-1]):
            lim = math.sqrt(6.0 / self.widths[i])
            p[ws] = rng.uniform(-lim, lim, ws.stop - ws.start).astype(np.float32)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        L.check(L.lib().esr_mlp_set_params(self.handle, _f32p(p)), "esr_mlp_set_params")
        self.t = 0

    def params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        L.check(L.lib().esr_mlp_get_params(self.handle, _f32p(p)), "esr_mlp_get_params")
        return p

    def _check_x(self, X):
        import torch
        if X.dtype != torch.bfloat16 or not X.is_cuda or X.dim() != 2 or X.shape[1] != self.widths[0]:
            raise ValueError(f"X must be a bfloat16 CUDA (rows, {self.widths[0]}) tensor")
        return X.contiguous()

    def forward(self, X):
        import torch
        X = self._check_x(X)
        Y = torch.empty((X.shape[0], self.widths[-1]), dtype=torch.float32, device=X.device)
        L.check(L.lib().esr_mlp_forward(self.handle, X.shape[0], C.c_void_p(X.data_ptr()),
                                        C.c_void_p(Y.data_ptr()), stream_ptr()), "esr_mlp_forward")
        return Y

    def loss_grad(self, X, T, want_grads=False):
        X = self._check_x(X)
        T = T.contiguous()
        if T.shape != (X.shape[0], self.widths[-1]):
            raise ValueError("T must be (rows, out) float32")
        loss = C.c_double()
        g = np.empty(self.n_params, np.float32) if want_grads else None
        L.check(L.lib().esr_mlp_loss_grad(self.handle, X.shape[0], C.c_void_p(X.data_ptr()),
                                          C.c_void_p(T.data_ptr()), C.byref(loss),
                                          _f32p(g) if want_grads else None, stream_ptr()), "esr_mlp_loss_grad")
        return (loss.value, g) if want_grads else loss.value

    def adamw(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=1e-5):
        self.t += 1
        L.check(L.lib().esr_mlp_adamw(self.handle, float(lr), float(beta1), float(beta2), float(eps),
                                      float(weight_decay), self.t, stream_ptr()), "esr_mlp_adamw")

    def fit(self, X, T, steps=2000, batch=16384, lr=1e-3, lr_min=1e-5, weight_decay=1e-5, warmup=50, seed=0,
            log_every=0):
        """AdamW over random mini-batches with a cosine learning-rate schedule
        (models/base.py TrainConfig: AdamW, cosine). Returns the loss history."""
        import torch
        X = self._check_x(X)
        rows = X.shape[0]
        batch = min(int(batch), rows, self.max_rows)
        g = torch.Generator(device=X.device)
        g.manual_seed(int(seed))
        hist = []
        for s in range(int(steps)):
            sel = torch.randint(0, rows, (batch,), generator=g, device=X.device)
            xb, tb = X.index_select(0, sel), T.index_select(0, sel)
            loss = self.loss_grad(xb, tb)
            if s < warmup:
                rate = lr * (s + 1) / warmup
            else:
                q = (s - warmup) / max(1, steps - warmup)
                rate = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * q))
            self.adamw(rate, weight_decay=weight_decay)
            hist.append(loss)
            if log_every and (s % log_every == 0 or s == steps - 1):
                print(f"[warmnet] step {s} loss {loss:.4e} lr {rate:.2e}", flush=True)
        return hist

    def save(self, path):
        np.savez(path, widths=np.asarray(self.widths, np.int32), params=self.params(), mu=self.mu, sig=self.sig)

    @classmethod
    def load(cls, path, max_rows=1 << 15):
        d = np.load(path, allow_pickle=False)
        w = [int(x) for x in d["widths"]]
        net = cls(in_dim=w[0], hidden=w[1:-1], out_dim=w[-1], max_rows=max_rows)
        net.set_params(d["params"])
        net.mu, net.sig = d["mu"].astype(np.float32), d["sig"].astype(np.float32)
        return net

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            L.lib().esr_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def primitives(W: np.ndarray, gamma=1.4) -> np.ndarray:
    """helper.py:38-46 with M = 1 (float64)."""
    W = np.asarray(W, dtype=np.float64)
    r = W[..., 0]
    u, v = W[..., 1] / r, W[..., 2] / r
    p = (W[..., 3] - 0.5 * r * (u * u + v * v)) * (gamma - 1.0)
    return np.stack([r, u, v, p], axis=-1)
