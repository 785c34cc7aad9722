"""Device residency of meshes and states.

PyTorch provides device memory and the current HIP stream; all arithmetic
runs in libeulersr.so. There is no CPU fallback: without a ROCm GPU every
entry point raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .mesh import rcm_seed_order

_torch = None


def torch_mod():
    global _torch
    if _torch is None:
        import torch
        _torch = torch
    return _torch


def require_gpu():
    torch = torch_mod()
    if not torch.cuda.is_available():
        raise RuntimeError("eulersr_amd runs on a ROCm GPU (gfx950); no GPU is visible")
    L.lib()
    return torch


def stream_ptr():
    torch = torch_mod()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def as_device_state(W):
    """(N,4) float32 contiguous device tensor view/copy of W and a flag telling
    whether W was a numpy array."""
    torch = require_gpu()
    if isinstance(W, torch.Tensor):
        t = W
        if t.device.type != "cuda" or t.dtype != torch.float32 or not t.is_contiguous():
            t = t.to(device="cuda", dtype=torch.float32).contiguous()
        return t, False
    a = np.ascontiguousarray(np.asarray(W, dtype=np.float32))
    return torch.from_numpy(a).to("cuda"), True


def back(t, was_numpy):
    if was_numpy:
        return t.cpu().numpy()
    return t


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def wall_markers_of(mesh, kw=None):
    md = getattr(mesh, "metadata", {}) or {}
    wm = (kw or {}).get("wall_markers", md.get("wall_markers", [2]))
    try:
        return tuple(int(m) for m in wm)
    except TypeError:
        return (int(wm),)


class DeviceMesh:
    """A mesh uploaded to the GPU (renumbered, structure of arrays)."""

    def __init__(self, mesh, wall_markers=(2,), renumber=True):
        require_gpu()
        md = getattr(mesh, "metadata", {}) or {}
        self.n_cells = int(np.asarray(mesh.tris).shape[0])
        keep = []

        def f32(x, shape=None):
            a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
            keep.append(a)
            return _ptr(a, C.c_float)

        def i32(x):
            a = np.ascontiguousarray(np.asarray(x, dtype=np.int32))
            keep.append(a)
            return _ptr(a, C.c_int32)

        wm = np.asarray(wall_markers, dtype=np.int32)
        d = L.MeshDesc()
        d.n_points = int(np.asarray(mesh.points).shape[0])
        d.n_cells = self.n_cells
        d.n_faces = int(np.asarray(mesh.faces).shape[0])
        d.points = f32(mesh.points)
        d.tris = i32(mesh.tris)
        d.neighbors = i32(mesh.neighbors)
        d.faces = i32(mesh.faces)
        d.face_markers = i32(mesh.face_markers)
        d.face_connectivity = i32(mesh.face_connectivity)
        fco = getattr(mesh, "face_connectivity_opposite", np.arange(d.n_faces))
        d.face_connectivity_opposite = i32(fco)
        d.barycenter = f32(mesh.barycenter)
        d.area = f32(mesh.area)
        d.surface = f32(mesh.surface)
        d.normals = f32(mesh.normals)
        d.inlet_subsonic = f32(getattr(mesh, "inlet_subsonic", [1.0, 1.0, 0.0, 1.0]))
        d.n_wall_markers = int(wm.size)
        d.wall_markers = i32(wm)
        d.force_marker = int(md.get("force_marker", 2))
        d.sponge = 0 if md.get("case") == "bump" else 1
        d.renumber = 1 if renumber else 0
        d.rcm_seed_order = i32(rcm_seed_order(mesh.neighbors)) if renumber else None
        h = C.c_void_p()
        L.check(L.lib().esr_mesh_create(C.byref(d), C.byref(h)), "esr_mesh_create")
        self.handle = h
        n = C.c_int32()
        nw = C.c_int32()
        nb = C.c_int64()
        L.check(L.lib().esr_mesh_info(h, C.byref(n), C.byref(nw), C.byref(nb)), "esr_mesh_info")
        self.n_wall = nw.value
        self.device_bytes = nb.value

    def permutation(self):
        p = np.empty(self.n_cells, dtype=np.int32)
        L.check(L.lib().esr_mesh_permutation(self.handle, _ptr(p, C.c_int32)), "esr_mesh_permutation")
        return p

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and L._lib is not None:
            L.lib().esr_mesh_destroy(h)
            self.handle = None


def device_mesh(mesh, kw=None, renumber=True) -> DeviceMesh:
    """DeviceMesh cached on the mesh object per (wall markers, renumbering)."""
    key = (wall_markers_of(mesh, kw), bool(renumber))
    cache = mesh.__dict__.setdefault("_esr_device", {})
    dm = cache.get(key)
    if dm is None:
        dm = DeviceMesh(mesh, wall_markers=key[0], renumber=renumber)
        cache[key] = dm
    return dm
