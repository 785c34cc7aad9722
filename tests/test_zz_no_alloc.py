"""The allocation-free step entry point (esr_step_noalloc) through the C ABI:
bit for bit the result of esr_step, on two calls in a row whose second mesh
is smaller than the first (the shared scratch array keeps the first call's
contents), for every time scheme, the HLLC / Roe class, Lax-Friedrichs (its
global wave speed) and both reconstructions, on both stage paths; no device
memory is allocated by the calls; meshes beyond the array's size and custom
sponge widths are refused with ESR_ERR_VALUE."""
import ctypes as C

import numpy as np
import pytest

from conftest import base_cfg, mesh_of, perturbed_state

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["cell", "face"])
def stage_path(request):
    from eulersr_amd.solve import set_stage_path
    set_stage_path(request.param)
    yield request.param
    set_stage_path("auto")


def _call(fn, mesh, kw, scheme, dt, W):
    import torch
    from eulersr_amd import _lib as L
    from eulersr_amd.device import device_mesh, stream_ptr
    from eulersr_amd.euler_solver import SCHEME, make_opts
    dm = device_mesh(mesh, kw)
    opts = make_opts(kw)
    Wt = torch.from_numpy(np.ascontiguousarray(W)).cuda()
    out = torch.empty_like(Wt)
    st = getattr(L.lib(), fn)(dm.handle, C.byref(opts), SCHEME[scheme], C.c_float(dt), C.c_void_p(Wt.data_ptr()),
                              C.c_void_p(out.data_ptr()), stream_ptr())
    torch.cuda.synchronize()
    return st, out.cpu().numpy()


SETS = [("SRK2", "hllc", "muscl"), ("SRK2", "roe", "muscl"), ("RK4", "hllc", "muscl"), ("EE", "roe", "constant"),
        ("RK2", "lf", "muscl"), ("SRK2", "rusanov", "constant")]


@pytest.mark.parametrize("scheme,flux,recon", SETS)
def test_noalloc_step_equals_step_twice_second_smaller(scheme, flux, recon):
    from oracle import fvm as F
    for geom, h in (("diamond", 0.1), ("diamond", 0.2)):       # 1,626 then 740 cells
        mesh = mesh_of(geom, h)
        cfg = base_cfg(Mach=1.8, aoa=3.0)
        W, inlet = perturbed_state(mesh, cfg, seed=21)
        kw = dict(gamma=1.4, M=1.0, reconstruction=recon, flux=flux, value=inlet, entropy=False)
        dt = float(F.get_dt(W, F.omesh(mesh, np.float32), CFL=0.5))
        print(f"case: noalloc {scheme} {flux} {recon} {geom} h={h} ({mesh.n_cells} cells)", flush=True)
        s0, ref = _call("esr_step", mesh, kw, scheme, dt, W)
        s1, got = _call("esr_step_noalloc", mesh, kw, scheme, dt, W)
        assert s0 == 0 and s1 == 0
        assert np.all(np.isfinite(got))
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_noalloc_step_allocates_no_device_memory():
    import torch
    mesh = mesh_of("diamond", 0.1)
    cfg = base_cfg(Mach=2.0)
    W, inlet = perturbed_state(mesh, cfg, seed=22)
    kw = dict(gamma=1.4, M=1.0, reconstruction="muscl", flux="hllc", value=inlet, entropy=False)
    from eulersr_amd import _lib as L
    from eulersr_amd.device import device_mesh, stream_ptr
    from eulersr_amd.euler_solver import SCHEME, make_opts
    dm = device_mesh(mesh, kw)
    opts = make_opts(kw)
    Wt = torch.from_numpy(W).cuda()
    out = torch.empty_like(Wt)
    args = (dm.handle, C.byref(opts), SCHEME["RK4"], C.c_float(1e-3), C.c_void_p(Wt.data_ptr()),
            C.c_void_p(out.data_ptr()), stream_ptr())
    assert L.lib().esr_step_noalloc(*args) == 0
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(20):
        assert L.lib().esr_step_noalloc(*args) == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    print(f"case: noalloc free memory {free0} -> {free1}", flush=True)
    assert free1 == free0


def test_noalloc_refuses_large_meshes_and_custom_sponge():
    from eulersr_amd import _lib as L
    from eulersr_amd.meshing import build_periodic_channel
    big = build_periodic_channel(300, 230)            # 138,000 cells
    assert big.n_cells > 131072
    W = np.tile(np.array([1.0, 1.0, 0.0, 3.0], np.float32), (big.n_cells, 1))
    kw = dict(gamma=1.4, M=1.0, reconstruction="muscl", flux="hllc", value=W[0], entropy=False)
    print(f"case: noalloc refuses {big.n_cells} cells", flush=True)
    st, _ = _call("esr_step_noalloc", big, kw, "SRK2", 1e-4, W)
    assert st == L.ESR_ERR_VALUE
    assert b"131072 cells" in L.lib().esr_last_error()
    mesh = mesh_of("diamond", 0.2)
    Wd, inlet = perturbed_state(mesh, base_cfg(), seed=23)
    kw = dict(gamma=1.4, M=1.0, reconstruction="muscl", flux="hllc", value=inlet, entropy=False, sponge_width=1.0)
    st, _ = _call("esr_step_noalloc", mesh, kw, "SRK2", 1e-4, Wd)
    assert st == L.ESR_ERR_VALUE
    assert b"sponge" in L.lib().esr_last_error()
