"""Parity on the GPU (through the C ABI) for the boundary paths the airfoil
and diamond meshes do not reach: the subsonic inlet (marker 5,
helper.py:179-194), periodic faces (marker 1, mesh.py:266-321: a boundary
in the LSQ gradient, helper.py:142, interior in MUSCL, euler_solver.py:82-93)
and the bump channel (meshing/bump.py:139-210: wall markers [2, 6], no
sponge, helper.py:266-267, forces on marker 2).

Tolerances as in test_gpu_parity.py: residuals and single steps within 4x
the oracle's own float32-vs-float64 difference (floor 1e-6 of the
magnitude); per-cell primitives within max relative 1e-5 after 100 SSP-RK2
steps; forces within 1e-6 of the wall sum's magnitude."""
import copy

import numpy as np
import pytest

from conftest import base_cfg, mesh_of, perturbed_state
from test_gpu_parity import assert_noise_close

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["cell", "face"])
def stage_path(request):
    from eulersr_amd.solve import set_stage_path
    set_stage_path(request.param)
    yield request.param
    set_stage_path("auto")


def _subsonic_mesh():
    return _cached("subsonic", _make_subsonic)


def _make_subsonic():
    """diamond h=0.2 whose outlet faces are, every other one, subsonic
    inlets (marker 5) with a ghost (rho, u, v) of its own."""
    m = copy.copy(mesh_of("diamond", 0.2))
    m.__dict__.pop("_esr_device", None)
    fm = np.asarray(m.face_markers).copy()
    out = np.nonzero(fm == 4)[0]
    fm[out[0::2]] = 5
    m.face_markers = fm
    m.inlet_subsonic = np.array([1.1, 1.6, 0.2, 1.0])
    return m


def _bump():
    return mesh_of("bump", 0.05)


def _channel():
    from eulersr_amd.meshing import build_periodic_channel
    return _cached("channel", lambda: build_periodic_channel(40, 20))


_MESHES = {}


def _cached(key, make):
    if key not in _MESHES:
        _MESHES[key] = make()
    return _MESHES[key]


CASES = {
    "subsonic_inlet": (_subsonic_mesh, dict(Mach=1.8, aoa=3.0)),
    "periodic": (_channel, dict(Mach=1.6, aoa=4.0, case="periodic_channel")),
    "bump_subsonic": (_bump, dict(Mach=0.6, aoa=0.0, case="bump")),
    "bump_supersonic": (_bump, dict(Mach=1.6, aoa=0.0, case="bump")),
}


def _kw(inlet, flux, recon):
    return dict(gamma=1.4, M=1.0, reconstruction=recon, flux=flux, value=inlet, entropy=False)


def _marker_counts(mesh):
    fm = np.asarray(mesh.face_markers)
    return {int(k): int((fm == k).sum()) for k in np.unique(fm)}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("flux", ["hllc", "roe", "ausm", "rusanov"])
@pytest.mark.parametrize("recon", ["constant", "muscl"])
def test_residual_matches_oracle(case, flux, recon):
    from eulersr_amd import euler_solver
    from oracle import fvm as F
    make, over = CASES[case]
    mesh = make()
    cfg = base_cfg(**over)
    W, inlet = perturbed_state(mesh, cfg, seed=11)
    kw = _kw(inlet, flux, recon)
    print(f"case: residual {case} ({mesh.n_cells} cells, markers {_marker_counts(mesh)}) flux={flux} "
          f"recon={recon}", flush=True)
    Rg = euler_solver.residual(W, mesh, **kw)
    R32 = F.residual(W, F.omesh(mesh, np.float32), **kw)
    R64 = F.residual(W.astype(np.float64), F.omesh(mesh, np.float64), **dict(kw, value=inlet.astype(np.float64)))
    assert_noise_close(Rg, R32, R64, f"residual {case} {flux}/{recon}")


def test_cases_reach_their_boundary_kinds():
    """The meshes carry the markers they are meant to test, and the bump
    case switches the sponge off."""
    sub, ch, bu = _subsonic_mesh(), _channel(), _bump()
    assert _marker_counts(sub).get(5, 0) > 0 and _marker_counts(sub).get(4, 0) > 0
    assert _marker_counts(ch).get(1, 0) > 0 and np.all(np.asarray(ch.neighbors)[
        np.isin(np.asarray(ch.face_connectivity), np.nonzero(np.asarray(ch.face_markers) == 1)[0])] >= 0)
    mb = _marker_counts(bu)
    assert mb.get(2, 0) > 0 and mb.get(6, 0) > 0 and bu.metadata["case"] == "bump"
    assert list(bu.metadata["wall_markers"]) == [2, 6]
    from eulersr_amd.device import device_mesh
    dm = device_mesh(bu)
    assert dm.n_wall == mb[2]            # forces on the bump faces only


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("flux", ["hllc", "roe"])
def test_100_steps_primitives_within_1e5(case, flux):
    import torch
    from eulersr_amd import euler_solver
    from oracle import fvm as F
    make, over = CASES[case]
    mesh = make()
    cfg = base_cfg(**over)
    W0, inlet = perturbed_state(mesh, cfg, amp=0.02, seed=12)
    kw = _kw(inlet, flux, "muscl")
    om = F.omesh(mesh, np.float32)
    dt = F.get_dt(W0, om, CFL=0.5)
    Wg = torch.from_numpy(W0).cuda()
    Wr = W0.copy()
    print(f"case: 100 SRK2 steps {case} {flux} dt={dt}", flush=True)
    for _ in range(100):
        Wg = euler_solver.time_step_RK2_SSP(Wg, mesh, dt, **kw)
        Wr = F.time_step_RK2_SSP(Wr, om, dt, **kw)
    pg = F.getPrimitive(Wg.cpu().numpy().astype(np.float64))
    pr = F.getPrimitive(Wr.astype(np.float64))
    U = np.hypot(inlet[1], inlet[2]) / inlet[0]
    scale = np.array([1.0, U, U, 1.0])
    rel = np.abs(pg - pr) / np.maximum(np.abs(pr), scale)
    print("max rel per primitive:", rel.max(axis=0))
    assert np.all(np.isfinite(pg))
    assert rel.max() <= 1e-5


@pytest.mark.parametrize("scheme", ["EE", "RK4"])
def test_bump_step_each_scheme(scheme):
    from eulersr_amd import euler_solver
    from oracle import fvm as F
    mesh = _bump()
    cfg = base_cfg(Mach=0.7, aoa=0.0, case="bump")
    W, inlet = perturbed_state(mesh, cfg, seed=13)
    kw = _kw(inlet, "hllc", "muscl")
    om32 = F.omesh(mesh, np.float32)
    dt = F.get_dt(W, om32, CFL=0.5)
    print(f"case: bump one {scheme} step", flush=True)
    Wg = euler_solver.time_step(scheme, W, mesh, dt, **kw)
    W32 = F.SCHEMES[scheme](W, om32, dt, **kw)
    W64 = F.SCHEMES[scheme](W.astype(np.float64), F.omesh(mesh, np.float64), np.float64(dt),
                            **dict(kw, value=inlet.astype(np.float64)))
    assert_noise_close(Wg - W, W32 - W, W64 - W.astype(np.float64), f"bump step {scheme} increment")


def test_bump_forces_match_oracle():
    from eulersr_amd import helper
    from oracle import fvm as F
    mesh = _bump()
    cfg = base_cfg(Mach=0.7, aoa=0.0, case="bump")
    W, _ = perturbed_state(mesh, cfg, seed=14)
    U = 0.7 * np.sqrt(1.4)
    cd, cl = helper.get_force_coefficients(W, mesh, 1.0, U, 1.0)
    om = F.omesh(mesh, np.float32)
    rcd, rcl = F.forces(W, om, 1.0, U, 1.0)
    fids, cells, locs = F.wall_faces(om)
    assert len(fids) == _marker_counts(mesh)[2]
    P = F.getPrimitive(W.astype(np.float64))[cells, 3]
    q = 0.5 * U ** 2
    sx = np.sum(np.abs(P * om.normals[cells, locs, 0] * om.surface[fids])) / q
    sy = np.sum(np.abs(P * om.normals[cells, locs, 1] * om.surface[fids])) / q
    print(f"case: bump forces gpu ({cd}, {cl}) oracle ({rcd}, {rcl})", flush=True)
    assert abs(cd - rcd) <= 1e-6 * sx
    assert abs(cl - rcl) <= 1e-6 * sy


def test_batched_boundary_cases_equal_solo_runs():
    """The three meshes in one batched solver (graph-replayed chunks) give
    bit-identical states to each case solved alone."""
    from eulersr_amd import main as M
    from eulersr_amd.solve import BatchSolver
    specs, inits = [], []
    for case in ("subsonic_inlet", "periodic", "bump_subsonic"):
        make, over = CASES[case]
        mesh = make()
        cfg = base_cfg(**dict(over, CFL=0.5, tf=10.0))
        W, inlet = M.initialize(mesh, cfg)
        specs.append(M.case_spec(mesh, inlet, cfg))
        inits.append(W)
    kwargs = dict(scheme="SRK2", threshold=0.0, check_every=5, patience=1, max_steps=60)
    bs = BatchSolver(specs, **kwargs)
    for i, W in enumerate(inits):
        bs.set_state(i, W)
    bs.prepare()
    bs.run()
    batched = [bs.get_state(i).cpu().numpy() for i in range(len(specs))]
    bs.close()
    for i, (sp, W) in enumerate(zip(specs, inits)):
        one = BatchSolver([sp], **kwargs)
        one.set_state(0, W)
        one.prepare()
        one.run()
        single = one.get_state(0).cpu().numpy()
        one.close()
        print(f"case {i}: batched vs solo", flush=True)
        assert np.all(np.isfinite(single))
        assert np.array_equal(batched[i].view(np.uint32), single.view(np.uint32))
