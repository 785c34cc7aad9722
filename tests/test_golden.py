"""Golden fixture tests (tests/golden/diamond_h0.3.npz, written by
tests/golden/make_golden.py from the CPU oracle; the reference has no golden
vectors of its own and cannot run here).

CPU: the mesher reproduces the stored mesh, the mesh derivation its geometry
and connectivity bit for bit, the C++ and oracle RCM its permutation, and the
oracle its residuals, steps, time step and forces. GPU (through the C ABI, on
both stage paths): the residual of every flux x reconstruction, the entropy
flag, one step of each scheme, get_dt and CD/CL against the stored float32
and float64 oracle vectors, with the parity tolerance of test_gpu_parity.py
(4x the oracle's own float32-vs-float64 difference, floor 1e-6 of the
magnitude)."""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

FIXTURE = Path(__file__).resolve().parent / "golden" / "diamond_h0.3.npz"
FLUXES = ["hllc", "roe", "rusanov", "lf", "ausm", "tadmor"]
RECONS = ["constant", "muscl"]
SCHEMES = ["EE", "RK2", "SRK2", "RK4"]
MACH = 1.8


@pytest.fixture(scope="module")
def gold():
    with np.load(FIXTURE, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def gmesh(gold):
    from eulersr_amd.mesh import Mesh
    md = json.loads(bytes(gold["mesh_metadata"]).decode())
    return Mesh.from_points(gold["mesh_points"], gold["mesh_tris"], gold["mesh_neighbors"], gold["mesh_faces"],
                            gold["mesh_face_markers"], metadata=md)


def _kw(gold, flux, recon, dtype=np.float32, **extra):
    kw = dict(gamma=1.4, M=1.0, reconstruction=recon, flux=flux, value=gold["inlet"].astype(dtype), entropy=False)
    kw.update(extra)
    return kw


def _bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _close(a, b, rel):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return a.shape == b.shape and float(np.max(np.abs(a - b))) <= rel * max(float(np.max(np.abs(b))), 1e-300)


# -- CPU ----------------------------------------------------------------------

def test_mesher_reproduces_the_stored_mesh(gold):
    from eulersr_amd.meshing import build_mesh
    m = build_mesh("diamond", 0.3)
    for k in ("points", "tris", "neighbors", "faces", "face_markers"):
        assert _bits(np.asarray(getattr(m, k)), gold[f"mesh_{k}"]), k


def test_geometry_and_connectivity_bit_exact(gold, gmesh):
    for k in ("barycenter", "area", "surface", "normals", "face_connectivity"):
        assert _bits(np.asarray(getattr(gmesh, k)), gold[f"geo_{k}"]), k


def test_rcm_permutation_bit_exact(gold, gmesh):
    from eulersr_amd import _lib as L
    from oracle import rcm as RCM
    # the fixture's order seeds at the lowest-index cell of least degree
    stable = np.argsort((np.asarray(gmesh.neighbors) >= 0).sum(axis=1), kind="stable")
    assert np.array_equal(np.asarray(RCM.rcm(gmesh.neighbors, stable), dtype=np.int32), gold["rcm_perm"])
    nb = np.ascontiguousarray(gmesh.neighbors, dtype=np.int32)
    perm = np.empty(gmesh.n_cells, dtype=np.int32)
    assert L.lib().esr_rcm(gmesh.n_cells, nb.ctypes.data_as(C.POINTER(C.c_int32)), None,
                           perm.ctypes.data_as(C.POINTER(C.c_int32))) == 0
    assert np.array_equal(perm, gold["rcm_perm"])


@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("recon", RECONS)
def test_oracle_reproduces_golden_residuals(gold, gmesh, flux, recon):
    from oracle import fvm as F
    W = gold["W"]
    R32 = F.residual(W, F.omesh(gmesh, np.float32), **_kw(gold, flux, recon))
    R64 = F.residual(W.astype(np.float64), F.omesh(gmesh, np.float64), **_kw(gold, flux, recon, np.float64))
    # float64: libm/SIMD transcendental paths may differ by an ulp between hosts
    assert _close(R64, gold[f"R64_{flux}_{recon}"], 1e-12)
    assert _close(R32, gold[f"R32_{flux}_{recon}"], 1e-5)


def test_oracle_reproduces_golden_steps_dt_forces(gold, gmesh):
    from oracle import fvm as F
    W = gold["W"]
    om32, om64 = F.omesh(gmesh, np.float32), F.omesh(gmesh, np.float64)
    dt = F.get_dt(W, om32, CFL=0.5)
    assert np.float32(dt) == gold["dt"]
    for sc in SCHEMES:
        s32 = F.SCHEMES[sc](W, om32, dt, **_kw(gold, "hllc", "muscl"))
        s64 = F.SCHEMES[sc](W.astype(np.float64), om64, np.float64(dt), **_kw(gold, "hllc", "muscl", np.float64))
        assert _close(s64, gold[f"S64_{sc}"], 1e-12), sc
        assert _close(s32, gold[f"S32_{sc}"], 1e-6), sc
    cd, cl = F.forces(W, om32, 1.0, MACH * np.sqrt(1.4), 1.0)
    assert np.allclose([cd, cl], gold["forces"], rtol=1e-12, atol=1e-14)


def test_golden_residuals_are_nontrivial(gold):
    """Every stored residual is finite and non-zero, and the fluxes differ
    from each other (a fixture of zeros or copies would pin nothing)."""
    seen = []
    for flux in FLUXES:
        for recon in RECONS:
            r = gold[f"R64_{flux}_{recon}"]
            assert np.all(np.isfinite(r)) and np.max(np.abs(r)) > 0
            for s in seen:
                assert not np.allclose(r, s)
            seen.append(r)


# -- GPU ----------------------------------------------------------------------

@pytest.fixture(params=["cell", "face"])
def stage_path(request):
    from eulersr_amd.solve import get_stage_path, set_stage_path
    set_stage_path(request.param)
    assert get_stage_path() == request.param
    yield request.param
    set_stage_path("auto")


@pytest.mark.gpu
@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("recon", RECONS)
def test_gpu_residual_matches_golden(gold, gmesh, stage_path, flux, recon):
    from eulersr_amd import euler_solver
    from test_gpu_parity import assert_noise_close
    print(f"case: golden residual diamond h=0.3 flux={flux} recon={recon} path={stage_path}", flush=True)
    Rg = euler_solver.residual(gold["W"], gmesh, **_kw(gold, flux, recon))
    assert_noise_close(Rg, gold[f"R32_{flux}_{recon}"], gold[f"R64_{flux}_{recon}"], f"golden {flux}/{recon}")


@pytest.mark.gpu
@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_gpu_entropy_residual_matches_golden(gold, gmesh, stage_path, flux):
    from eulersr_amd import euler_solver
    from test_gpu_parity import assert_noise_close
    print(f"case: golden entropy residual flux={flux} path={stage_path}", flush=True)
    Rg = euler_solver.residual(gold["W"], gmesh, **_kw(gold, flux, "muscl", entropy=True))
    assert_noise_close(Rg, gold[f"RE32_{flux}"], gold[f"RE64_{flux}"], f"golden entropy {flux}")


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", SCHEMES)
def test_gpu_step_matches_golden(gold, gmesh, stage_path, scheme):
    from eulersr_amd import euler_solver
    from test_gpu_parity import assert_noise_close
    W = gold["W"]
    dt = float(gold["dt"])
    print(f"case: golden {scheme} step dt={dt} path={stage_path}", flush=True)
    Wg = euler_solver.time_step(scheme, W, gmesh, np.float32(dt), **_kw(gold, "hllc", "muscl"))
    assert_noise_close(Wg - W, gold[f"S32_{scheme}"] - W, gold[f"S64_{scheme}"] - W.astype(np.float64),
                       f"golden step {scheme}")


@pytest.mark.gpu
def test_gpu_dt_and_forces_match_golden(gold, gmesh):
    from eulersr_amd import helper
    W = gold["W"]
    dg = float(helper.get_dt(W, gmesh, CFL=0.5, gamma=1.4, M=1.0))
    d0 = float(gold["dt"])
    print(f"case: golden get_dt gpu {dg!r} fixture {d0!r}", flush=True)
    assert abs(dg - d0) <= 2e-7 * abs(d0)
    U = MACH * np.sqrt(1.4)
    cd, cl = helper.get_force_coefficients(W, gmesh, 1.0, U, 1.0)
    rcd, rcl = gold["forces"]
    # rounding bound: 1e-6 of the sum of |P n ds| / (q L) (the wall sum cancels)
    from oracle import fvm as F
    om = F.omesh(gmesh, np.float32)
    fids, cells, locs = F.wall_faces(om)
    P = F.getPrimitive(W.astype(np.float64))[cells, 3]
    q = 0.5 * U ** 2
    sx = np.sum(np.abs(P * om.normals[cells, locs, 0] * om.surface[fids])) / q
    sy = np.sum(np.abs(P * om.normals[cells, locs, 1] * om.surface[fids])) / q
    print(f"case: golden forces gpu ({cd}, {cl}) fixture ({rcd}, {rcl}) scales ({sx}, {sy})", flush=True)
    assert abs(cd - rcd) <= 1e-6 * sx and abs(cl - rcl) <= 1e-6 * sy
