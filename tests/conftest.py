"""Shared fixtures. GPU tests are marked ``gpu``; CPU tests check the oracle
against closed forms and invariants, the host mesh code and the C ABI."""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (gfx950) and the built libeulersr.so")


_MESHES = {}


def mesh_of(geom, h, **kw):
    key = (geom, h, tuple(sorted(kw.items())))
    if key not in _MESHES:
        from eulersr_amd.meshing import build_mesh
        _MESHES[key] = build_mesh(geom, h, **kw)
    return _MESHES[key]


@pytest.fixture(scope="session")
def diamond02():
    return mesh_of("diamond", 0.2)


@pytest.fixture(scope="session")
def diamond01():
    return mesh_of("diamond", 0.1)


@pytest.fixture(scope="session")
def naca02():
    return mesh_of("naca0012", 0.2)


def base_cfg(**over):
    cfg = dict(rho_inf=1.0, p_inf=1.0, gamma=1.4, Mach=2.5, aoa=5.0, case="diamond", time_scheme="SRK2",
               flux="HLLC", reconstruction="MUSCL", CFL=0.6, tf=4.0, stationarity_threshold=1e-6,
               stationarity_check_every=5, stationarity_patience=1)
    cfg.update(over)
    return cfg


def perturbed_state(mesh, cfg, amp=0.05, seed=0):
    """Freestream plus a smooth deterministic perturbation (exercises every
    flux branch and the limiter), float32 conserved."""
    from oracle import fvm as F
    W, inlet = F.initialize(mesh.n_cells, cfg, np.float32)
    prim = F.getPrimitive(W.astype(np.float64), gamma=cfg["gamma"])
    x, y = mesh.barycenter[:, 0], mesh.barycenter[:, 1]
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * np.pi, size=(4, 2))
    for v in range(4):
        s = np.sin(2.1 * x + ph[v, 0]) * np.cos(1.7 * y + ph[v, 1])
        scale = 1.0 if v in (0, 3) else np.hypot(prim[0, 1], prim[0, 2])
        prim[:, v] = prim[:, v] + amp * scale * s
    W = F.getConserved(prim, gamma=cfg["gamma"]).astype(np.float32)
    return W, inlet
