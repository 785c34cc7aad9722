"""The drop-in boundary keeps JAX's functional semantics: no entry point
mutates its input state, whether it is a numpy array (copied to the device)
or a float32 CUDA tensor (used in place, zero copy) -- residual, every time
stepper, get_dt, the force coefficients, the relative change and main.run;
and a CUDA input gives the same result as the numpy one, bit for bit."""
import numpy as np
import pytest

from conftest import base_cfg, mesh_of, perturbed_state

pytestmark = pytest.mark.gpu


def _calls(mesh, inlet, dt):
    from eulersr_amd import euler_solver, helper
    kw = dict(gamma=1.4, M=1.0, reconstruction="MUSCL", flux="HLLC", value=inlet, entropy=False)
    U = 2.0 * np.sqrt(1.4)
    return {
        "residual": lambda W: euler_solver.residual(W, mesh, **kw),
        "EE": lambda W: euler_solver.time_step_Euler(W, mesh, dt, **kw),
        "RK2": lambda W: euler_solver.time_step_RK2(W, mesh, dt, **kw),
        "SRK2": lambda W: euler_solver.time_step_RK2_SSP(W, mesh, dt, **kw),
        "RK4": lambda W: euler_solver.time_step_RK4(W, mesh, dt, **kw),
        "get_dt": lambda W: np.float32(helper.get_dt(W, mesh, CFL=0.6)),
        "forces": lambda W: np.array(helper.get_force_coefficients(W, mesh, 1.0, U, 1.0)),
        "relative_change": lambda W: np.float64(helper.relative_change(W * 1.001, W, mesh)),
    }


def test_entry_points_do_not_mutate_inputs():
    import torch
    mesh = mesh_of("diamond", 0.2)
    cfg = base_cfg(Mach=2.0, aoa=3.0)
    W, inlet = perturbed_state(mesh, cfg, seed=11)
    dt = np.float32(1e-3)
    for name, fn in _calls(mesh, inlet, dt).items():
        print(f"case: {name}", flush=True)
        Wn = W.copy()
        out_np = fn(Wn)
        assert np.array_equal(Wn.view(np.uint32), W.view(np.uint32)), f"{name} mutated a numpy input"
        Wt = torch.from_numpy(W.copy()).cuda()
        ref = Wt.clone()
        out_t = fn(Wt)
        torch.cuda.synchronize()
        assert torch.equal(Wt, ref), f"{name} mutated a CUDA input"
        out_t = out_t.cpu().numpy() if isinstance(out_t, torch.Tensor) else out_t
        a, b = np.atleast_1d(np.asarray(out_np)), np.atleast_1d(np.asarray(out_t))
        assert a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8)), name


def test_run_does_not_mutate_its_initial_state():
    from eulersr_amd import main as M
    mesh = mesh_of("diamond", 0.2)
    cfg = base_cfg(Mach=2.0, aoa=3.0, tf=0.2, verbose=False, export={})
    W, inlet = M.initialize(mesh, cfg)
    W0 = W.copy()
    Wf = M.run(W, mesh, inlet, cfg, out_dirs=None)
    assert np.array_equal(W.view(np.uint32), W0.view(np.uint32))
    assert not np.array_equal(Wf, W0)
