"""Local time stepping (esr_solve_opts.local_dt, an acceleration mode with
no reference counterpart: every cell advances with its own CFL step from the
step's start state) against the oracle's restatement (oracle/fvm.py
local_dt fed to the reference steppers), one step of each scheme, with the
parity tolerance of test_gpu_parity.py."""
import numpy as np
import pytest

from conftest import base_cfg, mesh_of, perturbed_state
from test_gpu_parity import assert_noise_close

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme", ["EE", "RK2", "SRK2", "RK4"])
def test_local_dt_step_matches_oracle(scheme):
    from eulersr_amd import main as M
    from eulersr_amd.solve import BatchSolver
    from oracle import fvm as F
    mesh = mesh_of("diamond", 0.2)
    cfg = base_cfg(Mach=2.5, aoa=5.0)
    W, inlet = perturbed_state(mesh, cfg, seed=5)
    sp = M.case_spec(mesh, inlet, cfg)
    print(f"case: local-dt {scheme} step diamond h=0.2 cfl {sp.cfl}", flush=True)
    bs = BatchSolver([sp], scheme=scheme, threshold=0.0, check_every=5, patience=1, local_dt=True)
    try:
        bs.set_state(0, W)
        bs.prepare()
        bs.run(1)
        Wg = bs.get_state(0).cpu().numpy()
    finally:
        bs.close()
    kw = dict(gamma=1.4, M=1.0, reconstruction="MUSCL", flux="HLLC", value=inlet, entropy=False)
    om32, om64 = F.omesh(mesh, np.float32), F.omesh(mesh, np.float64)
    d32 = F.local_dt(W, om32, CFL=sp.cfl)
    W32 = F.SCHEMES[scheme](W, om32, d32, **kw)
    W64 = F.SCHEMES[scheme](W.astype(np.float64), om64, d32.astype(np.float64),
                            **dict(kw, value=inlet.astype(np.float64)))
    assert_noise_close(Wg - W, W32 - W, W64 - W.astype(np.float64), f"local-dt step {scheme} increment")
