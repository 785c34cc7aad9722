"""Mesh convergence of the diamond on the GPU (postproc.mesh_convergence,
the reference's euler/postproc/mesh_convergence_diamond.py campaign) on the
short r = 2 chain h = 0.1, 0.05, 0.025 for two points: CD and CL of every
level, the wall-profile error to the finest level, and one Richardson / GCI
triplet each; at M = 2 the wall Mach profile converges (its error to the
finest level falls from h = 0.1 to 0.05) and CD approaches its Richardson
limit.

Measured on MI355X: CD observed order 2.20 (M = 2, alpha = 0) and 2.00
(alpha = 3), GCI_fine 1.1 % and 1.5 % (the reference reports orders 2.24 and
2.52 for the shock angle at M = 1.5 and 2.5, README.md:52-56)."""
import math

import pytest

pytestmark = pytest.mark.gpu

CHAIN = (0.1, 0.05, 0.025)
POINTS = (("P6_mach2_symmetric", 2.0, 0.0), ("P5_incidence", 2.0, 3.0))


def test_mesh_convergence_chain():
    from eulersr_amd import postproc as PP
    from eulersr_amd.device import require_gpu
    require_gpu()
    out = PP.mesh_convergence(points=POINTS, chain=CHAIN)
    for label, m, a in POINTS:
        res = out[label]
        rows = res["rows"]
        assert [r["h"] for r in rows] == list(CHAIN)
        for r in rows:
            print(f"{label} h={r['h']} cells={r['n_cells']} CD={r['cd']:.5f} CL={r['cl']:.5f} "
                  f"wall err={r['wall_profile_rel_err']}", flush=True)
            assert math.isfinite(r["cd"]) and math.isfinite(r["cl"])
        assert rows[-1]["wall_profile_rel_err"] == 0.0
        assert rows[0]["wall_profile_rel_err"] > rows[1]["wall_profile_rel_err"] > 0.0
        g = res["gci_cd"]
        print(f"{label} GCI(CD): {g}", flush=True)
        assert len(g) == 1
        if g[0]["p_obs"] is not None:
            assert g[0]["p_obs"] > 0
            # the extrapolated CD lies beyond the finest value, on the side the sequence moves to
            f2, f3 = rows[1]["cd"], rows[2]["cd"]
            assert (g[0]["extrap"] - f3) * (f3 - f2) >= 0
    # symmetric point: no lift at any level
    assert all(abs(r["cl"]) < 5e-3 for r in out["P6_mach2_symmetric"]["rows"])
