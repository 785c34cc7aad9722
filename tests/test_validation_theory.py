"""CPU pins of the shock-expansion theory used to validate the diamond case
(eulersr_amd.validation): textbook values (Anderson, Modern Compressible
Flow, oblique-shock and Prandtl-Meyer tables for gamma = 1.4), inverse
round trips, the maximum-deflection limit, and the panel-pressure closed
form of the diamond's drag through the oracle's wall integration."""
import math

import numpy as np
import pytest

from eulersr_amd import validation as V


def test_prandtl_meyer_table_values():
    assert V.prandtl_meyer(1.0) == 0.0
    assert math.degrees(V.prandtl_meyer(2.0)) == pytest.approx(26.38, abs=0.01)
    assert math.degrees(V.prandtl_meyer(3.0)) == pytest.approx(49.76, abs=0.01)
    for M in (1.01, 1.3, 2.0, 3.7, 8.0):
        assert V.inverse_prandtl_meyer(V.prandtl_meyer(M)) == pytest.approx(M, rel=1e-12)
    with pytest.raises(ValueError):
        V.inverse_prandtl_meyer(math.radians(131.0))     # nu_max = 130.45 deg


def test_oblique_shock_table_values():
    # M = 2, theta = 10 deg: beta = 39.31 deg, p2/p1 = 1.7066, M2 = 1.6405
    b = V.shock_angle(2.0, math.radians(10.0))
    assert math.degrees(b) == pytest.approx(39.31, abs=0.01)
    M2, pr, ok, beta = V.turn(2.0, math.radians(10.0))
    assert ok and beta == b
    assert pr == pytest.approx(1.7066, abs=2e-4)
    assert M2 == pytest.approx(1.6405, abs=2e-3)
    # M = 3, theta = 20 deg: beta = 37.76 deg (weak), 82.15 deg (strong)
    assert math.degrees(V.shock_angle(3.0, math.radians(20.0))) == pytest.approx(37.76, abs=0.02)
    assert math.degrees(V.shock_angle(3.0, math.radians(20.0), weak=False)) == pytest.approx(82.15, abs=0.05)


def test_shock_angle_inverts_the_deflection_relation():
    for M in (1.2, 1.5, 2.0, 2.5, 4.0):
        for th_deg in (0.5, 2.0, 5.0, 8.0, 12.0):
            th = math.radians(th_deg)
            for weak in (True, False):
                b = V.shock_angle(M, th, weak=weak)
                if b is None:
                    continue
                assert V.deflection_of(M, b) == pytest.approx(th, abs=1e-10)
                assert math.asin(1.0 / M) < b < math.pi / 2


def test_maximum_deflection_and_mach_angle_limit():
    # M = 2: theta_max = 22.97 deg
    assert V.shock_angle(2.0, math.radians(22.9)) is not None
    assert V.shock_angle(2.0, math.radians(23.05)) is None
    # weak shock of vanishing strength -> Mach wave
    assert V.shock_angle(2.5, 1e-7) == pytest.approx(math.asin(1 / 2.5), abs=1e-6)
    # a brute-force maximum of theta(beta) agrees with the attach/detach edge
    betas = np.linspace(math.asin(1 / 1.6) + 1e-6, math.pi / 2 - 1e-6, 200001)
    tmax = max(V.deflection_of(1.6, b) for b in betas[::50])
    assert V.shock_angle(1.6, tmax * 0.999) is not None
    assert V.shock_angle(1.6, tmax * 1.001) is None


def test_expansion_is_isentropic():
    M2, pr, ok, beta = V.turn(2.0, -math.radians(10.0))
    assert ok and beta is None
    assert math.degrees(V.prandtl_meyer(M2)) == pytest.approx(36.38, abs=0.01)
    # p0 constant: p2/p1 = (T2/T1)^(g/(g-1))
    t = (1 + 0.2 * 4.0) / (1 + 0.2 * M2 * M2)
    assert pr == pytest.approx(t ** 3.5, rel=1e-12)


def test_diamond_panels_symmetry_and_cascade():
    p = V.diamond_panels(2.0, 0.0, 5.0)
    assert p["upper_front"].Cp == pytest.approx(p["lower_front"].Cp, rel=1e-14)
    assert p["upper_rear"].Cp == pytest.approx(p["lower_rear"].Cp, rel=1e-14)
    assert p["upper_front"].Cp > 0 > p["upper_rear"].Cp
    # the rear panel state is the front state expanded by 2 eps
    M2, pr, _, _ = V.turn(p["upper_front"].M, -math.radians(10.0))
    assert p["upper_rear"].M == pytest.approx(M2, rel=1e-14)
    assert p["upper_rear"].p_ratio == pytest.approx(p["upper_front"].p_ratio * pr, rel=1e-14)
    # incidence: more compression below, less above
    q = V.diamond_panels(2.0, 3.0, 5.0)
    assert q["lower_front"].Cp > p["lower_front"].Cp > q["upper_front"].Cp
    # alpha > eps: the upper front panel expands
    r = V.diamond_panels(2.5, 6.0, 5.0)
    assert r["upper_front"].deflection < 0 and r["upper_front"].beta is None
    # detached: a deflection above the maximum at M = 1.05
    assert not all(x.attached for x in V.diamond_panels(1.05, 0.0, 5.0).values())


def test_theory_drag_closed_form_through_wall_integration():
    """Uniform panel pressures on a symmetric diamond at alpha = 0: the wall
    integral of the theory state gives CD = (Cp_front - Cp_rear) t/c and
    CL = 0 (up to the faceted wall)."""
    from oracle import fvm as F
    from eulersr_amd.meshing import build_mesh
    mesh = build_mesh("diamond", 0.05)
    eps = V.diamond_eps_deg(mesh)
    assert eps == pytest.approx(5.0, abs=0.01)
    panels = V.diamond_panels(2.0, 0.0, eps)
    W = V.theory_state(mesh, panels)
    om = F.omesh(mesh, np.float64)
    U = 2.0 * math.sqrt(1.4)
    cd, cl = F.forces(W, om, 1.0, U, 1.0)
    t_c = mesh.metadata["height"] / mesh.metadata["chord"]
    assert cd == pytest.approx((panels["upper_front"].Cp - panels["upper_rear"].Cp) * t_c, rel=2e-3)
    assert abs(cl) < 1e-3 * abs(cd)
