"""Post-processing (eulersr_amd/postproc.py) on the CPU, pinned by closed
forms, brute force and invariants: Richardson / GCI recover the exact order
and limit of a power-law sequence; the wall profile of a uniform field is the
freestream everywhere along a monotone arc length over the lower wall faces;
the owner lookup matches a brute-force scan; total entropy of uniform fields
has a closed form."""
import math

import numpy as np
import pytest

from conftest import mesh_of
from eulersr_amd import postproc as PP


def test_gci_recovers_power_law_order_and_limit():
    # f(h) = f0 + C h^p is exactly what Richardson extrapolation assumes:
    # a dropped log, a swapped triplet or a wrong ratio breaks these values
    f0, C, p = 0.0123, 0.7, 2.0
    chain = [0.4, 0.2, 0.1, 0.05]
    vals = {h: f0 + C * h ** p for h in chain}
    out = PP.richardson_gci(vals, chain)
    assert len(out) == 2
    for row in out:
        assert row["p_obs"] == pytest.approx(p, abs=1e-9)
        assert row["extrap"] == pytest.approx(f0, abs=1e-12)
        f2, f3 = vals[row["h_mid"]], vals[row["h_fine"]]
        assert row["gci_fine_pct"] == pytest.approx(1.25 * abs((f3 - f2) / f3) / (2 ** p - 1) * 100, rel=1e-9)
    # p = 1.5 on a ratio-2 chain
    vals = {h: 1.0 - 0.2 * h ** 1.5 for h in chain}
    assert PP.richardson_gci(vals, chain)[0]["p_obs"] == pytest.approx(1.5, abs=1e-9)
    # oscillatory convergence has no observed order; missing levels are skipped
    osc = {0.4: 1.0, 0.2: 1.2, 0.1: 1.1, 0.05: 1.15}
    assert all(r["p_obs"] is None for r in PP.richardson_gci(osc, chain))
    assert PP.richardson_gci({0.4: 1.0, 0.2: 1.1}, chain) == []


def test_profile_error_closed_forms():
    s = np.linspace(0.0, 1.0, 50)
    ref = dict(s=s, mach=1.0 + s)
    assert PP.relative_profile_error(ref, dict(s=s, mach=1.0 + s)) == 0.0
    assert PP.relative_profile_error(ref, dict(s=s, mach=1.1 * (1.0 + s))) == pytest.approx(0.1, rel=1e-12)
    assert PP.relative_profile_error(ref, dict(s=s[:1], mach=s[:1])) is None
    assert PP.relative_profile_error(ref, dict(s=s + 2.0, mach=1.0 + s)) is None


@pytest.mark.parametrize("geom", ["diamond", "naca0012"])
def test_face_owner_matches_brute_force(geom):
    m = mesh_of(geom, 0.2)
    fc = np.asarray(m.face_connectivity)
    ids = np.arange(len(m.faces))[::7]
    cells, local = PP.face_owner_data(fc, ids)
    for f, c, k in zip(ids, cells, local):
        hits = np.argwhere(fc == f)
        assert (c, k) == tuple(hits[0])


@pytest.mark.parametrize("geom", ["diamond", "naca0012"])
def test_wall_profile_of_a_uniform_field(geom):
    m = mesh_of(geom, 0.1)
    M_inf, aoa = 1.7, 2.0
    c = math.sqrt(1.4 * 1.0 / 1.0)
    prim = np.array([1.0, M_inf * c * math.cos(math.radians(aoa)), M_inf * c * math.sin(math.radians(aoa)), 1.0])
    E = prim[3] / 0.4 + 0.5 * prim[0] * (prim[1] ** 2 + prim[2] ** 2)
    W = np.tile([prim[0], prim[0] * prim[1], prim[0] * prim[2], E], (m.n_cells, 1))
    wp = PP.get_wall_profile(W, m)
    ids = wp["face_ids"]
    assert ids.size > 0
    fm = np.asarray(m.face_markers)
    assert np.all(fm[ids] == 2)
    mid = np.asarray(m.points)[np.asarray(m.faces)[ids]].mean(axis=1)
    pts = np.asarray(m.points)
    assert np.all(mid[:, 1] <= 0.5 * (pts[:, 1].min() + pts[:, 1].max()))
    # every lower wall face, ordered by x, arc length along the midpoints
    lower = np.nonzero((fm == 2) & (np.asarray(m.points)[np.asarray(m.faces)].mean(axis=1)[:, 1]
                                    <= 0.5 * (pts[:, 1].min() + pts[:, 1].max())))[0]
    assert sorted(ids.tolist()) == sorted(lower.tolist())
    assert np.all(np.diff(wp["x"]) >= 0)
    assert wp["s"][0] == 0.0 and np.all(np.diff(wp["s"]) > 0)
    assert np.allclose(wp["mach"], M_inf, rtol=1e-12)
    assert np.allclose(wp["pressure"], 1.0, rtol=1e-12)


def test_entropy_of_uniform_fields():
    m = mesh_of("diamond", 0.2)
    A = float(np.sum(np.asarray(m.area, dtype=np.float64)))

    def uniform(rho, P):
        return np.tile([rho, 0.3 * rho, -0.1 * rho, P / 0.4 + 0.5 * rho * 0.1], (m.n_cells, 1))

    W0, W1 = uniform(1.0, 1.0), uniform(1.3, 1.9)
    assert PP.get_entropy_creation(W0, W0, m) == 0.0
    s1 = math.log(1.9 / 1.3 ** 1.4)
    assert PP.get_total_entropy(W1, m) == pytest.approx(1.3 * s1 * A / 0.4, rel=1e-12)
    assert PP.get_entropy_creation(W0, W1, m) == pytest.approx(1.3 * s1 * A / 0.4 - 0.0, rel=1e-12)
    prim = PP.primitives(W1)
    assert np.allclose(PP.get_mach_number(prim), math.hypot(0.3, 0.1) / math.sqrt(1.4 * 1.9 / 1.3))
