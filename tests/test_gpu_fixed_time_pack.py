"""Packed fixed-time solves (pipeline.fixed_time(pack=True)): the cases of a
kind run in sub-groups ordered by their step count, and sub-groups whose
cases are all done drop out of the launches between chunks. Every case must
end with the state, step count, CD and CL (bitwise) it ends with in the
unpacked solve, where every case is computed until the last one ends. A
ragged batch: diamond and NACA0012 at h = 0.1, M from 0.7 to 2.9 (so step
counts differ by about 2x within each kind), HLLC and Roe, with a sub-group
size and chunk small enough that several sub-groups drop out mid-solve."""
import numpy as np
import pytest

from conftest import base_cfg, mesh_of

pytestmark = pytest.mark.gpu


def test_packed_fixed_time_equals_unpacked(monkeypatch):
    from eulersr_amd import main as M
    from eulersr_amd import pipeline as P
    specs, inits = [], []
    rng = np.random.default_rng(3)
    for i in range(24):
        geom = ("diamond", "naca0012")[i % 2]
        mach = round(0.7 + 0.1 * int(rng.integers(0, 23)), 1)
        flux = "HLLC" if (i // 2) % 2 == 0 else "ROE"
        mesh = mesh_of(geom, 0.1)
        cfg = base_cfg(Mach=mach, aoa=float(rng.integers(-5, 6)), case=geom, flux=flux, tf=0.5)
        W, inlet = M.initialize(mesh, cfg)
        specs.append(M.case_spec(mesh, inlet, cfg))
        inits.append(W)
    monkeypatch.setattr(P, "PACK_SIZE", 3)
    monkeypatch.setattr(P, "PACK_CHUNK", 40)
    print(f"case: 24 cases, steps keys {[round(P._steps_key(sp), 2) for sp in specs[:6]]}...", flush=True)
    ra, sa = P.fixed_time(specs, inits)
    rb, sb = P.fixed_time(specs, inits, pack=True)
    steps = [r["steps_run"] for r in ra]
    print(f"steps run per case: min {min(steps)} max {max(steps)}", flush=True)
    assert max(steps) > 1.5 * min(steps)
    for i in range(len(specs)):
        assert rb[i]["steps_run"] == ra[i]["steps_run"] == ra[i]["N"]
        assert rb[i]["cd"] == ra[i]["cd"] and rb[i]["cl"] == ra[i]["cl"]
        assert np.array_equal(sa[i].cpu().numpy().view(np.uint32), sb[i].cpu().numpy().view(np.uint32)), i
