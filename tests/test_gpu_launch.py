"""Launch configuration and the autotuner: the block size of the stage
kernels never changes a result (bitwise), the two stage paths agree to
rounding, and the autotuner's choice reproduces the pinned configuration."""
import numpy as np
import pytest

from conftest import base_cfg, mesh_of

pytestmark = pytest.mark.gpu


def _cases():
    from eulersr_amd import main as M
    m1 = mesh_of("diamond", 0.2)
    m2 = mesh_of("naca0012", 0.2)
    specs, inits = [], []
    for mesh, mach, aoa, case in [(m1, 2.0, 3.0, "diamond"), (m1, 0.85, -2.0, "diamond"),
                                  (m2, 1.2, 1.0, "naca0012")]:
        cfg = base_cfg(Mach=mach, aoa=aoa, case=case)
        W, inlet = M.initialize(mesh, cfg)
        specs.append(M.case_spec(mesh, inlet, cfg))
        inits.append(W)
    return specs, inits


def _run(specs, inits, launch=None, autotune=False, steps=40):
    from eulersr_amd.solve import BatchSolver
    bs = BatchSolver(specs, scheme="SRK2", threshold=0.0, check_every=5, patience=1, max_steps=steps,
                     autotune=autotune)
    for i, W in enumerate(inits):
        bs.set_state(i, W)
    bs.prepare()
    if launch is not None:
        bs.set_launch(*launch)
    bs.run()
    out = [bs.get_state(i).cpu().numpy() for i in range(len(specs))]
    tuning = bs.tuning
    bs.close()
    return out, tuning


def test_block_size_and_order_are_bit_invariant():
    specs, inits = _cases()
    for path in ("cell", "face"):
        ref, _ = _run(specs, inits, launch=(path, 256, True))
        for block in (64, 128, 256):
            for xcd in (True, False):
                if (block, xcd) == (256, True):
                    continue
                print(f"path {path} block {block} xcd order {xcd}", flush=True)
                got, _ = _run(specs, inits, launch=(path, block, xcd))
                for a, b in zip(got, ref):
                    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_paths_agree_to_rounding():
    specs, inits = _cases()
    cell, _ = _run(specs, inits, launch=("cell", 256))
    face, _ = _run(specs, inits, launch=("face", 256))
    for a, b in zip(cell, face):
        scale = np.abs(b).max(axis=0)
        assert np.all(np.abs(a - b).max(axis=0) <= 1e-5 * scale)


def test_autotune_matches_pinned_choice_and_is_cached():
    specs, inits = _cases()
    got, tuning = _run(specs, inits, autotune=True)
    print(f"autotune: {tuning}", flush=True)
    assert tuning["path"] in ("cell", "face") and tuning["block"] in (64, 128, 256)
    assert tuning["us_per_stage"] > 0
    ref, _ = _run(specs, inits, launch=(tuning["path"], tuning["block"], tuning["xcd_order"]))
    for a, b in zip(got, ref):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # a second solver of the same launch shape reuses the cached choice
    _, again = _run(specs, inits, autotune=True)
    assert again == tuning


def test_set_launch_rejects_bad_values():
    from eulersr_amd.solve import BatchSolver
    specs, inits = _cases()
    bs = BatchSolver(specs[:1], threshold=0.0, max_steps=5)
    with pytest.raises(ValueError):
        bs.set_launch("cell", 96)
    with pytest.raises(KeyError):
        bs.set_launch("edge", 64)
    bs.close()


def test_kernel_timings_leave_the_state_alone():
    """esr_solver_time_kernels and esr_solver_time_kernel_runs return a
    positive time per stage kernel and leave every case's state and the
    solve that follows bit for bit as they were."""
    from eulersr_amd.solve import BatchSolver
    specs, inits = _cases()
    ref, _ = _run(specs, inits, launch=("cell", 64, True), steps=20)
    bs = BatchSolver(specs, scheme="SRK2", threshold=0.0, check_every=5, patience=1, max_steps=20)
    for i, W in enumerate(inits):
        bs.set_state(i, W)
    bs.prepare()
    bs.set_launch("cell", 64, True)
    bs.run(10)
    before = [bs.get_state(i).cpu().numpy() for i in range(len(specs))]
    a = bs.time_kernels(5)
    b = bs.time_kernel_runs(5)
    print(f"in stage {a}, back to back {b}", flush=True)
    assert all(x > 0 and np.isfinite(x) for x in a + b)
    after = [bs.get_state(i).cpu().numpy() for i in range(len(specs))]
    for x, y in zip(before, after):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    bs.run()
    got = [bs.get_state(i).cpu().numpy() for i in range(len(specs))]
    bs.close()
    for x, y in zip(got, ref):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
