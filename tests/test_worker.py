"""The subprocess worker (eulersr_amd.worker, counterpart of
eval/_fvm_solve_worker.py): its arguments and run configuration, and the
error paths of the reference, which are decided on the CPU before any solve
(a field of the wrong shape: ValueError; --probe-only / --forces-only
without a field: SystemExit)."""
import numpy as np
import pytest

from eulersr_amd import worker as WK

BASE = ["--case", "diamond", "--mach", "2.0", "--aoa", "3", "--tf", "0.5", "--stationarity-threshold", "1e-5"]


@pytest.fixture(scope="module")
def mesh_file(tmp_path_factory):
    from eulersr_amd.meshing import build_mesh
    m = build_mesh("diamond", 0.3)
    return m, m.save(tmp_path_factory.mktemp("mesh") / "diamond_h0.3.npz")


def _argv(mesh_path, scratch, *extra):
    return BASE + ["--mesh-path", str(mesh_path), "--scratch-dir", str(scratch), *extra]


def test_arguments_and_configuration(mesh_file, tmp_path):
    _, path = mesh_file
    a = WK.build_parser().parse_args(_argv(path, tmp_path))
    assert a.stationarity_patience == 1 and a.cd_tol == 0.001 and a.cl_tol == 0.001
    assert a.stationarity_check_every is None and a.engineering_patience is None
    cfg = WK.configure(a)
    assert cfg["stationarity_check_every"] == 10000 and cfg["verbose"] is False
    assert cfg["Mach"] == 2.0 and cfg["aoa"] == 3.0 and cfg["tf"] == 0.5 and cfg["stationarity_threshold"] == 1e-5
    assert "engineering_target_cd" not in cfg
    assert cfg["export"]["summary"] is False and cfg["export"]["bundle"] is False
    a = WK.build_parser().parse_args(_argv(path, tmp_path, "--stationarity-check-every", "5",
                                           "--stationarity-patience", "3", "--target-cd", "0.02",
                                           "--target-cl", "0.1", "--emit-summary"))
    cfg = WK.configure(a)
    assert cfg["stationarity_check_every"] == 5 and cfg["stationarity_patience"] == 3
    assert cfg["engineering_target_cd"] == 0.02 and cfg["engineering_target_cl"] == 0.1
    assert cfg["engineering_patience"] == 3 and cfg["export"]["summary"] is True
    # one target alone does not switch the engineering criterion on
    cfg = WK.configure(WK.build_parser().parse_args(_argv(path, tmp_path, "--target-cd", "0.02")))
    assert "engineering_target_cd" not in cfg


def test_required_arguments():
    with pytest.raises(SystemExit):
        WK.build_parser().parse_args(["--case", "diamond"])


@pytest.mark.parametrize("flag", ["--probe-only", "--forces-only"])
def test_probe_and_forces_need_a_field(mesh_file, tmp_path, flag):
    _, path = mesh_file
    with pytest.raises(SystemExit) as e:
        WK.main(_argv(path, tmp_path, flag))
    assert "--init-field" in str(e.value)


def test_field_of_the_wrong_shape_is_refused(mesh_file, tmp_path):
    m, path = mesh_file
    bad = tmp_path / "bad.npy"
    np.save(bad, np.ones((m.n_cells + 1, 4), dtype=np.float32))
    with pytest.raises(ValueError, match="init-field"):
        WK.main(_argv(path, tmp_path, "--init-field", str(bad), "--probe-only"))
