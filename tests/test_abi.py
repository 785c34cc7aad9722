"""The C-ABI library loads and exports every symbol that include/eulersr.h
declares; the product path has no CPU fallback."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "eulersr.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(esr_[a-z_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("esr_mesh_create", "esr_residual", "esr_step", "esr_get_dt", "esr_forces",
              "esr_solver_create", "esr_solver_run", "esr_last_error", "esr_rcm"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from eulersr_amd import _lib
    lib = _lib.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes table covers the header exactly
    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_version_and_error_string():
    from eulersr_amd import _lib
    L = _lib.lib()
    assert L.esr_version() >= 100
    # a host-only call that fails sets the error string
    st = L.esr_rcm(0, None, None, None)
    assert st == _lib.ESR_ERR_VALUE
    assert b"bad arguments" in L.esr_last_error()


def test_product_path_refuses_without_gpu(diamond02):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from eulersr_amd import euler_solver
    W = np.ones((diamond02.n_cells, 4), dtype=np.float32)
    with pytest.raises(RuntimeError):
        euler_solver.residual(W, diamond02, flux="hllc", reconstruction="muscl")


def test_option_errors_match_reference():
    from eulersr_amd.euler_solver import make_opts
    with pytest.raises(ValueError, match="Unknown reconstruction"):
        make_opts(dict(reconstruction="weno"))
    with pytest.raises(ValueError, match="Unknown flux"):
        make_opts(dict(flux="godunov"))
    o = make_opts(dict(flux="HLL", reconstruction="second_order"))
    from eulersr_amd import _lib
    assert o.flux == _lib.FLUX_HLLC and o.recon == _lib.RECON_MUSCL


def test_stage_path_setting_is_host_only_and_validated():
    from eulersr_amd import _lib
    from eulersr_amd.solve import get_stage_path, set_stage_path
    L = _lib.lib()
    assert L.esr_set_stage_path(7) == 1          # ESR_ERR_VALUE
    assert b"stage path" in L.esr_last_error()
    for p in ("face", "cell", "auto"):
        set_stage_path(p)
        assert get_stage_path() == ("cell" if p == "auto" else p)
