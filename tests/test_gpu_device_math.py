"""The square root and logarithm of the flux kernels (sound speeds, Roe
averages, entropy jumps) are the correctly rounded sqrtf and the library
logf, bit for bit, on every non-negative float of at least 2^-96: an
exhaustive comparison over all 2^31 non-negative bit patterns on the GPU
(esr_check_device_math). Below 2^-96 (tiny and denormal inputs, which those
quantities never take) they may differ."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def test_sqrt_and_log_match_libm_on_every_normal_input():
    import torch
    from eulersr_amd import _lib
    from eulersr_amd.device import stream_ptr
    torch.zeros(1, device="cuda")
    counts = (C.c_uint64 * 4)()
    print("case: exhaustive sqrt/log check over 2^31 patterns", flush=True)
    _lib.check(_lib.lib().esr_check_device_math(counts, stream_ptr()), "esr_check_device_math")
    print(f"mismatches: sqrt {counts[0]} (below 2^-96: {counts[1]}), log {counts[2]} (below 2^-96: {counts[3]})",
          flush=True)
    assert counts[0] == 0 and counts[2] == 0
    # the tiny range is where the two differ: the comparison ran there too
    assert counts[1] > 0 and counts[3] > 0
