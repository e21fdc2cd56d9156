"""GPU self test of the matrix-core lane layout the FP64 contraction assumes."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mfma_f64_16x16x4_lane_layout(gpu_lib):
    out = np.zeros(512)
    rc = gpu_lib.lib().bsahip_selftest_mfma_layout(out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    rows = out[:256].reshape(4, 64)
    cols = out[256:].reshape(4, 64)
    lane = np.arange(64)
    for t in range(4):
        np.testing.assert_array_equal(rows[t], (lane >> 4) + 4 * t)
        np.testing.assert_array_equal(cols[t], lane & 15)
