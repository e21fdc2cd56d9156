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


def test_mfma_bf16_16x16x32_lane_layout(gpu_lib):
    out = np.zeros(768, dtype=np.float32)
    rc = gpu_lib.lib().bsahip_selftest_mfma_bf16_layout(out.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == 0
    lane = np.arange(64)
    d1, d2, d3 = (out[256 * i:256 * (i + 1)].reshape(4, 64) for i in range(3))
    for t in range(4):
        np.testing.assert_array_equal(d1[t], 4 * (lane >> 4) + t)      # D rows
        np.testing.assert_array_equal(d2[t], lane & 15)                 # k = 0..15 order
        np.testing.assert_array_equal(d3[t], (lane & 15) + 16)          # k = 16..31 order


def test_mfma_f32_16x16x4_lane_layout(gpu_lib):
    out = np.zeros(512, dtype=np.float32)
    rc = gpu_lib.lib().bsahip_selftest_mfma_f32_layout(out.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == 0
    lane = np.arange(64)
    d1, d2 = out[:256].reshape(4, 64), out[256:].reshape(4, 64)
    for t in range(4):
        row = 4 * (lane >> 4) + t                  # register t of lane l: D[4(l/16)+t][l%16]
        col = lane & 15
        np.testing.assert_array_equal(d1[t], row + 16 * (col % 4))
        np.testing.assert_array_equal(d2[t], 6 + 16 * col)
