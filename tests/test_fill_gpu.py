"""The device-side seeded fill reproduces the oracle's splitmix64 fill bit for bit."""
import numpy as np
import pytest

from gemm_avx512f_amd.device import FILL_F64_RANGE, DeviceMatrix, fill_uniform
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fill_matches_oracle_stream(dtype):
    s = O.Stream(42)
    fill = s.f64 if dtype == np.float64 else s.f32
    a, b = fill(37, 23, 40), fill(50, 11)
    dA = DeviceMatrix.from_host(np.zeros((37, 23), dtype), ld=40)
    dB = DeviceMatrix(50, 11, dtype)
    nxt = fill_uniform(dA, 42, 0)
    fill_uniform(dB, 42, nxt)
    got_a = dA.to_host_padded()
    assert np.array_equal(got_a, a)                      # padding rows untouched (zero)
    assert np.array_equal(dB.to_host(), b)


def test_range_fill_matches_oracle():
    for r in (0, 8, 32):
        s = O.Stream(7 + r)
        x = s.f64(64, 33, r=r)
        d = DeviceMatrix(64, 33, np.float64)
        fill_uniform(d, 7 + r, 0, kind=FILL_F64_RANGE, range_r=r)
        assert np.array_equal(d.to_host(), x), r
