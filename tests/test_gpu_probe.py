"""The matrix-core throughput probe behind the bench's instruction ceiling:
both FP16 shapes report a positive rate below the published dense peak
(2.5 PFLOP/s), and bad shapes are refused."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def test_mfma_rate_probe():
    from moldy_amd import engine as E
    rates = {}
    for shape in (16, 32):
        v = ctypes.c_double()
        assert E.lib().mgpu_mfma_rate(0, shape, ctypes.byref(v)) == 0
        rates[shape] = v.value
        assert 50.0 < v.value < 2600.0
    v = ctypes.c_double()
    assert E.lib().mgpu_mfma_rate(0, 8, ctypes.byref(v)) != 0
    print("FP16 MFMA TFLOP/s:", rates)
