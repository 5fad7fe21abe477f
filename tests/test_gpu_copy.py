"""GPU check of the streaming measurement (biot_copy_bandwidth): a plausible HBM rate for the
copies and the velocity access pattern, and argument errors as status codes."""
import ctypes as C

import pytest

from biot_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
def test_copy_bandwidth_rate(variant):
    g = C.c_double()
    assert _lib.lib().biot_copy_bandwidth(0, 256 * 2 ** 20, 5, variant, C.byref(g)) == 0
    print(f"copy variant {variant}: {g.value:.0f} GB/s", flush=True)
    assert 500.0 < g.value < 16000.0  # between a slow PCIe-bound copy and twice the HBM peak


def test_copy_bandwidth_arguments():
    L, g = _lib.lib(), C.c_double()
    for args in ((0, 100, 5, 0), (0, 0, 5, 0), (0, 1 << 20, 0, 0), (0, 1 << 20, 5, 6), (-1, 1 << 20, 5, 0),
                 (0, (1 << 30) + 16, 5, 2)):
        assert L.biot_copy_bandwidth(*args, C.byref(g)) != 0
    assert L.biot_copy_bandwidth(0, 1 << 20, 5, 0, None) != 0
