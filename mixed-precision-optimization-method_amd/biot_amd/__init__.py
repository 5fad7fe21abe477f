"""MI355X-native solver of the 13-field staggered-grid Biot system (Python front end).

The compute path is lib/libbiot.so (hand-written gfx950 HIP kernels behind the C-ABI of
include/biot.h); this package only drives it.
"""
from ._lib import FIELDS, BiotError, lib, mask_of, ricker  # noqa: F401
from .solver import Solver  # noqa: F401
