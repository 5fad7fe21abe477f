"""GPU-side set-up (biot_set_medium_device) and the device conversion of the reference
arrays into the kernels' coefficient arrays.

* The 23 reference arrays built on the GPU equal the host builder's bit for bit (the host
  builder is itself pinned bit-exact against the oracle), on a ragged grid for homogeneous,
  layered, perturbed and fluid-bearing media.
* The stored kernel arrays equal the host conversion (1/dx as float, 0.01 folded) rounded to
  the storage type: exactly for FP64, bit for bit in float for FP32; the 16-bit storage scales
  match between the two builders.
"""
import numpy as np
import pytest

from biot_amd import _lib, media
from biot_amd.solver import Solver

pytestmark = pytest.mark.gpu

NX, NY, NZ = 37, 30, 41


def _medium(kind):
    shape = (NZ, NY, NX)
    if kind == "homogeneous":
        return media.homogeneous(shape)
    if kind == "layered":
        return media.layered(shape)
    if kind == "perturbed":
        return media.perturbed(shape, corr_cells=3.0, seed=9)
    m = media.layered(shape)
    jj, ii = np.meshgrid(np.arange(NY), np.arange(NX), indexing="ij")
    hole = ((ii - NX // 2) ** 2 + (jj - NY // 2) ** 2) <= 6
    for k in ("mu", "Kb"):
        m[k][:, hole] = 0.0
    return m


@pytest.mark.parametrize("kind", ["homogeneous", "layered", "perturbed", "borehole"])
@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp16"])
def test_device_builder_matches_host_builder(kind, precision):
    print(f"device builder {kind} {precision}", flush=True)
    m = _medium(kind)
    w = np.array([0.0, 1.0, -0.5], np.float32)
    a = Solver(NX, NY, NZ, 5.0, 4e-4, order=4, precision=precision, npml=6)
    b = Solver(NX, NY, NZ, 5.0, 4e-4, order=4, precision=precision, npml=6)
    for s in (a, b):
        s.set_source(s.cell(18, 15, 20), w)
    a.set_medium(m)
    b.set_medium_device(m)
    ra, rb = a.coefficients(), b.coefficients()
    for q in range(_lib.NCOEF):
        assert np.array_equal(ra[q].view(np.uint32), rb[q].view(np.uint32)), _lib.COEF_NAMES[q]
    assert a.field_scales() == b.field_scales()
    K = _lib.kernel_coefficients(5.0, ra)
    for k in range(_lib.NKCOEF):
        ka, kb = a.kernel_coefficient(k).reshape(-1), b.kernel_coefficient(k).reshape(-1)
        assert np.array_equal(ka, kb), _lib.KCOEF_NAMES[k]
        if precision == "fp64":
            assert np.array_equal(ka, K[k]), _lib.KCOEF_NAMES[k]
        elif precision == "fp32":
            assert np.array_equal(ka.astype(np.float32), K[k].astype(np.float32)), _lib.KCOEF_NAMES[k]
    a.close()
    b.close()
