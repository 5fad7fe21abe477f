"""GPU check of the allocation-free misfit (biot_misfit_l2_noalloc through the C-ABI): it
equals biot_misfit_l2 on the same traces (residuals bit for bit; J, an FP64 sum whose block
partials are added atomically in no fixed order, to 1e-12 relative), on two calls in a row with
the second on a smaller trace layout, on FP32 and FP64 handles; layouts beyond its maximum are
refused before the device is touched."""
import ctypes as C

import numpy as np
import pytest

from biot_amd import configs
from biot_amd._lib import fptr, lib
from biot_amd.solver import BiotError

pytestmark = pytest.mark.gpu


def _noalloc(s, obs):
    obs = np.ascontiguousarray(np.asarray(obs, np.float32))
    res = np.zeros_like(obs)
    J = C.c_double()
    s._check(lib().biot_misfit_l2_noalloc(s._h, fptr(obs), fptr(res), C.byref(J)), "biot_misfit_l2_noalloc")
    return J.value, res


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_misfit_equals_allocating_misfit(precision):
    print(f"no-alloc misfit {precision}", flush=True)
    cfg = configs.fwd64(order=4, nt=150)
    s = configs.make_solver(cfg, precision=precision, nt=150)
    s.run(150)
    rng = np.random.default_rng(4)
    for call, (nrec, nt) in enumerate(((len(cfg["rec"]), 150), (7, 90))):
        if call:  # a smaller trace layout for the second call
            s.set_receivers(configs.cells(cfg, cfg["rec"][:nrec]), cfg["rec_fields"], nt)
            s.reset()
            s.run(nt)
        d = s.traces()
        obs = (d * np.float32(1.02) + rng.standard_normal(d.shape).astype(np.float32) * np.abs(d).max() * 1e-3)
        J0, r0 = s.misfit(obs)
        J1, r1 = _noalloc(s, obs)
        print(f"  call {call}: {d.size} samples, J {J0:.12e} / {J1:.12e}", flush=True)
        assert r1.shape == d.shape and np.array_equal(r1.view(np.uint32), r0.view(np.uint32))
        assert J0 > 0 and abs(J1 - J0) <= 1e-12 * J0
    s.close()


def test_noalloc_misfit_refuses_layouts_beyond_its_maximum():
    cfg = configs.fwd64(order=2, nt=10)
    s = configs.make_solver(cfg, precision="fp32", nt=10)
    cap = 8388576  # BIOT_MISFIT_NOALLOC_MAX_SAMPLES
    nrec = cap // (3 * 10) + 1
    s.set_receivers(np.zeros(nrec, np.int64), cfg["rec_fields"], 10)
    obs = np.zeros((3, 10, nrec), np.float32)
    with pytest.raises(BiotError) as e:
        _noalloc(s, obs)
    assert e.value.status == 1 and "at most" in str(e.value)
    s.close()
