"""GPU parity of the x-vectorised stencil kernels (2 or 4 cells per lane).

* Random state on a ragged 68 x 37 x 41 grid (partial tiles in every direction, PML 6 so
  slabs, corners and the interior all appear), 3 steps: FP32 against the oracle (1e-5 per
  field) and against the one-cell-per-lane kernels (1e-6); FP16 / FP16C against the
  one-cell-per-lane kernels of the same storage (2e-3: only the compiler's choice of fused
  multiply-adds differs, which moves an occasional binary16 rounding by one ulp).
* 64^3, 500 steps: receiver traces of every variant against the oracle (FP32 1e-5; FP16C
  1e-2 and FP16 2.5e-2 against the FP64 oracle).
"""
import numpy as np
import pytest

import oracle
from biot_amd import configs, media
from biot_amd.solver import Solver
from conftest import max_comp_rel_l2, rel_l2

pytestmark = pytest.mark.gpu

XV_CASES = [("fp32", 2), ("fp32", 4), ("fp16", 2), ("fp16", 4), ("fp16c", 4)]


def _random_state(shape, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape).astype(np.float32).astype(np.float64) for _ in range(13)]


def _ragged_run(precision, order, xv, steps=3):
    nx, ny, nz, npml = 68, 37, 41, 6
    shape = (nz, ny, nx)
    med = media.perturbed(shape, corr_cells=3.0, seed=3)
    s = Solver(nx, ny, nz, 5.0, 4e-4, order=order, precision=precision, npml=npml)
    s.set_medium(med)
    # an all-zero wavelet: nothing is injected and the 16-bit storage scales take amplitude 1
    s.set_source(s.cell(30, 18, 20), np.zeros(4, np.float32))
    s.set_launch_shape(64, 2, 4, xv)
    # unit random values in each field's storage scale (1 for FP32), so 16-bit storage
    # holds them without overflow
    scales = s.field_scales()
    st = [v * scales[f] for f, v in zip(oracle.FIELDS, _random_state(shape, 5))]
    for f, v in zip(oracle.FIELDS, st):
        s.write_field(f, v)
    s.run(steps)
    out = {f: s.field(f, f64=True) for f in oracle.FIELDS}
    s.close()
    return out, med, st


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("precision,xv", XV_CASES)
def test_xv_ragged_random_state(precision, xv, order):
    print(f"ragged random state {precision} xv {xv} order {order}", flush=True)
    got, med, st = _ragged_run(precision, order, xv)
    ref1, _, _ = _ragged_run(precision, order, 1)
    worst_self = max(rel_l2(got[f], ref1[f]) for f in oracle.FIELDS)
    print(f"  vs one-cell kernels: {worst_self:.3e}")
    if precision == "fp32":
        o = oracle.OracleRun(68, 37, 41, 5.0, 4e-4, order=order, mode="f32", npml=6, medium=med)
        for f, v in zip(oracle.FIELDS, st):
            o.set_field(f, v)
        for _ in range(3):
            o.step_v()
            o.step_t()
        worst = max(rel_l2(got[f], o.field(f)) for f in oracle.FIELDS)
        o.close()
        print(f"  vs oracle: {worst:.3e}")
        assert worst <= 1e-5
        assert worst_self <= 1e-6
    else:
        assert worst_self <= 2e-3


@pytest.mark.parametrize("precision,xv", XV_CASES)
def test_xv_fwd64_traces(precision, xv):
    print(f"fwd64 {precision} xv {xv}", flush=True)
    cfg = configs.fwd64(order=4)
    nt = cfg["nt"]
    mode = "f32" if precision == "fp32" else "f64"
    o = oracle.OracleRun(cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["dt"], order=4,
                         mode=mode, npml=cfg["npml"], medium=cfg["medium"])
    w = oracle.ricker_series(nt, cfg["dt"], cfg["f0"])
    ref = o.run(0, nt, nt, w, configs.cells(cfg, cfg["src"])[0], cfg["src_fields"],
                configs.cells(cfg, cfg["rec"]), cfg["rec_fields"])
    o.close()
    s = configs.make_solver(cfg, precision=precision)
    s.set_launch_shape(64, 4, 8, xv)
    s.run(nt)
    tr = s.traces()
    s.close()
    err = max_comp_rel_l2(tr, ref)
    print(f"  traces rel L2 = {err:.3e}")
    assert np.isfinite(tr).all()
    bar = {"fp32": 1e-5, "fp16c": 1e-2, "fp16": 2.5e-2}[precision]
    assert err <= bar
