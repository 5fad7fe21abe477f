"""CPU checks of the FP64 host adjoint (oracle/biot_host_adjoint.c), the checker of the device
adjoint (the reference is forward-only).

* Its forward (folded C-PML carry, coefficients K with 1/dx and 0.01 folded in) reproduces
  the oracle's FP64 restatement of the reference step (two-array C-PML, reference
  coefficient arrays) to 1e-9 relative on receiver traces (about 3e-11 measured).
* Dot-product tests: <J dK, y> from central differences of the host forward equals
  <dK, J^T y> from the host adjoint (per coefficient array and in the six inversion
  parameters through the library's chain rule).
"""
import numpy as np
import pytest

import oracle
from biot_amd import _lib, configs, media
from biot_amd._lib import PARAM_NAMES
from conftest import rel_l2

DX, DT = configs.DX, configs.DT
EXPLOSIVE = (1 << 6) | (1 << 7) | (1 << 8) | (1 << 12)
COMP = [0, 2, 12]   # vx, vz, p


def _case(n=20, npml=4, nt=60, seed=5, order=4):
    shape = (n, n, n)
    m = media.perturbed(shape, amplitude=0.10, corr_cells=3.0, seed=seed)
    c = n // 2
    src = c + c * n + c * n * n
    ii, jj = np.meshgrid(np.arange(npml + 1, n - npml - 1, 3), np.arange(npml + 1, n - npml - 1, 3))
    rec = (ii.ravel() + jj.ravel() * n + (npml + 2) * n * n).astype(np.int64)
    w = oracle.ricker_series(nt, DT, 60.0)
    return dict(n=n, npml=npml, nt=nt, order=order, medium=m, src=src, rec=rec, w=w)


def _host(case, K):
    n = case["n"]
    prof = oracle.host_profiles(n, n, n, DX, DT, case["npml"])
    return oracle.HostAdjoint(n, n, n, case["order"], prof, K, case["src"], EXPLOSIVE, case["w"],
                              case["rec"], COMP)


@pytest.mark.parametrize("order", [2, 4])
def test_host_forward_matches_oracle_f64(order):
    case = _case(order=order)
    n, nt = case["n"], case["nt"]
    o = oracle.OracleRun(n, n, n, DX, DT, order=order, mode="f64", npml=case["npml"],
                         medium=case["medium"])
    tr_o = o.run(0, nt, nt, case["w"], case["src"], configs.SRC_EXPLOSIVE, case["rec"],
                 configs.REC_FIELDS)
    o.close()
    K = _lib.kernel_coefficients(DX, o.coef)
    tr_h = _host(case, K).forward(nt)
    errs = [rel_l2(tr_h[q], tr_o[q]) for q in range(3)]
    print("host forward vs oracle f64 (vx, vz, p):", " ".join(f"{e:.2e}" for e in errs))
    assert np.abs(tr_o).max() > 0
    assert max(errs) <= 1e-9


def test_host_adjoint_dot_product_in_coefficients():
    case = _case()
    n, nt = case["n"], case["nt"]
    params = media.params_from_medium(case["medium"])
    K = _lib.params_coefficients(n, n, n, DX, DT, params)
    ha = _host(case, K)
    y = np.random.default_rng(3).standard_normal((3, nt, len(case["rec"])))
    gK, tr = ha.adjoint(nt, y, every=7)
    assert np.array_equal(tr, ha.forward(nt))
    rng = np.random.default_rng(9)
    d = [rng.standard_normal(n ** 3) for _ in range(22)]
    eps = 1e-6

    def J(sign):
        Kp = [k * (1 + sign * eps * dk) for k, dk in zip(K, d)]
        return float(np.sum(_host(case, Kp).forward(nt) * y))
    lhs = (J(1) - J(-1)) / (2 * eps)
    rhs = float(sum(np.sum(g * k * dk) for g, k, dk in zip(gK, K, d)))
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"<J dK, y> = {lhs:.12e}  <dK, J^T y> = {rhs:.12e}  rel {rel:.2e}")
    assert rel <= 1e-7
    # coefficients outside the update box carry no gradient
    h = case["order"] // 2
    for g in gK:
        g3 = g.reshape(n, n, n)
        assert not g3[:h].any() and not g3[n - h:].any()


def test_host_adjoint_dot_product_in_parameters():
    case = _case(seed=8)
    n, nt = case["n"], case["nt"]
    p0 = media.params_from_medium(case["medium"])
    y = np.random.default_rng(4).standard_normal((3, nt, len(case["rec"])))
    gK, _ = _host(case, _lib.params_coefficients(n, n, n, DX, DT, p0)).adjoint(nt, y, every=13)
    gp = _lib.params_chain(n, n, n, DX, DT, p0, gK)
    d = {k: media.gaussian_field((n, n, n), 2.0, 31 + q) for q, k in enumerate(PARAM_NAMES)}
    eps = 1e-6

    def J(sign):
        p = dict(p0, **{k: p0[k] * (1 + sign * eps * d[k]) for k in PARAM_NAMES})
        return float(np.sum(_host(case, _lib.params_coefficients(n, n, n, DX, DT, p)).forward(nt) * y))
    lhs = (J(1) - J(-1)) / (2 * eps)
    rhs = float(sum(np.sum(g.reshape(n, n, n) * p0[k] * d[k]) for g, k in zip(gp, PARAM_NAMES)))
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"<J dp, y> = {lhs:.12e}  <dp, J^T y> = {rhs:.12e}  rel {rel:.2e}")
    assert rel <= 1e-7


def test_host_adjoint_argument_errors():
    case = _case(n=12, npml=2, nt=10)
    n = case["n"]
    K = [np.ones(n ** 3)] * 22
    ha = _host(case, K)
    with pytest.raises(ValueError):
        ha.adjoint(10, np.zeros((3, 9, len(case["rec"]))))
    with pytest.raises(ValueError):
        _host(case, K[:21])
