"""GPU checks of the hand-written adjoint (biot_adjoint through the C-ABI).

* Dot-product test: for random trace weights y and a random smooth perturbation dp of all
  six parameters, <J dp, y> from central differences of FP64 forward runs equals
  <dp, J^T y> from the adjoint (FP64 adjoint: 1e-7; FP32 adjoint: 1e-4).
* Finite differences at 8 seeded cells of the 128^3, 1000-step sphere workload: dJ/dp of
  every parameter against central differences of the FP64 forward misfit (bars in the
  test's docstring).
* The FP64 adjoint equals the FP64 host adjoint (oracle/biot_host_adjoint.c, scatter-form
  transposed differences, its own dot-product tests on the CPU) on a 24^3 perturbed case:
  traces 1e-9, parameter gradients 1e-8 (rel L2; the two differ only in summation order).
* The torch.autograd binding returns the same gradients as the direct call.
* Every step's state stored (every = 1, no segment recomputation) gives the gradients of
  50-step checkpoints bit for bit.
* FP16M's gradients (packed residual records in its states) against the FP32 ones.
"""
import math

import numpy as np
import pytest
import torch

from biot_amd import BiotError, autograd, configs, media
from biot_amd._lib import PARAM_NAMES
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _start_params(cfg_start):
    return media.params_from_medium(cfg_start["medium"])


def _forward_f64(s, params, nt):
    s.set_params(params)
    s.reset()
    s.run(nt)
    return s.traces_f64()


def _smooth_direction(shape, seed):
    rng = np.random.default_rng(seed)
    out = {}
    for q, k in enumerate(PARAM_NAMES):
        out[k] = media.gaussian_field(shape, 3.0, seed + 17 * q) * (0.5 + rng.random())
    return out


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-7), ("fp32", 1e-4)])
def test_adjoint_dot_product(precision, tol):
    cfg = configs.adjoint128(n=48, nt=300, every=40, sphere=False)
    nt, every = cfg["nt"], cfg["every"]
    print(f"dot product {precision} 48^3 {nt} steps", flush=True)
    p0 = _start_params(cfg)
    shape = p0["rho"].shape
    d = _smooth_direction(shape, 7)
    s64 = configs.make_solver(cfg, precision="fp64")
    eps = 1e-5
    plus = dict(p0, **{k: p0[k] * (1 + eps * d[k]) for k in PARAM_NAMES})
    minus = dict(p0, **{k: p0[k] * (1 - eps * d[k]) for k in PARAM_NAMES})
    dd = (_forward_f64(s64, plus, nt) - _forward_f64(s64, minus, nt)) / (2 * eps)
    y = np.random.default_rng(11).standard_normal(dd.shape).astype(np.float32)
    lhs = float(np.sum(dd * y.astype(np.float64)))
    s = s64 if precision == "fp64" else configs.make_solver(cfg, precision=precision)
    s.set_params(p0)
    s.forward_checkpointed(nt, every)
    g, _ = s.adjoint(y)
    rhs = float(sum(np.sum(g[k] * p0[k] * d[k]) for k in PARAM_NAMES))
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"  <J dp, y> = {lhs:.12e}  <dp, J^T y> = {rhs:.12e}  rel {rel:.2e}")
    assert rel <= tol


def _fd_cells(cfg, count, seed):
    rng = np.random.default_rng(seed)
    lo, hi = 24, cfg["nx"] - 24
    return [tuple(int(v) for v in rng.integers(lo, hi, 3)) for _ in range(count)]


def test_adjoint_matches_finite_differences_at_seeded_cells():
    """128^3, 1000 steps, checkpoints every 50: L2 misfit against traces of the +8 % Kb
    sphere model, gradients of the start model at 8 seeded cells for all six parameters,
    against fourth-order central differences (step 1e-3 of the value) of the FP64 forward
    misfit.

    Bars (relative, per cell and parameter): the FP64 adjoint 1e-3; the FP32 adjoint fed the
    FP64 residual 1e-3.  The FP32 run end to end (FP32 forward residual) is bounded by the
    FP32 forward's trace error relative to the residual, not by the adjoint: the residual is
    ~2e-3 of the data while FP32 traces carry ~3e-6 relative error, so the residual itself
    is off by ~2e-3 and single-cell gradients (time integrals with heavy cancellation) by up
    to a few 1e-2; that bar is 5e-2."""
    true = configs.adjoint128()
    start = configs.adjoint128(sphere=False)
    nt, every = start["nt"], start["every"]
    print("adjoint128 fd check", flush=True)
    s64 = configs.make_solver(start, precision="fp64")
    obs = _forward_f64(s64, media.params_from_medium(true["medium"]), nt).astype(np.float32)
    p0 = _start_params(start)
    s64.set_params(p0)
    s64.forward_checkpointed(nt, every)
    J0, res = s64.misfit(obs)
    g64, _ = s64.adjoint(res)
    s32 = configs.make_solver(start, precision="fp32")
    s32.set_params(p0)
    s32.forward_checkpointed(nt, every)
    _, res32 = s32.misfit(obs)
    g32r64, _ = s32.adjoint(res)
    g32, _ = s32.adjoint(res32)
    print(f"  J0 = {J0:.6e}", flush=True)
    worst = dict(fp64=0.0, fp32_res64=0.0, fp32=0.0)
    obs64 = obs.astype(np.float64)

    def dJ(da, db):
        # J(a) - J(b) = sum (da - db) ((da + db)/2 - obs), formed from the double traces so
        # the tiny single-cell change is not lost in the rounding of J itself
        return math.fsum(((da - db) * (0.5 * (da + db) - obs64)).ravel())

    for (i, j, k) in _fd_cells(start, 8, 2024):
        for name in PARAM_NAMES:
            h = 1e-3 * abs(p0[name][k, j, i])
            d = {}
            for m in (1, -1, 2, -2):
                p = {kk: v.copy() for kk, v in p0.items()}
                p[name][k, j, i] += m * h
                d[m] = _forward_f64(s64, p, nt)
            # fourth-order central difference
            fd = (8 * dJ(d[1], d[-1]) - dJ(d[2], d[-2])) / (12 * h)
            errs = {key: abs(g[name][k, j, i] - fd) / abs(fd)
                    for key, g in (("fp64", g64), ("fp32_res64", g32r64), ("fp32", g32))}
            for key, e in errs.items():
                worst[key] = max(worst[key], e)
            print(f"  ({i},{j},{k}) {name:6s} fd {fd:+.6e} adj64 {g64[name][k, j, i]:+.6e} "
                  + " ".join(f"{key} {e:.1e}" for key, e in errs.items()), flush=True)
    print("  worst: " + " ".join(f"{key} {e:.2e}" for key, e in worst.items()))
    assert worst["fp64"] <= 1e-3
    assert worst["fp32_res64"] <= 1e-3
    assert worst["fp32"] <= 5e-2


def test_fp64_adjoint_matches_host_adjoint():
    import oracle
    from biot_amd import _lib
    from biot_amd.solver import Solver

    n, npml, nt, every = 24, 4, 80, 20
    print(f"fp64 adjoint vs host adjoint {n}^3 {nt} steps", flush=True)
    m = media.perturbed((n, n, n), amplitude=0.10, corr_cells=3.0, seed=12)
    p0 = media.params_from_medium(m)
    c = n // 2
    src = c + c * n + c * n * n
    ii, jj = np.meshgrid(np.arange(npml + 1, n - npml - 1, 3), np.arange(npml + 1, n - npml - 1, 3))
    rec = (ii.ravel() + jj.ravel() * n + (npml + 2) * n * n).astype(np.int64)
    w = oracle.ricker_series(nt, configs.DT, 60.0)
    s = Solver(n, n, n, configs.DX, configs.DT, order=4, precision="fp64", npml=npml)
    s.set_params(p0)
    s.set_source(src, w, configs.SRC_EXPLOSIVE)
    s.set_receivers(rec, configs.REC_FIELDS, nt)
    s.forward_checkpointed(nt, every)
    d = s.traces_f64()
    y = np.random.default_rng(21).standard_normal(d.shape).astype(np.float32)
    g, _ = s.adjoint(y)
    s.close()
    K = _lib.params_coefficients(n, n, n, configs.DX, configs.DT, p0)
    prof = oracle.host_profiles(n, n, n, configs.DX, configs.DT, npml)
    mask = sum(1 << oracle.FIELDS.index(f) for f in configs.SRC_EXPLOSIVE)
    comp = [oracle.FIELDS.index(f) for f in oracle.FIELDS if f in configs.REC_FIELDS]
    ha = oracle.HostAdjoint(n, n, n, 4, prof, K, src, mask, w, rec, comp)
    gK, tr = ha.adjoint(nt, y.astype(np.float64), every=every)
    gp = _lib.params_chain(n, n, n, configs.DX, configs.DT, p0, gK)
    et = max(float(np.linalg.norm(tr[q] - d[q]) / np.linalg.norm(d[q])) for q in range(3))
    eg = {k: float(np.linalg.norm(g[k] - gp[q].reshape(n, n, n)) / np.linalg.norm(gp[q]))
          for q, k in enumerate(PARAM_NAMES)}
    print(f"  traces rel L2 {et:.2e}; gradients " + " ".join(f"{k} {e:.2e}" for k, e in eg.items()))
    assert et <= 1e-9
    assert max(eg.values()) <= 1e-8


@pytest.mark.parametrize("precision", ["fp32", "fp16", "fp16t", "fp16c", "fp16m", "fp64"])
def test_stored_states_equal_recomputed_segments(precision):
    cfg = configs.adjoint128(n=48, nt=260, every=50, sphere=False)
    nt = cfg["nt"]
    print(f"stored states vs 50-step segments {precision} 48^3 {nt} steps", flush=True)
    p0 = _start_params(cfg)
    s = configs.make_solver(cfg, precision=precision)
    s.set_params(p0)
    y = None
    out = {}
    for every in (50, 1):
        s.forward_checkpointed(nt, every)
        tr = s.traces()
        if y is None:
            y = np.random.default_rng(5).standard_normal(tr.shape).astype(np.float32)
        g, gk = s.adjoint(y, params=True, coefs=True)
        out[every] = (tr, g, gk)
        print(f"  every {every}: {s.last_run_ms():.1f} ms adjoint", flush=True)
    assert np.array_equal(out[50][0], out[1][0])
    for k in PARAM_NAMES:
        assert np.array_equal(out[50][1][k], out[1][1][k]), k
    for a, b in zip(out[50][2], out[1][2]):
        assert np.array_equal(a, b)
    assert any(np.abs(out[1][1][k]).max() > 0 for k in PARAM_NAMES)
    s.close()


def test_fp16m_gradient_against_fp32():
    """FP16M gradients against the FP32 ones, FP16T (the 16-bit policy the batched gradients
    use) beside them, on the 64^3 start model, 400 steps, 50-step checkpoints.  FP16M's
    checkpoints carry its residual records, the adjoint reads tau and p with their 4-bit
    residuals and the coefficients as FP16T stores them (binary16 from the parameters; the
    FP16M forward forms them in float from binary16 moduli).  Two misfits:

    * energy: observed traces zero, so the residual is the data and the 16-bit traces' own
      error stays small beside it.  Measured on MI355X: J 1.7e-5 from FP32; lambda, alpha, M
      1.4e-5, mu 3.6e-4, rho 8.7e-4, rho_f 2.8e-3 (FP16T 3.6e-5, 6.8e-4, 7.3e-4, 7.8e-3);
      held to 1e-4 (J; lambda, alpha, M), 3e-3 (mu, rho) and 1e-2 (rho_f);
    * model: observed traces from the FP64 run of a model with lambda +10 % and mu -5 %
      everywhere.  Its residual is a small part of the data late in the run, and both 16-bit
      policies sit well above their batched-gradient levels there (FP16T measured 3.4e-3 on
      lambda, 8e-2 on mu, 0.16 on rho_f; FP16M 8.4e-4, 4.3e-2, 5.8e-2, and 4.1e-2 on rho
      against FP16T's 3.5e-2).  FP16M is held to FP16T's level: within 2x per parameter and J
      within 2x FP16T's J error (plus 1 %)."""
    cfg = configs.adjoint128(n=64, nt=400, every=50, sphere=False)
    nt, every = cfg["nt"], cfg["every"]
    p0 = _start_params(cfg)
    p_true = {k: v.copy() for k, v in p0.items()}
    p_true["lambda"] *= 1.10
    p_true["mu"] *= 0.95
    s64 = configs.make_solver(cfg, precision="fp64")
    obs_model = _forward_f64(s64, p_true, nt).astype(np.float32)
    s64.close()
    res = {}
    for case, obs in (("energy", np.zeros_like(obs_model)), ("model", obs_model)):
        print(f"fp16m / fp16t gradients vs fp32, 64^3 400 steps, {case} misfit", flush=True)
        out = {}
        for prec in ("fp32", "fp16t", "fp16m"):
            s = configs.make_solver(cfg, precision=prec)
            out[prec] = autograd.misfit_gradient(p0, s, obs, nt, every)
            s.close()
        J32, g32 = out["fp32"]
        for prec in ("fp16t", "fp16m"):
            J, g = out[prec]
            assert all(np.isfinite(g[k]).all() for k in PARAM_NAMES), (case, prec)
            e = {k: rel_l2(g[k], g32[k]) for k in PARAM_NAMES}
            res[case, prec] = (abs(J / J32 - 1), e)
            print(f"  {prec}: J/J32 - 1 = {J / J32 - 1:.2e}; gradient rel L2 " +
                  " ".join(f"{k} {x:.2e}" for k, x in e.items()), flush=True)
    dJ, e = res["energy", "fp16m"]
    assert dJ <= 1e-4
    assert max(e[k] for k in ("lambda", "alpha", "M")) <= 1e-4
    assert max(e["mu"], e["rho"]) <= 3e-3
    assert e["rhof"] <= 1e-2
    (dJm, em), (dJt, et) = res["model", "fp16m"], res["model", "fp16t"]
    assert dJm <= 2 * dJt + 1e-2
    for k in PARAM_NAMES:
        assert em[k] <= 2 * et[k] + 1e-3, k


def test_autograd_binding_matches_direct_adjoint():
    cfg = configs.adjoint128(n=40, nt=200, every=30, sphere=False)
    nt, every = cfg["nt"], cfg["every"]
    true = configs.adjoint128(n=40, nt=200, every=30, sphere=True)
    s = configs.make_solver(cfg, precision="fp32")
    s.set_params(media.params_from_medium(true["medium"]))
    s.reset()
    s.run(nt)
    obs = s.traces()
    p0 = _start_params(cfg)
    J, g_direct = autograd.misfit_gradient(p0, s, obs, nt, every)
    tens = {k: torch.tensor(v, dtype=torch.float64, requires_grad=k in PARAM_NAMES)
            for k, v in p0.items()}
    d = autograd.traces(tens, s, nt, every)
    loss = 0.5 * ((d.double() - torch.from_numpy(obs).double()) ** 2).sum()
    loss.backward()
    assert abs(loss.item() - J) <= 1e-6 * J
    for k in PARAM_NAMES:
        a, b = tens[k].grad.numpy(), g_direct[k]
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max(), k


def test_adjoint_argument_errors():
    cfg = configs.adjoint128(n=32, nt=50, every=10, sphere=False)
    s = configs.make_solver(cfg, precision="fp32")
    with pytest.raises(BiotError):  # no checkpointed run yet
        s.adjoint(np.zeros((s.ncomp, s.nt_max, s.nrec), np.float32))
    with pytest.raises(BiotError):  # more steps than the receivers hold
        s.forward_checkpointed(cfg["nt"] + 1, 10)
    s16 = configs.make_solver(cfg, precision="fp16m")
    with pytest.raises(BiotError):  # FP16M's adjoint coefficients come from biot_set_params
        s16.forward_checkpointed(cfg["nt"], 10)
    s16.close()
    s.forward_checkpointed(cfg["nt"], 10)
    with pytest.raises(BiotError):  # parameter gradients need biot_set_params
        s.adjoint(np.zeros((s.ncomp, s.nt_max, s.nrec), np.float32), params=True)
    _, gk = s.adjoint(np.zeros((s.ncomp, s.nt_max, s.nrec), np.float32), params=False, coefs=True)
    assert all(not np.any(x) for x in gk)
