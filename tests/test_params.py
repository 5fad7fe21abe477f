"""Host checks of the inversion parameterization (no GPU).

* The double-precision kernel coefficients built from (lambda, mu, alpha, M, rho, rho_f, C1,
  tphi) agree with the reference's float coefficient builder (main.cpp:322-494, pinned
  bit-exact against the oracle in test_host_setup.py) to float rounding, and zero on the
  same cells, for homogeneous, layered, perturbed and fluid-bearing media.
* The chain rule dJ/dK -> dJ/dparams equals central differences of the coefficient builder
  along random directions (an independent check of every hand-derived partial).
"""
import numpy as np
import pytest

from biot_amd import _lib, media

DX, DT = 5.0, 4e-4
HINV = float(np.float32(1.0) / np.float32(DX))


def _medium(kind, shape):
    nz, ny, nx = shape
    if kind == "homogeneous":
        return media.homogeneous(shape)
    if kind == "layered":
        return media.layered(shape)
    if kind == "perturbed":
        return media.perturbed(shape, corr_cells=3.0, seed=11)
    m = media.layered(shape)  # fluid column: mu = Kb = 0 (outputfunc.cpp:176-193)
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    hole = ((ii - nx // 2) ** 2 + (jj - ny // 2) ** 2) <= 4
    for k in ("mu", "Kb"):
        m[k][:, hole] = 0.0
    return m


def _reference_kernel_coefs(ref):
    """The reference arrays in kernel order with 1/dx (float, kernel.cu:158) and the 0.01
    stress constants folded (kernel.cu:484-490)."""
    r = {n: ref[i].astype(np.float64).reshape(-1) for i, n in enumerate(_lib.COEF_NAMES)}
    out = []
    for n in _lib.KCOEF_NAMES:
        if n[:2] in ("U1", "W2", "W3"):
            out.append(r[n] * HINV)
        elif n[:2] in ("U2", "W1"):
            out.append(r[n])
        else:
            key = {"P1": "PP1", "P2": "PP2", "S1": "SP1", "S2": "SP2", "S3": "SP3",
                   "SXY": "SPXY", "SXZ": "SPXZ", "SYZ": "SPYZ"}[n]
            out.append(0.01 * r[key] * HINV)
    return out


@pytest.mark.parametrize("kind", ["homogeneous", "layered", "perturbed", "borehole"])
def test_params_coefficients_match_reference_builder(kind):
    nx, ny, nz = 14, 12, 16
    shape = (nz, ny, nx)
    m = _medium(kind, shape)
    p = media.params_from_medium(m)
    ours = _lib.params_coefficients(nx, ny, nz, DX, DT, p)
    ref = _reference_kernel_coefs(_lib.build_coefficients(nx, ny, nz, DT, m))
    for k, name in enumerate(_lib.KCOEF_NAMES):
        a, b = ours[k], ref[k]
        assert np.array_equal(a == 0, b == 0), f"{name}: zero pattern differs"
        scale = np.abs(b).max()
        if scale == 0:
            continue
        err = np.abs(a - b).max() / scale
        # the reference evaluates in float: a few float ulps of the largest entry
        assert err < 5e-7, f"{name}: {err:.3e}"


def test_params_chain_matches_central_differences():
    nx, ny, nz = 10, 9, 11
    shape = (nz, ny, nx)
    p = media.params_from_medium(media.perturbed(shape, corr_cells=2.0, seed=5))
    rng = np.random.default_rng(3)
    gk = [rng.standard_normal(nx * ny * nz) for _ in range(_lib.NKCOEF)]
    grads = _lib.params_chain(nx, ny, nz, DX, DT, p, gk)
    K0 = _lib.params_coefficients(nx, ny, nz, DX, DT, p)
    for q, name in enumerate(_lib.PARAM_NAMES):
        base = p[name].reshape(-1)
        d = rng.standard_normal(base.size) * np.abs(base).max()
        eps = 1e-6
        plus = dict(p, **{name: (base + eps * d).reshape(shape)})
        minus = dict(p, **{name: (base - eps * d).reshape(shape)})
        Kp = _lib.params_coefficients(nx, ny, nz, DX, DT, plus)
        Km = _lib.params_coefficients(nx, ny, nz, DX, DT, minus)
        fd = sum(float(np.dot(gk[k], (Kp[k] - Km[k]) / (2 * eps))) for k in range(_lib.NKCOEF))
        ad = float(np.dot(grads[q], d))
        assert abs(fd - ad) <= 1e-6 * abs(ad), f"{name}: fd {fd:.9e} vs chain {ad:.9e}"
    # coefficients outside the interior carry no gradient and are zero
    for k in range(_lib.NKCOEF):
        Kk = K0[k].reshape(shape)
        assert not Kk[0].any() and not Kk[-1].any() and not Kk[:, 0].any() and not Kk[:, :, -1].any()


def test_params_chain_rejects_missing_arrays():
    p = media.params_from_medium(media.homogeneous((5, 5, 5)))
    with pytest.raises(ValueError):
        _lib.params_coefficients(5, 5, 6, DX, DT, p)
