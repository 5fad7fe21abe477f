"""The reference's own compiled outputfunc.cpp (oracle/_ref/liboutputfunc.so, built from the
reference sources by `make -C oracle ref`) as the checker of two parts of the path:

* the Biot moduli (outputfunc.cpp:153-222, create_model_all's borehole model: a fluid column
  of radius 10 cells in a porous rock): the oracle's restatement, fed the moduli
  create_model_all forms from its velocities and densities, reproduces every stored array bit
  for bit, and the library's host coefficient builder on that medium equals the oracle's
  coefficients built from the reference's own moduli arrays;
* the text writers (outputfunc.cpp:35-50, 97-151): the library's biot_write_text_2d and
  biot_write_text_xz write byte-identical files to wfile, wfile3d and wfile1d.

The reference library exists only where the reference sources are (it is built here, not on
a machine without them), so these tests skip without it."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402
from biot_amd import _lib  # noqa: E402


@pytest.fixture(scope="module")
def ref():
    oracle.build()
    if oracle.build_ref() is None and not os.path.exists(oracle.REF_LIB_PATH):
        pytest.skip("reference sources absent: oracle/_ref not built")
    return oracle.RefOutputfunc()


def model_inputs(m, nz, ny, nx):
    """The moduli create_model_all forms per cell (outputfunc.cpp:170-217), in its own
    arithmetic: float*float products stay float (vp*vp, rhof*vf*vf), a double operand
    promotes the rest of the expression (C's left-to-right evaluation)."""
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    fluid2d = ((ii - nx // 2) ** 2 + (jj - ny // 2) ** 2) <= 100
    fluid = np.broadcast_to(fluid2d, (nz, ny, nx))
    vp, vs, vf, rhos, rhof = m["vp"], m["vs"], m["vf"], m["rhos"], m["rhof"]
    f = np.float32
    vp2 = (vp * vp).astype(f)                    # float product
    vs2 = (vs * vs).astype(f)
    kf = ((rhof * vf).astype(f) * vf).astype(f)  # rhof*vf*vf: float throughout
    ks_fluid = ((rhof * vp).astype(f) * vp).astype(f)
    bracket = vp2.astype(np.float64) - vs2.astype(np.float64) * 4.0 / 3.0
    por = np.where(fluid, 1.0, 0.1)
    ks_rock = rhos.astype(np.float64) * bracket
    kb_rock = rhos.astype(np.float64) * (1 - 0.1) * bracket
    mu_rock = (1 - 0.1) * rhos.astype(np.float64) * vs.astype(np.float64) * vs.astype(np.float64)
    return fluid, dict(
        Ks=np.where(fluid, ks_fluid.astype(np.float64), ks_rock),
        Kf=kf.astype(np.float64),
        Kb=np.where(fluid, 0.0, kb_rock),
        mu=np.where(fluid, 0.0, mu_rock),
        phi=por,
        rhos=rhos.astype(np.float64),
        rhof=rhof.astype(np.float64),
        eta=np.full((nz, ny, nx), 1.0 * pow(10.0, -3)),
        perm=np.full((nz, ny, nx), 2 * pow(10.0, -12)),
    )


@pytest.mark.parametrize("nz,ny,nx", [(4, 24, 24), (3, 31, 26)])
def test_moduli_bit_exact_vs_reference_create_model_all(ref, nz, ny, nx):
    m = ref.model(nz, ny, nx)
    fluid, medium = model_inputs(m, nz, ny, nx)
    assert fluid.any() and (~fluid).any()
    ours = oracle.biot_moduli(medium, (nz, ny, nx), tao=3.0, porousm=8.0)
    for k in ("M", "HH", "H2u", "C", "C1", "C2", "rho", "rhof", "mu"):
        a = ours[k].reshape(nz, ny, nx)
        assert np.array_equal(a.view(np.uint32), m[k].view(np.uint32)), k
    # the rock's closed form: Kb = (1 - phi) Ks makes alpha = phi, so M = Kf / phi and
    # C = alpha M = Kf; the fluid column has C = M = H = lambda_c = Kf, C1 = 0, C2 = rho_f
    rock = ~fluid
    assert np.allclose(m["M"][rock], medium["Kf"][rock] / 0.1, rtol=1e-6)
    assert np.allclose(m["C"][rock], medium["Kf"][rock], rtol=1e-6)
    for k in ("M", "C", "HH", "H2u"):
        assert np.all(m[k][fluid] == np.float32(medium["Kf"][fluid]))
    assert np.all(m["C1"][fluid] == 0) and np.all(m["C2"][fluid] == m["rhof"][fluid])


def test_coefficients_from_reference_moduli(ref):
    """The library's host coefficient builder (moduli in, main.cpp:322-494) on the borehole
    medium against the oracle's averaging and coefficient restatement applied to the moduli
    arrays the reference's create_model_all stored."""
    nz, ny, nx, dt = 6, 24, 22, 4e-4
    m = ref.model(nz, ny, nx)
    _, medium = model_inputs(m, nz, ny, nx)
    mod = {k: np.ascontiguousarray(m[k]).reshape(-1) for k in ("M", "HH", "H2u", "C", "C1", "C2", "rho",
                                                                "rhof", "mu")}
    av = oracle.stagger_average(mod, nx, ny, nz)
    want = oracle.coefficients(mod, av, nx, ny, nz, dt)
    ours = _lib.build_coefficients(nx, ny, nz, dt, medium)
    for a in range(23):
        assert np.array_equal(ours[a].view(np.uint32), want[a].view(np.uint32)), _lib.COEF_NAMES[a]


def _values(shape, seed):
    rng = np.random.default_rng(seed)
    v = (rng.standard_normal(shape) * 10.0 ** rng.integers(-30, 30, shape)).astype(np.float32)
    flat = v.reshape(-1)
    special = np.array([0.0, -0.0, 1e-40, -3.4e38, np.inf, -np.inf, 1.0, -1.5e-7], np.float32)
    flat[:min(flat.size, special.size)] = special[:flat.size]
    return v


@pytest.mark.parametrize("nr,nc", [(5, 7), (1, 1), (3, 0), (0, 4)])
def test_text_2d_byte_identical_to_wfile(ref, tmp_path, nr, nc):
    data = _values((nr, nc), 3)
    ref.write_2d(tmp_path / "ref.dat", data)
    _lib.write_text_2d(tmp_path / "ours.dat", data)
    assert (tmp_path / "ours.dat").read_bytes() == (tmp_path / "ref.dat").read_bytes()


@pytest.mark.parametrize("nz,ny,nx", [(4, 5, 6), (3, 2, 7), (2, 1, 3)])
def test_text_xz_byte_identical_to_wfile3d_and_wfile1d(ref, tmp_path, nz, ny, nx):
    vol = _values((nz, ny, nx), 11)
    ref.write_3d(tmp_path / "ref3d.dat", vol)
    ref.write_1d(tmp_path / "ref1d.dat", vol)
    _lib.write_text_xz(tmp_path / "ours.dat", vol)
    ours = (tmp_path / "ours.dat").read_bytes()
    assert ours == (tmp_path / "ref3d.dat").read_bytes()
    assert ours == (tmp_path / "ref1d.dat").read_bytes()
