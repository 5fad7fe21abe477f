"""CPU tests of the product's host side: the C-ABI library loads and exports every symbol its
header declares; its set-up stages (coefficients, PML profiles, Ricker) equal the oracle's
bit for bit; argument errors come back as status codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from biot_amd import _lib, media
from biot_amd._lib import BIOT_OK, Grid
from biot_amd.solver import BiotError, Solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(biot_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    decl = declared_functions()
    assert len(decl) >= 25
    missing = [n for n in decl if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.exported_symbols()) <= set(decl)
    assert L.biot_abi_version() == 2


def _oracle_coefs(medium, nx, ny, nz, dt):
    shape = (nz, ny, nx)
    mod = oracle.biot_moduli(medium, shape)
    av = oracle.stagger_average(mod, nx, ny, nz)
    return oracle.coefficients(mod, av, nx, ny, nz, dt)


@pytest.mark.parametrize("kind", ["homogeneous", "layered", "perturbed", "borehole"])
def test_coefficients_bit_exact_vs_oracle(kind):
    nx, ny, nz, dt = 20, 18, 24, 4e-4
    shape = (nz, ny, nx)
    if kind == "homogeneous":
        m = media.homogeneous(shape)
    elif kind == "layered":
        m = media.layered(shape)
    elif kind == "perturbed":
        m = media.perturbed(shape, corr_cells=3.0, seed=7)
    else:  # fluid column as in the reference's borehole model (mu = Kb = 0)
        m = media.layered(shape)
        jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        hole = ((ii - nx // 2) ** 2 + (jj - ny // 2) ** 2) <= 9
        for k in ("mu", "Kb"):
            m[k][:, hole] = 0.0
    ours = _lib.build_coefficients(nx, ny, nz, dt, m)
    ref = _oracle_coefs(m, nx, ny, nz, dt)
    for a in range(23):
        assert np.array_equal(ours[a].view(np.uint32), ref[a].view(np.uint32)), _lib.COEF_NAMES[a]


@pytest.mark.parametrize("n,npml,dx,dt", [(64, 10, 5.0, 4e-4), (256, 20, 5.0, 4e-4),
                                          (240, 30, 0.01, 1e-6), (33, 7, 2.5, 1e-4)])
def test_pml_profile_bit_exact_vs_oracle(n, npml, dx, dt):
    ours = _lib.pml_profile(n, dx, dt, npml)
    ref = oracle.pml_profile(n, dx, dt, npml)
    for a, b in zip(ours, ref):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("f0,dt,nt", [(25.0, 4e-4, 500), (20.0, 4e-4, 2000), (1e4, 1e-6, 3500)])
def test_ricker_bit_exact_vs_oracle(f0, dt, nt):
    ours = _lib.ricker(f0, dt, nt)
    ref = oracle.ricker_series(nt, dt, f0)
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))


def test_argument_errors_are_status_codes():
    L = _lib.lib()
    h = C.c_void_p()
    bad = Grid(nx=64, ny=64, nz=64, order=3, precision=0, npml=10, dx=5, dt=4e-4, pml_v0=4000,
               pml_R=1e-6, Cv=1e8, Cp=1, device=0)
    assert L.biot_create(C.byref(bad), C.byref(h)) == 1  # BIOT_ERR_ARG
    bad.order, bad.nx = 4, 4
    assert L.biot_create(C.byref(bad), C.byref(h)) == 1
    bad.nx, bad.npml = 64, 40
    assert L.biot_create(C.byref(bad), C.byref(h)) == 1
    for prec in (-1, 6, 7):  # BIOT_FP32 .. BIOT_FP16M are 0..5
        bad.npml, bad.precision = 10, prec
        assert L.biot_create(C.byref(bad), C.byref(h)) == 1
    assert L.biot_ricker(C.c_float(-1.0), C.c_float(1e-3), 4, None) == 1
    assert L.biot_status_string(4) == b"non-finite value detected"


@pytest.mark.parametrize("precision", [0, 1, 2, 3, 4, 5])
def test_create_without_device_fails_loudly(precision):
    """Without a GPU the library reports a HIP error instead of falling back to the CPU (every
    storage mode passes the argument checks and reaches the device)."""
    L = _lib.lib()
    h = C.c_void_p()
    g = Grid(nx=16, ny=16, nz=16, order=2, precision=precision, npml=2, dx=5, dt=4e-4, pml_v0=4000,
             pml_R=1e-6, Cv=1e8, Cp=1, device=0)
    rc = L.biot_create(C.byref(g), C.byref(h))
    if rc == BIOT_OK:
        L.biot_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == 2  # BIOT_ERR_HIP


def _create_refusal(nx, ny, nz, precision, npml, device=0):
    L = _lib.lib()
    h = C.c_void_p()
    g = Grid(nx=nx, ny=ny, nz=nz, order=4, precision=precision, npml=npml, dx=5, dt=4e-4, pml_v0=4000,
             pml_R=1e-6, Cv=1e8, Cp=1, device=device)
    rc = L.biot_create(C.byref(g), C.byref(h))
    assert rc != BIOT_OK and not h.value
    return rc, L.biot_last_error(None).decode()


def test_create_refuses_arrays_beyond_32bit_offsets():
    """The kernels form byte offsets in 32 bits from each array's base: a grid whose field
    array (n x storage width) or one axis's C-PML memory (8 carries x slab cells x carry width)
    reaches 4 GiB is refused with its reason before any device call, so the result does not
    depend on a GPU being present."""
    # FP64 fields: 1024*1024*512 cells * 8 B = 4 GiB exactly
    assert _create_refusal(1024, 1024, 512, 2, 10) == (1, "grid too large: one field array would exceed 4 GiB")
    # FP32 fields: 2^30 cells * 4 B
    assert _create_refusal(1024, 1024, 1024, 0, 10)[1].startswith("grid too large")
    # FP32 1000^3 (fields 4.0e9 B) with a 100-cell PML: each axis's carries 8 * 2.0e8 * 4 B
    rc, msg = _create_refusal(1000, 1000, 1000, 0, 100)
    assert rc == 1 and msg.startswith("absorbing slabs too large")
    # FP16 carries are binary16: the same slabs fit (8 * 2.0e8 * 2 B = 3.2e9 B)
    rc, msg = _create_refusal(1000, 1000, 1000, 1, 100, device=-1)
    assert msg in ("no HIP device", "device index out of range")
    # FP16C keeps float carries: refused again
    assert _create_refusal(1000, 1000, 1000, 3, 100)[1].startswith("absorbing slabs too large")
    # just inside both limits the checks pass and only the device index stops it
    rc, msg = _create_refusal(1000, 1000, 1000, 0, 10, device=-1)
    assert msg in ("no HIP device", "device index out of range")


def test_create_reports_argument_errors_through_null_handle():
    rc, msg = _create_refusal(64, 64, 64, 0, 40)
    assert rc == 1 and msg == "npml out of range for the grid"
    L = _lib.lib()
    g = Grid(nx=64, ny=64, nz=64, order=3, precision=0, npml=10, dx=5, dt=4e-4, pml_v0=4000,
             pml_R=1e-6, Cv=1e8, Cp=1, device=0)
    h = C.c_void_p()
    assert L.biot_create(C.byref(g), C.byref(h)) == 1
    assert L.biot_last_error(None) == b"order must be 2 or 4, got 3"
    with pytest.raises(BiotError, match="npml out of range"):
        Solver(64, 64, 64, 5.0, 4e-4, order=4, npml=40)


def test_batch_create_reports_argument_errors_through_null_handle():
    """biot_batch_create refuses bad configurations before any device call and leaves the
    reason in biot_batch_last_error(NULL); the Python front end raises with it."""
    from biot_amd._lib import BatchConfig
    from biot_amd.batch import Batch

    L = _lib.lib()
    base = dict(nx=32, ny=32, nz=32, order=4, precision=1, npml=4, dx=5, dt=4e-4, pml_v0=4000,
                pml_R=1e-6, Cv=1e8, Cp=1, device=0)
    cases = [(dict(nx=7), "at least 8 cells"), (dict(order=3), "order must be 2 or 4"),
             (dict(precision=-1), "BIOT_FP32, BIOT_FP16, BIOT_FP16T, BIOT_FP16C, BIOT_FP16M or BIOT_FP64"),
             (dict(precision=6), "BIOT_FP32, BIOT_FP16, BIOT_FP16T, BIOT_FP16C, BIOT_FP16M or BIOT_FP64"),
             (dict(dt=0.0), "dx and dt must be positive"), (dict(npml=-1), "npml >= 0")]
    for change, why in cases:
        c = BatchConfig(**{**base, **change})
        h = C.c_void_p()
        assert L.biot_batch_create(C.byref(c), C.byref(h)) == 1 and not h.value
        assert why in L.biot_batch_last_error(None).decode(), change
    with pytest.raises(BiotError, match="order must be 2 or 4"):
        Batch(32, 32, 32, 5.0, 4e-4, order=3)
