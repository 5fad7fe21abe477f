"""ctypes binding of the C-ABI in include/biot.h (lib/libbiot.so).

The library is the only compute path: if it is missing this module raises, there is no
Python or CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libbiot.so")

BIOT_OK = 0
STATUS = {0: "BIOT_OK", 1: "BIOT_ERR_ARG", 2: "BIOT_ERR_HIP", 3: "BIOT_ERR_STATE",
          4: "BIOT_ERR_NONFINITE", 5: "BIOT_ERR_RANGE", 6: "BIOT_ERR_NOMEM"}
FP32, FP16, FP64, FP16C, FP16T, FP16M = 0, 1, 2, 3, 4, 5
FIELDS = ["vx", "vy", "vz", "qx", "qy", "qz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "p"]
FIELD_INDEX = {n: i for i, n in enumerate(FIELDS)}
NCOEF = 23
COEF_NAMES = ["W1X", "W1Y", "W1Z", "W2X", "W2Y", "W2Z", "W3X", "W3Y", "W3Z",
              "U1X", "U1Y", "U1Z", "U2X", "U2Y", "U2Z",
              "PP1", "PP2", "SP1", "SP2", "SP3", "SPXY", "SPXZ", "SPYZ"]


def mask_of(fields):
    m = 0
    for f in fields:
        m |= 1 << (FIELD_INDEX[f] if isinstance(f, str) else int(f))
    return m


SRC_REFERENCE = mask_of(["txx", "tyy", "tzz"])


class BiotError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("order", C.c_int32), ("precision", C.c_int32), ("npml", C.c_int32),
                ("dx", C.c_float), ("dt", C.c_float), ("pml_v0", C.c_float),
                ("pml_R", C.c_float), ("Cv", C.c_float), ("Cp", C.c_float),
                ("device", C.c_int32)]


_DP = C.POINTER(C.c_double)


class Medium(C.Structure):
    _fields_ = [("Ks", _DP), ("Kf", _DP), ("Kb", _DP), ("mu", _DP), ("phi", _DP),
                ("rhos", _DP), ("rhof", _DP), ("eta", _DP), ("perm", _DP),
                ("tao", C.c_double), ("porousm", C.c_double)]


class Params(C.Structure):
    _fields_ = [("lambda_", _DP), ("mu", _DP), ("alpha", _DP), ("M", _DP), ("rho", _DP),
                ("rhof", _DP), ("C1", _DP), ("tphi", _DP)]


PARAM_NAMES = ["lambda", "mu", "alpha", "M", "rho", "rhof"]
FIXED_PARAM_NAMES = ["C1", "tphi"]
NKCOEF = 22
KCOEF_NAMES = ["U1X", "U1Y", "U1Z", "U2X", "U2Y", "U2Z", "W1X", "W1Y", "W1Z", "W2X", "W2Y",
               "W2Z", "W3X", "W3Y", "P1", "P2", "S1", "S2", "S3", "SXY", "SXZ", "SYZ"]

class BatchConfig(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("order", C.c_int32), ("precision", C.c_int32), ("npml", C.c_int32),
                ("dx", C.c_float), ("dt", C.c_float), ("pml_v0", C.c_float),
                ("pml_R", C.c_float), ("Cv", C.c_float), ("Cp", C.c_float),
                ("device", C.c_int32)]


class Shot(C.Structure):
    _fields_ = [("x0", C.c_int32), ("y0", C.c_int32), ("z0", C.c_int32),
                ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("src_cell", C.c_int64), ("src_mask", C.c_uint32),
                ("rec_cells", C.POINTER(C.c_int64)), ("nrec", C.c_int64),
                ("rec_mask", C.c_uint32)]


_lib = None


def lib():
    """Loads lib/libbiot.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with "
                          f"`python mixed-precision-optimization-method_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    H = C.c_void_p
    i32, i64, u32 = C.c_int32, C.c_int64, C.c_uint32
    FP = C.POINTER(C.c_float)
    sig = {
        "biot_create": ([C.POINTER(Grid), C.POINTER(H)], C.c_int),
        "biot_destroy": ([H], C.c_int),
        "biot_last_error": ([H], C.c_char_p),
        "biot_status_string": ([C.c_int], C.c_char_p),
        "biot_abi_version": ([], C.c_int),
        "biot_set_medium": ([H, C.POINTER(Medium)], C.c_int),
        "biot_set_coefficients": ([H, C.POINTER(FP)], C.c_int),
        "biot_get_coefficients": ([H, C.POINTER(FP)], C.c_int),
        "biot_get_pml_profile": ([H, i32, FP, FP, FP, FP], C.c_int),
        "biot_build_coefficients": ([i32, i32, i32, C.c_float, C.c_float, C.c_float,
                                     C.POINTER(Medium), C.POINTER(FP)], C.c_int),
        "biot_pml_profile_1d": ([i32, C.c_float, C.c_float, i32, C.c_float, C.c_float,
                                 FP, FP, FP, FP], C.c_int),
        "biot_set_source": ([H, i64, u32, FP, i32], C.c_int),
        "biot_ricker": ([C.c_float, C.c_float, i32, FP], C.c_int),
        "biot_set_receivers": ([H, C.POINTER(i64), i64, u32, i32], C.c_int),
        "biot_reset": ([H], C.c_int),
        "biot_run": ([H, i32], C.c_int),
        "biot_step_count": ([H, C.POINTER(i32)], C.c_int),
        "biot_synchronize": ([H], C.c_int),
        "biot_last_run_ms": ([H, _DP], C.c_int),
        "biot_profile_steps": ([H, i32, _DP, _DP], C.c_int),
        "biot_set_launch_shape": ([H, i32, i32, i32], C.c_int),
        "biot_get_launch_shape": ([H, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)], C.c_int),
        "biot_autotune": ([H, i32], C.c_int),
        "biot_read_traces": ([H, FP, i64], C.c_int),
        "biot_read_traces_f64": ([H, _DP, i64], C.c_int),
        "biot_read_field": ([H, i32, FP], C.c_int),
        "biot_read_field_f64": ([H, i32, _DP], C.c_int),
        "biot_write_field_f64": ([H, i32, _DP], C.c_int),
        "biot_snapshot_xz": ([H, i32, i32, FP], C.c_int),
        "biot_get_field_scales": ([H, FP], C.c_int),
        "biot_device_bytes": ([H, C.POINTER(i64)], C.c_int),
        "biot_set_params": ([H, C.POINTER(Params)], C.c_int),
        "biot_set_medium_device": ([H, C.POINTER(Medium)], C.c_int),
        "biot_kernel_coefficients": ([C.c_float, i64, C.POINTER(FP), C.POINTER(_DP)], C.c_int),
        "biot_read_kernel_coefficients": ([H, i32, _DP], C.c_int),
        "biot_write_text_2d": ([C.c_char_p, FP, i32, i32], C.c_int),
        "biot_write_text_xz": ([C.c_char_p, FP, i32, i32, i32], C.c_int),
        "biot_write_binary": ([C.c_char_p, FP, i64], C.c_int),
        "biot_run_stream_traces": ([H, i32, i32, FP], C.c_int),
        "biot_run_stream": ([H, i32, i32, FP, i32, i32, C.POINTER(C.c_int32), i32, FP], C.c_int),
        "biot_forward_checkpointed": ([H, i32, i32], C.c_int),
        "biot_misfit_l2": ([H, FP, FP, _DP], C.c_int),
        "biot_misfit_l2_noalloc": ([H, FP, FP, _DP], C.c_int),
        "biot_adjoint": ([H, FP, C.POINTER(_DP), C.POINTER(_DP)], C.c_int),
        "biot_params_coefficients": ([C.POINTER(Grid), C.POINTER(Params), C.POINTER(_DP)], C.c_int),
        "biot_params_chain": ([C.POINTER(Grid), C.POINTER(Params), C.POINTER(_DP), C.POINTER(_DP)],
                              C.c_int),
        "biot_batch_create": ([C.POINTER(BatchConfig), C.POINTER(H)], C.c_int),
        "biot_batch_destroy": ([H], C.c_int),
        "biot_batch_last_error": ([H], C.c_char_p),
        "biot_batch_add_shot": ([H, C.POINTER(Shot), C.POINTER(Params), FP, i32], C.c_int),
        "biot_batch_shot_count": ([H, C.POINTER(i32)], C.c_int),
        "biot_batch_set_model": ([H, C.POINTER(Params)], C.c_int),
        "biot_batch_shot_handle": ([H, i32, C.POINTER(H)], C.c_int),
        "biot_batch_set_launch": ([H, i32, i32, i32, i32], C.c_int),
        "biot_batch_get_launch": ([H, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                   C.POINTER(i32)], C.c_int),
        "biot_batch_autotune": ([H, i32], C.c_int),
        "biot_batch_forward": ([H, i32, i32], C.c_int),
        "biot_batch_read_traces": ([H, i32, FP, i64], C.c_int),
        "biot_batch_gradient": ([H, i32, i32, C.POINTER(FP), _DP, C.POINTER(_DP)], C.c_int),
        "biot_batch_last_ms": ([H, _DP], C.c_int),
        "biot_batch_gradient_groups": ([H, C.POINTER(i32)], C.c_int),
        "biot_batch_release": ([H], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def exported_symbols():
    """Names the binding expects the library to export."""
    return ["biot_create", "biot_destroy", "biot_last_error", "biot_status_string",
            "biot_abi_version", "biot_set_medium", "biot_set_coefficients",
            "biot_get_coefficients", "biot_get_pml_profile", "biot_build_coefficients",
            "biot_pml_profile_1d", "biot_set_source", "biot_ricker",
            "biot_set_receivers", "biot_reset", "biot_run", "biot_step_count",
            "biot_synchronize", "biot_last_run_ms", "biot_profile_steps",
            "biot_set_launch_shape", "biot_get_launch_shape", "biot_autotune",
            "biot_read_traces", "biot_read_traces_f64", "biot_read_field",
            "biot_read_field_f64", "biot_write_field_f64", "biot_snapshot_xz",
            "biot_get_field_scales", "biot_device_bytes", "biot_write_text_2d",
            "biot_write_text_xz", "biot_write_binary", "biot_run_stream_traces", "biot_run_stream",
            "biot_set_params",
            "biot_set_medium_device", "biot_kernel_coefficients", "biot_read_kernel_coefficients",
            "biot_forward_checkpointed", "biot_misfit_l2", "biot_misfit_l2_noalloc", "biot_adjoint",
            "biot_params_coefficients", "biot_params_chain", "biot_batch_create",
            "biot_batch_destroy", "biot_batch_last_error", "biot_batch_add_shot",
            "biot_batch_shot_count", "biot_batch_set_model", "biot_batch_shot_handle",
            "biot_batch_set_launch",
            "biot_batch_get_launch", "biot_batch_autotune", "biot_batch_forward",
            "biot_batch_read_traces", "biot_batch_gradient", "biot_batch_last_ms",
            "biot_batch_gradient_groups",
            "biot_batch_release"]


def fptr(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_float))


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def i64ptr(a):
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def ricker(f0, dt, nt):
    """Reference Ricker wavelet (kernel.cu:1898) computed by the library."""
    out = np.zeros(nt, dtype=np.float32)
    rc = lib().biot_ricker(C.c_float(f0), C.c_float(dt), int(nt), fptr(out))
    if rc != BIOT_OK:
        raise BiotError(rc, "biot_ricker")
    return out


def build_coefficients(nx, ny, nz, dt, medium, Cv=1e8, Cp=1.0, tao=3.0, porousm=8.0):
    """Host-only coefficient builder (main.cpp:322-494) of the library; no GPU needed.
    medium: dict of scalars or (nz, ny, nx) arrays.  Returns 23 float32 arrays."""
    shape = (nz, ny, nx)
    keys = ("Ks", "Kf", "Kb", "mu", "phi", "rhos", "rhof", "eta", "perm")
    arrs = {k: np.ascontiguousarray(np.broadcast_to(np.asarray(medium[k], np.float64), shape)).reshape(-1)
            for k in keys}
    m = Medium(**{k: dptr(arrs[k]) for k in keys}, tao=tao, porousm=porousm)
    out = [np.zeros(nx * ny * nz, np.float32) for _ in range(NCOEF)]
    ptrs = (C.POINTER(C.c_float) * NCOEF)(*[fptr(a) for a in out])
    rc = lib().biot_build_coefficients(nx, ny, nz, dt, Cv, Cp, C.byref(m), ptrs)
    if rc != BIOT_OK:
        raise BiotError(rc, "biot_build_coefficients")
    return out


def pml_profile(n, dx, dt, npml, v0=4000.0, R=1e-6):
    """Host-only single-domain C-PML profile (main.cpp:537-566): d, d2, e, e2."""
    arrs = [np.zeros(n, np.float32) for _ in range(4)]
    rc = lib().biot_pml_profile_1d(n, dx, dt, npml, v0, R, *[fptr(a) for a in arrs])
    if rc != BIOT_OK:
        raise BiotError(rc, "biot_pml_profile_1d")
    return arrs


def params_struct(params, n):
    """Params structure over a dict of the eight per-cell arrays (kept alive by the caller
    through the returned tuple's second element)."""
    keep = {k: np.ascontiguousarray(np.asarray(params[k], np.float64).reshape(-1))
            for k in PARAM_NAMES + FIXED_PARAM_NAMES}
    for k, a in keep.items():
        if a.size != n:
            raise ValueError(f"parameter {k} has {a.size} cells, expected {n}")
    st = Params(lambda_=dptr(keep["lambda"]), mu=dptr(keep["mu"]), alpha=dptr(keep["alpha"]),
                M=dptr(keep["M"]), rho=dptr(keep["rho"]), rhof=dptr(keep["rhof"]),
                C1=dptr(keep["C1"]), tphi=dptr(keep["tphi"]))
    return st, keep


def _host_grid(nx, ny, nz, dx, dt, Cv, Cp):
    g = Grid()
    g.nx, g.ny, g.nz, g.dx, g.dt, g.Cv, g.Cp = nx, ny, nz, dx, dt, Cv, Cp
    return g


def params_coefficients(nx, ny, nz, dx, dt, params, Cv=1e8, Cp=1.0):
    """Host-only: the 22 kernel coefficient arrays (double, kernel order) of a parameter set."""
    n = nx * ny * nz
    st, keep = params_struct(params, n)
    out = [np.zeros(n, np.float64) for _ in range(NKCOEF)]
    ptrs = (_DP * NKCOEF)(*[dptr(a) for a in out])
    g = _host_grid(nx, ny, nz, dx, dt, Cv, Cp)
    rc = lib().biot_params_coefficients(C.byref(g), C.byref(st), ptrs)
    if rc != BIOT_OK:
        raise BiotError(rc, "biot_params_coefficients")
    return out


def params_chain(nx, ny, nz, dx, dt, params, grad_coef, Cv=1e8, Cp=1.0):
    """Host-only: dJ/dK (22 arrays) -> dJ/d(lambda, mu, alpha, M, rho, rho_f)."""
    n = nx * ny * nz
    st, keep = params_struct(params, n)
    gk = [np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)) for a in grad_coef]
    out = [np.zeros(n, np.float64) for _ in range(6)]
    g = _host_grid(nx, ny, nz, dx, dt, Cv, Cp)
    rc = lib().biot_params_chain(C.byref(g), C.byref(st), (_DP * NKCOEF)(*[dptr(a) for a in gk]),
                                 (_DP * 6)(*[dptr(a) for a in out]))
    if rc != BIOT_OK:
        raise BiotError(rc, "biot_params_chain")
    return out


def write_text_2d(path, data):
    """The reference's text format (outputfunc.cpp:35-50) for a 2-D array (rows x columns)."""
    a = np.ascontiguousarray(np.asarray(data, np.float32))
    if a.ndim != 2:
        raise ValueError("2-D array expected")
    rc = lib().biot_write_text_2d(os.fsencode(path), fptr(a), a.shape[0], a.shape[1])
    if rc != BIOT_OK:
        raise BiotError(rc, f"biot_write_text_2d({path})")


def write_text_xz(path, volume):
    """The y = ny/2 slice of a (nz, ny, nx) volume in the reference's text format."""
    a = np.ascontiguousarray(np.asarray(volume, np.float32))
    if a.ndim != 3:
        raise ValueError("(nz, ny, nx) volume expected")
    rc = lib().biot_write_text_xz(os.fsencode(path), fptr(a), *a.shape)
    if rc != BIOT_OK:
        raise BiotError(rc, f"biot_write_text_xz({path})")


def write_binary(path, data):
    a = np.ascontiguousarray(np.asarray(data, np.float32)).reshape(-1)
    rc = lib().biot_write_binary(os.fsencode(path), fptr(a), a.size)
    if rc != BIOT_OK:
        raise BiotError(rc, f"biot_write_binary({path})")


def kernel_coefficients(dx, coefs):
    """Host-only: the 22 unscaled kernel coefficient arrays (double) of the 23 reference arrays."""
    arrs = [np.ascontiguousarray(np.asarray(c, np.float32).reshape(-1)) for c in coefs]
    n = arrs[0].size
    out = [np.zeros(n, np.float64) for _ in range(NKCOEF)]
    rc = lib().biot_kernel_coefficients(C.c_float(dx), n, (C.POINTER(C.c_float) * NCOEF)(*[fptr(a) for a in arrs]),
                                        (_DP * NKCOEF)(*[dptr(a) for a in out]))
    if rc != BIOT_OK:
        raise BiotError(rc, "biot_kernel_coefficients")
    return out
