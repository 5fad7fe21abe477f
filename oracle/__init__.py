"""CPU oracle of the Biot time step (TEST INFRASTRUCTURE ONLY).

ctypes wrapper of oracle/liboracle.so, the plain-C restatement of the reference's
arithmetic (see biot_oracle.h for the file:line of each restated part).  Only tests/, the
smoke check and the benchmark's CPU-baseline leg use it, as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MODE = {"f32": 0, "f16": 1, "f64": 2}
FIELDS = ["vx", "vy", "vz", "qx", "qy", "qz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "p"]
NCOEF = 23
FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB_PATH


class OrConfig(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("order", C.c_int),
                ("mode", C.c_int), ("DT", C.c_float), ("Hinv", C.c_float),
                ("coef", FP * NCOEF), ("pd", FP * 3), ("pd2", FP * 3), ("pe", FP * 3),
                ("pe2", FP * 3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_f32_to_f16.argtypes = [C.c_float]
        L.or_f32_to_f16.restype = C.c_uint16
        L.or_f64_to_f16.argtypes = [C.c_double]
        L.or_f64_to_f16.restype = C.c_uint16
        L.or_f16_to_f32.argtypes = [C.c_uint16]
        L.or_f16_to_f32.restype = C.c_float
        L.or_biot_moduli.argtypes = [C.c_int64] + [DP] * 9 + [C.c_double, C.c_double] + [FP] * 9
        L.or_stagger_average.argtypes = [C.c_int] * 3 + [FP] * 5 + [FP] * 15
        L.or_coefficients.argtypes = ([C.c_int] * 3 + [C.c_float] * 3 + [FP] * 19 + [FP * NCOEF])
        L.or_pml_profile.argtypes = [C.c_int, C.c_float, C.c_float, C.c_int, C.c_float,
                                     C.c_float, FP, FP, FP, FP]
        L.or_ricker.argtypes = [C.c_int, C.c_float, C.c_float]
        L.or_ricker.restype = C.c_float
        L.or_state_create.argtypes = [C.POINTER(OrConfig)]
        L.or_state_create.restype = C.c_void_p
        L.or_state_destroy.argtypes = [C.c_void_p]
        L.or_state_zero.argtypes = [C.c_void_p]
        L.or_state_get_field.argtypes = [C.c_void_p, C.c_int, FP]
        L.or_state_get_field_f64.argtypes = [C.c_void_p, C.c_int, DP]
        L.or_state_set_field_f64.argtypes = [C.c_void_p, C.c_int, DP]
        L.or_source.argtypes = [C.c_void_p, C.c_int64, C.c_uint, C.c_float]
        L.or_step_v.argtypes = [C.c_void_p]
        L.or_step_t.argtypes = [C.c_void_p]
        L.or_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, FP, C.c_int64, C.c_uint,
                             C.c_int, C.POINTER(C.c_int64), C.c_uint, FP]
        L.or_run_f64.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, FP, C.c_int64, C.c_uint,
                                 C.c_int, C.POINTER(C.c_int64), C.c_uint, DP]
        L.or_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(FP)


def _d(a):
    return a.ctypes.data_as(DP)


def mask_of(fields):
    m = 0
    for f in fields:
        m |= 1 << FIELDS.index(f)
    return m


def biot_moduli(medium, shape, tao=3.0, porousm=8.0):
    """outputfunc.cpp:205-217 -> dict of float32 arrays M HH H2u C C1 C2 rho rhof mu."""
    n = int(np.prod(shape))
    ins = {k: np.ascontiguousarray(np.broadcast_to(np.asarray(medium[k], np.float64), shape)).reshape(-1)
           for k in ("Ks", "Kf", "Kb", "mu", "phi", "rhos", "rhof", "eta", "perm")}
    names = ["M", "HH", "H2u", "C", "C1", "C2", "rho", "rhof", "mu"]
    outs = {k: np.zeros(n, np.float32) for k in names}
    lib().or_biot_moduli(n, *[_d(ins[k]) for k in ("Ks", "Kf", "Kb", "mu", "phi", "rhos", "rhof",
                                                     "eta", "perm")],
                         tao, porousm, *[_f(outs[k]) for k in names])
    return outs


def stagger_average(mod, nx, ny, nz):
    n = nx * ny * nz
    names = ["muxz", "muxy", "muyz", "rhox", "rhoy", "rhoz", "rhofx", "rhofy", "rhofz",
             "C1x", "C1y", "C1z", "C2x", "C2y", "C2z"]
    outs = {k: np.zeros(n, np.float32) for k in names}
    lib().or_stagger_average(nx, ny, nz, *[_f(mod[k]) for k in ("mu", "rho", "rhof", "C1", "C2")],
                             *[_f(outs[k]) for k in names])
    return outs


def coefficients(mod, av, nx, ny, nz, DT, Cv=1e8, Cp=1.0):
    """main.cpp:431-494 with the moduli masked to the interior like main.cpp:355-358."""
    n = nx * ny * nz
    interior = np.zeros((nz, ny, nx), bool)
    interior[1:-1, 1:-1, 1:-1] = True
    interior = interior.reshape(-1)
    masked = {k: np.where(interior, mod[k], 0).astype(np.float32) for k in ("M", "C", "HH", "H2u")}
    coef = [np.zeros(n, np.float32) for _ in range(NCOEF)]
    arr = (FP * NCOEF)(*[_f(c) for c in coef])
    lib().or_coefficients(nx, ny, nz, DT, Cv, Cp,
                          *[_f(av[k]) for k in ("muxz", "muxy", "muyz", "rhox", "rhoy", "rhoz",
                                                 "rhofx", "rhofy", "rhofz", "C1x", "C1y", "C1z",
                                                 "C2x", "C2y", "C2z")],
                          *[_f(masked[k]) for k in ("M", "C", "HH", "H2u")], arr)
    return coef


def pml_profile(n, H, DT, NP, v0=4000.0, RC=1e-6):
    arrs = [np.zeros(n, np.float32) for _ in range(4)]
    lib().or_pml_profile(n, H, DT, NP, v0, RC, *[_f(a) for a in arrs])
    return arrs  # d, d2, e, e2


def ricker(it, DT, F0):
    return lib().or_ricker(int(it), DT, F0)


def ricker_series(nt, DT, F0):
    return np.array([ricker(i, DT, F0) for i in range(nt)], np.float32)


class OracleRun:
    """One oracle state: set-up from a medium exactly as the reference's main.cpp does it,
    then the time loop of callCUDAKernel on the CPU."""

    def __init__(self, nx, ny, nz, dx, dt, order=2, mode="f32", npml=10, medium=None,
                 coef=None, Cv=1e8, Cp=1.0, v0=4000.0, RC=1e-6, tao=3.0, porousm=8.0):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.shape = (nz, ny, nx)
        self.n = nx * ny * nz
        self.mode = mode
        if coef is None:
            mod = biot_moduli(medium, self.shape, tao, porousm)
            av = stagger_average(mod, nx, ny, nz)
            coef = coefficients(mod, av, nx, ny, nz, dt, Cv, Cp)
        self.coef = [np.ascontiguousarray(np.asarray(c, np.float32).reshape(-1)) for c in coef]
        self.prof = []
        for n_ in (nx, ny, nz):
            if npml > 0:
                d, d2, e, e2 = pml_profile(n_, dx, dt, npml, v0, RC)
            else:
                d = d2 = e = e2 = np.zeros(n_, np.float32)
            pack = [np.ascontiguousarray((a * np.float32(0.01)).astype(np.float32))
                    for a in (d, d2, e, e2)]
            self.prof.append(pack)
        cfg = OrConfig()
        cfg.nx, cfg.ny, cfg.nz, cfg.order, cfg.mode = nx, ny, nz, order, MODE[mode]
        cfg.DT = dt
        cfg.Hinv = np.float32(1.0) / np.float32(dx)
        for a in range(NCOEF):
            cfg.coef[a] = _f(self.coef[a])
        for ax in range(3):
            cfg.pd[ax] = _f(self.prof[ax][0])
            cfg.pd2[ax] = _f(self.prof[ax][1])
            cfg.pe[ax] = _f(self.prof[ax][2])
            cfg.pe2[ax] = _f(self.prof[ax][3])
        self._cfg = cfg
        self._st = lib().or_state_create(C.byref(cfg))
        if not self._st:
            raise ValueError("or_state_create rejected the configuration")

    def close(self):
        if self._st:
            lib().or_state_destroy(self._st)
            self._st = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, it0, nsteps, nt_total, wavelet, src_cell, src_fields, rec_cells, rec_fields):
        w = np.ascontiguousarray(np.asarray(wavelet, np.float32))
        rec = np.ascontiguousarray(np.asarray(rec_cells, np.int64).reshape(-1))
        ncomp = len([f for f in FIELDS if f in rec_fields])
        dt = np.float64 if self.mode == "f64" else np.float32
        tr = np.zeros((ncomp, nt_total, len(rec)), dt)
        fn = lib().or_run_f64 if self.mode == "f64" else lib().or_run
        fn(self._st, it0, nsteps, nt_total, _f(w), int(src_cell), mask_of(src_fields), len(rec),
           rec.ctypes.data_as(C.POINTER(C.c_int64)), mask_of(rec_fields),
           tr.ctypes.data_as(DP if self.mode == "f64" else FP))
        return tr

    def step_v(self):
        lib().or_step_v(self._st)

    def step_t(self):
        lib().or_step_t(self._st)

    def source(self, cell, fields, amp):
        lib().or_source(self._st, int(cell), mask_of(fields), float(amp))

    def field(self, name):
        out = np.zeros(self.n, np.float64)
        lib().or_state_get_field_f64(self._st, FIELDS.index(name), _d(out))
        return out.reshape(self.shape)

    def set_field(self, name, values):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(values, np.float64), self.shape)).reshape(-1)
        lib().or_state_set_field_f64(self._st, FIELDS.index(name), _d(v))


def num_threads():
    return lib().or_num_threads()
