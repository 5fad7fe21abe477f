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
DEFAULT_LIB_PATH = os.path.join(HERE, "liboracle.so")
# BIOT_ORACLE_LIB names another build of the same sources (the mutant checks of
# tests/test_oracle_mutants.py load deliberately broken builds through it)
LIB_PATH = os.environ.get("BIOT_ORACLE_LIB") or DEFAULT_LIB_PATH
MODE = {"f32": 0, "f16": 1, "f64": 2}
FIELDS = ["vx", "vy", "vz", "qx", "qy", "qz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "p"]
NCOEF = 23
FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return DEFAULT_LIB_PATH


REF_SRC = "/root/reference"
REF_LIB_PATH = os.path.join(HERE, "_ref", "liboutputfunc.so")


def build_ref():
    """Compile the reference's own outputfunc.cpp into _ref/liboutputfunc.so (`make ref`)
    where the reference sources are present; returns the library path or None."""
    if not os.path.isfile(os.path.join(REF_SRC, "outputfunc.cpp")):
        return None
    subprocess.run(["make", "-C", HERE, "-s", "ref", f"REF={REF_SRC}"], check=True)
    return REF_LIB_PATH


class RefOutputfunc:
    """ctypes binding of the reference's compiled outputfunc.cpp (C++ linkage, so the symbols
    carry their mangled names): the float*** allocator space3d, create_model_all and the text
    writers wfile / wfile3d / wfile1d (outputfunc.cpp:14-151, 153-222)."""

    F1 = C.POINTER(C.c_float)
    F2 = C.POINTER(F1)
    F3 = C.POINTER(F2)
    # create_model_all's 14 volumes in argument order (outputfunc.cpp:153-154)
    MODEL = ("vp", "vs", "rhos", "vf", "vfg", "rho", "rhof", "M", "C", "C1", "C2", "HH", "H2u", "mu")

    def __init__(self, path=REF_LIB_PATH):
        L = C.CDLL(path)
        i = C.c_int
        self.space2d = L["_Z7space2dii"]
        self.space2d.argtypes, self.space2d.restype = [i, i], self.F2
        self.free_space2d = L["_Z12free_space2dPPfi"]
        self.free_space2d.argtypes, self.free_space2d.restype = [self.F2, i], None
        self.space3d = L["_Z7space3diii"]
        self.space3d.argtypes, self.space3d.restype = [i, i, i], self.F3
        self.free_space3d = L["_Z12free_space3dPPPfii"]
        self.free_space3d.argtypes, self.free_space3d.restype = [self.F3, i, i], None
        self.create_model_all = L["_Z16create_model_allPPPfS1_S1_S1_S1_S1_S1_S1_S1_S1_S1_S1_S1_S1_iii"]
        self.create_model_all.argtypes, self.create_model_all.restype = [self.F3] * 14 + [i, i, i], None
        self.wfile = L["_Z5wfilePcPPfii"]
        self.wfile.argtypes, self.wfile.restype = [C.c_char_p, self.F2, i, i], None
        self.wfile3d = L["_Z7wfile3dPcPPPfiii"]
        self.wfile3d.argtypes, self.wfile3d.restype = [C.c_char_p, self.F3, i, i, i], None
        self.wfile1d = L["_Z7wfile1dPcPfiii"]
        self.wfile1d.argtypes, self.wfile1d.restype = [C.c_char_p, self.F1, i, i, i], None

    def model(self, nz, ny, nx):
        """create_model_all on an nz x ny x nx grid: dict of (nz, ny, nx) float32 arrays."""
        vols = {k: self.space3d(nz, ny, nx) for k in self.MODEL}
        try:
            self.create_model_all(*[vols[k] for k in self.MODEL], nz, ny, nx)
            return {k: self.to_numpy(v, nz, ny, nx) for k, v in vols.items()}
        finally:
            for v in vols.values():
                self.free_space3d(v, nz, ny)

    @staticmethod
    def to_numpy(a, nz, ny, nx):
        out = np.empty((nz, ny, nx), np.float32)
        for k in range(nz):
            for j in range(ny):
                C.memmove(out[k, j].ctypes.data, a[k][j], nx * 4)
        return out

    def write_2d(self, path, data):
        """wfile on a space2d copy of a 2-D array."""
        nr, nc = data.shape
        a = self.space2d(nr, nc)
        try:
            for r in range(nr):
                row = np.ascontiguousarray(data[r], np.float32)
                C.memmove(a[r], row.ctypes.data, nc * 4)
            self.wfile(os.fsencode(path), a, nr, nc)
        finally:
            self.free_space2d(a, nr)

    def write_3d(self, path, vol):
        """wfile3d on a space3d copy of an (nz, ny, nx) volume."""
        nz, ny, nx = vol.shape
        a = self.space3d(nz, ny, nx)
        try:
            for k in range(nz):
                for j in range(ny):
                    row = np.ascontiguousarray(vol[k, j], np.float32)
                    C.memmove(a[k][j], row.ctypes.data, nx * 4)
            self.wfile3d(os.fsencode(path), a, nz, ny, nx)
        finally:
            self.free_space3d(a, nz, ny)

    def write_1d(self, path, vol):
        """wfile1d on a flat (nz, ny, nx) volume."""
        a = np.ascontiguousarray(vol, np.float32)
        self.wfile1d(os.fsencode(path), a.ctypes.data_as(self.F1), *a.shape)


class OrConfig(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("order", C.c_int),
                ("mode", C.c_int), ("DT", C.c_float), ("Hinv", C.c_float),
                ("coef", FP * NCOEF), ("pd", FP * 3), ("pd2", FP * 3), ("pe", FP * 3),
                ("pe2", FP * 3)]


_lib = None


class HaProblem(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("order", C.c_int),
                ("prof", (DP * 4) * 3), ("K", DP * 22), ("src_cell", C.c_int64),
                ("src_mask", C.c_uint), ("wavelet", FP), ("nt_w", C.c_int),
                ("rec", C.POINTER(C.c_int64)), ("nrec", C.c_int64), ("comp", C.POINTER(C.c_int)),
                ("ncomp", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            if LIB_PATH != DEFAULT_LIB_PATH:
                raise FileNotFoundError(f"oracle library {LIB_PATH} (BIOT_ORACLE_LIB) is missing")
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
        L.ha_forward.argtypes = [C.POINTER(HaProblem), C.c_int, DP]
        L.ha_adjoint.argtypes = [C.POINTER(HaProblem), C.c_int, C.c_int, DP, DP * 22, DP]
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


# ---- FP64 host adjoint (biot_host_adjoint.c; the reference has no adjoint) ----

def host_profiles(nx, ny, nz, dx, dt, npml, v0=4000.0, RC=1e-6):
    """Per axis (x, y, z) the four double profiles of the folded C-PML carry on the device's
    FP64 path: e and a = -dt d / 2 at integer nodes, then at half nodes, with d and e taken
    through the reference's x0.01 float storage (main.cpp:1512-1523) and back x100."""
    out = []
    c = np.float32(0.01)
    for n in (nx, ny, nz):
        d, d2, e, e2 = pml_profile(n, dx, dt, npml, v0, RC)
        ds, d2s, es, e2s = [(a * c).astype(np.float32).astype(np.float64) for a in (d, d2, e, e2)]
        out.append([es * 100.0, -float(dt) * ds * 100.0 * 0.5,
                    e2s * 100.0, -float(dt) * d2s * 100.0 * 0.5])
    return out


class HostAdjoint:
    """The FP64 host forward and adjoint of one configuration: grid, order, profiles
    (host_profiles), the 22 kernel coefficient arrays K (double, enum biot_kcoef order),
    source (cell, field mask, float wavelet) and receivers (cells, recorded field indices)."""

    def __init__(self, nx, ny, nz, order, prof, K, src_cell, src_mask, wavelet, rec, comp):
        self.n = nx * ny * nz
        self._keep = []

        def keep(a, dt):
            a = np.ascontiguousarray(np.asarray(a, dt).reshape(-1))
            self._keep.append(a)
            return a
        P = HaProblem()
        P.nx, P.ny, P.nz, P.order = nx, ny, nz, order
        for ax in range(3):
            for q in range(4):
                P.prof[ax][q] = keep(prof[ax][q], np.float64).ctypes.data_as(DP)
        if len(K) != 22:
            raise ValueError("22 coefficient arrays expected")
        for q in range(22):
            a = keep(K[q], np.float64)
            if a.size != self.n:
                raise ValueError(f"coefficient {q} has {a.size} cells, expected {self.n}")
            P.K[q] = a.ctypes.data_as(DP)
        P.src_cell, P.src_mask = int(src_cell), int(src_mask)
        w = keep(wavelet, np.float32)
        P.wavelet, P.nt_w = w.ctypes.data_as(FP), w.size
        r = keep(rec, np.int64)
        P.rec, P.nrec = r.ctypes.data_as(C.POINTER(C.c_int64)), r.size
        cp = keep(comp, np.int32)
        P.comp, P.ncomp = cp.ctypes.data_as(C.POINTER(C.c_int)), cp.size
        self.P, self.nrec, self.ncomp = P, r.size, cp.size

    def forward(self, nt):
        tr = np.zeros((self.ncomp, nt, self.nrec), np.float64)
        if lib().ha_forward(C.byref(self.P), int(nt), tr.ctypes.data_as(DP)) != 0:
            raise RuntimeError("ha_forward failed")
        return tr

    def adjoint(self, nt, adj, every=10):
        """dJ/dK (list of 22 arrays) for dJ/d(traces) = adj, and the forward traces."""
        a = np.ascontiguousarray(np.asarray(adj, np.float64))
        if a.shape != (self.ncomp, nt, self.nrec):
            raise ValueError(f"adjoint source shape {a.shape}, expected {(self.ncomp, nt, self.nrec)}")
        gK = [np.zeros(self.n, np.float64) for _ in range(22)]
        tr = np.zeros((self.ncomp, nt, self.nrec), np.float64)
        rc = lib().ha_adjoint(C.byref(self.P), int(nt), int(every), a.ctypes.data_as(DP),
                              (DP * 22)(*[g.ctypes.data_as(DP) for g in gK]), tr.ctypes.data_as(DP))
        if rc != 0:
            raise RuntimeError(f"ha_adjoint failed ({rc})")
        return gK, tr
