"""Host-side driver of the Biot solver over the C-ABI.

`Solver` mirrors the reference's driver surface (main.cpp + callCUDAKernel): grid and medium
set-up, one source, a receiver layout, the time loop, traces and xz snapshots.  All
compute runs in lib/libbiot.so on the GPU.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (BIOT_OK, FIELD_INDEX, FIELDS, NCOEF, BiotError, Grid, Medium, dptr, fptr,
                   i64ptr, lib, mask_of)

PRECISION = {"fp32": _lib.FP32, "fp16": _lib.FP16, "fp64": _lib.FP64, "fp16c": _lib.FP16C,
             "fp16t": _lib.FP16T, "fp16m": _lib.FP16M}
MEDIUM_KEYS = ("Ks", "Kf", "Kb", "mu", "phi", "rhos", "rhof", "eta", "perm")


class Solver:
    """One handle: a grid on one GPU with its medium, source, receivers and wavefields."""

    def __init__(self, nx, ny, nz, dx, dt, order=2, precision="fp32", npml=10,
                 pml_v0=4000.0, pml_R=1e-6, Cv=1e8, Cp=1.0, device=0):
        self.shape = (int(nz), int(ny), int(nx))  # numpy view order (k, j, i)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.n = self.nx * self.ny * self.nz
        self.precision = precision
        g = Grid(nx=nx, ny=ny, nz=nz, order=order, precision=PRECISION[precision], npml=npml,
                 dx=dx, dt=dt, pml_v0=pml_v0, pml_R=pml_R, Cv=Cv, Cp=Cp, device=device)
        self.grid = g
        h = C.c_void_p()
        rc = lib().biot_create(C.byref(g), C.byref(h))
        if rc != BIOT_OK:
            msg = lib().biot_last_error(None)
            raise BiotError(rc, "biot_create failed for grid %dx%dx%d order %d: %s"
                            % (nx, ny, nz, order, msg.decode() if msg else ""))
        self._h = h
        self.ncomp = 0
        self.nrec = 0
        self.nt_max = 0
        self.rec_fields = []
        self.ckpt_generation = 0  # counts checkpointed forward runs (the adjoint's state)

    # -- helpers -----------------------------------------------------------------------
    def _check(self, rc, what):
        if rc != BIOT_OK:
            msg = lib().biot_last_error(self._h)
            raise BiotError(rc, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().biot_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def cell(self, i, j, k):
        """Linear index of cell (i, j, k), x fastest (main.cpp:149)."""
        return int(i) + int(j) * self.nx + int(k) * self.nx * self.ny

    # -- medium ------------------------------------------------------------------------
    def set_medium(self, medium, tao=3.0, porousm=8.0):
        """medium: dict with Ks Kf Kb mu phi rhos rhof eta perm (scalars or per-cell arrays
        of shape (nz, ny, nx))."""
        arrs = {}
        for k in MEDIUM_KEYS:
            v = np.asarray(medium[k], dtype=np.float64)
            arrs[k] = np.ascontiguousarray(np.broadcast_to(v, self.shape)).reshape(-1)
        m = Medium(**{k: dptr(arrs[k]) for k in MEDIUM_KEYS}, tao=tao, porousm=porousm)
        self._check(lib().biot_set_medium(self._h, C.byref(m)), "biot_set_medium")

    def set_medium_device(self, medium, tao=3.0, porousm=8.0):
        """set_medium with the moduli, averages and coefficients built on the GPU."""
        arrs = {}
        for k in MEDIUM_KEYS:
            v = np.asarray(medium[k], dtype=np.float64)
            arrs[k] = np.ascontiguousarray(np.broadcast_to(v, self.shape)).reshape(-1)
        m = Medium(**{k: dptr(arrs[k]) for k in MEDIUM_KEYS}, tao=tao, porousm=porousm)
        self._check(lib().biot_set_medium_device(self._h, C.byref(m)), "biot_set_medium_device")

    def kernel_coefficient(self, k):
        """Stored kernel coefficient array k (enum biot_kcoef order) as double."""
        out = np.zeros(self.n, np.float64)
        self._check(lib().biot_read_kernel_coefficients(self._h, int(k), dptr(out)),
                    "biot_read_kernel_coefficients")
        return out.reshape(self.shape)

    def set_coefficients(self, coefs):
        arrs = [np.ascontiguousarray(np.asarray(c, np.float32).reshape(-1)) for c in coefs]
        ptrs = (C.POINTER(C.c_float) * NCOEF)(*[fptr(a) for a in arrs])
        self._check(lib().biot_set_coefficients(self._h, ptrs), "biot_set_coefficients")

    def coefficients(self):
        out = [np.zeros(self.n, np.float32) for _ in range(NCOEF)]
        ptrs = (C.POINTER(C.c_float) * NCOEF)(*[fptr(a) for a in out])
        self._check(lib().biot_get_coefficients(self._h, ptrs), "biot_get_coefficients")
        return [a.reshape(self.shape) for a in out]

    def pml_profile(self, axis):
        n = (self.nx, self.ny, self.nz)[axis]
        arrs = [np.zeros(n, np.float32) for _ in range(4)]
        self._check(lib().biot_get_pml_profile(self._h, axis, *[fptr(a) for a in arrs]),
                    "biot_get_pml_profile")
        return arrs  # d, d2, e, e2

    # -- source / receivers --------------------------------------------------------------
    def set_source(self, cell, wavelet, fields=("txx", "tyy", "tzz")):
        w = np.ascontiguousarray(np.asarray(wavelet, np.float32))
        self._check(lib().biot_set_source(self._h, int(cell), mask_of(fields), fptr(w), len(w)),
                    "biot_set_source")

    def set_receivers(self, cells, fields, nt_max):
        c = np.ascontiguousarray(np.asarray(cells, np.int64).reshape(-1))
        self._check(lib().biot_set_receivers(self._h, i64ptr(c), len(c), mask_of(fields),
                                             int(nt_max)), "biot_set_receivers")
        m = mask_of(fields)
        self.rec_fields = [f for f in FIELDS if m & (1 << FIELD_INDEX[f])]
        self.ncomp, self.nrec, self.nt_max = len(self.rec_fields), len(c), int(nt_max)

    # -- time loop -----------------------------------------------------------------------
    def reset(self):
        self._check(lib().biot_reset(self._h), "biot_reset")

    def run(self, nsteps):
        self._check(lib().biot_run(self._h, int(nsteps)), "biot_run")

    def step(self):
        it = C.c_int32()
        self._check(lib().biot_step_count(self._h, C.byref(it)), "biot_step_count")
        return it.value

    def last_run_ms(self):
        ms = C.c_double()
        self._check(lib().biot_last_run_ms(self._h, C.byref(ms)), "biot_last_run_ms")
        return ms.value

    def set_launch_shape(self, bx, by, zchunk):
        self._check(lib().biot_set_launch_shape(self._h, bx, by, zchunk), "biot_set_launch_shape")

    def launch_shape(self):
        """(bx, by, zchunk) of the stencil kernels."""
        v = [C.c_int32() for _ in range(3)]
        self._check(lib().biot_get_launch_shape(self._h, *[C.byref(x) for x in v]),
                    "biot_get_launch_shape")
        return tuple(x.value for x in v)

    def autotune(self, trial_steps=5):
        """Picks the fastest tile shape for this grid (cached per GPU arch); resets the
        wavefields."""
        self._check(lib().biot_autotune(self._h, int(trial_steps)), "biot_autotune")
        return self.launch_shape()

    def profile_steps(self, nsteps):
        """Advances nsteps with HIP events around every stencil launch; returns the mean
        (velocity, stress) kernel durations in ms."""
        a, b = C.c_double(), C.c_double()
        self._check(lib().biot_profile_steps(self._h, int(nsteps), C.byref(a), C.byref(b)),
                    "biot_profile_steps")
        return a.value, b.value

    # -- outputs -------------------------------------------------------------------------
    def traces(self):
        """Receiver traces, shape (ncomp, nt_max, nrec) in recorded-field order."""
        out = np.zeros((self.ncomp, self.nt_max, self.nrec), np.float32)
        self._check(lib().biot_read_traces(self._h, fptr(out), out.size), "biot_read_traces")
        return out

    def field(self, name, f64=False):
        f = FIELD_INDEX[name] if isinstance(name, str) else int(name)
        if f64:
            out = np.zeros(self.n, np.float64)
            self._check(lib().biot_read_field_f64(self._h, f, dptr(out)), "biot_read_field_f64")
        else:
            out = np.zeros(self.n, np.float32)
            self._check(lib().biot_read_field(self._h, f, fptr(out)), "biot_read_field")
        return out.reshape(self.shape)

    def write_field(self, name, values):
        f = FIELD_INDEX[name] if isinstance(name, str) else int(name)
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(values, np.float64), self.shape)).reshape(-1)
        self._check(lib().biot_write_field_f64(self._h, f, dptr(v)), "biot_write_field_f64")

    def snapshot_xz(self, name, j):
        f = FIELD_INDEX[name] if isinstance(name, str) else int(name)
        out = np.zeros((self.nz, self.nx), np.float32)
        self._check(lib().biot_snapshot_xz(self._h, f, int(j), fptr(out)), "biot_snapshot_xz")
        return out

    def field_scales(self):
        out = np.zeros(13, np.float32)
        self._check(lib().biot_get_field_scales(self._h, fptr(out)), "biot_get_field_scales")
        return dict(zip(FIELDS, out.tolist()))

    def device_bytes(self):
        b = C.c_int64()
        self._check(lib().biot_device_bytes(self._h, C.byref(b)), "biot_device_bytes")
        return b.value

    def run_stream_traces(self, nsteps, chunk=100):
        """nsteps like run(), the traces copied to the host on a side stream chunk steps at a
        time while the loop continues; returns them, shape (ncomp, nt_max, nrec)."""
        out = np.zeros((self.ncomp, self.nt_max, self.nrec), np.float32)
        self._check(lib().biot_run_stream_traces(self._h, int(nsteps), int(chunk), fptr(out)),
                    "biot_run_stream_traces")
        return out

    def run_stream(self, nsteps, chunk=100, snap_field="txx", snap_j=None, snap_steps=(),
                   traces=True):
        """nsteps like run(), with the traces (if traces) and xz snapshots of snap_field at
        y = snap_j (default ny/2) after the listed steps (counted from the current step) copied
        to the host on a side stream; returns (traces or None, snapshots (nsnap, nz, nx))."""
        steps = np.ascontiguousarray(np.asarray(snap_steps, np.int32))
        snaps = np.zeros((steps.size, self.nz, self.nx), np.float32)
        tr = np.zeros((self.ncomp, self.nt_max, self.nrec), np.float32) if traces else None
        f = FIELD_INDEX[snap_field] if isinstance(snap_field, str) else int(snap_field)
        j = self.ny // 2 if snap_j is None else int(snap_j)
        rc = lib().biot_run_stream(self._h, int(nsteps), int(chunk), fptr(tr) if tr is not None else None,
                                   f, j, steps.ctypes.data_as(C.POINTER(C.c_int32)), steps.size,
                                   fptr(snaps))
        self._check(rc, "biot_run_stream")
        return tr, snaps

    def write_traces_text(self, path, comp=0, traces=None):
        """One recorded component in the reference's trace-file layout (sisx*.dat,
        kernel.cu:2557-2571): one row per receiver, nt_max columns."""
        tr = self.traces() if traces is None else traces
        _lib.write_text_2d(path, tr[comp].T)

    def write_snapshot_text(self, path, name):
        """The y = ny/2 xz slice of a field in the reference's snapshot format
        (txxpro*.dat, kernel.cu:2247-2385)."""
        _lib.write_text_2d(path, self.snapshot_xz(name, self.ny // 2))

    def traces_f64(self):
        """Receiver traces in double (FP64 handles only), shape (ncomp, nt_max, nrec)."""
        out = np.zeros((self.ncomp, self.nt_max, self.nrec), np.float64)
        self._check(lib().biot_read_traces_f64(self._h, dptr(out), out.size), "biot_read_traces_f64")
        return out

    # -- inversion -----------------------------------------------------------------------
    def set_params(self, params):
        """Per-cell inversion parameters: dict with lambda mu alpha M rho rhof (inverted) and
        C1 tphi (fixed), arrays of shape (nz, ny, nx) or scalars."""
        full = {k: np.broadcast_to(np.asarray(params[k], np.float64), self.shape)
                for k in _lib.PARAM_NAMES + _lib.FIXED_PARAM_NAMES}
        st, keep = _lib.params_struct(full, self.n)
        self._check(lib().biot_set_params(self._h, C.byref(st)), "biot_set_params")

    def forward_checkpointed(self, nt, every=50):
        """Reset, nt recorded steps, a device checkpoint every `every` steps."""
        self.ckpt_generation += 1
        self._check(lib().biot_forward_checkpointed(self._h, int(nt), int(every)),
                    "biot_forward_checkpointed")

    def misfit(self, observed):
        """(J, residual): J = 1/2 sum (d - observed)^2 reduced in FP64 on the device."""
        obs = np.ascontiguousarray(np.asarray(observed, np.float32))
        if obs.shape != (self.ncomp, self.nt_max, self.nrec):
            raise ValueError(f"observed traces {obs.shape}, expected "
                             f"{(self.ncomp, self.nt_max, self.nrec)}")
        res = np.zeros_like(obs)
        J = C.c_double()
        self._check(lib().biot_misfit_l2(self._h, fptr(obs), fptr(res), C.byref(J)), "biot_misfit_l2")
        return J.value, res

    def adjoint(self, adj_source, params=True, coefs=False):
        """Adjoint of the last checkpointed run for the trace-space gradient adj_source.
        Returns (dict of dJ/d parameter arrays or None, list of 22 dJ/dK arrays or None)."""
        a = np.ascontiguousarray(np.asarray(adj_source, np.float32))
        if a.shape != (self.ncomp, self.nt_max, self.nrec):
            raise ValueError(f"adjoint source {a.shape}, expected "
                             f"{(self.ncomp, self.nt_max, self.nrec)}")
        gp = [np.zeros(self.n, np.float64) for _ in range(6)] if params else None
        gk = [np.zeros(self.n, np.float64) for _ in range(_lib.NKCOEF)] if coefs else None
        pp = (C.POINTER(C.c_double) * 6)(*[dptr(x) for x in gp]) if params else None
        kp = (C.POINTER(C.c_double) * _lib.NKCOEF)(*[dptr(x) for x in gk]) if coefs else None
        self._check(lib().biot_adjoint(self._h, fptr(a), pp, kp), "biot_adjoint")
        outp = {k: v.reshape(self.shape) for k, v in zip(_lib.PARAM_NAMES, gp)} if params else None
        outk = [v.reshape(self.shape) for v in gk] if coefs else None
        return outp, outk
