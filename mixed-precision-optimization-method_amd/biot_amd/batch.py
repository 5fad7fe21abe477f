"""Batched shots over the C-ABI (biot_batch_*): many windows of one model grid advanced in
one persistent, graph-captured launch per half step, and their summed misfit gradient."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (BIOT_OK, FIELD_INDEX, FIELDS, PARAM_NAMES, BatchConfig, BiotError, Shot, dptr,
                   fptr, i64ptr, lib, mask_of)
from .solver import PRECISION


class Batch:
    """Shots cut from a model of shape (nz, ny, nx)."""

    def __init__(self, nx, ny, nz, dx, dt, order=4, precision="fp16", npml=20, pml_v0=4000.0,
                 pml_R=1e-6, Cv=1e8, Cp=1.0, device=0):
        self.model_shape = (int(nz), int(ny), int(nx))
        self.n_model = int(nx) * int(ny) * int(nz)
        c = BatchConfig(nx=nx, ny=ny, nz=nz, order=order, precision=PRECISION[precision],
                        npml=npml, dx=dx, dt=dt, pml_v0=pml_v0, pml_R=pml_R, Cv=Cv, Cp=Cp,
                        device=device)
        h = C.c_void_p()
        rc = lib().biot_batch_create(C.byref(c), C.byref(h))
        if rc != BIOT_OK:
            msg = lib().biot_batch_last_error(None)
            raise BiotError(rc, f"biot_batch_create: {msg.decode() if msg else ''}")
        self._h = h
        self.precision = precision
        self.shots = []  # (ncomp, nt, nrec, window shape, origin)

    def _check(self, rc, what):
        if rc != BIOT_OK:
            msg = lib().biot_batch_last_error(self._h)
            raise BiotError(rc, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().biot_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_shot(self, origin, size, params, src_ijk, wavelet, rec_ijk, rec_fields,
                 src_fields=("txx", "tyy", "tzz", "p")):
        """origin (x0, y0, z0) and size (nx, ny, nz) of the window; params: the window's
        eight parameter arrays (nz, ny, nx); src_ijk / rec_ijk in window coordinates."""
        x0, y0, z0 = (int(v) for v in origin)
        wx, wy, wz = (int(v) for v in size)
        n = wx * wy * wz
        st, keep = _lib.params_struct(params, n)
        rec = np.asarray(rec_ijk, np.int64).reshape(-1, 3)
        cells = np.ascontiguousarray(rec[:, 0] + rec[:, 1] * wx + rec[:, 2] * wx * wy)
        si, sj, sk = (int(v) for v in src_ijk)
        w = np.ascontiguousarray(np.asarray(wavelet, np.float32))
        shot = Shot(x0=x0, y0=y0, z0=z0, nx=wx, ny=wy, nz=wz, src_cell=si + sj * wx + sk * wx * wy,
                    src_mask=mask_of(src_fields), rec_cells=i64ptr(cells), nrec=len(cells),
                    rec_mask=mask_of(rec_fields))
        self._check(lib().biot_batch_add_shot(self._h, C.byref(shot), C.byref(st), fptr(w), len(w)),
                    "biot_batch_add_shot")
        m = mask_of(rec_fields)
        ncomp = sum(1 for f in FIELDS if m & (1 << FIELD_INDEX[f]))
        self.shots.append(dict(ncomp=ncomp, nt=len(w), nrec=len(cells), size=(wx, wy, wz),
                               origin=(x0, y0, z0)))

    def set_model(self, params):
        """Model-grid parameters (dict of the eight arrays, shape (nz, ny, nx)): every shot's
        window is cut out and its coefficients rebuilt."""
        st, keep = _lib.params_struct(params, self.n_model)
        self._check(lib().biot_batch_set_model(self._h, C.byref(st)), "biot_batch_set_model")

    def set_launch(self, bx, by, zchunk, blocks):
        self._check(lib().biot_batch_set_launch(self._h, bx, by, zchunk, blocks), "biot_batch_set_launch")

    def launch(self):
        v = [C.c_int32() for _ in range(4)]
        self._check(lib().biot_batch_get_launch(self._h, *[C.byref(x) for x in v]), "biot_batch_get_launch")
        return tuple(x.value for x in v)

    def autotune(self, trial_steps=3):
        self._check(lib().biot_batch_autotune(self._h, int(trial_steps)), "biot_batch_autotune")
        return self.launch()

    def forward(self, nt, graph_steps=50):
        self._check(lib().biot_batch_forward(self._h, int(nt), int(graph_steps)), "biot_batch_forward")

    def traces(self, shot):
        d = self.shots[shot]
        out = np.zeros((d["ncomp"], d["nt"], d["nrec"]), np.float32)
        self._check(lib().biot_batch_read_traces(self._h, int(shot), fptr(out), out.size),
                    "biot_batch_read_traces")
        return out

    def gradient(self, nt, observed, every=50):
        """(J, dict of six model-grid gradient arrays (nz, ny, nx))."""
        obs = [np.ascontiguousarray(np.asarray(o, np.float32)) for o in observed]
        if len(obs) != len(self.shots):
            raise ValueError("one observed array per shot")
        for i, (o, d) in enumerate(zip(obs, self.shots)):
            # the library reads the shot's whole trace layout, ncomp x (the steps its receivers
            # hold) x nrec samples, from each array
            want = (d["ncomp"], d["nt"], d["nrec"])
            if o.shape != want:
                raise ValueError(f"shot {i}: observed traces have shape {o.shape}, expected {want}")
        optr = (C.POINTER(C.c_float) * len(obs))(*[fptr(o) for o in obs])
        grads = [np.zeros(self.n_model, np.float64) for _ in range(6)]
        gptr = (C.POINTER(C.c_double) * 6)(*[dptr(g) for g in grads])
        J = C.c_double()
        self._check(lib().biot_batch_gradient(self._h, int(nt), int(every), optr, C.byref(J), gptr),
                    "biot_batch_gradient")
        return J.value, {k: g.reshape(self.model_shape) for k, g in zip(PARAM_NAMES, grads)}

    def last_ms(self):
        ms = C.c_double()
        self._check(lib().biot_batch_last_ms(self._h, C.byref(ms)), "biot_batch_last_ms")
        return ms.value

    def gradient_groups(self):
        g = C.c_int32()
        self._check(lib().biot_batch_gradient_groups(self._h, C.byref(g)), "biot_batch_gradient_groups")
        return g.value

    def release(self):
        self._check(lib().biot_batch_release(self._h), "biot_batch_release")
