"""The standard workloads of this solver, with their sizes.

fwd64   : 64^3, dx 5 m, dt 0.4 ms, 500 steps, PML 10, homogeneous porous rock, 25 Hz
          Ricker on txx/tyy/tzz/p at (32, 32, 32), 32 receivers on x = 16..47, y = 32,
          z = 10 (the first plane inside the PML) recording vx, vz, p.
layered256 : 256^3, 2000 steps, PML 20, three equal z-layers, 20 Hz Ricker at the centre,
          256 x 256 receivers on the plane z = 20 recording vx, vz, p.
perturbed512 : 512 x 512 x 384, 3000 steps, FP16 storage, the layered medium with +-10 %
          Gaussian-correlated (50 m) perturbations of Kb, mu, phi; 128 x 128 receivers (every
          4th cell) on z = 20.
adjoint128 : 128^3, 1000 steps, checkpoints every 50 steps.
batch16 : 16 shots in cubic windows of 96..192 cells cropped from perturbed512.
"""
import numpy as np

from . import media

DX = 5.0
DT = 4e-4
SRC_EXPLOSIVE = ("txx", "tyy", "tzz", "p")
REC_FIELDS = ("vx", "vz", "p")


def fwd64(order=4, nt=500):
    n = 64
    shape = (n, n, n)
    rec = [(i, 32, 10) for i in range(16, 48)]
    return dict(name="fwd64", nx=n, ny=n, nz=n, dx=DX, dt=DT, order=order, npml=10, nt=nt,
                f0=25.0, src=(32, 32, 32), src_fields=SRC_EXPLOSIVE,
                medium=media.homogeneous(shape), rec=rec, rec_fields=REC_FIELDS)


def layered256(order=4, nt=2000, n=256):
    shape = (n, n, n)
    npml = 20
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    rec = np.stack([ii.ravel(), jj.ravel(), np.full(n * n, npml)], axis=1)
    return dict(name="layered256", nx=n, ny=n, nz=n, dx=DX, dt=DT, order=order, npml=npml,
                nt=nt, f0=20.0, src=(n // 2, n // 2, n // 2), src_fields=SRC_EXPLOSIVE,
                medium=media.layered(shape), rec=rec, rec_fields=REC_FIELDS)


def perturbed512(order=4, nt=3000, nx=512, ny=512, nz=384, seed=2024, stride=4):
    shape = (nz, ny, nx)
    npml = 20
    ii, jj = np.meshgrid(np.arange(0, nx, stride), np.arange(0, ny, stride), indexing="xy")
    rec = np.stack([ii.ravel(), jj.ravel(), np.full(ii.size, npml)], axis=1)
    return dict(name="perturbed512", nx=nx, ny=ny, nz=nz, dx=DX, dt=DT, order=order,
                npml=npml, nt=nt, f0=20.0, src=(nx // 2, ny // 2, nz // 2),
                src_fields=SRC_EXPLOSIVE, medium=media.perturbed(shape, seed=seed), rec=rec,
                rec_fields=REC_FIELDS)


def adjoint128(order=4, nt=1000, n=128, every=50, sphere=True):
    """True model: homogeneous rock with a +8 % Kb sphere of radius 30 m (6 cells) at the
    centre; the start model has no sphere (make it with sphere=False).  Explosive 25 Hz
    source above the sphere, receivers every 8 cells on the planes z = 16 and z = n - 16
    recording vx, vz, p."""
    shape = (n, n, n)
    npml = 10
    m = media.homogeneous(shape)
    if sphere:
        c = n // 2
        kk, jj, ii = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        ball = (ii - c) ** 2 + (jj - c) ** 2 + (kk - c) ** 2 <= 6 ** 2
        m["Kb"] = np.where(ball, m["Kb"] * 1.08, m["Kb"])
    g = np.arange(4, n - 3, 8)
    ii, jj = np.meshgrid(g, g, indexing="xy")
    rec = np.concatenate([np.stack([ii.ravel(), jj.ravel(), np.full(ii.size, z)], axis=1)
                          for z in (16, n - 16)])
    return dict(name="adjoint128", nx=n, ny=n, nz=n, dx=DX, dt=DT, order=order, npml=npml,
                nt=nt, every=every, f0=25.0, src=(n // 2, n // 2, 24), src_fields=SRC_EXPLOSIVE,
                medium=m, rec=rec, rec_fields=REC_FIELDS)


def batch16(nshots=16, nt=1000, model=(512, 512, 384), side=(96, 192), seed=7, npml=20,
            order=4, rec_stride=8, medium=None):
    """Uneven cubic windows (side ~ U[side], seeded) at seeded positions inside the model
    grid (nx, ny, nz); the medium is the 512 x 512 x 384 perturbed one unless given (arrays
    (nz, ny, nx)).  Each shot: 20 Hz explosive source at the window centre, receivers every
    rec_stride cells on the window plane z = npml recording vx, vz, p."""
    nx, ny, nz = model
    rng = np.random.default_rng(seed)
    sides = rng.integers(side[0], side[1] + 1, nshots)
    if medium is None:
        medium = media.perturbed((nz, ny, nx), seed=2024)
    shots = []
    for s in sides:
        s = int(s)
        x0 = int(rng.integers(0, nx - s + 1))
        y0 = int(rng.integers(0, ny - s + 1))
        z0 = int(rng.integers(0, nz - s + 1))
        g = np.arange(npml, s - npml, rec_stride)
        ii, jj = np.meshgrid(g, g, indexing="xy")
        rec = np.stack([ii.ravel(), jj.ravel(), np.full(ii.size, npml)], axis=1)
        shots.append(dict(origin=(x0, y0, z0), size=(s, s, s), src=(s // 2, s // 2, s // 2), rec=rec))
    return dict(name="batch16", model=model, medium=medium, shots=shots, dx=DX, dt=DT, order=order,
                npml=npml, nt=nt, f0=20.0, src_fields=SRC_EXPLOSIVE, rec_fields=REC_FIELDS)


def window_params(cfg, shot):
    """The inversion parameters of one shot's window, (nz, ny, nx) arrays."""
    x0, y0, z0 = shot["origin"]
    wx, wy, wz = shot["size"]
    sub = media.crop(cfg["medium"], z0, y0, x0, (wz, wy, wx))
    return media.params_from_medium(sub)


def make_batch(cfg, precision="fp16", params=None):
    """A Batch with every shot of a batch workload added (params: per-shot parameter dicts,
    default the workload's medium)."""
    from ._lib import ricker
    from .batch import Batch

    nx, ny, nz = cfg["model"]
    b = Batch(nx, ny, nz, cfg["dx"], cfg["dt"], order=cfg["order"], precision=precision,
              npml=cfg["npml"])
    w = ricker(cfg["f0"], cfg["dt"], cfg["nt"])
    for i, sh in enumerate(cfg["shots"]):
        p = window_params(cfg, sh) if params is None else params[i]
        b.add_shot(sh["origin"], sh["size"], p, sh["src"], w, sh["rec"], cfg["rec_fields"],
                   cfg["src_fields"])
    return b


def cells(cfg, ijk):
    ijk = np.asarray(ijk, dtype=np.int64).reshape(-1, 3)
    return ijk[:, 0] + ijk[:, 1] * cfg["nx"] + ijk[:, 2] * cfg["nx"] * cfg["ny"]


def make_solver(cfg, precision="fp32", order=None, nt=None, with_receivers=True):
    """Builds a Solver for a workload dict (medium, source, receivers set)."""
    from ._lib import ricker
    from .solver import Solver

    nt = cfg["nt"] if nt is None else nt
    s = Solver(cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["dt"],
               order=cfg["order"] if order is None else order, precision=precision,
               npml=cfg["npml"])
    w = ricker(cfg["f0"], cfg["dt"], nt)
    s.set_source(int(cells(cfg, cfg["src"])[0]), w, cfg["src_fields"])
    s.set_medium_device(cfg["medium"])  # the same bits as set_medium, built on the GPU
    if with_receivers:
        s.set_receivers(cells(cfg, cfg["rec"]), cfg["rec_fields"], nt)
    return s
