"""Synthetic porous media: homogeneous, z-layered, and Gaussian-correlated perturbations.

Each builder returns a dict of per-cell float64 arrays of shape (nz, ny, nx) with the keys
the C-ABI's biot_medium takes (Ks Kf Kb mu phi rhos rhof eta perm).
"""
import numpy as np

# the homogeneous porous rock of the 64^3 forward case
BASE = dict(rhos=2650.0, rhof=1040.0, phi=0.2, perm=1e-13, eta=1e-3,
            Ks=36e9, Kf=2.25e9, Kb=9.6e9, mu=5.1e9)
# three z-layers of the layered cases (top, middle, bottom)
LAYERS = dict(Kb=(6e9, 9.6e9, 14e9), mu=(3e9, 5.1e9, 8e9), phi=(0.3, 0.2, 0.12))


def homogeneous(shape, **over):
    p = dict(BASE, **over)
    return {k: np.full(shape, float(v)) for k, v in p.items()}


def layer_index(nz, nlayers=3):
    """Layer id per z index, equal thicknesses (top layer first)."""
    bounds = [round(nz * (i + 1) / nlayers) for i in range(nlayers)]
    return np.searchsorted(np.asarray(bounds), np.arange(nz), side="right")


def layered(shape, layers=LAYERS, **over):
    nz = shape[0]
    m = homogeneous(shape, **over)
    lid = layer_index(nz, len(next(iter(layers.values()))))
    for key, vals in layers.items():
        col = np.asarray(vals, dtype=np.float64)[lid]
        m[key] = np.broadcast_to(col[:, None, None], shape).copy()
    return m


def gaussian_field(shape, corr_cells, seed):
    """Zero-mean field of Gaussian correlation exp(-r^2 / L^2) (L = corr_cells), scaled to
    max |value| = 1.  White noise smoothed by a separable Gaussian of std L/sqrt(2)."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape, dtype=np.float32)
    f = gaussian_filter(w, sigma=corr_cells / np.sqrt(2.0), mode="wrap")
    f -= f.mean()
    f /= np.abs(f).max()
    return f.astype(np.float64)


def perturbed(shape, amplitude=0.10, corr_cells=10.0, seed=2024, keys=("Kb", "mu", "phi"),
              base=None):
    """Layered medium with one seeded Gaussian-correlated field scaling `keys` by
    (1 + amplitude * field), field in [-1, 1]."""
    m = layered(shape) if base is None else base
    f = gaussian_field(shape, corr_cells, seed)
    for k in keys:
        m[k] = m[k] * (1.0 + amplitude * f)
    return m


def crop(medium, k0, j0, i0, size):
    return {k: np.ascontiguousarray(v[k0:k0 + size[0], j0:j0 + size[1], i0:i0 + size[2]])
            for k, v in medium.items()}


def params_from_medium(medium, tao=3.0, porousm=8.0):
    """Inversion parameters of a medium dict, in double: drained Lame lambda = Kb - 2 mu / 3,
    mu, alpha = 1 - Kb/Ks, M = Kf Ks / (phi Ks + (alpha - phi) Kf), rho = (1 - phi) rho_s +
    phi rho_f, rho_f, and the fixed C1 = eta/perm, tphi = (1 + 2/porousm) tao / phi
    (outputfunc.cpp:205-217).  Fluid cells (mu == 0 and Kb == 0, outputfunc.cpp:176-193)
    map to lambda = mu = 0, alpha = 1, M = Kf, rho = rho_f, C1 = 0, tphi = 1."""
    g = {k: np.asarray(v, np.float64) for k, v in medium.items()}
    shape = np.broadcast_shapes(*[v.shape for v in g.values()])
    g = {k: np.broadcast_to(v, shape) for k, v in g.items()}
    fluid = (g["mu"] == 0) & (g["Kb"] == 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        alpha = 1.0 - g["Kb"] / g["Ks"]
        M = g["Kf"] * g["Ks"] / (g["phi"] * g["Ks"] + (alpha - g["phi"]) * g["Kf"])
        out = dict(lam=g["Kb"] - 2.0 * g["mu"] / 3.0, mu=g["mu"], alpha=alpha, M=M,
                   rho=(1 - g["phi"]) * g["rhos"] + g["phi"] * g["rhof"], rhof=g["rhof"],
                   C1=g["eta"] / g["perm"], tphi=(1 + 2 / porousm) * tao / g["phi"])
    fl = dict(lam=0.0, mu=0.0, alpha=1.0, M=g["Kf"], rho=g["rhof"], rhof=g["rhof"], C1=0.0,
              tphi=1.0)
    res = {}
    for k, v in out.items():
        f = np.broadcast_to(np.asarray(fl[k], np.float64), shape)
        res["lambda" if k == "lam" else k] = np.where(fluid, f, v).astype(np.float64)
    return res
