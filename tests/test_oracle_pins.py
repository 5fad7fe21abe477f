"""Pins of the CPU oracle against facts that do not restate its own formulas.

 - binary16 conversion against numpy's IEEE implementation (exhaustive decode, RNE encode);
 - Biot moduli against Gassmann's equation in its textbook form and hand-derived values;
 - staggered averages against scipy's harmonic mean;
 - PML profile against the reflection coefficient it is designed for, R = exp(-2/v0 int d);
 - the C-PML memory recursion against its closed-form solution for a constant derivative,
   for every derivative of both half steps (the profile node follows the derivative's);
 - the assembled velocity/stress updates against the exact discrete plane-wave eigenmode of
   staggered leapfrog (P and S waves along every axis and polarisation, orders 2 and 4),
   whose dispersion relation
   sin(w dt/2) = (c dt/dx) * S(k dx) and impedance v = -tau/(rho c) are textbook facts;
 - the fluid-velocity coefficients against Darcy's law (steady state of the damped update);
 - the Biot fluid coupling (inertial term m - rho_f^2/rho, damping eta/kappa, coupling
   rho_f/rho) against the exact solution of the Biot equations' relaxation ODE, momentum
   conservation, and the transient toward Darcy's flux along each axis;
 - the pressure/stress coefficients against the Biot constitutive law on linear fields;
 - order-4 differences are exact on cubics;
 - the Ricker wavelet against its closed-form peak, zeros and symmetry.
"""
import math

import numpy as np
import pytest
from scipy.stats import hmean

import oracle

BASE = dict(rhos=2650.0, rhof=1040.0, phi=0.2, perm=1e-13, eta=1e-3,
            Ks=36e9, Kf=2.25e9, Kb=9.6e9, mu=5.1e9)


# ---------------------------------------------------------------------------------------
# binary16
# ---------------------------------------------------------------------------------------
def test_f16_decode_exhaustive():
    L = oracle.lib()
    codes = np.arange(65536, dtype=np.uint32)
    ours = np.array([L.or_f16_to_f32(int(c)) for c in codes], np.float32)
    ref = codes.astype(np.uint16).view(np.float16).astype(np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(ours), nan)
    assert np.array_equal(ours[~nan].view(np.uint32), ref[~nan].view(np.uint32))


def _f16_cases(rng):
    # random magnitudes across the whole binary16 range and beyond, plus exact ties,
    # subnormal boundaries and overflow boundaries
    mags = 2.0 ** rng.uniform(-30, 17, 20000)
    x = mags * rng.choice([-1.0, 1.0], mags.size)
    halves = np.arange(0, 0x7c00, 7, dtype=np.uint16).view(np.float16).astype(np.float64)
    nxt = np.nextafter(halves.astype(np.float16), np.float16(np.inf)).astype(np.float64)
    ties = (halves + nxt) / 2.0
    special = np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 65536.0, 1e10, 2.0 ** -24,
                        2.0 ** -25, 2.0 ** -25 * 1.0000001, 2.0 ** -26, 6.0975552e-05,
                        6.103515625e-05, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11])
    return np.concatenate([x, ties, -ties, special, -special])


def test_f16_encode_f32_rne():
    L = oracle.lib()
    x = _f16_cases(np.random.default_rng(1)).astype(np.float32)
    ours = np.array([L.or_f32_to_f16(float(v)) for v in x], np.uint16)
    ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref)


def test_f16_encode_f64_rne():
    L = oracle.lib()
    x = _f16_cases(np.random.default_rng(2))
    ours = np.array([L.or_f64_to_f16(float(v)) for v in x], np.uint16)
    ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref)


# ---------------------------------------------------------------------------------------
# moduli, averages, coefficients
# ---------------------------------------------------------------------------------------
def _moduli(**over):
    m = dict(BASE, **over)
    return oracle.biot_moduli({k: np.array([v]) for k, v in m.items()}, (1,))


def test_moduli_match_gassmann_textbook_form():
    rng = np.random.default_rng(3)
    for _ in range(200):
        p = dict(Ks=rng.uniform(20e9, 60e9), Kf=rng.uniform(1e9, 3e9), phi=rng.uniform(0.05, 0.4),
                 mu=rng.uniform(1e9, 20e9))
        p["Kb"] = rng.uniform(0.1, 0.9) * p["Ks"]
        out = _moduli(**p)
        # Gassmann: K_sat = Kb + (1 - Kb/Ks)^2 / (phi/Kf + (1-phi)/Ks - Kb/Ks^2)
        ks, kb, kf, phi, mu = p["Ks"], p["Kb"], p["Kf"], p["phi"], p["mu"]
        ksat = kb + (1 - kb / ks) ** 2 / (phi / kf + (1 - phi) / ks - kb / ks ** 2)
        H = float(out["HH"][0])
        assert abs((H - 4 * mu / 3) - ksat) / ksat < 2e-6
        assert abs(float(out["H2u"][0]) - (ksat - 2 * mu / 3)) / ksat < 4e-6  # lambda_c
        # C = alpha * M and M is the Biot modulus 1/(phi/Kf + (alpha-phi)/Ks)
        alpha = 1 - kb / ks
        Mb = 1 / (phi / kf + (alpha - phi) / ks)
        assert abs(float(out["M"][0]) - Mb) / Mb < 1e-6
        assert abs(float(out["C"][0]) - alpha * Mb) / Mb < 2e-6


def test_moduli_hand_values_fwd64_rock():
    out = _moduli()
    # alpha = 1 - 9.6/36 = 11/15; M = 2.25e9*36e9 / (0.2*36e9 + (11/15 - 0.2)*2.25e9) = 8.1e19/8.4e9
    assert abs(float(out["M"][0]) - 8.1e19 / 8.4e9) / 1e9 < 1e-6
    assert float(out["rho"][0]) == 2328.0          # 0.8*2650 + 0.2*1040
    assert float(out["C2"][0]) == 19500.0          # 3.75 * 1040 / 0.2
    assert float(out["C1"][0]) == np.float32(1e10)  # eta / kappa
    H = float(out["HH"][0])
    assert abs(H - ((11 / 15) ** 2 * 8.1e19 / 8.4e9 + 9.6e9 + 4 * 5.1e9 / 3)) / H < 1e-6


def test_moduli_fluid_cell():
    out = _moduli(mu=0.0, Kb=0.0)
    kf = np.float32(2.25e9)
    for k in ("M", "HH", "H2u", "C"):
        assert out[k][0] == kf
    assert out["C1"][0] == 0 and out["C2"][0] == np.float32(1040.0)


def test_averages_homogeneous_harmonic_and_zero_guard():
    nx, ny, nz = 6, 5, 4
    rng = np.random.default_rng(4)
    mu = rng.uniform(1e9, 9e9, (nz, ny, nx)).astype(np.float32)
    mu[2, 2, 3] = 0.0
    one = np.ones((nz, ny, nx), np.float32)
    mod = {"mu": mu.reshape(-1), "rho": (2000 * one).reshape(-1), "rhof": (1000 * one).reshape(-1),
           "C1": (5 * one).reshape(-1), "C2": (7 * one).reshape(-1)}
    av = {k: v.reshape(nz, ny, nx) for k, v in oracle.stagger_average(mod, nx, ny, nz).items()}
    for k, want in (("rhox", 2000), ("rhoy", 2000), ("rhoz", 2000), ("rhofx", 1000),
                    ("C1y", 5), ("C2z", 7)):
        assert np.all(av[k][1:-1, 1:-1, 1:-1] == want)
        assert np.all(av[k][0] == 0) and np.all(av[k][:, :, -1] == 0)
    for k in range(1, nz - 1):
        for j in range(1, ny - 1):
            for i in range(1, nx - 1):
                quad = {"muxz": [mu[k, j, i], mu[k, j, i + 1], mu[k + 1, j, i], mu[k + 1, j, i + 1]],
                        "muyz": [mu[k, j, i], mu[k, j + 1, i], mu[k + 1, j, i], mu[k + 1, j + 1, i]],
                        "muxy": [mu[k, j, i], mu[k, j, i + 1], mu[k, j + 1, i], mu[k, j + 1, i + 1]]}
                for name, q in quad.items():
                    got = av[name][k, j, i]
                    if min(q) == 0:
                        assert got == 0
                    else:
                        assert abs(got - hmean(np.asarray(q, np.float64))) <= 1e-6 * got


def test_pml_profile_reflection_and_shape():
    n, NP, dx, dt, v0, R = 200, 20, 5.0, 4e-4, 4000.0, 1e-6
    d, d2, e, e2 = oracle.pml_profile(n, dx, dt, NP, v0, R)
    L = NP * dx
    d0 = 3 * v0 * math.log(1 / R) / (2 * L)
    # nonzero exactly on the two faces: left [0, NP), right [n-1-NP, n-1] (d2 there)
    assert np.all(d[:NP] > 0) and np.all(d[NP:n - 1 - NP + 1] == 0) and np.all(d[n - NP:] > 0)
    assert np.all(d2[:NP] > 0) and np.all(d2[NP:n - 1 - NP] == 0) and np.all(d2[n - 1 - NP:] > 0)
    assert abs(d[0] - d0) / d0 < 1e-6
    # mirror symmetry: integer node i <-> n-1-i, half node i+1/2 <-> (n-2-i)+1/2
    assert np.allclose(d[:NP], d[::-1][:NP], rtol=1e-5)
    assert np.allclose(d2[:NP], d2[n - 2 - np.arange(NP)], rtol=1e-5)
    # design reflection: exp(-(2/v0) * integral of d across the layer) ~= R
    integral = 0.5 * (d[0] + d[NP]) * dx + d[1:NP].sum() * dx
    assert abs(math.log(math.exp(-2 * integral / v0)) / math.log(R) - 1) < 0.02
    assert np.allclose(e[:NP], np.exp(-d[:NP].astype(np.float64) * dt), rtol=1e-6)
    assert np.all(e[NP:n - 1 - NP] == 0)  # outside the layer e stays at its zero fill


def test_ricker_closed_form():
    f, dt = 25.0, 1e-5
    T0 = 1.2 / f
    it0 = int(round(T0 / dt))
    assert abs(oracle.ricker(it0, dt, f) - (-1.0)) < 1e-6
    w = oracle.ricker_series(int(2 * T0 / dt) + 1, dt, f)
    # symmetric about the peak, zero crossings at |t - T0| = 1/(pi f sqrt(2))
    k = np.arange(1, it0)
    assert np.allclose(w[it0 + k], w[it0 - k], atol=2e-6)
    tz = 1 / (math.pi * f * math.sqrt(2))
    iz = int(round((T0 + tz) / dt))
    assert abs(w[iz]) < 3 * math.pi * f * dt * 2  # within one sample's slope of zero
    assert abs(w.sum() * dt) < 1e-4 * T0


# ---------------------------------------------------------------------------------------
# the stepping arithmetic
# ---------------------------------------------------------------------------------------
def _run(nx, ny, nz, order, medium, npml=0, dx=5.0, dt=4e-4, coef=None, mode="f64", **kw):
    return oracle.OracleRun(nx, ny, nz, dx, dt, order=order, mode=mode, npml=npml,
                            medium=medium, coef=coef, **kw)


# Every C-PML memory variable lives where its derivative lives: a forward difference at a
# half node takes the half-node profile of its axis, a backward difference at an integer node
# the integer-node profile.  One case per derivative of the velocity half step that feeds a
# solid velocity (kernel.cu:172-185, 188-216): (axis, half node, stress field, velocity,
# reference U1 array).
_V_PML_CASES = [
    (0, True, "txx", "vx", 9),   # x1
    (1, True, "tyy", "vy", 10),  # y1
    (2, True, "tzz", "vz", 11),  # z1
    (1, False, "txy", "vx", 9),  # x2
    (2, False, "txz", "vx", 9),  # x3
    (0, False, "txy", "vy", 10),  # y2
    (2, False, "tyz", "vy", 10),  # y3
    (0, False, "txz", "vz", 11),  # z2
    (1, False, "tyz", "vz", 11),  # z3
]
# ... and per derivative of the stress half step (kernel.cu:430-445, 449-482): (axis, half
# node, velocity field, stress field, reference coefficient array, sign of the update)
_T_PML_CASES = [
    (0, False, "vx", "p", 15, -1.0),    # uxx -> p via PP1
    (1, False, "vy", "p", 15, -1.0),    # uyy
    (2, False, "vz", "p", 15, -1.0),    # uzz
    (0, False, "qx", "p", 16, -1.0),    # wx -> p via PP2
    (1, False, "qy", "p", 16, -1.0),    # wy
    (2, False, "qz", "p", 16, -1.0),    # wz
    (1, True, "vx", "txy", 20, 1.0),    # uxy -> txy via SPXY
    (0, True, "vy", "txy", 20, 1.0),    # uyx
    (2, True, "vx", "txz", 21, 1.0),    # uxz -> txz via SPXZ
    (0, True, "vz", "txz", 21, 1.0),    # uzx
    (2, True, "vy", "tyz", 22, 1.0),    # uyz -> tyz via SPYZ
    (1, True, "vz", "tyz", 22, 1.0),    # uzy
]


def _psi_closed_form(prof, c, half, dt, D, nsteps):
    """Memory values psi_1..psi_n of the recursion for a derivative D constant from step 1 (with
    D_prev = 0 before it): psi_1 = a D, psi_m = e psi_(m-1) + a (1 + e) D, i.e.
    psi_m = a D e^(m-1) + a (1 + e) D (1 - e^(m-1)) / (1 - e), with e and a = -dt d / 2 taken
    from the stored x0.01 profile values (the FP64 mode keeps the products in double)."""
    es = float(prof[3 if half else 2][c])
    ds = float(prof[1 if half else 0][c])
    e = es * 100.0
    a = -float(np.float32(dt)) * ds * 100.0 * 0.5
    return [a * D * e ** (m - 1) + a * (1 + e) * D * (1 - e ** (m - 1)) / (1 - e)
            for m in range(1, nsteps + 1)]


def _ramp(n, axis):
    r = np.arange(n, dtype=np.float64)
    return [r[None, None, :], r[None, :, None], r[:, None, None]][axis]


def _at(n, axis, c):
    return [(n // 2, n // 2, c), (n // 2, c, n // 2), (c, n // 2, n // 2)][axis]


@pytest.mark.parametrize("axis,half,tau,vel,U", _V_PML_CASES)
def test_cpml_recursion_closed_form(axis, half, tau, vel, U):
    """Constant derivative D inside the layer (a unit ramp of one stress along one axis):
    the solid velocity accumulates U1 (D + psi_m) over the steps, psi_m in closed form."""
    n, NP, dx, dt = 24, 6, 5.0, 4e-4
    shape = (n, n, n)
    zero = [np.zeros(n ** 3, np.float32) for _ in range(23)]
    zero[U][:] = 1.0
    o = _run(n, n, n, 2, None, npml=NP, dx=dx, dt=dt, coef=zero)
    o.set_field(tau, np.broadcast_to(_ramp(n, axis), shape))
    nsteps = 7
    for _ in range(nsteps):
        o.step_v()
        o.step_t()
    v = o.field(vel)
    D = float(np.float32(1.0) / np.float32(dx))
    for c in (1, 2, NP - 1, n - 3):
        want = sum(D + p for p in _psi_closed_form(o.prof[axis], c, half, dt, D, nsteps))
        got = v[_at(n, axis, c)]
        assert abs(got - want) <= 1e-12 * abs(want), (c, got, want)
    # outside the layer psi stays 0: v = nsteps * D exactly
    mid = (n // 2, n // 2, n // 2)
    assert v[mid] == nsteps * D


@pytest.mark.parametrize("axis,half,vel,tau,K,sign", _T_PML_CASES)
def test_cpml_recursion_closed_form_stress_side(axis, half, vel, tau, K, sign):
    """Stress half steps alone on a unit velocity ramp: the stress or pressure accumulates
    sign * 0.01 * K (D + psi_m) (kernel.cu:484-490 with one coefficient array set to 1)."""
    n, NP, dx, dt = 24, 6, 5.0, 4e-4
    shape = (n, n, n)
    zero = [np.zeros(n ** 3, np.float32) for _ in range(23)]
    zero[K][:] = 1.0
    o = _run(n, n, n, 2, None, npml=NP, dx=dx, dt=dt, coef=zero)
    o.set_field(vel, np.broadcast_to(_ramp(n, axis), shape))
    nsteps = 7
    for _ in range(nsteps):
        o.step_t()
    t = o.field(tau)
    D = float(np.float32(1.0) / np.float32(dx))
    for c in (1, 2, NP - 1, n - 3):
        want = sign * 0.01 * sum(D + p for p in _psi_closed_form(o.prof[axis], c, half, dt, D, nsteps))
        got = t[_at(n, axis, c)]
        assert abs(got - want) <= 1e-12 * abs(want), (c, got, want)
    mid = (n // 2, n // 2, n // 2)
    assert abs(t[mid] - sign * 0.01 * nsteps * D) <= 1e-14 * nsteps * D


def _elastic_medium():
    m = dict(BASE)
    m["eta"] = 0.0  # C1 = 0: no relative flow (the reference zeroes W1..W3, main.cpp:431-444)
    return m


def _S(order, x):
    """Symbol of the staggered difference: 2 i S(k dx) e^{...}."""
    if order == 2:
        return math.sin(x / 2)
    return 9 / 8 * math.sin(x / 2) - 1 / 24 * math.sin(3 * x / 2)


_SHEAR = {(0, 1): "txy", (1, 0): "txy", (0, 2): "txz", (2, 0): "txz", (1, 2): "tyz", (2, 1): "tyz"}


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("axis,pol", [(a, b) for a in range(3) for b in range(3)])
def test_plane_wave_eigenmode(order, axis, pol):
    """Exact discrete plane wave of the undrained elastic limit propagating along `axis` with
    particle motion along `pol` (P when they agree, S otherwise): every interior cell must
    follow it to FP64 rounding.  Velocity c = sqrt(H/rho) (P) or sqrt(mu/rho) (S) with the
    Gassmann-pinned moduli, dispersion sin(w dt/2) = (c dt/dx) S(k dx), amplitude ratio from
    the impedance rho c.  The nine cases drive every difference of both half steps along its
    own axis (kernel.cu:172-185, 430-445): a difference taken along the wrong stride, or
    forward where it should be backward, breaks the mode."""
    dims = [64, 64, 64]
    dims[axis] = 96
    nx, ny, nz = dims
    dx, dt = 5.0, 4e-4
    Cv, Cp = 1e8, 1.0
    med = _elastic_medium()
    mod = _moduli(eta=0.0)
    rho = float(mod["rho"][0])
    P = axis == pol
    Mmod = float(mod["HH"][0]) if P else float(mod["mu"][0])
    c = math.sqrt(Mmod / rho)
    k = 2 * math.pi / (12 * dx)
    s = (c * dt / dx) * _S(order, k * dx)
    w = 2 / dt * math.asin(s)
    T = 1.0
    V = -(Cv / Cp) * T / (rho * c)  # v = -tau / Z in the reference representation
    n_a = dims[axis]
    x_int = np.arange(n_a) * dx
    x_half = (np.arange(n_a) + 0.5) * dx
    shape = (nz, ny, nx)

    def along(vals):
        return np.broadcast_to(vals.reshape([(1, 1, -1), (1, -1, 1), (-1, 1, 1)][axis]), shape)

    o = _run(nx, ny, nz, order, med, Cv=Cv, Cp=Cp, dx=dx, dt=dt)
    v_name = ["vx", "vy", "vz"][pol]
    if P:  # normal stress at integer nodes, velocity at half nodes along the axis
        tau_name, tau_x, v_x = ["txx", "tyy", "tzz"][axis], x_int, x_half
    else:  # shear stress at half nodes, transverse velocity at integer nodes
        tau_name, tau_x, v_x = _SHEAR[(axis, pol)], x_half, x_int
    # tau at t=0, v at t=-dt/2 (the velocity is half a step behind)
    o.set_field(tau_name, along(T * np.cos(k * tau_x)))
    o.set_field(v_name, along(V * np.cos(k * v_x + w * dt / 2)))
    if P:  # the matching lateral stresses of the plane P wave (lambda_c / H ratio)
        lam = float(mod["H2u"][0]) / Mmod
        for b in range(3):
            if b != axis:
                o.set_field(["txx", "tyy", "tzz"][b], along(lam * T * np.cos(k * tau_x)))
    N = 6
    for _ in range(N):
        o.step_v()
        o.step_t()
    h = order // 2
    m = 2 * h * (N + 1)
    sl = (slice(m, nz - m), slice(m, ny - m), slice(m, nx - m))
    tau = o.field(tau_name)[sl]
    v = o.field(v_name)[sl]
    tau_want = along(T * np.cos(k * tau_x - w * N * dt))[sl]
    v_want = along(V * np.cos(k * v_x - w * (N - 0.5) * dt))[sl]
    # the coefficient arrays are float (as the reference stores them), so the phase carries
    # their ~1e-7 relative rounding; a wrong staggering, sign or scale is an O(1) miss
    assert np.max(np.abs(tau - tau_want)) < 1e-7 * T
    assert np.max(np.abs(v - v_want)) < 1e-7 * abs(V)
    # nothing leaks into the other velocity components
    for b in range(3):
        if b != pol:
            assert np.max(np.abs(o.field(["vx", "vy", "vz"][b])[sl])) < 1e-9 * abs(V)


def test_darcy_steady_state_of_fluid_velocity():
    """Velocity half steps alone under a fixed uniform pressure gradient G and stress
    gradient S: q converges to Darcy's flux -(kappa/eta) (G + (rho_f/rho) S) (reference
    representation: times Cv/Cp)."""
    n, dx, dt = 12, 5.0, 4e-4
    Cv, Cp = 1e8, 1.0
    med = dict(BASE)
    o = _run(n, n, n, 2, med, Cv=Cv, Cp=Cp, dx=dx, dt=dt)
    x = np.arange(n) * dx
    G, Sg = 3.0, 2.0  # Pa / m
    o.set_field("p", np.broadcast_to(G * x[None, None, :], (n, n, n)))
    o.set_field("txx", np.broadcast_to(Sg * x[None, None, :], (n, n, n)))
    for _ in range(3000):
        o.step_v()
    q = o.field("qx")[n // 2, n // 2, n // 2]
    mod = _moduli()
    rho, rhof = float(mod["rho"][0]), float(mod["rhof"][0])
    want = -(Cv / Cp) * (BASE["perm"] / BASE["eta"]) * (G + rhof / rho * Sg)
    assert abs(q - want) / abs(want) < 1e-5, (q, want)


# The Biot system in the (v, q) form the reference integrates (Biot 1956, low-frequency):
#     rho  dv/dt + rho_f dq/dt              = div tau
#     rho_f dv/dt + m dq/dt + (eta/kappa) q = -grad p,   m = C2 (tortuosity mass, pinned by
# the hand value of C2 above).  Eliminating dv/dt along one axis gives
#     (m - rho_f^2/rho) dq/dt + C1 q = -(dp/dx + (rho_f/rho) dtau/dx),   C1 = eta/kappa,
# an ODE per cell whose exact solution is known; the trapezoidal (Crank-Nicolson) rule of the
# reference reproduces it to (dt/T)^2/12 relative, T = (m - rho_f^2/rho)/C1.  A permeable
# rock (kappa = 1e-9 m^2) makes T = 19 ms, about 50 steps, so the pins see the inertial
# term m - rho_f^2/rho, the damping C1 and the coupling rho_f/rho directly.
def _permeable():
    med = dict(BASE)
    med["perm"] = 1e-9
    mod = _moduli(perm=1e-9)
    rho, rhof, C1, C2 = (float(mod[k][0]) for k in ("rho", "rhof", "C1", "C2"))
    return med, rho, rhof, C1, C2


def test_fluid_relaxation_follows_biot_ode():
    """No stress or pressure gradient: q decays as q0 exp(-t/T), and the total momentum
    rho v + rho_f q is conserved, so v = (rho_f/rho) (q0 - q).  One step is exactly the
    trapezoidal rule q1 = q0 (m' - dt C1/2) / (m' + dt C1/2), m' = m - rho_f^2/rho."""
    n, dx, dt = 8, 5.0, 4e-4
    med, rho, rhof, C1, m = _permeable()
    mp = m - rhof ** 2 / rho
    Tr = mp / C1
    o = _run(n, n, n, 2, med, Cv=1e8, Cp=1.0, dx=dx, dt=dt)
    q0 = [1.0, -2.0, 0.5]
    for name, q in zip(("qx", "qy", "qz"), q0):
        o.set_field(name, np.full((n, n, n), q))
    c = (n // 2, n // 2, n // 2)
    o.step_v()
    for name, q in zip(("qx", "qy", "qz"), q0):
        want = q * (mp - 0.5 * dt * C1) / (mp + 0.5 * dt * C1)
        assert abs(o.field(name)[c] - want) <= 1e-6 * abs(q), name
    nsteps = 101  # t = 2.1 T
    for _ in range(nsteps - 1):
        o.step_v()
    t = nsteps * float(np.float32(dt))
    for qn, vn, q in zip(("qx", "qy", "qz"), ("vx", "vy", "vz"), q0):
        got_q, got_v = o.field(qn)[c], o.field(vn)[c]
        assert abs(got_q - q * math.exp(-t / Tr)) <= 5e-5 * abs(q), (qn, got_q)
        assert abs(got_v - (rhof / rho) * (q - got_q)) <= 1e-6 * abs(q), (vn, got_v)
        # every cell of the update box follows the same ODE (uniform state)
        assert np.allclose(o.field(qn)[1:-1, 1:-1, 1:-1], got_q, rtol=1e-12, atol=0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fluid_transient_under_gradients(axis):
    """Velocity half steps from rest under a fixed pressure gradient G and normal-stress
    gradient S along one axis: q rises as q_inf (1 - exp(-t/T)) toward Darcy's flux
    q_inf = -(G + (rho_f/rho) S) / C1, and v = S t / rho - (rho_f/rho) q (reference units:
    times Cv/Cp).  The other components stay zero."""
    n, dx, dt = 8, 5.0, 4e-4
    Cv, Cp = 1e8, 1.0
    med, rho, rhof, C1, m = _permeable()
    Tr = (m - rhof ** 2 / rho) / C1
    o = _run(n, n, n, 2, med, Cv=Cv, Cp=Cp, dx=dx, dt=dt)
    G, S = 3.0, -2.0  # Pa / m
    x = _ramp(n, axis) * dx
    o.set_field("p", np.broadcast_to(G * x, (n, n, n)))
    o.set_field(["txx", "tyy", "tzz"][axis], np.broadcast_to(S * x, (n, n, n)))
    nsteps = 101
    for _ in range(nsteps):
        o.step_v()
    t = nsteps * float(np.float32(dt))
    q_inf = -(Cv / Cp) * (G + rhof / rho * S) / C1
    c = (n // 2, n // 2, n // 2)
    q = o.field(["qx", "qy", "qz"][axis])[c]
    v = o.field(["vx", "vy", "vz"][axis])[c]
    assert abs(q - q_inf * (1 - math.exp(-t / Tr))) <= 5e-5 * abs(q_inf), (q, q_inf)
    v_want = (Cv / Cp) * S * t / rho - (rhof / rho) * q
    assert abs(v - v_want) <= 1e-6 * abs(v_want), (v, v_want)
    for b in range(3):
        if b != axis:
            assert o.field(["qx", "qy", "qz"][b])[c] == 0.0
            assert o.field(["vx", "vy", "vz"][b])[c] == 0.0


def test_constitutive_law_on_linear_fields():
    """One stress half step on linear velocity fields: dp = -dt (C div v + M div q),
    dtau_xx = dt (lambda_c div v + 2 mu dvx/dx + C div q), dtau_xy = dt mu (dvx/dy + dvy/dx)."""
    n, dx, dt = 10, 5.0, 4e-4
    Cv, Cp = 1e8, 1.0
    o = _run(n, n, n, 2, dict(BASE), Cv=Cv, Cp=Cp, dx=dx, dt=dt)
    k_, j_, i_ = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    gx, gy, gz, gq, gxy, gyx = 2.0, -1.0, 0.5, 3.0, 0.7, -0.4  # gradients in 1/s per (1/Cv)
    gqy, gqz = -1.5, 0.25
    o.set_field("vx", (gx * i_ + gxy * j_) * dx)
    o.set_field("vy", (gy * j_ + gyx * i_) * dx)
    o.set_field("vz", gz * k_ * dx)
    o.set_field("qx", gq * i_ * dx)
    o.set_field("qy", gqy * j_ * dx)
    o.set_field("qz", gqz * k_ * dx)
    o.step_t()
    mod = _moduli()
    M, C, lam, mu = (float(mod[k][0]) for k in ("M", "C", "H2u", "mu"))
    divv = (gx + gy + gz) / Cv  # physical 1/s
    divq = (gq + gqy + gqz) / Cv
    c = (n // 2, n // 2, n // 2)
    dp = o.field("p")[c]
    assert abs(dp - (-dt * Cp * (C * divv + M * divq))) <= 1e-6 * abs(dp)
    dtxx = o.field("txx")[c]
    want = dt * Cp * (lam * divv + 2 * mu * gx / Cv + C * divq)
    assert abs(dtxx - want) <= 1e-6 * abs(want)
    dtxy = o.field("txy")[c]
    want = dt * Cp * mu * (gxy + gyx) / Cv
    assert abs(dtxy - want) <= 1e-6 * abs(want)


def test_order4_exact_on_cubics():
    """The fourth-order staggered difference differentiates cubics exactly: one stress half
    step with vx = x^3 (x at the half nodes) gives dp = -dt C 3 x_i^2 / Cv."""
    n, dx, dt = 16, 5.0, 4e-4
    Cv, Cp = 1e8, 1.0
    o = _run(n, n, n, 4, dict(BASE), Cv=Cv, Cp=Cp, dx=dx, dt=dt)
    xh = (np.arange(n) + 0.5) * dx
    o.set_field("vx", np.broadcast_to((xh ** 3 * 1e-3)[None, None, :], (n, n, n)))
    o.step_t()
    C = float(_moduli()["C"][0])
    xi = np.arange(n) * dx
    want = -dt * Cp * C * 3e-3 * xi ** 2 / Cv
    got = o.field("p")[n // 2, n // 2, :]
    sl = slice(2, n - 2)
    # exact up to the float rounding of the stored coefficient (an order-2 stencil would be
    # off by dx^2 x''' / 24 relative, ~1e-2 here)
    assert np.allclose(got[sl], want[sl], rtol=1e-6, atol=0)


def test_source_and_receiver_index_exact():
    n = 8
    o = _run(n, n, n, 2, dict(BASE), mode="f32")
    cell = 3 + 5 * n + 6 * n * n
    o.source(cell, ["txx", "p"], 0.75)
    txx, p, tyy = o.field("txx"), o.field("p"), o.field("tyy")
    assert txx[6, 5, 3] == np.float32(0.75) and p[6, 5, 3] == np.float32(0.75)
    assert np.count_nonzero(txx) == 1 and np.count_nonzero(p) == 1 and np.count_nonzero(tyy) == 0
