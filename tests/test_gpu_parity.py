"""GPU parity of the HIP path (through the C-ABI) against the CPU oracle.

Bars: FP32 storage <= 1e-5 relative L2 on receiver traces (max over components) against the
oracle's FP32 restatement of the reference arithmetic; FP64 <= 1e-10 against the oracle's
FP64 mode; FP16 storage <= 1e-2 against the oracle's FP64 mode.  Source and receiver
indexing are checked bit for bit at the full 256^3 size.
"""
import numpy as np
import pytest

import oracle
from biot_amd import BiotError, configs, media, ricker
from biot_amd.solver import Solver
from conftest import max_comp_rel_l2, rel_l2

pytestmark = pytest.mark.gpu


def _oracle_traces(cfg, order, mode, nt):
    o = oracle.OracleRun(cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["dt"], order=order,
                         mode=mode, npml=cfg["npml"], medium=cfg["medium"])
    w = oracle.ricker_series(nt, cfg["dt"], cfg["f0"])
    tr = o.run(0, nt, nt, w, configs.cells(cfg, cfg["src"])[0], cfg["src_fields"],
               configs.cells(cfg, cfg["rec"]), cfg["rec_fields"])
    return tr, o


def _gpu_traces(cfg, order, precision, nt):
    s = configs.make_solver(cfg, precision=precision, order=order, nt=nt)
    s.run(nt)
    return s.traces(), s


@pytest.mark.parametrize("order", [2, 4])
def test_fp32_fwd64_traces_match_oracle(order):
    """64^3, 500 steps: every field everywhere at step 250 (wave inside the grid, PML
    active) and the receiver traces over all 500 steps."""
    print(f"fwd64 fp32 order {order}", flush=True)
    cfg = configs.fwd64(order=order)
    nt, half = cfg["nt"], cfg["nt"] // 2
    o = oracle.OracleRun(cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["dt"], order=order,
                         mode="f32", npml=cfg["npml"], medium=cfg["medium"])
    w = oracle.ricker_series(nt, cfg["dt"], cfg["f0"])
    src, rec = configs.cells(cfg, cfg["src"])[0], configs.cells(cfg, cfg["rec"])
    args = (w, src, cfg["src_fields"], rec, cfg["rec_fields"])
    ref = o.run(0, half, nt, *args)
    s = configs.make_solver(cfg, precision="fp32", order=order, nt=nt)
    s.run(half)
    errs = {f: rel_l2(s.field(f), o.field(f)) for f in oracle.FIELDS}
    print("  fields at step %d: %s" % (half, " ".join("%s %.1e" % kv for kv in errs.items())))
    ref = ref + o.run(half, nt - half, nt, *args)
    s.run(nt - half)
    err = max_comp_rel_l2(s.traces(), ref)
    print(f"  traces rel L2 = {err:.3e}")
    s.close()
    o.close()
    assert err <= 1e-5
    # Whole fields: the relative fluid velocity is the small residual of two nearly
    # cancelling terms (pressure gradient against the inertial coupling, W2 sum(dtau) vs
    # W3 dp), which amplifies FP32 rounding differences ~1e3-fold; the pressure inherits a
    # little of it through M div q.  The FP64 test below holds every field to 1e-10.
    bars = {f: (5e-3 if f in ("qx", "qy", "qz") else 2e-4 if f == "p" else 3e-5) for f in errs}
    assert all(errs[f] <= bars[f] for f in errs), errs


@pytest.mark.parametrize("order", [2, 4])
def test_fp64_fwd64_matches_oracle_f64(order):
    """Same arithmetic in double on both sides: traces and every field agree to 1e-10
    (the folded C-PML carry and the kernels' coefficient folding are exact rewrites)."""
    print(f"fwd64 fp64 order {order}", flush=True)
    cfg = configs.fwd64(order=order)
    nt = 300
    ref, o = _oracle_traces(cfg, order, "f64", nt)
    tr, s = _gpu_traces(cfg, order, "fp64", nt)
    err = max_comp_rel_l2(tr, ref)
    errs = {f: rel_l2(s.field(f, f64=True), o.field(f)) for f in oracle.FIELDS}
    print(f"  traces rel L2 = {err:.3e}; worst field {max(errs.values()):.3e}")
    s.close()
    o.close()
    assert err <= 1e-7  # the trace buffer is float32 on every path
    assert max(errs.values()) <= 1e-10, errs


@pytest.mark.parametrize("precision,order,bar", [("fp16c", 4, 1e-2), ("fp16t", 4, 1e-2), ("fp16m", 4, 1e-2),
                                                 ("fp16", 4, 2.5e-2), ("fp16c", 2, 1e-2), ("fp16t", 2, 1e-2),
                                                 ("fp16m", 2, 1e-2)])
def test_half_precision_fwd64_vs_fp64_oracle(precision, order, bar):
    """Receiver traces of the 16-bit storage modes against the FP64 oracle, 500 steps, the
    workloads' order 4 and the reference's order 2.  fp16c (binary16 + 8-bit residual on v,
    tau, p), fp16t (residual on tau, p) and fp16m (tau, p binary16 + a packed 4-bit residual; binary16 v, q and moduli)
    must meet the 1e-2 target; plain binary16 storage (every array binary16, as the
    reference's `half` arrays) misses it: 2.5e-2 is its measured level (1.6e-2) with
    headroom, not the target."""
    print(f"fwd64 {precision} order {order}", flush=True)
    cfg = configs.fwd64(order=order)
    nt = cfg["nt"]
    ref, _ = _oracle_traces(cfg, order, "f64", nt)
    tr, s = _gpu_traces(cfg, order, precision, nt)
    errs = [rel_l2(tr[c], ref[c]) for c in range(3)]
    print(f"  {precision} traces rel L2 vs fp64 (vx, vz, p) = " + " ".join(f"{e:.3e}" for e in errs))
    s.close()
    assert np.isfinite(tr).all()
    assert max(errs) <= bar


@pytest.mark.parametrize("precision,bar", [("fp16c", 1e-2), ("fp16t", 1e-2), ("fp16m", 1e-2), ("fp16", 5e-2)])
def test_half_precision_perturbed_medium_vs_gpu_fp64(precision, bar):
    """Layered medium with +-10 % Gaussian-correlated perturbations of Kb, mu, phi (the
    512 x 512 x 384 workload's medium statistics) on a 192 x 192 x 160 grid, 1200 steps,
    64 x 64 surface receivers: 16-bit modes against the GPU FP64 path.  The target 1e-2 holds
    for fp16c, fp16t and fp16m; plain fp16's 5e-2 is its measured level (1.6e-2) with
    headroom, not the target."""
    print(f"perturbed proxy {precision}", flush=True)
    cfg = configs.perturbed512(nt=1200, nx=192, ny=192, nz=160, stride=3)
    ref_s = configs.make_solver(cfg, precision="fp64")
    ref_s.run(cfg["nt"])
    ref = ref_s.traces()
    ref_s.close()
    s = configs.make_solver(cfg, precision=precision)
    s.run(cfg["nt"])
    tr = s.traces()
    s.close()
    errs = [rel_l2(tr[c], ref[c]) for c in range(3)]
    print(f"  {precision} traces rel L2 vs fp64 (vx, vz, p) = " + " ".join(f"{e:.3e}" for e in errs))
    assert np.isfinite(tr).all()
    assert max(errs) <= bar


def test_half_precision_full_size_perturbed512():
    """The 512 x 512 x 384, 3000-step perturbed-medium workload at full size, on two receiver
    sets in one run: the 128 x 128 surface receivers (vx, vz, p) and the reference's own layout,
    the vertical line through the source column (kernel.cu:2027-2031; the source cell left
    out), which records the reference's txx and here also vx, vz and p.  FP16C (binary16 +
    8-bit residual on v, tau, p), FP16T (8-bit residual on tau, p) and FP16M (tau, p binary16
    + a packed 4-bit residual; binary16 v, q and moduli) within 1e-2 of the GPU FP64 path on
    every trace of both sets (the near-source pressure is the hardest: each step's increment
    is small against its value); plain FP16 is reported (it misses 1e-2: 2.0e-2 measured)."""
    import time

    from biot_amd._lib import FIELD_INDEX

    print("perturbed512 full size", flush=True)
    t0 = time.time()
    cfg = configs.perturbed512()
    nsurf = len(cfg["rec"])
    sx, sy, sz = cfg["src"]
    column = [(sx, sy, k) for k in range(cfg["npml"], cfg["nz"] - cfg["npml"]) if k != sz]
    fields = ("vx", "vz", "p", "txx")
    cfg = dict(cfg, rec=np.concatenate([np.asarray(cfg["rec"]), np.asarray(column)]), rec_fields=fields)
    order = sorted(FIELD_INDEX[f] for f in fields)  # traces come in field-index order
    ci = {f: order.index(FIELD_INDEX[f]) for f in fields}
    print(f"  medium built in {time.time() - t0:.0f} s", flush=True)
    traces = {}
    for prec in ("fp64", "fp16c", "fp16t", "fp16m", "fp16"):
        t0 = time.time()
        s = configs.make_solver(cfg, precision=prec)
        s.run(cfg["nt"])
        traces[prec] = s.traces()
        ms = s.last_run_ms()
        s.close()
        print(f"  {prec}: {ms / 1e3:.1f} s loop, {time.time() - t0:.0f} s total", flush=True)
    ref = traces["fp64"]
    surf = {p: [rel_l2(traces[p][ci[f], :, :nsurf], ref[ci[f], :, :nsurf]) for f in ("vx", "vz", "p")]
            for p in ("fp16c", "fp16t", "fp16m", "fp16")}
    cols = {p: {f: rel_l2(traces[p][ci[f], :, nsurf:], ref[ci[f], :, nsurf:]) for f in fields} for p in surf}
    col = {p: max(cols[p].values()) for p in surf}
    for p in surf:
        print(f"  {p} surface traces rel L2 vs fp64 (vx, vz, p) = " + " ".join(f"{x:.3e}" for x in surf[p]) +
              "; source column (txx, vx, vz, p) = " +
              " ".join(f"{cols[p][f]:.3e}" for f in ("txx", "vx", "vz", "p")))
    assert all(np.isfinite(traces[p]).all() for p in ("fp16c", "fp16t", "fp16m", "fp16"))
    for p in ("fp16c", "fp16t", "fp16m"):
        assert max(surf[p]) <= 1e-2 and col[p] <= 1e-2


def _random_state(shape, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape).astype(np.float32).astype(np.float64) for _ in range(13)]


@pytest.mark.parametrize("precision,mode,tol", [("fp32", "f32", 1e-5), ("fp64", "f64", 1e-11)])
def test_full_size_random_state_matches_oracle(precision, mode, tol):
    """256^3, order 4, layered medium, PML 20: every field everywhere (PML slabs and edges
    included) after 4 steps from a random state."""
    print(f"layered256 random state {precision}", flush=True)
    n, npml, order = 256, 20, 4
    shape = (n, n, n)
    med = media.layered(shape)
    s = Solver(n, n, n, 5.0, 4e-4, order=order, precision=precision, npml=npml)
    s.set_medium(med)
    o = oracle.OracleRun(n, n, n, 5.0, 4e-4, order=order, mode=mode, npml=npml, medium=med)
    st = _random_state(shape, 11)
    names = oracle.FIELDS
    for f, v in zip(names, st):
        s.write_field(f, v)
        o.set_field(f, v)
    del st
    s.run(4)
    for _ in range(4):
        o.step_v()
        o.step_t()
    worst = 0.0
    for f in names:
        e = rel_l2(s.field(f, f64=True), o.field(f))
        worst = max(worst, e)
        assert e <= tol, (f, e)
    print(f"  worst field rel L2 = {worst:.3e}")
    o.close()
    s.close()


@pytest.mark.parametrize("precision,tol", [("fp16", 3e-2), ("fp16c", 1e-3), ("fp16t", 1e-2), ("fp16m", 1e-2)])
def test_storage_scale_change_mid_run(precision, tol):
    """A new source amplitude changes the 16-bit storage scales, and the handle re-expresses
    every stored field and C-PML carry in the new units.  A source that is silent after step
    80 is replaced at step 120 (after the wave has entered the absorbing layers) by 16x itself:
    the physics is unchanged, so the run with the replacement must keep the accuracy of the
    run without it against the FP64 run (a wrongly scaled carry would not: the late traces
    carry the absorbing layers' response)."""
    print(f"scale change {precision}", flush=True)
    n, npml, nt = 48, 8, 240
    m = media.homogeneous((n, n, n))
    c = n // 2
    src = c + c * n + c * n * n
    w = np.zeros(nt, np.float32)
    w[:80] = ricker(40.0, 4e-4, 80)
    rec = np.array([c + c * n + (npml + 2) * n * n, 3 + c * n + c * n * n, c + 4 * n + c * n * n], np.int64)
    out = {}
    for name, prec, swap in (("ref", "fp64", False), ("plain", precision, False), ("swap", precision, True)):
        s = Solver(n, n, n, 5.0, 4e-4, order=4, precision=prec, npml=npml)
        s.set_medium(m)
        s.set_source(src, w, ("txx", "tyy", "tzz", "p"))
        s.set_receivers(rec, ("vx", "vz", "p"), nt)
        if swap:
            s.run(120)
            s.set_source(src, w * 16, ("txx", "tyy", "tzz", "p"))
            s.run(nt - 120)
        else:
            s.run(nt)
        out[name] = s.traces().astype(np.float64)
        s.close()
    ref = out["ref"]
    e_plain = max(rel_l2(out["plain"][q], ref[q]) for q in range(3))
    e_swap = max(rel_l2(out["swap"][q], ref[q]) for q in range(3))
    e_late = max(rel_l2(out["swap"][q][150:], ref[q][150:]) for q in range(3))
    print(f"  vs fp64: without the change {e_plain:.2e}, with it {e_swap:.2e} (after step 150 {e_late:.2e})")
    assert np.abs(ref[:, 150:]).max() > 0
    assert e_swap <= max(1.5 * e_plain, tol) and e_late <= 2 * tol


@pytest.mark.parametrize("precision", ["fp32", "fp16", "fp64"])
def test_source_and_receiver_indexing_bit_exact_full_size(precision):
    """256^3 with 65536 surface receivers: zero coefficients leave the fields as written,
    so every trace sample must equal the written pattern at its cell, bit for bit, and the
    source cell must accumulate exactly the injected samples."""
    print(f"indexing {precision}", flush=True)
    cfg = configs.layered256(order=4, nt=4)
    n = cfg["nx"]
    s = Solver(n, n, n, 5.0, 4e-4, order=4, precision=precision, npml=20)
    zeros = [np.zeros(n ** 3, np.float32) for _ in range(23)]
    s.set_coefficients(zeros)
    # the source fixes the FP16 storage scales (peak |w| = 4): set it before writing
    src = configs.cells(cfg, [(37, 91, 200)])[0]
    w = np.array([1.0, 2.0, -4.0, 0.5], np.float32)
    s.set_source(src, w, ("txx", "p"))
    k_, j_, i_ = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    pat = ((i_ * 7 + j_ * 3 + k_ * 5) % 1024).astype(np.float64)  # pat/4 exact in binary16
    for f in ("vx", "vz", "p", "txx", "tyy", "tzz"):
        s.write_field(f, pat)
    rec = configs.cells(cfg, cfg["rec"])
    s.set_receivers(rec, ("vx", "vz", "p"), 4)
    s.run(4)
    tr = s.traces()
    flat = pat.reshape(-1)
    bad = [(c, it, int(np.count_nonzero(tr[c, it] != flat[rec].astype(np.float32))))
           for c in range(3) for it in range(4)]
    assert all(b[2] == 0 for b in bad), bad
    p = s.field("p").reshape(-1)
    txx = s.field("txx").reshape(-1)
    assert p[src] == flat[src] + w.sum() and txx[src] == flat[src] + w.sum()
    mask = np.ones(n ** 3, bool)
    mask[src] = False
    assert int(np.count_nonzero(p[mask] != flat[mask].astype(np.float32))) == 0
    assert int(np.count_nonzero(s.field("tyy").reshape(-1) != flat.astype(np.float32))) == 0
    s.close()


def test_snapshot_reset_determinism_and_errors():
    cfg = configs.fwd64(order=4, nt=120)
    s = configs.make_solver(cfg, nt=120)
    s.run(120)
    t1 = s.traces()
    snap = s.snapshot_xz("txx", 32)
    assert np.array_equal(snap, s.field("txx")[:, 32, :])
    s.reset()
    assert not np.any(s.field("vx"))
    s.run(120)
    assert np.array_equal(t1, s.traces())
    with pytest.raises(BiotError) as e:
        s.set_receivers([64 ** 3], ("p",), 10)
    assert e.value.status == 1
    with pytest.raises(BiotError) as e:
        s.snapshot_xz("p", 64)
    assert e.value.status == 1
    s.close()
    t = Solver(32, 32, 32, 5.0, 4e-4, order=2)
    with pytest.raises(BiotError) as e:
        t.run(1)
    assert e.value.status == 3
    t.close()


def _fluid_column_medium(shape):
    """Layered rock with a fluid-filled column (mu = Kb = 0: a fluid cell, C1 = 0) through the
    grid along z, next to the source."""
    m = media.layered(shape)
    nz, ny, nx = shape
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    hole = ((ii - nx // 2 - 6) ** 2 + (jj - ny // 2) ** 2) <= 4
    for k in ("mu", "Kb"):
        m[k][:, hole] = 0.0
    return m


@pytest.mark.parametrize("setup", ["medium", "medium_device", "params"])
def test_fp16m_coefficients_from_moduli(setup):
    """FP16M forms the staggered averages and all 22 coefficients in the kernels from the
    stored cell moduli.  On a layered medium with a fluid column (harmonic shear means with a
    zero mu, the C1 = 0 guard of the fluid-velocity coefficients) and a ragged 72 x 61 x 66
    grid, 400 steps, its receiver traces stay within the same 1e-2 of the GPU FP64 path as
    FP16T, which reads the reference's coefficient arrays (FP16T is measured beside it); the
    three set-up paths give the same storage."""
    print(f"fp16m coefficients from moduli via {setup}", flush=True)
    nx, ny, nz, nt = 72, 61, 66, 400
    shape = (nz, ny, nx)
    m = _fluid_column_medium(shape)
    src = (nx // 2, ny // 2, nz // 2)
    rec = [(i, j, 12) for i in range(8, nx - 8, 4) for j in range(8, ny - 8, 4)]
    rec += [(nx // 2 + 6, ny // 2, k) for k in range(10, nz - 10, 3)]  # inside the fluid column
    cfg = dict(nx=nx, ny=ny, nz=nz, dx=5.0, dt=4e-4, order=4, npml=10, f0=25.0, src=src,
               src_fields=configs.SRC_EXPLOSIVE, medium=m, rec=rec, rec_fields=configs.REC_FIELDS)
    out = {}
    for prec in ("fp64", "fp16t", "fp16m"):
        s = configs.make_solver(cfg, precision=prec, nt=nt)
        if prec == "fp16m" and setup == "medium":
            s.set_medium(m)
        elif prec == "fp16m" and setup == "params":
            s.set_params(media.params_from_medium(m))
        s.run(nt)
        out[prec] = s.traces()
        s.close()
    errs = {p: [rel_l2(out[p][c], out["fp64"][c]) for c in range(3)] for p in ("fp16t", "fp16m")}
    for p, e in errs.items():
        print(f"  {p} vs fp64 (vx, vz, p): " + " ".join(f"{x:.3e}" for x in e))
    assert np.isfinite(out["fp16m"]).all() and np.abs(out["fp64"]).max() > 0
    assert max(errs["fp16m"]) <= 1e-2


def test_fp16m_refusals():
    """FP16M needs the moduli and has no adjoint or stored coefficient arrays."""
    s = Solver(24, 24, 24, 5.0, 4e-4, order=4, precision="fp16m", npml=4)
    with pytest.raises(BiotError) as e:
        s.set_coefficients([np.ones(24 ** 3, np.float32)] * 23)
    assert e.value.status == 3
    s.set_medium(media.homogeneous((24, 24, 24)))
    with pytest.raises(BiotError) as e:
        s.kernel_coefficient(0)
    assert e.value.status == 3
    s.set_receivers([100], ("p",), 8)
    s.set_source(s.cell(12, 12, 12), np.ones(8, np.float32), ("p",))
    with pytest.raises(BiotError):
        s.forward_checkpointed(8, 4)
    s.run(8)
    assert np.isfinite(s.traces()).all()
    s.close()


def _fp16m_numpy(x32):
    """FP16M's stored value of a float32 array: the nearest binary16 h and the 4-bit residual
    r = rint((x - h) / u) clamped to +-7, u = ulp(h)/16 (2^-28 for subnormal h), value
    h + r u."""
    h = x32.astype(np.float16)
    e = np.maximum((h.view(np.uint16).astype(np.int64) >> 10) & 31, 1)
    u = np.ldexp(1.0, e - 29)
    r = np.clip(np.rint((x32.astype(np.float64) - h.astype(np.float64)) / u), -7, 7)
    return h.astype(np.float64) + r * u


def test_fp16m_storage_round_trip():
    """A written stress field reads back as its stored value (binary16 of value / scale plus a
    4-bit residual in units of ulp/16) times the scale: within 2^-14 of the value in binary16's
    normal range, within 2^-28 of the scale below it; writing one field keeps the residuals the
    other six fields share its records with; the velocity as binary16."""
    n = 24
    s = Solver(n, n, n, 5.0, 4e-4, order=4, precision="fp16m", npml=4)
    s.set_source(s.cell(12, 12, 12), np.ones(8, np.float32), ("p",))
    s.set_medium(media.homogeneous((n, n, n)))
    sc = s.field_scales()
    rng = np.random.default_rng(3)
    v = rng.standard_normal(n ** 3) * 10.0 ** rng.uniform(-9, 3, n ** 3) * sc["txx"]
    # values a residual of exactly +-8 units away from their binary16 (clamped to 7)
    half = (np.float16(1.0) + np.arange(64) * np.float16(2.0 ** -10)).astype(np.float32)
    v[:64] = (half + np.float32(2.0 ** -11)).astype(np.float64) * sc["txx"]
    s.write_field("tyy", -0.5 * v.reshape(n, n, n))
    tyy = s.field("tyy", f64=True)
    s.write_field("txx", v.reshape(n, n, n))
    assert np.array_equal(s.field("tyy", f64=True), tyy)
    got = s.field("txx", f64=True).reshape(-1)
    x32 = (v * (1.0 / sc["txx"])).astype(np.float32)
    want = _fp16m_numpy(x32) * sc["txx"]
    assert np.array_equal(got, want)
    x = x32.astype(np.float64) * sc["txx"]
    normal = np.abs(x32) >= 2.0 ** -14
    assert normal.sum() > n ** 3 // 2 and (~normal).sum() > 100
    assert np.all(np.abs(got - x)[normal] <= 2.0 ** -14 * np.abs(x)[normal])
    assert np.all(np.abs(got - x)[~normal] <= 2.0 ** -28 * sc["txx"])
    s.write_field("vx", (v / sc["txx"] * sc["vx"] * 1e-3).reshape(n, n, n))
    gv = s.field("vx", f64=True).reshape(-1)
    wv = (v / sc["txx"] * sc["vx"] * 1e-3 * (1.0 / sc["vx"])).astype(np.float32).astype(np.float16).astype(np.float64) * sc["vx"]
    assert np.array_equal(gv, wv)
    s.close()
