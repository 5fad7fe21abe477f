#!/usr/bin/env python3
"""Benchmark of the 13-field Biot time step on one MI355X.

Workload (default): 256^3 grid, 4th-order staggered space, PML 20, three-layer porous
medium, 20 Hz Ricker at the centre, 65,536 surface receivers recording vx, vz, p; FP32
storage (the drop-in path), 2000 steps timed after a warm-up.  The same grid is then run
with FP16 storage.  A "step" is one time step over the whole grid (source, velocity half
step, stress half step, receiver sampling), inputs resident in HBM.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mixed-precision-optimization-method_amd")
sys.path.insert(0, PKG)

METRIC = "Mcell-updates/s for the 13-field Biot step, FP16 vs FP32, plus adjoint grad ms/shot"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def slab_counts(n, L, R, h):
    """Updated cells (h..n-1-h) that lie in the low or the high absorbing slab of one axis
    (a cell in both counts once, as the kernels treat it: in the low slab)."""
    lo = max(0, min(L, n - h) - h)
    hi = max(0, (n - h) - max(n - R, h, L))
    return lo + hi


# element bytes per storage mode: field, coefficient, C-PML carry, residual (FP16C on the 10
# compensated fields v, tau, p; FP16T on the 7 of tau, p; FP16M's residual record of its tau,
# p); FP16 keeps every array in binary16, FP16C, FP16T and FP16M keep float carries
WIDTHS = {"fp32": (4, 4, 4, 0), "fp16": (2, 2, 2, 0), "fp16c": (2, 2, 4, 1), "fp16t": (2, 2, 4, 1),
          "fp16m": (2, 2, 4, 1), "fp64": (8, 8, 8, 0)}
# residual bytes per updated cell (velocity kernel, stress kernel, one fused step): FP16C and
# FP16T one per access of a compensated field; FP16M the cell's 4-byte residual record of its
# tau, p (read by the velocity kernel, read and written by the stress kernel)
RESIDUAL_ACCESSES = {"fp16c": (13, 17, 20), "fp16t": (7, 14, 14), "fp16m": (4, 8, 8)}
# medium bytes per updated cell where they are not the coefficient arrays: FP16M reads the
# cell's moduli records (velocity: rho, rho_f, C1, C2 in 8 bytes; stress: C, M, lambda_c, H in
# 8 bytes and the cell's mu in 2, the shear means formed from the mu of neighbouring cells;
# the neighbours' records and mu the kernels use are re-reads of the same arrays)
MEDIUM_BYTES = {"fp16m": (8, 10, 18)}


def algorithmic_bytes(nx, ny, nz, order, npml_cells, mode):
    """Compulsory HBM bytes per launch of the velocity and stress kernels and per step.

    Velocity kernel per updated cell: reads 7 stress/pressure + 6 velocity fields, writes 6,
    reads 14 coefficient arrays; in an absorbing slab of one axis it reads and writes 4
    carries.  Stress kernel: reads 6 + 7, writes 7, reads 8 coefficients, 4 carries per
    slab.  The fused-step figure is 26 field + 22 coefficient accesses per cell plus
    16 carry accesses per slab membership (one read and one write of 8 carries); FP16C adds
    one residual byte per access of v, tau, p (V: 10 reads + 3 writes, T: 10 + 7, step 20),
    FP16T per access of tau, p (V: 7 reads, T: 7 + 7, step 14); FP16M reads its 4-byte
    residual record in V and reads and writes it in T (step 8)."""
    wf, wk, wc, wr = WIDTHS[mode]
    rv, rt, rstep = RESIDUAL_ACCESSES.get(mode, (0, 0, 0))
    kv, kt, kstep = MEDIUM_BYTES.get(mode, (14 * wk, 8 * wk, 22 * wk))
    h = order // 2
    ni, nj, nk = nx - 2 * h, ny - 2 * h, nz - 2 * h
    upd = ni * nj * nk
    Lx, Rx, Ly, Ry, Lz, Rz = npml_cells
    memb = (slab_counts(nx, Lx, Rx, h) * nj * nk + slab_counts(ny, Ly, Ry, h) * ni * nk +
            slab_counts(nz, Lz, Rz, h) * ni * nj)
    v = upd * (19 * wf + kv + rv * wr) + memb * 8 * wc
    t = upd * (20 * wf + kt + rt * wr) + memb * 8 * wc
    step = upd * (26 * wf + kstep + rstep * wr) + memb * 16 * wc
    return v, t, step, upd, memb


def adjoint_bytes(nx, ny, nz, order, npml_cells, nt, every):
    """Compulsory HBM bytes of the adjoint sweep of an FP32 handle (nt steps, checkpoints every
    `every`): per segment a state load (every field and C-PML carry read and written); per
    step a state save, the recomputed forward step (the two stencil kernels) and the four
    adjoint passes; per segment a flush of the float partial sums into the FP64 sums.

    Adjoint passes per updated cell (4-byte adjoint fields, derivative adjoints G, coefficients
    and partial sums; neighbours' values are re-reads): stress-local reads 7 lambda(tau, p),
    8 coefficients, 6 forward v/q, 8 partial sums and writes 12 G and 8 sums (49 accesses);
    stress-gather reads 6 lambda(v, q) and 12 G, writes 6 lambda (24); velocity-local reads
    14 coefficients, 6 lambda(v, q), 6 forward q (two states), 7 forward tau/p, 14 partial sums
    and writes 3 lambda(q), 12 G, 14 sums (76); velocity-gather reads 7 lambda(tau, p) and 12 G,
    writes 7 lambda (26): 175 accesses = 700 B.  Per slab membership and adjoint half step: 4
    forward carries read, 4 carry adjoints read and written (12 accesses, 24 per step = 96 B).
    Flush per updated cell: 22 float partial sums read and cleared, 22 FP64 sums read and
    written (22 x (4 + 4 + 8 + 8) = 528 B).

    every = 1 (every step's state kept by the forward run): no state load or save and no
    recomputed step in the sweep, a flush every 50 steps."""
    h = order // 2
    v, t, _, upd, memb = algorithmic_bytes(nx, ny, nz, order, npml_cells, "fp32")
    n = nx * ny * nz
    Lx, Rx, Ly, Ry, Lz, Rz = npml_cells
    slab_cells = (Lx + Rx) * ny * nz + nx * (Ly + Ry) * nz + nx * ny * (Lz + Rz)
    state = 13 * n * 4 + 8 * slab_cells * 4
    del h
    if every == 1:
        per_step = 700 * upd + 96 * memb
        return nt * per_step + ((nt + 49) // 50) * 528 * upd, per_step, upd
    nseg = (nt + every - 1) // every
    per_step = 2 * state + v + t + 700 * upd + 96 * memb
    total = nseg * 2 * state + nt * per_step + nseg * 528 * upd
    return total, per_step, upd


PROFILE = os.path.join(ROOT, "profiles", "stencil_kernels_layered256.md")


def profiled_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (counters scaled by a 1 GiB copy calibration), or None."""
    try:
        for line in open(PROFILE):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 4 and cells[0] == kernel:
                return float(cells[3])
    except (OSError, ValueError):
        pass
    return None


def slab_widths(s):
    """(L, R) per axis from the PML profile the handle holds."""
    out = []
    for ax, n in enumerate((s.nx, s.ny, s.nz)):
        d, d2, _, _ = s.pml_profile(ax)
        nzr = [(d[i] != 0 or d2[i] != 0) for i in range(n)]
        L = max([i + 1 for i in range(n // 2) if nzr[i]], default=0)
        R = max([n - i for i in range(n // 2, n) if nzr[i]], default=0)
        out += [L, R]
    return tuple(out)


def run_precision(cfg, precision, steps, warmup):
    from biot_amd import configs

    nt = warmup + steps
    t0 = time.time()
    s = configs.make_solver(cfg, precision=precision, nt=nt)
    log(f"[bench] {precision}: set-up {time.time() - t0:.1f} s, device {s.device_bytes() / 2**30:.2f} GiB")
    t0 = time.time()
    shape = s.autotune(5)
    log(f"[bench] {precision}: tile {shape} ({time.time() - t0:.1f} s autotune)")
    s.run(warmup)  # biot_run returns after the device has finished
    t0 = time.perf_counter()
    s.run(steps)
    wall = time.perf_counter() - t0
    ev_ms = s.last_run_ms()
    tr = s.traces()
    import numpy as np

    finite = bool(np.isfinite(tr).all())
    # per-kernel durations over a separate profiled stretch (events around each launch)
    pv, pt = s.profile_steps(min(100, steps))
    widths = slab_widths(s)
    s.close()
    return dict(wall_s=wall, event_ms=ev_ms, kernel_ms_v=pv, kernel_ms_t=pt, widths=widths,
                finite=finite, shape=list(shape))


def host_cpu():
    """The host the CPU baseline ran on: logical CPUs (`nproc` of the machine), the CPUs this
    process may run on, and the processor model from /proc/cpuinfo (None where absent)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return dict(nproc=os.cpu_count(), usable_cpus=usable, cpu_model=model)


def cpu_baseline(cfg, nsteps):
    """The oracle (the reference update arithmetic restated in C, OpenMP over z-planes)
    timed on the host cores over a bounded number of steps of the same workload."""
    sys.path.insert(0, ROOT)
    import numpy as np

    import oracle
    from biot_amd import configs

    oracle.build()
    o = oracle.OracleRun(cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["dt"], order=cfg["order"],
                         mode="f32", npml=cfg["npml"], medium=cfg["medium"])
    w = oracle.ricker_series(nsteps, cfg["dt"], cfg["f0"])
    rec = configs.cells(cfg, cfg["rec"])
    t0 = time.perf_counter()
    o.run(0, nsteps, nsteps, w, configs.cells(cfg, cfg["src"])[0], cfg["src_fields"], rec,
          cfg["rec_fields"])
    dt = time.perf_counter() - t0
    n = cfg["nx"] * cfg["ny"] * cfg["nz"]
    o.close()
    return dict(value=n * nsteps / dt / 1e6, unit="Mcell-updates/s", cores=oracle.num_threads(),
                kind="port",
                sample=f"{nsteps} steps of the same 256^3 order-4 FP32 workload, oracle in C "
                       f"(-O3, OpenMP), {dt:.1f} s", host=host_cpu())


def adjoint_shot(repeats=2, fd_cells=2):
    """Adjoint gradient of the 128^3, 1000-step sphere workload (FP32 wavefields, FP64
    gradient sums): forward keeping states, FP64 misfit, adjoint sweep, per shot, in the two
    state modes: every step's state kept (1001 states, 140 GB: the sweep recomputes nothing;
    the headline figures) and 50-step checkpoints with segment recomputation (`segments`).
    Observed traces come from the FP64 run of the true model.  The sweep's roofline uses
    adjoint_bytes; the FP32 gradient is checked against fourth-order central differences of the
    FP64 forward misfit (step 1e-3 of the value) at `fd_cells` seeded cells for lambda and
    rho."""
    import math

    import numpy as np

    from biot_amd import configs, media

    true, start = configs.adjoint128(), configs.adjoint128(sphere=False)
    nt, every = start["nt"], start["every"]
    s64 = configs.make_solver(start, precision="fp64")
    s64.set_params(media.params_from_medium(true["medium"]))
    s64.reset()
    s64.run(nt)
    obs = s64.traces_f64().astype(np.float32)
    p0 = media.params_from_medium(start["medium"])
    s = configs.make_solver(start, precision="fp32")
    s.set_params(p0)
    widths = slab_widths(s)
    modes = {}
    for mode_every in (1, every):
        best = None
        for _ in range(repeats):
            t0 = time.perf_counter()
            s.forward_checkpointed(nt, mode_every)
            f_ms = s.last_run_ms()
            J, res = s.misfit(obs)
            g, _ = s.adjoint(res)
            a_ms = s.last_run_ms()
            wall = (time.perf_counter() - t0) * 1e3
            r = dict(ms_per_shot=round(f_ms + a_ms, 2), forward_checkpointed_ms=round(f_ms, 2),
                     adjoint_ms=round(a_ms, 2), wall_ms_per_shot=round(wall, 1), misfit=J,
                     finite=bool(all(np.isfinite(v).all() for v in g.values())), grad=g)
            if best is None or r["ms_per_shot"] < best["ms_per_shot"]:
                best = r
        total, per_step, upd = adjoint_bytes(start["nx"], start["ny"], start["nz"], start["order"], widths, nt,
                                             mode_every)
        ach = total / (best["adjoint_ms"] * 1e-3) / 1e9
        what = ("four adjoint passes and flushes over the stored states" if mode_every == 1 else
                "state copies, recomputed forward steps, four adjoint passes, flushes")
        best["roofline"] = {"bound": "hbm", "kernel": f"adjoint sweep ({what})", "achieved": round(ach, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                            "traffic": None, "algorithmic_bytes": total, "bytes_per_step": per_step,
                            "bytes_per_cell_step": round(per_step / upd, 1)}
        best["ratio_to_forward"] = round(best["ms_per_shot"] / best["forward_checkpointed_ms"], 2)
        modes[mode_every] = best
        log(f"[bench] adjoint every {mode_every}: {best['ms_per_shot']:.0f} ms/shot "
            f"(forward {best['forward_checkpointed_ms']:.0f} ms, sweep {best['adjoint_ms']:.0f} ms)")
    # The residual of the FP32 forward is a small difference (the sphere's scattered field, a
    # few 1e-3 of the traces), so it carries the FP32 traces' error magnified: the same FP32
    # states and adjoint are run again on the residual of the FP64 forward to separate the two
    obs64 = obs.astype(np.float64)
    s64.set_params(p0)
    s64.reset()
    s64.run(nt)
    res64 = s64.traces_f64() - obs64
    _, res32 = s.misfit(obs)
    g_r64, _ = s.adjoint(res64.astype(np.float32))
    s.close()
    best = modes[1]
    g = best.pop("grad")
    seg = modes[every]
    seg.pop("grad")
    best["segments"] = dict(seg, every=every)
    # the FP64 adjoint of the same residual, for the normwise agreement
    s64.forward_checkpointed(nt, every)
    g64, _ = s64.adjoint(res64.astype(np.float32))
    normwise = {name: {"fp32_residual": float(np.linalg.norm(g[name] - g64[name]) / np.linalg.norm(g64[name])),
                       "fp64_residual": float(np.linalg.norm(g_r64[name] - g64[name]) / np.linalg.norm(g64[name]))}
                for name in g64}
    # central differences of the FP64 misfit at seeded cells
    rng = np.random.default_rng(2024)
    worst = worst64 = 0.0
    checks = []
    for _ in range(fd_cells):
        i, j, k = (int(v) for v in rng.integers(24, start["nx"] - 24, 3))
        for name in ("lambda", "rho"):
            hstep = 1e-3 * abs(p0[name][k, j, i])
            d = {}
            for m in (1, -1, 2, -2):
                p = {kk: v.copy() for kk, v in p0.items()}
                p[name][k, j, i] += m * hstep
                s64.set_params(p)
                s64.reset()
                s64.run(nt)
                d[m] = s64.traces_f64()

            def dJ(a, b):  # J(a) - J(b) from the double traces, without the rounding of J itself
                return math.fsum(((d[a] - d[b]) * (0.5 * (d[a] + d[b]) - obs64)).ravel())

            fd = (8 * dJ(1, -1) - dJ(2, -2)) / (12 * hstep)  # fourth-order central difference
            e = abs(g[name][k, j, i] - fd) / abs(fd)
            e64 = abs(g_r64[name][k, j, i] - fd) / abs(fd)
            worst = max(worst, e)
            worst64 = max(worst64, e64)
            checks.append({"cell": [i, j, k], "param": name, "fd": fd, "adjoint_fp32": float(g[name][k, j, i]),
                           "rel": round(e, 5), "adjoint_fp32_fp64_residual": float(g_r64[name][k, j, i]),
                           "rel_fp64_residual": round(e64, 7)})
    s64.close()
    res_err = float(np.linalg.norm(np.asarray(res32, np.float64) - res64) / np.linalg.norm(res64))
    best["fd_check"] = {"worst_rel": round(worst, 5), "worst_rel_fp64_residual": round(worst64, 7), "points": checks,
                        "residual_rel_to_traces": float(np.linalg.norm(res64) / np.linalg.norm(obs64)),
                        "fp32_residual_rel_error": res_err,
                        "normwise_vs_fp64_adjoint": normwise,
                        "note": "FP32 states, FP32 adjoint, FP64 gradient sums against fourth-order central "
                                "differences of the FP64 forward misfit (step 1e-3 of the value), with the "
                                "residual of the FP32 forward (rel) and of the FP64 forward "
                                "(rel_fp64_residual); normwise: against the FP64 adjoint of the FP64 residual"}
    best.update(workload="adjoint128: 128^3, order 4, PML 10, 1000 steps, every step's state kept "
                         "(segments: checkpoint every 50 and recompute), FP32 wavefields, FP64 gradient "
                         "sums, 512 receivers x (vx, vz, p)",
                unit="ms/shot", note="ms_per_shot: HIP-event time of the forward run that keeps the "
                "states plus the adjoint sweep and FP64 gradient sums; wall adds misfit, gradient "
                "download and the host chain rule to the six parameters")
    return best


def batch_shots(nt=1000, every=50, graph_steps=50):
    """16 shots in cubic windows of 96..192 cells (seeded) cut from the 512 x 512 x 384
    perturbed medium, FP16 wavefields: the graph-captured batched forward (throughput over
    all windows) and the batched misfit gradient (checkpoints every 50 steps, FP64 sums,
    per-shot gradients added into the model grid on the device); then the batched forward
    on FP16M wavefields."""
    import numpy as np

    from biot_amd import configs

    t0 = time.time()
    cfg = configs.batch16(nt=nt)
    b = configs.make_batch(cfg, precision="fp16")
    cells = sum(int(np.prod(sh["size"])) for sh in cfg["shots"])
    log(f"[bench] batch: {len(cfg['shots'])} shots, {cells / 1e6:.1f} M cells, set-up {time.time() - t0:.1f} s")
    launch = b.autotune(3)
    b.forward(nt, graph_steps=graph_steps)
    obs = [b.traces(i) for i in range(len(cfg["shots"]))]
    b.forward(nt, graph_steps=graph_steps)
    f_ms = b.last_ms()
    # the same tile shape on persistent grids (blocks taking tiles from a work queue) beside
    # the tuned launch: the measurement behind the autotuner's choice
    persistent = {}
    for blocks in (1024, 2048, 4096):
        b.set_launch(launch[0], launch[1], launch[2], blocks)
        b.forward(nt, graph_steps=graph_steps)
        persistent[str(blocks)] = round(cells * nt / b.last_ms() / 1e3, 1)
    b.set_launch(*launch)
    # observed data: the same shots with a slower, softer start model would need a second
    # model; the gradient's cost does not depend on the data, so the recorded traces scaled
    # by 1.01 stand in for observations
    obs = [o * np.float32(1.01) for o in obs]
    J, g = b.gradient(nt, obs, every=every)
    g_ms = b.last_ms()
    groups = b.gradient_groups()
    finite = bool(all(np.isfinite(v).all() for v in g.values()))
    b.close()
    # the same forward on FP16M wavefields (binary16 v, q; tau, p with a packed 4-bit
    # residual), the 16-bit policy within 1e-2 of FP64 on the 512 x 512 x 384 workload
    bm = configs.make_batch(cfg, precision="fp16m")
    launch_m = bm.autotune(3)
    bm.forward(nt, graph_steps=graph_steps)
    bm.forward(nt, graph_steps=graph_steps)
    m_ms = bm.last_ms()
    finite = finite and bool(all(np.isfinite(bm.traces(i)).all() for i in range(len(cfg["shots"]))))
    bm.close()
    return dict(workload=f"batch16: 16 shots, cubic windows {min(sh['size'][0] for sh in cfg['shots'])}.."
                         f"{max(sh['size'][0] for sh in cfg['shots'])} cells cut from the 512x512x384 "
                         f"perturbed medium, order 4, PML 20, {nt} steps, FP16 wavefields",
                shots=len(cfg["shots"]), cells=cells, nt=nt,
                forward_ms=round(f_ms, 1), forward_Mcell_updates_per_s=round(cells * nt / f_ms / 1e3, 1),
                graph_steps=graph_steps, launch_bx_by_zchunk_blocks=list(launch),
                grid_kind=(f"persistent ({launch[3]} blocks)" if launch[3] in (1024, 2048, 4096)
                           else f"one block per tile ({launch[3]} blocks)"),
                persistent_grid_Mcell_updates_per_s=persistent, policy="fp16 (binary16 wavefields)",
                gradient_ms=round(g_ms, 1), gradient_ms_per_shot=round(g_ms / len(cfg["shots"]), 1),
                checkpoint_every=every, gradient_groups=groups, misfit=J,
                fp16m_forward_Mcell_updates_per_s=round(cells * nt / m_ms / 1e3, 1),
                fp16m_launch_bx_by_zchunk_blocks=list(launch_m), finite=finite)


def stream_copy_rate(mib=1024):
    """HBM rate (GB/s, bytes read + written) of torch's device copy of `mib` MiB: what a plain
    streaming kernel reaches on this GPU.  Measured in a child process (`--copy-calib`) before
    the solver library starts: torch ships its own HIP runtime, and the two do not share one
    process's GPU."""
    import subprocess

    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-calib", str(mib)],
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        log(f"[bench] copy reference failed: {r.stderr.strip()[-300:]}")
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])["copy_GBs"]


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB (torch's copy kernel), repeated 5 times:
    the HBM-traffic counters of this dispatch calibrate FETCH_SIZE / WRITE_SIZE."""
    import torch

    n = mib * 2 ** 20 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    # the same copy timed (20 repetitions after the 5 above): the HBM rate a plain
    # streaming kernel reaches on this GPU, for comparison with the stencil kernels
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 20
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4,
                      "copy_ms": round(ms, 4), "copy_GBs": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1)}))


def step_roofline(r, bv, bt, bstep, upd, kernels):
    """Roofline of one time step: the step's algorithmic bytes over the two stencil kernels'
    mean launch durations (HIP events on the solver's stream), with the per-kernel figures and
    the calibrated traffic of both kernels from the committed profile."""
    tv, tt = r["kernel_ms_v"], r["kernel_ms_t"]
    ach = bstep / ((tv + tt) * 1e-3) / 1e9
    trv, trt = profiled_traffic(kernels[0]), profiled_traffic(kernels[1])
    return {"bound": "hbm", "kernel": f"{kernels[0]} + {kernels[1]} (one time step)",
            "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": (trv + trt) if trv is not None and trt is not None else None,
            "traffic_source": os.path.relpath(PROFILE, ROOT),
            "algorithmic_bytes_per_step": bstep, "bytes_per_cell_update": round(bstep / upd, 2),
            "mean_step_kernel_ms": round(tv + tt, 5),
            "kernels": {
                "velocity": {"kernel": kernels[0], "mean_launch_ms": round(tv, 5), "algorithmic_bytes": bv,
                             "achieved_GBs": round(bv / (tv * 1e-3) / 1e9, 1),
                             "frac": round(bv / (tv * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": trv},
                "stress": {"kernel": kernels[1], "mean_launch_ms": round(tt, 5), "algorithmic_bytes": bt,
                           "achieved_GBs": round(bt / (tt * 1e-3) / 1e9, 1),
                           "frac": round(bt / (tt * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": trt}}}


KERNEL_TAG = {"fp32": "P32", "fp16": "P16", "fp16c": "P16C", "fp16t": "P16T", "fp16m": "P16M", "fp64": "P64"}


def policy_line(cfg, mode, r, ref_value, steps, n):
    v = n * steps / r["wall_s"] / 1e6
    bv, bt, bstep, upd, _ = algorithmic_bytes(cfg["nx"], cfg["ny"], cfg["nz"], cfg["order"], r["widths"], mode)
    tag = KERNEL_TAG[mode]
    roof = step_roofline(r, bv, bt, bstep, upd, (f"k_velocity<{tag},{cfg['order']}>", f"k_stress<{tag},{cfg['order']}>"))
    return {"value": round(v, 1), "unit": "Mcell-updates/s", "speedup_vs_fp32": round(v / ref_value, 3),
            "ms_per_step": round(r["wall_s"] * 1e3 / steps, 4), "step_achieved_GBs": roof["achieved"],
            "step_frac": roof["frac"], "bytes_per_cell_update": roof["bytes_per_cell_update"],
            "kernel_ms_v": round(r["kernel_ms_v"], 5), "kernel_ms_t": round(r["kernel_ms_t"], 5),
            "finite": r["finite"], "tile_bx_by_zchunk": r["shape"]}


def perturbed512_section(steps, warmup, modes=("fp32", "fp16m", "fp16t", "fp16c", "fp16")):
    """The 512 x 512 x 384 perturbed-medium grid (the FP16-storage workload): throughput of the
    FP32 drop-in and the 16-bit policies over `steps` timed steps after `warmup`."""
    from biot_amd import configs

    t0 = time.time()
    cfg = configs.perturbed512(nt=warmup + steps)
    log(f"[bench] 512x512x384 medium built in {time.time() - t0:.0f} s")
    n = cfg["nx"] * cfg["ny"] * cfg["nz"]
    out = {"workload": f"perturbed512: 512x512x384 cells, order 4, PML 20, layered medium with seeded "
                       f"+-10 % Gaussian-correlated (50 m) perturbations of Kb, mu, phi, 20 Hz Ricker, "
                       f"{len(cfg['rec'])} receivers x (vx, vz, p); {steps} timed steps after {warmup}",
           "cells": n}
    ref = None
    for mode in modes:
        r = run_precision(cfg, mode, steps, warmup)
        line = policy_line(cfg, mode, r, ref or 1.0, steps, n)
        if mode == "fp32":
            ref = line["value"]
            line.pop("speedup_vs_fp32")
        out[mode] = line
        log(f"[bench] 512x512x384 {mode}: {line['value']:.0f} Mcell/s"
            + (f" ({line['speedup_vs_fp32']:.2f}x fp32)" if mode != "fp32" else ""))
    return out


def perturbed512_accuracy(modes=("fp16m",)):
    """The 512 x 512 x 384 workload's full 3000 steps: receiver-trace relative L2 error of the
    16-bit policies against the GPU FP64 path on the 128 x 128 surface receivers (vx, vz, p),
    the accuracy half of the FP16-storage target (<= 1e-2)."""
    import numpy as np

    from biot_amd import configs

    cfg = configs.perturbed512()
    traces = {}
    for mode in ("fp64",) + tuple(modes):
        t0 = time.time()
        s = configs.make_solver(cfg, precision=mode)
        s.run(cfg["nt"])
        traces[mode] = s.traces_f64() if mode == "fp64" else s.traces().astype(np.float64)
        s.close()
        log(f"[bench] 512x512x384 {mode}: {cfg['nt']} steps for the accuracy check ({time.time() - t0:.0f} s)")
    ref = traces["fp64"]
    out = {"steps": cfg["nt"], "reference": "fp64 on the GPU", "receivers": f"{len(cfg['rec'])} surface x (vx, vz, p)"}
    for mode in modes:
        errs = [float(np.linalg.norm(traces[mode][c] - ref[c]) / np.linalg.norm(ref[c])) for c in range(3)]
        out[mode] = {"rel_l2_vx_vz_p": [round(e, 6) for e in errs], "max_rel_l2": round(max(errs), 6)}
    return out


def stream_overlap(cfg, nsteps, chunk=100):
    """Wall time of nsteps with every receiver trace copied to the host: biot_run followed by
    biot_read_traces, against biot_run_stream_traces (copies of each chunk of trace rows on a
    side stream while the loop continues); the two outputs must agree bit for bit."""
    import numpy as np

    from biot_amd import configs

    s = configs.make_solver(cfg, precision="fp32", nt=nsteps)
    s.autotune(5)
    s.run(20)
    s.reset()
    t0 = time.perf_counter()
    s.run(nsteps)
    a = s.traces()
    plain = time.perf_counter() - t0
    s.reset()
    t0 = time.perf_counter()
    b = s.run_stream_traces(nsteps, chunk)
    streamed = time.perf_counter() - t0
    same = bool(np.array_equal(a, b))
    s.close()
    return {"workload": f"{cfg['name']} FP32, {nsteps} steps, {len(cfg['rec'])} receivers x 3 components "
                        f"({a.nbytes / 1e9:.2f} GB of traces)",
            "run_then_read_s": round(plain, 3), "streamed_s": round(streamed, 3),
            "hidden_s": round(plain - streamed, 3), "chunk_steps": chunk, "outputs_equal": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB")
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-512", action="store_true")
    ap.add_argument("--steps-512", type=int, default=300)
    ap.add_argument("--no-accuracy", action="store_true", help="skip the 3000-step FP16M accuracy run of the 512 grid")
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=24)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-graph-steps", type=int, default=50,
                    help="steps per captured HIP graph in the batched forward (0: launch step by step)")
    args = ap.parse_args()
    if args.copy_calib:
        copy_calibration(args.copy_calib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the solver holds the whole grid on a single MI355X "
              f"and has no multi-GPU path; got --gpus {args.gpus}", file=sys.stderr)
        sys.exit(2)

    from biot_amd import configs

    copy_gbs = stream_copy_rate()  # before this process touches the GPU
    log(f"[bench] plain copy (torch): {copy_gbs} GB/s")
    cfg = configs.layered256(order=4, nt=args.warmup + args.steps, n=args.size)
    n = cfg["nx"] * cfg["ny"] * cfg["nz"]
    r32 = run_precision(cfg, "fp32", args.steps, args.warmup)
    value = n * args.steps / r32["wall_s"] / 1e6
    bv, bt, bstep, upd, memb = algorithmic_bytes(cfg["nx"], cfg["ny"], cfg["nz"], cfg["order"],
                                                 r32["widths"], "fp32")
    roof = step_roofline(r32, bv, bt, bstep, upd, ("k_velocity<P32,4>", "k_stress<P32,4>"))
    log(f"[bench] fp32: {value:.0f} Mcell/s, V {r32['kernel_ms_v']:.3f} ms T {r32['kernel_ms_t']:.3f} ms, "
        f"step {roof['frac']:.3f} of HBM peak")
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["wall_s"] * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"layered{args.size}: {cfg['nx']}^3 cells, order 4, PML 20, "
                               f"3-layer Biot medium, 20 Hz Ricker, {len(cfg['rec'])} receivers x "
                               f"(vx, vz, p), FP32 storage",
                   "grid": [cfg["nx"], cfg["ny"], cfg["nz"]], "order": 4, "npml": 20,
                   "dx_m": cfg["dx"], "dt_s": cfg["dt"], "tile_bx_by_zchunk": r32["shape"]},
        "roofline": roof,
        "stream_copy_GBs": copy_gbs,
        "finite": r32["finite"],
    }
    for mode in ([] if args.no_fp16 else ["fp16m", "fp16t", "fp16c", "fp16"]):
        r16 = run_precision(cfg, mode, args.steps, args.warmup)
        out[mode] = policy_line(cfg, mode, r16, value, args.steps, n)
        log(f"[bench] {mode}: {out[mode]['value']:.0f} Mcell/s ({out[mode]['speedup_vs_fp32']:.2f}x fp32), "
            f"step {out[mode]['step_frac']:.3f} of HBM peak")
    if not args.no_512:
        log("[bench] 512x512x384 ...")
        sec = perturbed512_section(args.steps_512, args.warmup)
        if not args.no_accuracy:
            acc = perturbed512_accuracy()
            sec["accuracy_3000_steps"] = acc
            sp, err = sec["fp16m"]["speedup_vs_fp32"], acc["fp16m"]["max_rel_l2"]
            sec["fp16_storage_target"] = {"target": ">= 1.7x FP32 Mcell-updates/s at <= 1e-2 receiver-trace "
                                                    "rel L2 against FP64", "policy": "fp16m",
                                          "speedup_vs_fp32": sp, "max_rel_l2": err,
                                          "met": bool(sp >= 1.7 and err <= 1e-2)}
            log(f"[bench] 512x512x384 fp16m: {sp:.2f}x fp32 at {err:.2e} against fp64")
        out["perturbed512"] = sec
    if not args.no_stream:
        log("[bench] streamed trace output ...")
        out["stream_overlap"] = stream_overlap(cfg, args.steps)
        log(f"[bench] streamed: {out['stream_overlap']['streamed_s']:.2f} s vs run+read "
            f"{out['stream_overlap']['run_then_read_s']:.2f} s")
    if not args.no_adjoint:
        log("[bench] adjoint 128^3 x 1000 steps ...")
        out["adjoint"] = adjoint_shot()
        log(f"[bench] adjoint: {out['adjoint']['ms_per_shot']:.0f} ms/shot")
    if not args.no_batch:
        log("[bench] batched shots ...")
        out["batch"] = batch_shots(graph_steps=args.batch_graph_steps)
        log(f"[bench] batch: forward {out['batch']['forward_Mcell_updates_per_s']:.0f} Mcell/s, "
            f"gradient {out['batch']['gradient_ms']:.0f} ms, FP16M forward "
            f"{out['batch']['fp16m_forward_Mcell_updates_per_s']:.0f} Mcell/s")
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_steps)
        log(f"[bench] cpu: {out['cpu_baseline']['value']:.1f} Mcell/s on {out['cpu_baseline']['cores']} threads")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
