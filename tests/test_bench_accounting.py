"""CPU checks of the benchmark's algorithmic-byte accounting (the numerator of every roofline
fraction it reports): the slab membership count against a brute-force enumeration of the
updated cells, and the per-cell byte figures against the kernels' access lists written out by
hand (reads and writes per updated cell and per slab membership)."""
import itertools
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


@pytest.mark.parametrize("n,L,R,h", [(32, 6, 6, 2), (32, 0, 5, 1), (20, 12, 12, 2), (9, 2, 3, 1), (64, 20, 20, 2)])
def test_slab_counts_brute_force(n, L, R, h):
    cells = range(h, n - h)
    brute = sum(1 for i in cells if i < L or i >= n - R)
    assert bench.slab_counts(n, L, R, h) == brute


def test_algorithmic_bytes_closed_form():
    nx, ny, nz, order = 24, 20, 16, 4
    h = order // 2
    pml = (4, 4, 3, 5, 2, 6)
    upd = (nx - 2 * h) * (ny - 2 * h) * (nz - 2 * h)
    # slab memberships: every updated cell once per axis whose slab holds it
    memb = 0
    for i, j, k in itertools.product(range(h, nx - h), range(h, ny - h), range(h, nz - h)):
        memb += (i < pml[0] or i >= nx - pml[1]) + (j < pml[2] or j >= ny - pml[3]) + \
                (k < pml[4] or k >= nz - pml[5])
    # velocity kernel per cell: reads 7 tau/p + 6 v/q, writes 6 v/q (19 field accesses),
    # 14 coefficients; stress kernel: reads 6 v/q + 7 tau/p, writes 7 (20), 8 coefficients;
    # one fused step: 13 reads + 13 writes, 22 coefficients; 4 carries read + written per
    # slab membership and half step
    field_access = {"v": 7 + 6 + 6, "t": 6 + 7 + 7, "step": 13 + 13}
    coef_access = {"v": 14, "t": 8, "step": 22}
    carry_access = {"v": 8, "t": 8, "step": 16}
    residual_access = {"fp16c": {"v": 10 + 3, "t": 10 + 7, "step": 20},   # v, tau, p carry residuals
                       "fp16t": {"v": 7, "t": 7 + 7, "step": 14},         # tau, p carry residuals
                       "fp16m": {"v": 4, "t": 4 + 4, "step": 8}}          # tau, p residual record
    # FP16M reads binary16 moduli instead of coefficients: velocity half step rho, rho_f, C1,
    # C2 (4 x 2 B); stress half step C, M, lambda_c, H (4 x 2 B) and mu (2 B)
    medium_bytes = {"fp16m": {"v": 4 * 2, "t": 4 * 2 + 2, "step": 4 * 2 + 4 * 2 + 2}}
    for mode, (wf, wk, wc, wr) in {"fp32": (4, 4, 4, 0), "fp16": (2, 2, 2, 0), "fp16c": (2, 2, 4, 1),
                                   "fp16t": (2, 2, 4, 1), "fp16m": (2, 2, 4, 1), "fp64": (8, 8, 8, 0)}.items():
        v, t, step, u, m = bench.algorithmic_bytes(nx, ny, nz, order, pml, mode)
        assert (u, m) == (upd, memb)
        for key, got in (("v", v), ("t", t), ("step", step)):
            ra = residual_access.get(mode, {}).get(key, 0)
            kb = medium_bytes[mode][key] if mode in medium_bytes else coef_access[key] * wk
            want = upd * (field_access[key] * wf + kb + ra * wr) + memb * carry_access[key] * wc
            assert got == want, (mode, key)


def test_fp32_step_bytes_per_interior_cell():
    """Without absorbing slabs the fused FP32 step moves 26 x 4 + 22 x 4 = 192 B per cell."""
    v, t, step, upd, memb = bench.algorithmic_bytes(16, 16, 16, 2, (0, 0, 0, 0, 0, 0), "fp32")
    assert memb == 0 and step == 192 * upd and v == 132 * upd and t == 112 * upd


def test_bench_refuses_more_than_one_gpu():
    """The benchmark measures one GPU; asked for more it exits with status 2 and says so,
    before it touches a device."""
    import subprocess

    r = subprocess.run([sys.executable, bench.__file__, "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout.strip()


def test_host_cpu_record():
    h = bench.host_cpu()
    assert set(h) == {"nproc", "usable_cpus", "cpu_model"}
    assert h["nproc"] >= 1 and 1 <= h["usable_cpus"] <= h["nproc"]
    assert h["cpu_model"] is None or isinstance(h["cpu_model"], str)


def test_adjoint_bytes_closed_form():
    """The adjoint sweep's byte model from its access lists: 4-byte values, per updated cell and
    adjoint step 49 + 24 + 76 + 26 accesses in the four passes, 24 carry accesses per slab
    membership; the recomputed forward step; a state save per step and a state load per
    segment (every field of every cell and every C-PML carry, read and written); a flush per
    segment (22 partial sums read and cleared, 22 FP64 sums read and written)."""
    nx, ny, nz, order, nt, every = 20, 18, 16, 4, 37, 10
    pml = (3, 3, 2, 4, 3, 2)
    v, t, _, upd, memb = bench.algorithmic_bytes(nx, ny, nz, order, pml, "fp32")
    n = nx * ny * nz
    slab = (pml[0] + pml[1]) * ny * nz + nx * (pml[2] + pml[3]) * nz + nx * ny * (pml[4] + pml[5])
    state = (13 * n + 8 * slab) * 4
    nseg = 4
    want = nseg * 2 * state + nt * (2 * state + v + t + (49 + 24 + 76 + 26) * 4 * upd + 24 * 4 * memb) \
        + nseg * 22 * (4 + 4 + 8 + 8) * upd
    total, per_step, u = bench.adjoint_bytes(nx, ny, nz, order, pml, nt, every)
    assert u == upd and total == want
    assert per_step == 2 * state + v + t + 700 * upd + 96 * memb
    # every step's state stored: the adjoint passes only, a flush every 50 steps (37 steps: one)
    total, per_step, u = bench.adjoint_bytes(nx, ny, nz, order, pml, nt, 1)
    assert per_step == (49 + 24 + 76 + 26) * 4 * upd + 24 * 4 * memb
    assert total == nt * per_step + 22 * (4 + 4 + 8 + 8) * upd
