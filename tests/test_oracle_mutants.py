"""The oracle's pins catch plausible mistakes: each case plants one mistake in a copy of the
oracle's sources (a sign flip in the inertial term m - rho_f^2/rho, an inverted rho_f/rho, a
difference taken along the wrong stride or in the wrong direction, a C-PML memory on the wrong
profile node), builds that copy, and runs tests/test_oracle_pins.py against it through
BIOT_ORACLE_LIB.  The pins must fail on every mutant.  The mistakes live only in temporary
copies; the oracle itself is untouched.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SOURCES = ("biot_oracle.c", "biot_oracle_core.inc", "biot_host_adjoint.c", "biot_oracle.h")
CFLAGS = ["-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11", "-w", "-shared"]

# (name, file, text, replacement); the text must occur exactly once
MUTANTS = [
    # the fluid coupling (main.cpp:447-481)
    ("W1: +rho_f^2/rho", "biot_oracle.c",
     "(Cvwp1 * (k2 - (r * r / q) - 0.5 * DT * k1) / (k2 - (r * r / q) + 0.5 * DT * k1))",
     "(Cvwp1 * (k2 + (r * r / q) - 0.5 * DT * k1) / (k2 + (r * r / q) + 0.5 * DT * k1))"),
    ("W3 denominator: +rho_f^2/rho", "biot_oracle.c",
     "(Cp * (k2 - r * r / q + 0.5 * DT * k1))", "(Cp * (k2 + r * r / q + 0.5 * DT * k1))"),
    ("W2 denominator: +rho_f^2/rho", "biot_oracle.c",
     "(Cp * (k2 - (r * r / q) + 0.5 * DT * k1))", "(Cp * (k2 + (r * r / q) + 0.5 * DT * k1))"),
    ("W1: damping term sign", "biot_oracle.c",
     "(Cvwp1 * (k2 - (r * r / q) - 0.5 * DT * k1)", "(Cvwp1 * (k2 - (r * r / q) + 0.5 * DT * k1)"),
    ("W2: rho/rho_f", "biot_oracle.c", "Cv * DT * Cvwp2 * (r / q)", "Cv * DT * Cvwp2 * (q / r)"),
    ("U2 inverted", "biot_oracle.c", "rf[d][c] * Cvup2 / (rh[d][c])", "rh[d][c] * Cvup2 / (rf[d][c])"),
    # velocity half step directions (kernel.cu:172-185)
    ("x2 along z", "biot_oracle_core.inc", "x2 = DB_(txy, c, sy)", "x2 = DB_(txy, c, sz)"),
    ("x3 along y", "biot_oracle_core.inc", "x3 = DB_(txz, c, sz)", "x3 = DB_(txz, c, sy)"),
    ("y1 along z", "biot_oracle_core.inc", "y1 = DF_(tyy, c, sy)", "y1 = DF_(tyy, c, sz)"),
    ("y3 along x", "biot_oracle_core.inc", "y3 = DB_(tyz, c, sz)", "y3 = DB_(tyz, c, sx)"),
    ("s2 along z", "biot_oracle_core.inc", "s2 = DF_(p, c, sy)", "s2 = DF_(p, c, sz)"),
    ("z2 along y", "biot_oracle_core.inc", "z2 = DB_(txz, c, sx)", "z2 = DB_(txz, c, sy)"),
    ("z3 along x", "biot_oracle_core.inc", "z3 = DB_(tyz, c, sy)", "z3 = DB_(tyz, c, sx)"),
    ("s3 along y", "biot_oracle_core.inc", "s3 = DF_(p, c, sz)", "s3 = DF_(p, c, sy)"),
    ("x2 forward", "biot_oracle_core.inc", "x2 = DB_(txy, c, sy)", "x2 = DF_(txy, c, sy)"),
    ("z1 backward", "biot_oracle_core.inc", "z1 = DF_(tzz, c, sz)", "z1 = DB_(tzz, c, sz)"),
    # stress half step directions (kernel.cu:430-445)
    ("uyy along z", "biot_oracle_core.inc", "uyy = DB_(vy, c, sy)", "uyy = DB_(vy, c, sz)"),
    ("wz along y", "biot_oracle_core.inc", "wz = DB_(qz, c, sz)", "wz = DB_(qz, c, sy)"),
    ("uzy along x", "biot_oracle_core.inc", "uzy = DF_(vz, c, sy)", "uzy = DF_(vz, c, sx)"),
    ("uxz along y", "biot_oracle_core.inc", "uxz = DF_(vx, c, sz)", "uxz = DF_(vx, c, sy)"),
    # C-PML memory on the wrong profile node (kernel.cu:188-216, 449-482)
    ("x2 memory at half nodes", "biot_oracle_core.inc",
     "PML_(P[1], Q[1], x2, eyj, dyj)", "PML_(P[1], Q[1], x2, eyj2, dyj2)"),
    ("uzy memory at integer nodes", "biot_oracle_core.inc",
     "PML_(P[11], Q[11], uzy, eyj2, dyj2)", "PML_(P[11], Q[11], uzy, eyj, dyj)"),
]


def _build(tmp, name, fname, text, repl):
    d = os.path.join(tmp, "m%02d" % [m[0] for m in MUTANTS].index(name))
    os.makedirs(d)
    for s in SOURCES:
        shutil.copy(os.path.join(ORACLE, s), d)
    path = os.path.join(d, fname)
    src = open(path).read()
    assert src.count(text) == 1, (name, src.count(text))
    with open(path, "w") as f:
        f.write(src.replace(text, repl))
    lib = os.path.join(d, "liboracle.so")
    subprocess.run(["gcc", *CFLAGS, "-o", lib, os.path.join(d, "biot_oracle.c"),
                    os.path.join(d, "biot_host_adjoint.c"), "-lm"], check=True)
    return lib


def _pins_pass(lib):
    env = dict(os.environ, BIOT_ORACLE_LIB=lib, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_oracle_pins.py"),
                        "-x", "-q", "-p", "no:cacheprovider"], env=env, capture_output=True, text=True,
                       cwd=ROOT)
    last = (r.stdout.strip().splitlines() or [""])[-1]
    return r.returncode == 0, last


@pytest.mark.slow
def test_pins_reject_every_mutant(tmp_path):
    with ThreadPoolExecutor(max_workers=8) as ex:
        libs = list(ex.map(lambda m: _build(str(tmp_path), *m), MUTANTS))
        results = list(ex.map(_pins_pass, libs))
    survivors = [(m[0], last) for m, (ok, last) in zip(MUTANTS, results) if ok]
    assert not survivors, f"mutants the pins did not catch: {survivors}"


def test_unmutated_copy_passes(tmp_path):
    """The same copy-and-build recipe without a mutation passes the pins (so a mutant's failure
    comes from its mistake, not from the recipe)."""
    d = os.path.join(str(tmp_path), "clean")
    os.makedirs(d)
    for s in SOURCES:
        shutil.copy(os.path.join(ORACLE, s), d)
    lib = os.path.join(d, "liboracle.so")
    subprocess.run(["gcc", *CFLAGS, "-o", lib, os.path.join(d, "biot_oracle.c"),
                    os.path.join(d, "biot_host_adjoint.c"), "-lm"], check=True)
    ok, last = _pins_pass(lib)
    assert ok, last
