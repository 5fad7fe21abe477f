"""The reference-side driver code in INTEGRATION.md compiles against include/biot.h and links
against the library: the first C++ block of section 1 is wrapped in a function, with the
reference driver's own variables (main.cpp) declared, and built with g++."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "mixed-precision-optimization-method_amd", "lib")

PRELUDE = """#include <cstdint>
#include <cstdio>
#include <vector>
%s
// the reference driver's variables the block uses (main.cpp)
static int NX_ext = 64, NY_ext = 64, NZ_ext = 64, NP = 10, NT = 500, sn = 32 + 32 * 64 + 32 * 64 * 64;
static int sxnew = 32, synew = 32;
static float H = 5.0f, DT = 4e-4f, F0 = 25.0f;
static const double *Ks, *Kf, *Kb, *mu, *phi, *rhos, *rhof, *eta, *perm;
void driver()
{
%s
}
int main() { return 0; }
"""


def _snippet():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, re.S)
    assert blocks, "no C++ block in INTEGRATION.md"
    lines = blocks[0].splitlines()
    inc = [ln.split("//")[0] for ln in lines if ln.startswith("#include")]
    body = [ln for ln in lines if not ln.startswith("#include")]
    return "\n".join(inc), "\n".join(body)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_integration_snippet_compiles_and_links(tmp_path):
    inc, body = _snippet()
    src = tmp_path / "driver.cpp"
    src.write_text(PRELUDE % (inc, body))
    exe = tmp_path / "driver"
    cmd = ["g++", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           str(src), "-o", str(exe)]
    lib = os.path.join(LIBDIR, "libbiot.so")
    if os.path.exists(lib):
        cmd += ["-L", LIBDIR, "-lbiot", "-Wl,--allow-shlib-undefined"]
    else:
        cmd[cmd.index("-o"):cmd.index("-o") + 2] = ["-c", "-o", str(tmp_path / "driver.o")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
