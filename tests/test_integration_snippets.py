"""The reference-side code in INTEGRATION.md compiles: every ```cpp block that
forms a declaration or a function body is compiled with g++ against
include/sphswe.h and the SimVars of the C++ layer (whose field names are the
reference's, src/include/SimVars.hpp), and linked against libsphswe.so.
No GPU is used: the programs are only built."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DOC = os.path.join(ROOT, "INTEGRATION.md")

PRELUDE = r"""
#include <complex>
#include <cstdlib>
#include <iostream>
#include <vector>
#include "sphswe.h"
#include "SimVars.hpp"
SimVars simVars;
"""

# the reference-side names the function-body snippets use
BODY_ARGS = ("SPHConfigStub *cfg, sphswe_rexi *rexi_obj, double *d_state, int n, int nsteps, double a, double g, "
             "double omega, double h0")


def _blocks():
    txt = open(DOC).read()
    return re.findall(r"```cpp\n(.*?)```", txt, re.S)


def _program(block):
    if "struct SPHConfig" in block:
        return PRELUDE + block + "\nint main() { SPHConfig c; (void)c; return 0; }\n"
    if "SPHData::" in block:
        return None   # member definitions of the reference's own class
    return (PRELUDE + "struct SPHConfigStub { sphswe_ctx *gpu; };\n"
            "void snippet(" + BODY_ARGS + ") {\n" +
            ("" if "sphswe_rexi *rexi" in block else "sphswe_rexi *rexi = rexi_obj;\n") + block + "\n(void)n; (void)nsteps; (void)a; (void)g;"
            " (void)omega; (void)h0; (void)d_state; (void)rexi_obj;\n}\nint main() { return 0; }\n")


def test_integration_has_snippets():
    assert len(_blocks()) >= 4


@pytest.mark.parametrize("k", range(8))
def test_snippet_compiles(tmp_path, k):
    blocks = _blocks()
    if k >= len(blocks):
        pytest.skip("no block %d" % k)
    prog = _program(blocks[k])
    if prog is None:
        pytest.skip("member definitions of the reference's SPHData")
    src = tmp_path / ("s%d.cpp" % k)
    src.write_text(prog)
    lib = os.path.join(ROOT, "sph_swe_amd", "lib")
    cmd = ["g++", "-std=c++14", "-Wall", "-Werror", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "sph_swe_amd", "host", "include"),
           str(src), "-L" + lib, "-lsphswe", "-o", str(tmp_path / ("s%d" % k))]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, prog + "\n" + r.stderr[-3000:]
