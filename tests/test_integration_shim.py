"""The reference-side binding shown in INTEGRATION.md compiles against the reference's
own headers (src/include/we_op.hpp, injector.hpp, param.hpp, vecReg.hpp) and the
engine's include/fwi3d_hip.h, and the C++ shown there is that compiled code."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "tests", "integration", "we_op_shim.cpp")
REF_INC = "/root/reference/src/include"


@pytest.mark.skipif(not os.path.isdir(REF_INC), reason="needs the reference checkout")
def test_shim_compiles_against_reference_headers():
    cmd = ["g++", "-std=c++11", "-fopenmp", "-fsyntax-only", "-Wall", "-Werror", "-isystem", REF_INC,
           "-I" + os.path.join(ROOT, "include"), SHIM]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


def test_integration_snippets_are_the_compiled_shim():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    shim = open(SHIM).read()
    bindings = text.split("## 1.", 1)[1].split("## 2b.", 1)[0]  # sections 1 and 2: the reference-side binding
    blocks = re.findall(r"```cpp\n(.*?)```", bindings, flags=re.S)
    assert blocks, "INTEGRATION.md shows no C++"
    shim_lines = {l.strip() for l in shim.splitlines()}
    for b in blocks:
        for line in b.splitlines():
            s = line.strip()
            if s and s != "..." and not s.startswith("//"):
                assert s in shim_lines, f"INTEGRATION.md line not in the compiled shim: {s}"
    # every fenced block is closed (an odd number of fences would swallow the text after it)
    assert text.count("```") % 2 == 0
