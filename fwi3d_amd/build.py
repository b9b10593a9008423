"""Build the engine library fwi3d_amd/libfwi3d_amd.so for gfx950 with hipcc, the
drop-in executables fwi3d_amd/bin/WE_MODELING_3D.x and fwi3d_amd/bin/FWI3D.x on top of
it, and the host library of their parameter parsing and file formats
(fwi3d_amd/libfwi3d_cli.so).

Usage: python -m fwi3d_amd.build   (also called by __graft_entry__.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfwi3d_amd.so")
CLI_LIB = os.path.join(HERE, "libfwi3d_cli.so")
BIN = os.path.join(HERE, "bin")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("engine.hip", "operator.hip", "lbfgs.hip")]
CLI = os.path.join(HERE, "csrc", "cli")
EXECUTABLES = {"WE_MODELING_3D.x": "we_modeling.cpp", "FWI3D.x": "fwi3d.hip"}
ARCH = os.environ.get("FWI3D_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17",
    # packed FP32 instructions stay off (hipcc would otherwise form v_pk_* from float code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed ({r.returncode}):\n{out}")


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    _run([hipcc] + FLAGS + ["-fPIC", "-shared"] + SOURCES + ["-o", tmp], verbose)
    os.replace(tmp, LIB)
    tmp = CLI_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", os.path.join(CLI, "cli_api.cpp"), "-o", tmp], verbose)
    os.replace(tmp, CLI_LIB)
    os.makedirs(BIN, exist_ok=True)
    for exe, src in EXECUTABLES.items():
        out = os.path.join(BIN, exe)
        _run([hipcc] + FLAGS + [os.path.join(CLI, src), "-o", out + ".tmp", "-L" + HERE, "-lfwi3d_amd",
                                "-Wl,-rpath,$ORIGIN/.."], verbose)
        os.replace(out + ".tmp", out)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
