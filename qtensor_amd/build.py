"""Build ``libqtensor_amd.so`` in-tree for gfx950 with hipcc.

``python -m qtensor_amd.build`` compiles ``csrc/qtensor_amd.hip`` (kernels and
the C ABI), ``csrc/noalloc.hip`` (the executor entry point that allocates
nothing) and ``csrc/planner.cpp`` (host-only ordering code) into
``qtensor_amd/libqtensor_amd.so`` (always recompiles).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "qtensor_amd.hip")
NOALLOC_SRC = os.path.join(HERE, "csrc", "noalloc.hip")   # qb_execute_noalloc and its static workspace
HOST_SRC = os.path.join(HERE, "csrc", "planner.cpp")    # host-only code (elimination ordering)
OUT = os.path.join(HERE, "libqtensor_amd.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-shared", "-fPIC",
    # no packed-FP32 instructions anywhere in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def build(verbose=False):
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", INCLUDE, "-o", tmp, SRC, NOALLOC_SRC, HOST_SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}) building {SRC}")
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
