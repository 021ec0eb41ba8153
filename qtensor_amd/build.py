"""Build ``libqtensor_amd.so`` in-tree for gfx950 with hipcc.

``python -m qtensor_amd.build`` compiles the library's translation units in
parallel into ``qtensor_amd/_build/`` and links ``qtensor_amd/libqtensor_amd.so``:

* ``csrc/qtensor_amd.hip``       kernels and the C ABI
* ``csrc/noalloc.hip``           the executor entry point that allocates nothing
* ``csrc/planner.cpp``           host-only ordering code
* ``csrc/lightcone_planner.cpp`` the host-only lightcone planner

It always recompiles every unit; ``--incremental`` recompiles only units whose
source (or a header) is newer than their object.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ("qtensor_amd.hip", "noalloc.hip", "planner.cpp", "lightcone_planner.cpp")
SRC = os.path.join(CSRC, "qtensor_amd.hip")
OUT = os.path.join(HERE, "libqtensor_amd.so")
OBJ_DIR = os.path.join(HERE, "_build")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HEADERS = (os.path.join(INCLUDE, "qtensor_amd.h"), os.path.join(CSRC, "internal.h"))

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions anywhere in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def _run(cmd, what):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}) building {what}")
    return res.stderr


def build(verbose=False, incremental=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    newest_header = max(os.path.getmtime(h) for h in HEADERS)
    jobs, objs = [], []
    for name in SOURCES:
        src = os.path.join(CSRC, name)
        obj = os.path.join(OBJ_DIR, name + ".o")
        objs.append(obj)
        if incremental and os.path.exists(obj) and \
                os.path.getmtime(obj) >= max(os.path.getmtime(src), newest_header):
            continue
        jobs.append(([hipcc()] + HIPCC_FLAGS + ["-I", INCLUDE, "-c", "-o", obj, src], name))
    with ThreadPoolExecutor(max(1, min(len(jobs), os.cpu_count() or 1))) as ex:
        logs = list(ex.map(lambda j: _run(*j), jobs))
    tmp = OUT + ".tmp"
    logs.append(_run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, OUT))
    if verbose:
        sys.stderr.write("".join(logs))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
