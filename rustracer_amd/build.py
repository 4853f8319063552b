"""Build recipe for librtamd.so (gfx950) — run by __graft_entry__.build().

The library is several HIP translation units: the host code (rt_lib.hip) and
the kernel units (k_*.hip), compiled in parallel with hipcc for MI355X and
linked into one shared object.  The host unit launches kernels defined in the
other units through their device stubs; each unit's constant tables are
initialised in place.  -ffp-contract=off keeps every f32 expression unfused,
like the reference's Rust code; packed-FP32 instruction selection is disabled.

-fno-slp-vectorize keeps every float scalar.  With packed FP32 arithmetic off,
the SLP vectorizer's <2 x float> values buy nothing, and the gfx950 backend
miscompiled one of them: a hit position merged from x/y lanes of a 64-bit
register pair across a divergent dielectric reflect/refract branch kept the y
lane only on the refract side, so reflected rays left an SSS sphere from a
garbage origin (NaN radiance).  Scalar code merges each component after the
branches reconverge.

build(force=True) recompiles every unit; build(force=False) recompiles only
units older than a source or header they may include (any file of csrc/ or
include/rt_api.h).
"""
import glob
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "librtamd.so")
API = os.path.join(os.path.dirname(HERE), "include", "rt_api.h")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fno-slp-vectorize",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-function", "-Wno-unused-result",
]


def units():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def _newest_dep():
    deps = glob.glob(os.path.join(CSRC, "*")) + [API, os.path.abspath(__file__)]
    return max(os.path.getmtime(p) for p in deps if os.path.exists(p))


def build(verbose=True, force=True, jobs=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    newest = _newest_dep()
    todo = []
    objs = []
    for src in units():
        obj = os.path.join(OBJ, os.path.splitext(os.path.basename(src))[0] + ".o")
        objs.append(obj)
        if force or not os.path.exists(obj) or os.path.getmtime(obj) < newest:
            todo.append((src, obj))

    def compile_one(job):
        src, obj = job
        tmp = obj + ".tmp"
        cmd = [hipcc] + HIPCC_FLAGS + ["-c", "-o", tmp, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, obj)

    workers = jobs or min(len(todo), max(1, min(os.cpu_count() or 1, 16))) or 1
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(compile_one, todo))
    if todo or not os.path.exists(OUT):
        tmp = OUT + ".tmp"
        cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    import sys
    build(force="--force" in sys.argv)
