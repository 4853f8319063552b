"""Build recipe for librtamd.so (gfx950) — run by __graft_entry__.build().

The whole library is one HIP translation unit compiled with hipcc for
MI355X.  -ffp-contract=off keeps every f32 expression unfused, like the
reference's Rust code; packed-FP32 instruction selection is disabled.

-fno-slp-vectorize keeps every float scalar.  With packed FP32 arithmetic off,
the SLP vectorizer's <2 x float> values buy nothing, and the gfx950 backend
miscompiled one of them: a hit position merged from x/y lanes of a 64-bit
register pair across a divergent dielectric reflect/refract branch kept the y
lane only on the refract side, so reflected rays left an SSS sphere from a
garbage origin (NaN radiance).  Scalar code merges each component after the
branches reconverge.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "rt_lib.hip")
OUT = os.path.join(HERE, "librtamd.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fno-slp-vectorize",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-function", "-Wno-unused-result",
]


def build(verbose=True):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
