"""Builds libav1pf.so (the gfx950 post-filter library) in-tree with hipcc.

    python -m goav1_amd.build        # or goav1_amd.build.build()

The library is the product: the C ABI of include/av1pf.h over hand-written
HIP kernels.  It is compiled for gfx950 only.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libav1pf.so")
# chain.hip compiles deblock.hip, cdef.hip, lr.hip and superres.hip into its
# own translation unit (the fused chain kernel inlines their tile bodies)
SOURCES = ["av1pf_api.hip", "chain.hip", "border.hip", "lfbuild.hip", "grain.hip", "search.hip",
           "lr_search.hip", "nn.hip", "search_finish.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libav1pf.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
