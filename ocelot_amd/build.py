"""Build recipe for the HIP library libocsc.so (gfx950), in-tree.

    python -m ocelot_amd.build          # or __graft_entry__.build()

hipcc compiles one translation unit, ocelot_amd/csrc/ocsc.hip, with
floating-point contraction off (the particle arithmetic reproduces numpy's
rounding order; fused multiply-adds are written explicitly where wanted) and
packed-FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "ocsc.hip")
OUT = os.path.join(HERE, "libocsc.so")
ARCH = os.environ.get("OCSC_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


def command(out=OUT, extra=()):
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result",
            "-I", os.path.join(REPO, "include"), "-I", os.path.join(HERE, "csrc"),
            *extra, "-o", out, SRC]


def build(verbose=True):
    tmp = OUT + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
