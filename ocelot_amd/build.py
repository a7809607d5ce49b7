"""Build recipe for the HIP library libocsc.so (gfx950), in-tree.

    python -m ocelot_amd.build          # or __graft_entry__.build()

hipcc compiles the translation unit (ocsc.hip: kernels and C ABI) with
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
SRCS = [os.path.join(HERE, "csrc", "ocsc.hip")]
OBJDIR = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libocsc.so")
ARCH = os.environ.get("OCSC_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def compile_cmd(src, obj, extra=()):
    return [hipcc(), f"--offload-arch={ARCH}", *FLAGS, "-I", os.path.join(REPO, "include"),
            "-I", os.path.join(HERE, "csrc"), *extra, "-c", "-o", obj, src]


def link_cmd(objs, out):
    return [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out, *objs]


def build(verbose=True, extra=()):
    """Compile every translation unit (in parallel) and link libocsc.so."""
    os.makedirs(OBJDIR, exist_ok=True)
    procs, objs = [], []
    for src in SRCS:
        obj = os.path.join(OBJDIR, os.path.splitext(os.path.basename(src))[0] + ".o")
        cmd = compile_cmd(src, obj, extra)
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
    rcs = [p.wait() for p in procs]
    if any(rcs):
        raise subprocess.CalledProcessError(max(rcs), "hipcc")
    tmp = OUT + ".tmp"
    cmd = link_cmd(objs, tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
