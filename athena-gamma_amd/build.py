"""Build recipe of the gfx950 library: hipcc -> athena-gamma_amd/lib/libathena_gamma_amd.so.

Run as `python athena-gamma_amd/build.py` or through __graft_entry__.build().
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libathena_gamma_amd.so")
SOURCES = ["agx_mhd.hip", "agx_mg.hip", "agx_refine.hip"]

# -ffp-contract=off: the FP64 path rounds every expression as written (bitwise parity
# with the CPU restatement).  -fno-hip-fp32-correctly-rounded-divide-sqrt: FP32 division
# and square root (used only by the mixed-precision reconstruction/Riemann kernels) take
# the hardware reciprocal/sqrt (about 1-2.5 ulp); FP64 code is unaffected.  Packed FP32
# instructions are disabled for all HIP code.
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-fno-hip-fp32-correctly-rounded-divide-sqrt",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPFLAGS + ["-shared", "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
