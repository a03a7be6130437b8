"""Build recipe of the gfx950 library: hipcc -> athena-gamma_amd/lib/libathena_gamma_amd.so
(objects in athena-gamma_amd/lib/obj/, compiled in parallel).

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
# every source of csrc/ (HIP kernels in .hip, host-only code in .cpp), largest first so
# the longest compile starts first
SOURCES = sorted((f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp"))),
                 key=lambda f: -os.path.getsize(os.path.join(CSRC, f)))

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


def build(verbose=False, jobs=None):
    """Compile every source to an object in parallel (one hipcc per file), then link."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [cc] + HIPFLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        return obj, r

    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, SOURCES))
    bad = [r for _, r in results if r.returncode != 0]
    if bad:
        for r in bad:
            sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    cmd = [cc] + HIPFLAGS + ["-shared", "-o", tmp] + [o for o, _ in results]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
