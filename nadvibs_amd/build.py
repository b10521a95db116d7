"""Builds libnadvibs.so (HIP, gfx950) in the package directory.

    python -m nadvibs_amd.build          # library only
The oracle has its own recipe (oracle/Makefile); the Fortran driver builds
with driver/Makefile.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnadvibs.so")

SOURCES = ["nv_operator.cpp", "nv_jit.cpp", "nv_kernels.hip", "nv_pro.hip", "nv_fcstart.hip", "nv_roots.hip", "nv_adjoint.hip",
           "nv_api.hip"]

# Packed FP32 instructions are disabled for every device pass.
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-value",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def build(verbose=True, jobs=None):
    """Compiles every source to an object in csrc/_build and links the .so."""
    out = os.path.join(CSRC, "_build")
    os.makedirs(out, exist_ok=True)
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(out, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [hipcc()] + HIPFLAGS + inc + ["-x", "hip" if src.endswith(".hip") else "c++",
                                          "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [hipcc(), "-O3", "-std=c++17", "-fPIC"] + inc + [
                "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        # the host pass does not know the packed-fp32 target feature: harmless
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
        if p.returncode != 0:
            failed = True
            print(" ".join(cmd) + "\n" + "\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("compilation failed")
    tmp = LIB + ".tmp"
    link = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lhiprtc"]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.check_call(link)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
