"""Compile libh3d.so (HIP kernels for gfx950 + host C ABI + model builder).

Usage: python -m h2_amd.build   (or h2_amd.build.build() from Python)

The library is written next to this file so that it travels with the source
tree.  Flags:
  -ffp-contract=off   keep every a*b+c as two rounded operations, as the
                      reference (Julia) computes them;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keep the compiler from emitting packed FP32 math.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libh3d.so")
SOURCES = ["h3d_kernels.hip", "h3d_rb.hip", "h3d_adjoint.hip", "h3d_api.hip", "h3d_noalloc.hip", "h3d_resident.hip", "h3d_model.cpp"]
ARCH = "gfx950"


def hipcc_path():
    p = shutil.which("hipcc")
    if p:
        return p
    p = "/opt/rocm/bin/hipcc"
    if os.path.exists(p):
        return p
    raise RuntimeError("hipcc not found (ROCm is required to build libh3d)")


def build(verbose=False):
    """Compile every source into h2_amd/libh3d.so; raises on failure."""
    tmp = LIB + ".tmp%d" % os.getpid()
    cmd = [hipcc_path(), "--offload-arch=" + ARCH, "-O3", "-fPIC", "-shared", "-std=c++17",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-function", "-I", INCLUDE, "-I", CSRC, "-o", tmp, "-lpthread"]
    cmd += [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise RuntimeError("hipcc failed:\n" + out)
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
