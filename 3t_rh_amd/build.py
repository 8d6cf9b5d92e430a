"""Builds librh3t.so in-tree: the host Field/GhostGrid library (g++) and the
gfx950 HIP kernels + device context (hipcc), linked into one shared object.

Run as ``python 3t_rh_amd/build.py`` or through ``__graft_entry__.build()``.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librh3t.so")

ARCH = os.environ.get("RH3T_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off everywhere: the kernels keep the reference's rounding
# sequence (Rust never contracts a*b+c).  Packed FP32 instructions are
# disabled for all device code.
HIPCC_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
]
CXX_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    if verbose and res.stdout.strip():
        print(res.stdout)
    return res.stdout


def build(verbose=False, resource_usage=False):
    os.makedirs(BUILD, exist_ok=True)
    inc = [f"-I{INCLUDE}", f"-I{CSRC}"]
    host_o = os.path.join(BUILD, "problem.o")
    _run(["g++", *CXX_FLAGS, *inc, "-c", os.path.join(CSRC, "problem.cpp"), "-o", host_o], verbose)
    extra = ["-Rpass-analysis=kernel-resource-usage"] if resource_usage else []
    out = ""
    dev_objs = []
    for name in ("engine", "adjoint", "noalloc"):
        obj = os.path.join(BUILD, f"{name}.o")
        out += _run([_hipcc(), *HIPCC_FLAGS, *inc, *extra, "-c", os.path.join(CSRC, f"{name}.hip"),
                     "-o", obj], verbose)
        dev_objs.append(obj)
    tmp = LIB + ".tmp"
    _run([_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", host_o, *dev_objs, "-o", tmp,
          "-Wl,-rpath,/opt/rocm/lib"], verbose)
    os.replace(tmp, LIB)
    return out if resource_usage else LIB


if __name__ == "__main__":
    print(build(verbose=True, resource_usage="--usage" in sys.argv))
