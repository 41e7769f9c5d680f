"""Build recipe for libmfv.so: every HIP/C++ source under csrc/ compiled with hipcc for gfx950 and
linked in-tree (brae_amd/libmfv.so), so the library travels with the source tree."""
import concurrent.futures
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmfv.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no silent FMA contraction, so the FP64 kernels keep the reference's rounding
# (a face loop's a*b + c is two roundings there). Packed-FP32 instructions are disabled.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-Wall", "-Wno-unused-function"]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + (["-x", "hip"] if src.endswith(".hip") else []) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = max(1, min(len(srcs), os.cpu_count() or 4, 16))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_cpp_callers(verbose=False):
    """The C++ callers under tests/cpp (brae's call sites through include/brae_mfv.hpp), linked against
    the in-tree libmfv.so with an $ORIGIN-relative rpath so they run wherever the tree is copied."""
    out = []
    for src in sorted(glob.glob(os.path.join(ROOT, "tests", "cpp", "*.cpp"))):
        exe = src[:-4] + ".bin"
        cmd = [HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", exe,
               "-L", PKG, "-lmfv", "-Wl,-rpath,$ORIGIN/../../brae_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
        out.append(exe)
        if verbose:
            print("built", exe)
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_callers(verbose=True)
    sys.exit(0)
