"""Builds libspiffe_amd.so (the HIP kernels and the C ABI) for gfx950, in-tree.

Usage: python -m spiffe_amd.build   (or spiffe_amd.build.build() from Python)
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libspiffe_amd.so")
SOURCES = ["spf_api.hip", "spf_api_poisson.hip", "spf_fdtd.hip", "spf_push.hip", "spf_deposit.hip",
           "spf_poisson.hip", "spf_sort.hip", "spf_util.hip", "spf_cathode.hip", "spf_solenoid.hip",
           "spf_noalloc.hip"]
# host-only C++ (no HIP), built with the system C++ compiler
HOST_SOURCES = ["spf_emit.cpp", "spf_namelist.cpp", "spf_geometry.cpp", "spf_sdds.cpp", "spf_deck.cpp"]
# the command-line deck runner, linked against the library
EXE = os.path.join(HERE, "spiffe-amd")

# -ffp-contract=off: no FMA contraction, each FP64 operation rounds as written;
# -packed-fp32-ops: no packed FP32 instructions are generated.
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-result",
            "-Wno-unused-function"]


# host C++: no FMA contraction, and sin/cos stay separate libm calls (no
# sincos fusion) so the emission draws are the libm results of each call
CXXFLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fno-builtin-sin",
            "-fno-builtin-cos", "-Wall"]


def _hipcc():
    for cand in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in cand or os.path.exists(cand):
            return cand
    return "hipcc"


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):
        cmd = [os.environ.get("CXX", "g++")] + CXXFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [_hipcc()] + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed for %s:\n%s\n%s" % (src, r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = SOURCES + HOST_SOURCES
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    cmd = [os.environ.get("CXX", "g++")] + CXXFLAGS + [os.path.join(CSRC, "spf_main.cpp"), "-o", EXE + ".tmp",
                                                        "-L" + HERE, "-lspiffe_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("runner link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(EXE + ".tmp", EXE)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
