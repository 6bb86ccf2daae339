"""Build recipe for the native libraries of qtnh_amd.

    python -m qtnh_amd.build            # everything
    python -m qtnh_amd.build hip        # libqtnh_hip.so only

Outputs (in-tree, git-ignored):
    qtnh_amd/lib/libqtnh_hip.so   gfx950 kernels + C ABI (include/qtnh_hip.h)
    qtnh_amd/lib/libqtnh.so       C++ drop-in API (include/qtnh.hpp) over the C ABI
    tests/cpp/bin/*               C++ tests of the drop-in API
"""
import concurrent.futures
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
LIB_DIR = os.path.join(PKG, "lib")
OBJ_DIR = os.path.join(PKG, "_build")
HIP_LIB = os.path.join(LIB_DIR, "libqtnh_hip.so")
CPP_LIB = os.path.join(LIB_DIR, "libqtnh.so")
TEST_SRC = os.path.join(ROOT, "tests", "cpp")
TEST_BIN = os.path.join(TEST_SRC, "bin")

ARCH = "gfx950"
# -ffp-contract=off: products and sums are rounded separately (the reference's
# arithmetic); FMA is used only where written explicitly.  Packed FP32 ops
# are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-I", INC]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I", INC]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the gfx950 kernels cannot be built")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def _jobs():
    try:
        n = int(os.environ.get("MAX_JOBS", "0"))
    except ValueError:
        n = 0
    return max(1, min(n or os.cpu_count() or 4, 16))


def build_hip(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    hipcc = _hipcc()
    srcs = sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))
    objs = [os.path.join(OBJ_DIR, os.path.basename(s) + ".o") for s in srcs]
    with concurrent.futures.ThreadPoolExecutor(_jobs()) as ex:
        futs = [ex.submit(_run, [hipcc] + HIP_FLAGS + ["-c", s, "-o", o]) for s, o in zip(srcs, objs)]
        for f in futs:
            out = f.result()
            if verbose and out.strip():
                print(out)
    tmp = HIP_LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_cpp(verbose=False):
    """C++ drop-in library (namespace qtnh) linked against libqtnh_hip.so."""
    srcs = sorted(glob.glob(os.path.join(PKG, "cpp", "*.cpp")))
    if not srcs:
        return None
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = [os.path.join(OBJ_DIR, "cpp_" + os.path.basename(s) + ".o") for s in srcs]
    with concurrent.futures.ThreadPoolExecutor(_jobs()) as ex:
        futs = [ex.submit(_run, ["g++"] + CXX_FLAGS + ["-c", s, "-o", o]) for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = CPP_LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp] + objs + ["-L" + LIB_DIR, "-lqtnh_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CPP_LIB)
    return CPP_LIB


def build_cpp_tests():
    srcs = sorted(glob.glob(os.path.join(TEST_SRC, "test_*.cpp")))
    if not srcs:
        return []
    os.makedirs(TEST_BIN, exist_ok=True)
    _run([sys.executable, os.path.join(TEST_SRC, "gen_golden.py"), os.path.join(TEST_BIN, "golden_gen.hpp")])
    outs = []

    def one(src):
        exe = os.path.join(TEST_BIN, os.path.splitext(os.path.basename(src))[0])
        _run(["g++"] + CXX_FLAGS + ["-I", TEST_SRC, "-I", TEST_BIN, src, "-o", exe, "-L" + LIB_DIR, "-lqtnh", "-lqtnh_hip",
                                    "-Wl,-rpath,$ORIGIN/../../../qtnh_amd/lib"])
        return exe

    with concurrent.futures.ThreadPoolExecutor(_jobs()) as ex:
        outs = list(ex.map(one, srcs))
    return outs


def build_all(verbose=False):
    build_hip(verbose)
    build_cpp(verbose)
    build_cpp_tests()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what == "hip":
        build_hip(True)
    elif what == "cpp":
        build_cpp(True)
        build_cpp_tests()
    else:
        build_all(True)
    print("built", what)
