"""Build libecl_wavelet_hip.so in-tree with hipcc for gfx950.

Each HIP translation unit under csrc/ is compiled to an object in
``_build/`` (in parallel) and linked into ``ecl_wavelet_amd/
libecl_wavelet_hip.so``.  Called by ``__graft_entry__.build()``.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libecl_wavelet_hip.so")
SOURCES = ["capi.hip", "forest.hip", "server.hip", "kernels_build.hip", "kernels_tree.hip",
           "kernels_query.hip", "probe.hip", "noalloc.hip",
           "textio.hip"]
ARCH = os.environ.get("ECL_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def common_flags():
    return [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
        # packed FP32 ops stay off (no float math here, but every HIP unit of
        # this project is built the same way)
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        "-Wall", "-Wno-unused-function", "-Wno-unused-parameter",
    ]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *common_flags(), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


DROPIN_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")
DROPIN_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_dropin")


def build_dropin_test():
    """The C++ drop-in test program (tests/cpp/test_dropin.cpp), linked
    against the in-tree library with an rpath to it."""
    cmd = [hipcc(), "-O2", "-std=c++20", f"--offload-arch={ARCH}",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(ROOT, "include"), DROPIN_TEST_SRC,
           "-o", DROPIN_TEST_BIN + ".tmp", "-L", PKG, "-lecl_wavelet_hip",
           "-Wl,-rpath,$ORIGIN/../../ecl_wavelet_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for test_dropin.cpp:\n{r.stdout}\n{r.stderr}")
    os.replace(DROPIN_TEST_BIN + ".tmp", DROPIN_TEST_BIN)
    return DROPIN_TEST_BIN


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_dropin_test())
