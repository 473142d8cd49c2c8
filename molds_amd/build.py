"""Builds libnddo.so (HIP kernels for gfx950 + host engine + C ABI) in-tree.

    python -m molds_amd.build        # or molds_amd.build.build()

Every source is compiled with hipcc for --offload-arch=gfx950 and linked into
molds_amd/lib/libnddo.so, which is what the ctypes binding loads.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "lib", "obj")
LIB = os.path.join(PKG, "lib", "libnddo.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["kernels_setup.hip", "kernels_scf.hip", "kernels_eig.hip", "kernels_xl.hip", "kernels_xlh.hip",
           "kernels_props.hip", "engine.cpp", "server.cpp", "params.cpp"]
# Packed FP32 instructions (v_pk_*_f32) are disabled for all device code: some of
# their forms have given silently wrong results on this hardware generation.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wno-unused-command-line-argument", "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]


# The per-geometry setup kernels follow the reference's operation order without
# fused multiply-adds (cancellation-sensitive multipole and overlap formulas).
PER_FILE = {"kernels_setup.hip": ["-ffp-contract=off"]}


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + PER_FILE.get(src, []) + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


ADAPTER_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_es_adapter.cpp")
ADAPTER_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_es_adapter")


def build_adapter_test():
    """Host-only C++ program over include/nddo_es.hpp, linked against libnddo.so."""
    os.makedirs(os.path.dirname(ADAPTER_TEST_BIN), exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), ADAPTER_TEST_SRC,
           "-L" + os.path.dirname(LIB), "-lnddo", "-Wl,-rpath," + os.path.dirname(LIB), "-o", ADAPTER_TEST_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for the adapter test:\n{r.stdout}\n{r.stderr}")
    return ADAPTER_TEST_BIN


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_adapter_test()
    if verbose:
        print("built", LIB, "and", ADAPTER_TEST_BIN)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
