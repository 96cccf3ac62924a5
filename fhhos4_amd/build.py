"""Builds libfhgpu.so (HIP kernels for gfx950 + C++ host + C ABI) in-tree.

    python -m fhhos4_amd.build        # or fhhos4_amd.build.build()

Every source is compiled with hipcc for --offload-arch=gfx950 into
fhhos4_amd/_build/ and linked into fhhos4_amd/libfhgpu.so.  Packed FP32
instructions are disabled in device code.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfhgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OMP_DIR = "/opt/rocm/llvm/lib"

SOURCES = ["fh_kernels.hip", "fh_elem.hip", "fh_hho_kernels.hip", "fh_nn_kernels.hip", "fh_condense.hip", "fh_matrix.cpp", "fh_io.cpp", "fh_solver.cpp", "fh_nn.cpp",
           "fh_uamg.cpp", "fh_hho.cpp", "fh_capi.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-fopenmp",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + COMMON + lang + ["-I", os.path.join(PKG, "..", "include"), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-fopenmp", "-o", tmp] + objs + [
        "-L" + OMP_DIR, "-Wl,-rpath," + OMP_DIR, "-lomp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
