"""Build librtx.so (HIP kernels + host runtime + C ABI) for gfx950.

Every source is compiled with hipcc; the FP flags keep the traversal's FP32
arithmetic identical to the reference expressions:
  -ffp-contract=off                         no FMA fusion
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE division and sqrt
  -fno-gpu-flush-denormals-to-zero          keep FP32 denormals
and packed FP32 instructions are disabled for device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "librtx.so")

SOURCES = ["api.cpp", "adapter.cpp", "builder.cpp", "scene.cpp", "formats.cpp", "traverse.hip", "partition.hip",
           "mlp.hip", "gbuild.hip", "static_scratch.hip", "coherence.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RTX_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-I", INCLUDE, "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-lambda-capture",
]
DEVICE = [
    "-x", "hip", f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


CXX = os.environ.get("CXX", "g++")
ROCM_INCLUDE = "/opt/rocm/include"


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + DEVICE + COMMON + ["-c", path, "-o", obj]
    else:
        # host-only C++: the system compiler, HIP runtime API headers only
        cmd = [CXX] + COMMON + ["-I", ROCM_INCLUDE, "-D__HIP_PLATFORM_AMD__", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
