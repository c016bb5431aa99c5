"""Builds liblgx.so (HIP kernels for gfx950 + C++ host code) in-tree.

Every HIP source is compiled with -ffp-contract=off: the kernels place
fused multiply-adds explicitly where they want them, and keep the shader's
separately rounded operations where texel selection must match the
reference arithmetic.  Packed-FP32 instruction selection is disabled.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblgx.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LGX_ARCH", "gfx950")

BASE = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-parameter",
]
# HIP sources: device code for gfx950, packed-FP32 selection off
HIP_FLAGS = [*BASE, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# host-only C++ sources (render graph, renderer): HIP runtime API headers only
CXX_FLAGS = [*BASE, "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include"]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith(".hip") or name.endswith(".cpp"):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, f"--offload-arch={ARCH}", *HIP_FLAGS, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *CXX_FLAGS, "-x", "c++", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


EXAMPLE_SRC = os.path.join(ROOT, "tests", "csrc", "integration_example.cpp")
EXAMPLE_BIN = os.path.join(ROOT, "tests", "csrc", "integration_example")


def build_integration_example(verbose=False):
    """Compiles the C++ consumer of the boundary headers (the code
    INTEGRATION.md shows) against liblgx.so; the GPU tests run it."""
    os.makedirs(BUILD, exist_ok=True)
    obj = os.path.join(BUILD, "integration_example.o")
    cmd = [HIPCC, *CXX_FLAGS, "-x", "c++", "-c", EXAMPLE_SRC, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    tmp = EXAMPLE_BIN + ".tmp"
    cmd = [HIPCC, obj, "-o", tmp, "-L", HERE, "-llgx", "-Wl,-rpath,$ORIGIN/../../legitengine_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, EXAMPLE_BIN)
    if verbose:
        print(f"built {EXAMPLE_BIN}")
    return EXAMPLE_BIN


if __name__ == "__main__":
    build(verbose=True)
    build_integration_example(verbose=True)
    sys.exit(0)
