"""Build libpfw_mi355x.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc.

Usage: ``python potentialframework.jl_amd/build.py`` or ``build_library()`` from Python.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpfw_mi355x.so")
WL_LIB = os.path.join(HERE, "libpfw_workloads.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for all device code: the compiler
# otherwise forms v_pk_{add,mul,fma}_f32 from scalar float code.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-variable",
]


def _sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    out = os.path.join(OBJ, src[:-4] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out, r.stderr


def build_library(verbose=False):
    """Compile every csrc/*.hip for gfx950 and link libpfw_mi355x.so. Always rebuilds."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = _sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for o, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    # host-only workload generator (synthetic fitting sets for tests and the benchmark)
    tmp = WL_LIB + ".tmp"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", tmp,
                        os.path.join(CSRC, "workloads.cpp")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"workload generator build failed:\n{r.stderr}")
    os.replace(tmp, WL_LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
