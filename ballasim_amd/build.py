"""Builds the library in-tree: ballasim_amd/lib/libbsx.so (HIP, gfx950) from ballasim_amd/csrc.

Every translation unit is compiled by hipcc under the floating-point policy of csrc/core/hd.h
(-ffp-contract=off, no fast-math, correctly rounded fp32 division and sqrt) with the packed-fp32
instructions disabled, then linked into one shared library. Objects go to ballasim_amd/_build.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libbsx.so")
ARCH = os.environ.get("BSX_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "capi.cpp",
    "sensor/params.cpp",
    "sensor/maps.cpp",
    "sensor/noise_tables.cpp",
    "sensor/model.cpp",
    "sensor/chain.cpp",
    "sensor/rig.cpp",
    "mlp/proposer.cpp",
    "output/output.cpp",
    "output/stream.cpp",
    "output/output.hip",
    "mlp/mlp.hip",
    "sensor/kernels.hip",
    "sensor/scratch.hip",
    "scene/scene_host.cpp",
    "scene/scene.hip",
    "scene/learned.hip",
]

COMMON = [
    "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the bsx library needs ROCm's hipcc to build")


def _compile(src):
    out = os.path.join(OBJ, src.replace("/", "_") + ".o")
    cmd = [hipcc(), "-c", os.path.join(CSRC, src), "-o", out] + COMMON
    if src.endswith(".hip"):
        cmd += DEVICE + ["-x", "hip"]
    else:
        cmd += DEVICE + ["--offload-host-only"]  # host code only: the kernels live in the .hip files
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    jobs = min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
