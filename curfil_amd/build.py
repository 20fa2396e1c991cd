"""Builds libcurfil_hip.so (HIP kernels + C ABI + host driver) for gfx950.

Every translation unit goes through hipcc; device code is compiled without
FP contraction and without packed-FP32 instructions (the training decisions
are bit-exact restatements of IEEE operation sequences).
"""
import concurrent.futures
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcurfil_hip.so")

SOURCES = [
    "k_integral.hip",
    "k_features.hip",
    "k_eval.hip",
    "k_bins.hip",
    "k_sort.hip",
    "k_partition.hip",
    "k_classify.hip",
    "predict_noalloc.hip",
    "capi.cpp",
    "forest_pack.cpp",
    "trainer.cpp",
    "tuner.cpp",
    "sampling.cpp",
    "synth.cpp",
    "server.cpp",
    "tree_json.cpp",
    "png_io.cpp",
    "dataset.cpp",
    "metrics.cpp",
    "tool_train.cpp",
    "tool_predict.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CURFIL_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wall",
    "-Wno-unused-result",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-flush-denormals-to-zero",
]


# A scalar-memory load whose immediate offset is not a multiple of 4 means the
# compiler built a byte-misaligned scalar base address; the scalar unit drops
# the low address bits, so such a load reads the wrong dwords.  Reject it.
_MISALIGNED_SMEM = re.compile(r"^\s*s_(?:load|buffer_load)_dword\w*\s+\S+,\s*s\[[0-9:]+\],\s*(-?0x[0-9a-f]+|-?[0-9]+)\s*$",
                              re.M)


def _check_device_asm(path):
    cmd = [HIPCC] + COMMON + DEVICE + ["--offload-device-only", "-S", "-x", "hip", path, "-o", "-"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"device assembly failed: {' '.join(cmd)}\n{r.stderr}")
    bad = [m.group(0).strip() for m in _MISALIGNED_SMEM.finditer(r.stdout) if int(m.group(1), 0) % 4 != 0]
    if bad:
        raise RuntimeError(f"{os.path.basename(path)}: misaligned scalar loads generated: {bad[:4]}")


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if src.endswith(".hip"):
        _check_device_asm(path)
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for o, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_tools()
    return LIB


TOOLS = ["curfil_train", "curfil_predict"]
BIN = os.path.join(HERE, "bin")


def build_tools():
    """The command-line tools (host C++ over the C ABI), linked against the library next to them."""
    os.makedirs(BIN, exist_ok=True)
    include = os.path.join(os.path.dirname(HERE), "include")
    for t in TOOLS:
        out = os.path.join(BIN, t)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", include, os.path.join(CSRC, "tools", t + ".cpp"), "-o",
               out + ".tmp", "-L", HERE, "-lcurfil_hip", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"tool build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    print(build(verbose=True))
