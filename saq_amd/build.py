"""Builds the gfx950 library saq_amd/libsaqgpu.so from saq_amd/csrc with hipcc.

Every HIP translation unit is compiled for gfx950 only, with FP contraction off
(the reference builds with -ffp-contract=off and the kernels restate its rounding)
and packed-FP32 instruction selection disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsaqgpu.so")
SOURCES = ["api.cpp", "plan.cpp", "query_prep.hip", "probe.hip", "encode.hip", "slots.hip", "batched.hip", "estimate.hip", "slice_scan.hip", "single.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-I" + os.path.join(os.path.dirname(HERE), "include")]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC] + FLAGS + lang + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{errs}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-L/opt/rocm/lib", "-lrocblas"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_host_tools(verbose)
    return LIB


def build_host_tools(verbose=False):
    """Host programs over the C ABI (saq_amd/host): plain g++, linked against libsaqgpu.so."""
    host = os.path.join(HERE, "host")
    inc = os.path.join(os.path.dirname(HERE), "include")
    for src in ("ivf_run.cpp",):
        out = os.path.join(host, src[:-4])
        cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-I" + inc, "-I" + host,
               os.path.join(host, src), "-o", out + ".tmp", "-L" + HERE, "-lsaqgpu", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed on {src}:\n{r.stderr}")
        os.replace(out + ".tmp", out)
        if verbose:
            print("built", out)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
