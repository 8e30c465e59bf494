"""Compiles the package's native libraries in-tree.

libgst.so    HIP kernels + C ABI for gfx950 (hipcc), loaded by gpu4gst._lib
libgstgen.so seeded workload generators (host only, g++)
gst_run      the reference harnesses' command lines over libgst.so (host only, g++)
_build/adapter_demo_gpu  tests/cpp/adapter_demo.cpp (reference-shaped C++ adapters)
             linked against libgst.so, run by the GPU tests

Every call recompiles from source; objects go to _build/ (git-ignored).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "gpu4gst")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("GST_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
EMU_LIB = os.path.join(BUILD, "libgst_emu.so")
ADAPTER_DEMO = os.path.join(BUILD, "adapter_demo_gpu")

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")


def build(verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in ("gst_kernels.hip", "gst_api.cpp"):
        obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = os.path.join(OUT, "libgst.so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, os.path.join(OUT, "libgst.so"))
    # the same kernel source built for the CPU (csrc/emu/hip_cpu_emu.h): used by the
    # CPU test suite to check kernel logic against the oracle without a GPU
    tmp = os.path.join(BUILD, "libgst_emu.so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-DGST_CPU_EMU",
          "-x", "c++", os.path.join(CSRC, "gst_kernels.hip"),
          "-x", "c++", os.path.join(CSRC, "gst_api.cpp"), "-o", tmp], verbose)
    os.replace(tmp, EMU_LIB)
    INC = os.path.join(os.path.dirname(HERE), "include")
    tmp = os.path.join(OUT, "gst_run.tmp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INC, os.path.join(CSRC, "gst_cli.cpp"),
          os.path.join(OUT, "libgst.so"), "-Wl,-rpath,$ORIGIN", "-o", tmp], verbose)
    os.replace(tmp, os.path.join(OUT, "gst_run"))
    # the reference-shaped C++ adapters (include/gst_reference_adapters.hpp) driven by
    # tests/cpp/adapter_demo.cpp over the gfx950 library, for the GPU test suite
    tmp = os.path.join(BUILD, "adapter_demo_gpu.tmp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INC,
          os.path.join(os.path.dirname(HERE), "tests", "cpp", "adapter_demo.cpp"),
          os.path.join(OUT, "libgst.so"), "-Wl,-rpath,$ORIGIN/../gpu4gst", "-o", tmp], verbose)
    os.replace(tmp, ADAPTER_DEMO)
    tmp = os.path.join(OUT, "libgstgen.so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
          os.path.join(CSRC, "gen.cpp"), "-o", tmp], verbose)
    os.replace(tmp, os.path.join(OUT, "libgstgen.so"))


if __name__ == "__main__":
    build(verbose=True)
