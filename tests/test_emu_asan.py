"""The CPU emulation build under AddressSanitizer: the kernel source
(csrc/emu/emu_kernels.cpp includes gst_kernels.hip) and the host code are
compiled with -fsanitize=address into _build/libgst_emu_asan.so, and
tests/emu_asan_cases.py drives every solver path through it with libasan
preloaded. Emulated device buffers are heap blocks with redzones, so an
out-of-bounds access by a kernel or by the host code ends the run with
AddressSanitizer's report and its source line.

Recipe (what the test runs):
    g++ -O1 -g -fno-omit-frame-pointer -fsanitize=address -std=c++17 -fPIC -shared
        -DGST_CPU_EMU csrc/emu/emu_kernels.cpp csrc/gst_api.cpp -o _build/libgst_emu_asan.so
    GPU4GST_LIB=_build/libgst_emu_asan.so LD_PRELOAD=$(g++ -print-file-name=libasan.so)
        ASAN_OPTIONS=detect_leaks=0 python tests/emu_asan_cases.py
(leak checking is off: the Python interpreter itself is not built with the
sanitizer and keeps its allocations until exit).
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpu4gst-sigmod-fullexp_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "_build", "libgst_emu_asan.so")


def _libasan():
    r = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    if r.returncode != 0 or not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("libasan not available to g++")
    return path


def test_emulated_kernels_under_address_sanitizer():
    asan = _libasan()
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    cc = subprocess.run(["g++", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address",
                         "-std=c++17", "-fPIC", "-shared", "-DGST_CPU_EMU",
                         os.path.join(CSRC, "emu", "emu_kernels.cpp"),
                         os.path.join(CSRC, "gst_api.cpp"), "-o", tmp],
                        capture_output=True, text=True, timeout=600)
    assert cc.returncode == 0, cc.stderr[-3000:]
    os.replace(tmp, LIB)
    env = dict(os.environ, GPU4GST_LIB=LIB, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "emu_asan_cases.py")],
                       env=env, capture_output=True, text=True, timeout=900)
    assert "ERROR: AddressSanitizer" not in r.stderr, r.stderr[-6000:]
    assert r.returncode == 0, r.stderr[-3000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["cases"] == 94
    assert out["fails"] == [], out["fails"]
