"""Build libgeye_amd.so (hand-written HIP for gfx950 + C++ host runtime) in-tree.

Every source is compiled with hipcc for gfx950 only.  -ffp-contract=off keeps the
reference's operation order exact (no fused multiply-add unless the source writes one),
and packed FP32 instructions are disabled for device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import concurrent.futures as cf

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgeye_amd.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = ["dpm_kernels.hip", "dpm_plan.cpp", "dpm_runtime.cpp"]
EXTRA_SOURCES = ["dpm_mfma.hip", "dpm_nms.hip", "dpm_ascore.hip", "dpm_rects.hip", "dpm_serve.cpp",
                 "dpm_noalloc.hip"]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "--offload-arch=gfx950",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [hipcc()] + COMMON + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {src}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = SOURCES + [s for s in EXTRA_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    build_compat()
    if verbose:
        print("built", OUT)
    return OUT


COMPAT_OUT = os.path.join(HERE, "libgeye_dpm_compat.so")
ROOT = os.path.dirname(HERE)
COMPAT_DEMO = os.path.join(ROOT, "tests", "compat_detect.cpp")
COMPAT_DEMO_OUT = os.path.join(ROOT, "tests", "_bin", "compat_detect")


def build_compat() -> None:
    """The reference-named C++ layer (host code over the C ABI) and its test driver."""
    gxx = shutil.which("g++") or "g++"
    inc = os.path.join(ROOT, "include")
    cmd = [gxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-I", inc,
           os.path.join(CSRC, "dpm_compat.cpp"), "-L", HERE, "-lgeye_amd",
           "-Wl,-rpath,$ORIGIN", "-o", COMPAT_OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("compat library build failed")
    os.replace(COMPAT_OUT + ".tmp", COMPAT_OUT)
    if os.path.exists(COMPAT_DEMO):
        os.makedirs(os.path.dirname(COMPAT_DEMO_OUT), exist_ok=True)
        cmd = [gxx, "-O2", "-std=c++17", "-I", inc, COMPAT_DEMO, "-L", HERE,
               "-lgeye_dpm_compat", "-lgeye_amd", "-Wl,-rpath," + HERE,
               "-Wl,-rpath,$ORIGIN/../../geye_amd", "-o", COMPAT_DEMO_OUT]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("compat test driver build failed")


if __name__ == "__main__":
    build(verbose=True)
