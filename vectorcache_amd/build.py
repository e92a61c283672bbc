"""Builds libvectorcache_hip.so in-tree for gfx950.

    python -m vectorcache_amd.build          # or vectorcache_amd.build.build()

* vc_kernels.hip, vc_ivf.hip, vc_probe.hip, vc_capi.cpp, vc_bucketed.cpp, vc_server.cpp, vc_multi.cpp,
  vc_prepared.cpp -> hipcc --offload-arch=gfx950
* vc_host.cpp, vc_landmark_host.cpp, vc_io.cpp -> g++ (host precompute, contraction off; file formats)

Floating-point flags are part of the numerics contract: -ffp-contract=off
everywhere (fused operations are explicit in the source), and packed FP32
instructions are disabled for device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "libvectorcache_hip.so"
BUILD = PKG / "_build"

ARCH = os.environ.get("VC_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}", f"-I{CSRC}", "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-mfma", "-ffp-contract=off", "-fno-fast-math", f"-I{CSRC}", "-Wall"]


def _run(cmd):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []
    for src in ("vc_kernels.hip", "vc_ivf.hip", "vc_probe.hip", "vc_capi.cpp", "vc_bucketed.cpp", "vc_server.cpp",
                "vc_multi.cpp", "vc_prepared.cpp"):
        obj = BUILD / (src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", CSRC / src, "-o", obj])
        objs.append(obj)
    for src in ("vc_host.cpp", "vc_landmark_host.cpp", "vc_io.cpp"):
        obj = BUILD / (src + ".o")
        _run(["g++", *HOST_FLAGS, "-c", CSRC / src, "-o", obj])
        objs.append(obj)
    tmp = OUT.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs, "-lrocblas"])
    tmp.replace(OUT)
    build_cpp_check()
    return OUT


CPP_CHECK_SRC = ROOT / "tests" / "cpp" / "test_turbo_quant_index.cpp"
CPP_CHECK_BIN = ROOT / "tests" / "cpp" / "test_turbo_quant_index"


def build_cpp_check() -> Path:
    """The C++ API check program (g++ against the C ABI only)."""
    _run(["g++", "-O2", "-std=c++20", f"-I{ROOT / 'include'}", f"-I{CSRC / 'include'}", CPP_CHECK_SRC,
          "-o", CPP_CHECK_BIN, f"-L{PKG}", "-lvectorcache_hip", "-Wl,-rpath,$ORIGIN/../../vectorcache_amd"])
    return CPP_CHECK_BIN


if __name__ == "__main__":
    build()
    sys.exit(0)
