"""Build recipe for libpyr3hip.so (gfx950) — run on the build host, never on import.

One hipcc invocation compiles the kernels and the C ABI into an in-tree
shared library next to this file. Flags:
  --offload-arch=gfx950              MI355X only
  -ffp-contract=off                  a*b+c stays two roundings (reference op order)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                     no packed-FP32 VALU forms
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libpyr3hip.so"
SOURCES = ["p3_api.cpp", "p3_chaos.hip", "p3_image.hip"]
HEADERS = ["p3_device.h", "p3_launch.h", "p3_variations.h"]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return str(Path(rocm) / "bin" / "hipcc")


def command(out: Path = LIB):
    return [
        hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
        "-I", str(PKG.parent / "include"),
        "-o", str(out),
    ] + [str(CSRC / s) for s in SOURCES]


def build(verbose: bool = True) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print("[pyr3_amd.build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
