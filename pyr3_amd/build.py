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
SOURCES = ["p3_api.cpp", "p3_batch.cpp", "p3_rtc.cpp", "p3_image.hip", "p3_denoise.hip",
           "p3_nn.hip"]
# Device headers embedded as text for the run-time chaos specialisation, in
# dependency order (their #include lines are dropped).
RTC_HEADERS = ["p3_args.h", "p3_device.h", "p3_variations.h", "p3_variations_ext.h",
               "p3_variation_dispatch.h", "p3_chaos_kernel.h"]
RTC_INC = CSRC / "p3_rtc_source.inc"
RTC_PRELUDE = ("typedef unsigned int uint32_t;\ntypedef int int32_t;\n"
               "typedef unsigned long long uint64_t;\n"
               "typedef unsigned short uint16_t;\ntypedef unsigned char uint8_t;\n"
               "typedef unsigned long size_t;\n")


def write_rtc_source() -> Path:
    """p3_rtc_source.inc: kRtcSource, the device headers as one string."""
    text = RTC_PRELUDE
    for h in RTC_HEADERS:
        for line in (CSRC / h).read_text().splitlines():
            if line.startswith("#include") or line.startswith("#pragma once"):
                continue
            text += line + "\n"
    chunks = [text[i:i + 16000] for i in range(0, len(text), 16000)]
    body = "\n".join('R"P3RTC(' + c + ')P3RTC"' for c in chunks)
    out = ("// Generated by pyr3_amd/build.py from " + ", ".join(RTC_HEADERS) + ".\n"
           "static const char kRtcSource[] =\n" + body + ";\n")
    RTC_INC.write_text(out)
    return RTC_INC


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return str(Path(rocm) / "bin" / "hipcc")


FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function"]


def compile_command(src: str, obj: Path):
    return [hipcc()] + FLAGS + ["-I", str(PKG.parent / "include"), "-c", "-o", str(obj),
                                str(CSRC / src)]


def link_command(objs, out: Path = LIB):
    return [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(out)] + \
        [str(o) for o in objs] + ["-lhiprtc"]


def build(verbose: bool = True) -> Path:
    """Compile every source to an object (in parallel), then link the
    shared library."""
    write_rtc_source()
    objdir = PKG / "_objs"
    objdir.mkdir(exist_ok=True)
    objs = [objdir / (Path(s).stem + ".o") for s in SOURCES]
    procs = []
    for s, o in zip(SOURCES, objs):
        cmd = compile_command(s, o)
        if verbose:
            print("[pyr3_amd.build]", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    failed = [s for s, p in zip(SOURCES, procs) if p.wait() != 0]
    if failed:
        raise RuntimeError(f"pyr3_amd.build: compile failed for {', '.join(failed)}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = link_command(objs, tmp)
    if verbose:
        print("[pyr3_amd.build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
