"""Builds libmptc_hip.so in-tree (mptc_amd/libmptc_hip.so) for gfx950.

Numerics flags: -ffp-contract=off (no fused multiply-add: every fp32 multiply
and add rounds separately, as in the reference's x86-64 SSE build),
correctly rounded fp32 division, no fast-math.  Packed-FP32 instruction
selection is disabled for all HIP code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libmptc_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "mptc_hip.hip")]
HEADERS = [os.path.join(HERE, "csrc", f) for f in ("dxt_device.h", "batch_kernels.h", "framegen.h",
                                                     "mlp_kernels.h", "net_infer.h", "prune_kernels.h")] + [
    os.path.join(ROOT, "include", "mptc.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


HOST_DIR = os.path.join(HERE, "host")
HOST_LIB = os.path.join(HOST_DIR, "libmptc_host.so")
CODEC_LIB = os.path.join(HOST_DIR, "libmptc_codec.so")
HOST_BIN = os.path.join(HOST_DIR, "encode_seq")
ENCODE_BIN = os.path.join(HOST_DIR, "mptc_encode")
CXX = os.environ.get("CXX", "g++")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"{cmd[0]} failed")


def build_codec(verbose: bool = False) -> str:
    """The CPU entropy stage and file format (mptc_amd/host/codec.cpp,
    include/mptc_codec.h): plain C++, no GPU dependency."""
    inc = os.path.join(ROOT, "include")
    tmp = f"{CODEC_LIB}.tmp{os.getpid()}"
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", inc,
          os.path.join(HOST_DIR, "codec.cpp"), "-o", tmp], verbose)
    os.replace(tmp, CODEC_LIB)
    return CODEC_LIB


def build_host(verbose: bool = False) -> str:
    """The C++ host layer (mptc_amd/host) over libmptc_hip.so, and its
    sequence driver; both find libmptc_hip.so through their run path."""
    build_codec(verbose)
    inc = os.path.join(ROOT, "include")
    tmp = f"{HOST_LIB}.tmp{os.getpid()}"
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", inc,
          os.path.join(HOST_DIR, "mptc_host.cpp"), os.path.join(HOST_DIR, "png.cpp"),
          "-L", HOST_DIR, "-lmptc_codec", "-L", HERE, "-lmptc_hip", "-lz",
          "-Wl,-rpath,$ORIGIN:$ORIGIN/..", "-o", tmp], verbose)
    os.replace(tmp, HOST_LIB)
    for src, out in (("encode_seq.cpp", HOST_BIN), ("mptc_encode.cpp", ENCODE_BIN)):
        tmpb = f"{out}.tmp{os.getpid()}"
        _run([CXX, "-O2", "-std=c++17", "-Wall", "-I", inc, os.path.join(HOST_DIR, src),
              "-L", HOST_DIR, "-lmptc_host", "-lmptc_codec", "-L", HERE, "-lmptc_hip",
              "-Wl,-rpath,$ORIGIN:$ORIGIN/..", "-o", tmpb], verbose)
        os.replace(tmpb, out)
    return HOST_LIB


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    tmp = f"{LIB}.tmp{os.getpid()}"
    cmd = [HIPCC, *FLAGS, *(extra or []), "-o", tmp, *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    if verbose and r.stderr.strip():
        # the host pass does not know the packed-fp32 target feature; that warning is expected
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l]
        if any(l.strip() for l in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
