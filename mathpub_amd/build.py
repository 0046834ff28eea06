"""Builds the gfx950 HIP engine (libwf.so) and the C host program (wf) in-tree.

hipcc cross-compiles for gfx950 without a GPU, so this runs on any build machine.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libwf.so")
HOST = os.path.join(PKG, "wf")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
    # packed FP32 instructions are switched off for every HIP file of this project
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib():
    src = os.path.join(PKG, "csrc", "wf_engine.hip")
    tmp = LIB + ".tmp"
    run([HIPCC] + HIP_FLAGS + ["-I" + INC, "-I" + os.path.join(PKG, "csrc"), "-o", tmp, src])
    os.replace(tmp, LIB)


def build_probe():
    src = os.path.join(PKG, "csrc", "wf_probe.hip")
    flags = [f for f in HIP_FLAGS if f not in ("-shared", "-fPIC")]
    run([HIPCC] + flags + ["-o", os.path.join(PKG, "wf_probe"), src])


HOST_SRC = [os.path.join(PKG, "host", f) for f in ("wf_main.c", "wf_text.c")]


def build_host(out=HOST, extra=(), rpath="$ORIGIN"):
    """The wf command (C, linked against libwf.so, found through rpath); extra = additional
    gcc flags (the sanitizer test builds a copy with -fsanitize=address,undefined)."""
    run(["gcc", "-O2", "-std=gnu11", "-Wall", "-Wextra", "-pthread", "-I" + INC] + list(extra) + ["-o", out] + HOST_SRC +
        ["-L" + PKG, "-lwf", "-Wl,-rpath," + rpath, "-lm"])


def main():
    build_lib()
    build_host()
    build_probe()


if __name__ == "__main__":
    sys.exit(main())
