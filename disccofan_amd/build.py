"""Build the native libraries in-tree.

    python -m disccofan_amd.build            # libdcf.so (HIP, gfx950) + libdcf_synth.so (C)

libdcf.so is compiled with hipcc for gfx950 only, with FP contraction off (the FP64 attribute
formulas must round exactly like the reference's) and packed-FP32 instruction selection off.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]

SOURCES = ["dcf_build.hip", "dcf_flood.hip", "dcf_ops.hip", "dcf_tiles.hip", "dcf_api.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build_synth() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libdcf_synth.so")
    _run(["gcc", "-O2", "-fPIC", "-fopenmp", "-shared", "-std=gnu11", "-o", out,
          os.path.join(HERE, "synth", "dcf_synth.c"), "-lm"])
    return out


def build_dcf(jobs: int = 4) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    cmds = []
    for s in SOURCES:
        o = os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o")
        objs.append(o)
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, s), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(LIBDIR, "libdcf.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_synth()
    build_dcf()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(LIBDIR, "libdcf.so"), os.path.join(LIBDIR, "libdcf_synth.so"), file=sys.stderr)
