"""Builds the in-tree shared libraries of fastfusion_amd.

libffgpu.so   -- the HIP fusion library (gfx950), C ABI in include/ff_abi.h
libffsynth.so -- host-only synthetic RGB-D scene generator

Run as ``python -m fastfusion_amd.build`` or through ``__graft_entry__.build()``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-Wno-unused-value", "-Wno-unused-result",
    # packed FP32 VALU ops stay off (plain VALU ops only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HIP_SOURCES = ["ff_fusion.hip", "ff_mesh.hip", "ff_learned.hip", "ff_batch.hip", "ff_mapio.hip", "ff_loop.hip"]
HOST_SOURCES = ["ff_input.cpp"]  # host-only parts of libffgpu.so (zlib for PNG)


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_gpu():
    os.makedirs(LIB, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(LIB, src.replace(".hip", ".o"))
        _run(["hipcc"] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(LIB, src.replace(".cpp", ".o"))
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-I", os.path.join(os.path.dirname(HERE), "include"), "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out = os.path.join(LIB, "libffgpu.so")
    _run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out] + objs + ["-lz"])
    return out


def build_synth():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libffsynth.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", os.path.join(CSRC, "synth.c"), "-o", out, "-lm"])
    return out


def build_all():
    build_synth()
    build_gpu()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
