"""Builds the oracle (test infrastructure): oracle/lib/libfforacle.so.

The oracle is a plain-C restatement of the reference's integration path; only
the tests, the smoke check and the benchmark's CPU-baseline leg load it.
The reference itself is not built: its hot-path translation units need
OpenCV and Boost headers that this image does not have.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib")

CFLAGS = ["-O2", "-fPIC", "-shared", "-std=c11", "-frounding-math", "-ffp-contract=off",
          "-fno-fast-math", "-msse2", "-mno-fma", "-fno-math-errno"]


def build():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libfforacle.so")
    srcs = [os.path.join(HERE, f) for f in ("ff_oracle.c", "ff_oracle_mesh.c", "ff_oracle_loop.c") if os.path.exists(os.path.join(HERE, f))]
    cmd = ["gcc"] + CFLAGS + srcs + ["-o", out, "-lm"]
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


if __name__ == "__main__":
    build()
