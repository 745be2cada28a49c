"""Build recipe for the oracle (test infrastructure only).

Compiles oracle/rama_oracle.cpp with g++ into oracle/_build/librama_oracle.so.
-ffp-contract=off keeps float expressions unfused, matching the HIP library,
which is compiled with the same flag.

The reference itself (oracle/_ref) is not built: its hot path is CUDA C++ that
needs nvcc, Thrust/CUB fetched over the network (CCCL v2.8.3) and CLI11, none
of which exist in this image.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "rama_oracle.cpp")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "librama_oracle.so")


def build(force: bool = True) -> str:
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= os.path.getmtime(SRC):
        return LIB
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp%d" % os.getpid()
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-o", tmp, SRC]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv or True))
