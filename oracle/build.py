"""Build recipe of the oracle (test infrastructure): oracle/_build/liboracle.so.

Compiled with g++ -O3 -ffp-contract=off (no -march=native, no fast-math),
OpenMP for the CPU baseline timings.  A second build with glibc's log/sin/cos
(-DORACLE_GLIBC_MATH) measures how often the deterministic libm and glibc
lead to different histories.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "liboracle.so")
LIB_GLIBC = os.path.join(OUT, "liboracle_glibc.so")


def _cc(out, extra):
    os.makedirs(OUT, exist_ok=True)
    cmd = ["g++", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
           "-shared", "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "oracle.cpp"),
           "-o", out + ".tmp"] + extra
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    os.replace(out + ".tmp", out)
    return out


def build():
    return _cc(LIB, []), _cc(LIB_GLIBC, ["-DORACLE_GLIBC_MATH"])


if __name__ == "__main__":
    print(build())
