"""Build recipes: libopenmc.so (HIP for gfx950 + C++ host) and the oracle.

Both are built in-tree so the shared objects travel with the repository to
the GPU machine.  Compilation never contracts floating-point expressions
(-ffp-contract=off) on either side, which the bit-exact histories rely on.
"""
import concurrent.futures
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "openmc_amd", "csrc")
LIBDIR = os.path.join(ROOT, "openmc_amd", "lib")
LIB = os.path.join(LIBDIR, "libopenmc.so")
OBJDIR = os.path.join(ROOT, "_build", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]

SOURCES = ["xml.cpp", "host_model.cpp", "hdf5_io.cpp", "capi.cpp", "capi_model.cpp", "transport.hip", "xs_lookup.hip", "learned.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def _compile(src):
    os.makedirs(OBJDIR, exist_ok=True)
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC] + COMMON + DEVICE + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_library(jobs=None):
    """Compile every HIP/C++ source for gfx950 and link libopenmc.so."""
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    sys.path.insert(0, ROOT)
    from oracle import build as ob
    return ob.build()


def build_all():
    lib = build_library()
    ora = build_oracle()
    return lib, ora


if __name__ == "__main__":
    print(build_all())
