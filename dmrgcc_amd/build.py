"""Build the native library ``dmrgcc_amd/libdmrgcc_amd.so`` for gfx950.

Every source under ``dmrgcc_amd/csrc`` is compiled by hipcc (device code for
gfx950 only) into ``dmrgcc_amd/_build`` and linked into one shared library
that exports the C ABI of ``include/dmrgcc_hip.h``.  Runs on a machine
without a GPU (hipcc cross-compiles).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdmrgcc_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["common.cpp", "kernels.hip", "model.cpp", "input.cpp", "dmrg.cpp", "capi.cpp"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-pthread",
    "-Wno-unused-result",
    # keep the compiler from forming packed FP32 ops on its own
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src):
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(SRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, msg))
    return obj, msg


ROOT = os.path.dirname(HERE)
DROPIN_SRC = os.path.join(ROOT, "tests", "cxx", "dropin_test.cpp")
DROPIN_LIB = os.path.join(ROOT, "tests", "cxx", "libdropin_test.so")


def build_dropin_test():
    """The test harness of the drop-in class headers (include/dmrgcc/):
    host C++ only, linked against the library built above."""
    if not os.path.exists(DROPIN_SRC):
        return None
    tmp = DROPIN_LIB + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-fPIC", "-shared", "-I" + ROOT, DROPIN_SRC, "-o", tmp,
           "-L" + HERE, "-ldmrgcc_amd", "-Wl,-rpath,$ORIGIN/../../dmrgcc_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("drop-in test harness failed to build:\n%s" % (r.stdout + r.stderr))
    os.replace(tmp, DROPIN_LIB)
    return DROPIN_LIB


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    if verbose:
        for obj, msg in results:
            if msg:
                print(msg)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-pthread", "-o", tmp] + [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % (r.stdout + r.stderr))
    os.replace(tmp, LIB)
    build_dropin_test()
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
