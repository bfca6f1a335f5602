"""Builds libnemaktis_bpm.so (HIP kernels for gfx950 + C++ host code + C ABI)
in-tree with hipcc, and the FP64 CPU oracle used by the tests.

    python -m nemaktis_amd.build            # library + oracle
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libnemaktis_bpm.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NMK_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include")]
# device code: no packed FP32 instructions (compiler-generated v_pk_* FP32 ops are disabled)
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]

HIP_SOURCES = ["kernels.hip", "viewer_kernels.hip", "noalloc.hip"]
CXX_SOURCES = ["json.cpp", "expr.cpp", "settings.cpp", "vti.cpp", "staging.cpp", "engine.cpp", "server.cpp",
               "viewer.cpp", "capi.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_library(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        jobs.append([HIPCC] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", o])
    for src in CXX_SOURCES:
        # host code; hipcc still runs a device pass over the HIP headers: for gfx950 only, with the same flags
        o = os.path.join(OBJ, src + ".o")
        jobs.append([HIPCC] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", o])
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [os.path.join(OBJ, s + ".o") for s in HIP_SOURCES + CXX_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B"])
    return os.path.join(ROOT, "oracle", "liboracle_bpm.so")


def main():
    lib = build_library(verbose="-v" in sys.argv)
    print("built", lib)
    print("built", build_oracle())


if __name__ == "__main__":
    main()
