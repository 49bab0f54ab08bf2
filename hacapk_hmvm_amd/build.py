"""Builds the native libraries in-tree.

    python hacapk_hmvm_amd/build.py

Outputs (git-ignored, shipped with the tree):
    hacapk_hmvm_amd/libhmvm_hip.so   loaders/writers, generator, schedule, HIP kernels, C ABI
    oracle/liboracle.so              CPU oracle (test infrastructure, built by oracle/Makefile)
Device code is compiled for gfx950 only.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hacapk_hmvm_amd")
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(ROOT, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
LIB = os.path.join(PKG, "libhmvm_hip.so")

CXXFLAGS = ["-O3", "-fPIC", "-fopenmp", "-std=c++17", "-Wall", "-I" + INC]
HIPFLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950", "-I" + INC, "-I" + CSRC,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]

CPP_SOURCES = ["hacapk_io.cpp", "hacapk_io64.cpp", "hgen.cpp", "hmvm_plan.cpp", "hmvm_cpp_api.cpp"]
HIP_SOURCES = ["hmvm_runtime.hip", "hmvm_krylov.hip", "hmvm_mlp.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    base = os.path.splitext(os.path.basename(src))[0]
    obj = os.path.join(OBJ, base + ".o")
    if src.endswith(".hip"):
        cmd = [os.path.join(ROCM, "bin", "hipcc"), *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = ["g++", *CXXFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    _run(cmd)
    return obj


def build_library():
    os.makedirs(OBJ, exist_ok=True)
    srcs = CPP_SOURCES + HIP_SOURCES
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp, *objs, "-fopenmp", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
          "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-Wl,--no-undefined"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir("/root/reference/src_hmvm"):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build():
    build_library()
    build_oracle()


if __name__ == "__main__":
    build()
    print("built", LIB)
