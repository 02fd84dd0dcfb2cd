"""Compile the HIP library ``libmantlepar_amd.so`` for gfx950 in-tree.

Usage: ``python -m mantlepar_amd.build`` (or ``build()`` from Python).  Every
source under ``csrc/`` is compiled with hipcc (device code for gfx950 only)
and linked into ``mantlepar_amd/libmantlepar_amd.so``.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmantlepar_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    # keep hipcc from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(line for line in r.stderr.splitlines()
                    if "packed-fp32-ops" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{err}")
    return obj, err


CPP_CHECK_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "facade_check.cpp")
CPP_CHECK = os.path.join(os.path.dirname(HERE), "tests", "cpp", "facade_check")


def build_cpp_check():
    """The compiled C++ caller of include/mantlepar_amd.hpp (host code only:
    g++ against the library, no device code)."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra",
           "-I", os.path.join(os.path.dirname(HERE), "include"), CPP_CHECK_SRC,
           "-L", HERE, "-lmantlepar_amd", "-Wl,-rpath,$ORIGIN/../../mantlepar_amd",
           "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib", "-o", CPP_CHECK + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on {CPP_CHECK_SRC}:\n{r.stderr}")
    os.replace(CPP_CHECK + ".tmp", CPP_CHECK)
    return CPP_CHECK


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    objs = []
    for obj, err in results:
        objs.append(obj)
        if verbose and err.strip():
            print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cpp_check()
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
