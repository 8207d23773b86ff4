"""Build recipe for the native library (libfgx.so) and the fgx-run executable.

Compiles the HIP sources for gfx950 with hipcc and the host C++ with g++,
in-tree, so the built files sit next to the package.  Run as
``python -m cudaprotoplanetarydisksfloat_amd.build`` or call ``build()``.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfgx.so")
EXE = os.path.join(HERE, "fgx-run")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FGX_OFFLOAD_ARCH", "gfx950")
# packed FP32 instructions are kept out of every kernel
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wno-unused-function"]

HIP_SOURCES = ["engine.hip"]
CXX_SOURCES = ["params.cpp", "setup.cpp", "capi.cpp", "output.cpp", "sg.cpp", "server.cpp", "tune.cpp", "group.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    name = os.path.splitext(os.path.basename(src))[0]
    obj = os.path.join(OBJ, name + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
    else:
        cmd = ["g++"] + CXX_FLAGS + ["-c", src, "-o", obj]
    _run(cmd)
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES + CXX_SOURCES]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "-pthread", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    run_obj = _compile(os.path.join(CSRC, "fgx_run.cpp"))
    _run(["g++", "-o", EXE, run_obj, "-L" + HERE, "-lfgx", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print("built", LIB, EXE)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
