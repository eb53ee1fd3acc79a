"""Build recipe for the in-tree native library and CLI.

Compiles the gfx950 kernels and host engine with hipcc, the plain host C++
with g++, and links ``fastmd_amd/libfastmd.so`` and ``fastmd_amd/bin/fastmd``.
Objects go to ``fastmd_amd/_build/``.  Used by ``__graft_entry__.build()``;
also runnable as ``python -m fastmd_amd.build``.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfastmd.so")
CLI = os.path.join(PKG, "bin", "fastmd")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["kernels.hip", "engine.hip", "server.hip", "capi.hip"]
CXX_SOURCES = ["host/topo.cpp"]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _obj(src):
    return os.path.join(OUT, src.replace("/", "_").rsplit(".", 1)[0] + ".o")


def build(verbose=False, jobs=8):
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    cmds = []
    for s in HIP_SOURCES:
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in CXX_SOURCES:
        cmds.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", _obj(s)])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [_obj(s) for s in HIP_SOURCES + CXX_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    _run(["g++"] + CXX_FLAGS + [os.path.join(CSRC, "fastmd_main.cpp"), "-o", CLI,
          f"-L{PKG}", "-lfastmd", "-Wl,-rpath,$ORIGIN/..",
          f"-L{os.path.join(ROCM, 'lib')}", "-lamdhip64", "-lpthread"])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
