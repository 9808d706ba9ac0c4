"""Builds the gfx950 library and the drop-in executables in-tree:

  crispy_amd/lib/libcrispy_hip.so   HIP kernels + the C ABI (include/crispy_hip.h),
                                    including the job server
  bin/preprocess, bin/kmerDist,     the reference's command lines over that ABI
  bin/genDist

Usage: python -m crispy_amd.build   (or crispy_amd.build.build()).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "lib", "obj")
BINDIR = os.path.join(ROOT, "bin")
LIB = os.path.join(LIBDIR, "libcrispy_hip.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["api.hip", "kmer.hip", "nw.hip", "preprocess.hip", "nw_fam_reg0.hip", "nw_fam_reg1.hip",
               "nw_fam_reg2.hip", "nw_fam_reg3.hip", "nw_fam_sliced24.hip",
               "nw_fam_sliced32.hip"]
HOST_SOURCES = ["host.cpp", "serve.cpp"]
CLI = {"kmerDist": "cli_kmer.cpp", "genDist": "cli_gen.cpp", "preprocess": "cli_preprocess.cpp"}

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # fp32 '/' must round correctly: distances are compared bit for bit
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # no packed fp32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose=False, jobs=None, save_temps=False):
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJDIR, src + ".o")
        extra = ["--save-temps"] if save_temps else []
        cmds.append([HIPCC, *HIP_FLAGS, *extra, "-I", INC, "-I", CSRC, "-c",
                     os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(OBJDIR, src + ".o")
        cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-I", INC, "-c",
                     os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-pthread", f"--offload-arch={ARCH}", "-o", tmp, *objs,
          f"-Wl,-rpath,{ROCM}/lib"], verbose)
    os.replace(tmp, LIB)
    for exe, src in CLI.items():
        out = os.path.join(BINDIR, exe)
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INC, os.path.join(CSRC, src),
              "-o", out + ".tmp", "-L", LIBDIR, "-lcrispy_hip",
              "-Wl,-rpath,$ORIGIN/../crispy_amd/lib", f"-Wl,-rpath,{ROCM}/lib"], verbose)
        os.replace(out + ".tmp", out)
    return LIB


def clean():
    shutil.rmtree(LIBDIR, ignore_errors=True)
    for exe in CLI:
        p = os.path.join(BINDIR, exe)
        if os.path.exists(p):
            os.remove(p)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv, save_temps="--save-temps" in sys.argv)
    print(LIB)
