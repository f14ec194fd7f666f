"""Builds fvstuff_amd/libdgsem.so for gfx950 with hipcc (in-tree).

One object per polynomial degree (dgsem_inst.hip with -DDG_P=p) plus the C-ABI
unit, compiled in parallel, linked into one shared library.  Then the C++
host driver fvstuff_amd/driver/dgsem_run (plain g++ against include/dgsem.h,
linked to libdgsem.so and the HIP runtime).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdgsem.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ROOT = os.path.dirname(HERE)
DRIVER_SRC = os.path.join(HERE, "driver", "dgsem_run.cpp")
DRIVER = os.path.join(HERE, "driver", "dgsem_run")
ARCH = "gfx950"
DEGREES = range(1, 8)

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a*b+c is fused only where it is written as one expression, so every
    # inlined copy of a helper rounds alike (a node converted in two places
    # gives the same bits, which keeps results independent of the tiling)
    "-ffp-contract=on",
    # LLVM's memory-clause scheduling strategy: the shock-capturing FP32
    # kernel at p = 3 takes 96 instead of 122 VGPRs (5 waves per SIMD instead
    # of 4) and runs 1.1-1.7 % faster (KH 1024^2 mixed); vortex p = 7, the
    # batched ensemble and the SD sibling within +-0.5 % (FP64 SD 1 % faster);
    # the same arithmetic, so the same bits.  (max-ilp: FP64 KH 21 % slower.)
    "-mllvm", "-amdgpu-sched-strategy=max-memory-clause",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    units = [(os.path.join(CSRC, "dgsem_inst.hip"), os.path.join(BUILD, f"dgsem_p{p}.o"), [f"-DDG_P={p}"])
             for p in DEGREES]
    for real in ("float", "double"):
        units.append((os.path.join(CSRC, "dgsem_batch.hip"), os.path.join(BUILD, f"dgsem_batch_{real}.o"),
                      [f"-DDG_REAL={real}"]))
    units.append((os.path.join(CSRC, "dgsem_lib.hip"), os.path.join(BUILD, "dgsem_lib.o"), []))
    cmds = [[HIPCC, *FLAGS, *extra, "-c", src, "-o", obj] for src, obj, extra in units]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(jobs) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + [u[1] for u in units])
    os.replace(tmp, LIB)
    build_driver()
    return LIB


def build_driver():
    """The C++ host driver: a plain host compiler, the ABI header, the
    library (found next to the binary's parent folder at run time) and the
    HIP runtime."""
    tmp = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
          "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROCM, "include"),
          DRIVER_SRC, "-o", tmp, "-L", HERE, "-ldgsem", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
          "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
