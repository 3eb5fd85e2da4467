"""Build the native library medusa_amd/lib/libmds.so for gfx950.

  host C   medusa_amd/csrc/host/*.c   gcc, strict IEEE (no FMA contraction)
  HIP      medusa_amd/csrc/hip/*.hip  hipcc --offload-arch=gfx950
  driver   medusa_amd/csrc/driver/mds_driver.c -> medusa_amd/lib/mds_driver, a C
           program in the shape of the reference main loop, linked to libmds.so

FP64 kernels are compiled with -ffp-contract=off so they round exactly like
the reference CPU build; the FP32-bulk functions re-enable contraction locally
with `#pragma clang fp contract(fast)`.  Packed FP32 instructions are disabled
(-packed-fp32-ops).

Usage:  python -m medusa_amd.build   (or medusa_amd.build.build())
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmds.so")
ARCH = os.environ.get("MDS_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CC = shutil.which("gcc") or "cc"

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-result", "-Wno-unused-value"]
C_FLAGS = ["-O2", "-std=c99", "-fPIC", "-fno-fast-math", "-ffp-contract=off", "-fopenmp", "-Wall",
           "-Wno-unused-function"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def sources():
    host = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host")) if f.endswith(".c"))
    hip = sorted(os.path.join(CSRC, "hip", f) for f in os.listdir(os.path.join(CSRC, "hip")) if f.endswith(".hip"))
    return host, hip


def build(verbose=True, jobs=8):
    """Compile every source (always) and link libmds.so."""
    host, hip = sources()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    procs = []
    for src in host:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        cmd = [CC] + C_FLAGS + ["-c", src, "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for src in hip:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [c for c, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError("compile failed: " + " ".join(failed[0]))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lgomp", "-lm"], verbose)
    os.replace(tmp, LIB)
    # the C host program in the shape of the reference main loop, on the C-ABI only
    drv = os.path.join(CSRC, "driver", "mds_driver.c")
    exe = os.path.join(LIBDIR, "mds_driver")
    _run([CC, "-O2", "-std=c99", "-Wall", drv, "-o", exe + ".tmp", f"-L{LIBDIR}", "-lmds", "-lm",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
