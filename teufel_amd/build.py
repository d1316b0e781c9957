"""Build recipe for the in-tree HIP library teufel_amd/lib/libteufel_gpu.so.

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 instructions are
disabled (-packed-fp32-ops): the kernels are written for scalar FP32/FP64 VALU.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libteufel_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TG_OFFLOAD_ARCH", "gfx950")

SOURCES = [os.path.join(HERE, "csrc", "tg_api.hip")]
DEPS = SOURCES + [os.path.join(HERE, "csrc", f) for f in ("lw_kernels.hip", "vay_kernel.hip", "tg_device.h")] + [
    os.path.join(ROOT, "include", "teufel_gpu.h")]

# -amdgpu-sched-strategy=max-ilp: the machine scheduler favours instruction-
# level parallelism over register pressure; the radiation kernels are bound
# by VALU issue and dependent chains, not by registers (1-2 % faster on c2).
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-amdgpu-sched-strategy=max-ilp",
         "-Wall", "-Wno-unused-command-line-argument"]


LIBS = ["-lrocfft"]   # per-pixel spectra (tg_screen_spectra)


def command(out=LIB_PATH):
    return [HIPCC] + FLAGS + ["-o", out] + SOURCES + LIBS


def build(verbose=False):
    """Compile the library (always recompiles)."""
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + err)
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
