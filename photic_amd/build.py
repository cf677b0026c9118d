"""Build recipe for libphotic.so (gfx950 kernels + C host code).

Compiles every HIP source for gfx950 with hipcc, the C host code with gcc against the HIP
runtime headers, and links one shared library in-tree: photic_amd/libphotic.so.  The FP64
path depends on -ffp-contract=off (no fused multiply-add the reference does not make); all HIP
code is built without packed FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libphotic.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIPCC_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + CSRC, "-I" + os.path.join(ROOT, "include"), "-Wno-unused-result",
]
CC_FLAGS = [
    "-std=c11", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-D__HIP_PLATFORM_AMD__",
    "-fopenmp",  # the drop-in's host-side gather/scatter loops (libgomp)
    "-I" + os.path.join(ROCM, "include"), "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
]
HIP_SOURCES = ["ph_kernels.hip", "ph_mlp.hip", "ph_stack.hip", "ph_stack_native.hip",
               "ph_region.hip", "ph_order.hip"]
C_SOURCES = ["photic_capi.c", "photic_samodel.c", "photic_io.c", "photic_order.c"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:  # the device compiles are independent: run them side by side
        obj = os.path.join(BUILD, src + ".o")
        cmd = [os.path.join(ROCM, "bin", "hipcc")] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src),
                                                                  "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, pr in procs if pr.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, "hipcc " + " ".join(failed))
    for src in C_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run(["gcc"] + CC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=" + ARCH, "-shared", "-fPIC",
          "-o", tmp] + objs + ["-lgomp", "-lm"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
