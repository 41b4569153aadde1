"""Build recipe for the HIP library (gfx950) and the oracle.

    python -m harm_ramt_amd.build          # product library only
    python -m harm_ramt_amd.build --all    # plus oracle/liboracle.so

The product is one shared library, harm_ramt_amd/_lib/libharm_amd.so, holding
the HIP kernels, the C-ABI of include/harm_amd.h, the host-side geometry /
initial-state generators and the restart-file reader/writer.  It is compiled with -ffp-contract=off so that the
FP64 arithmetic keeps the reference's operation order (no fused multiply-add
contraction), which makes the GPU step bit-identical to the CPU step.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIBDIR, "libharm_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
C_FLAGS = ["-O2", "-fPIC", "-ffp-contract=off", "-std=c11", "-D_GNU_SOURCE"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib():
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    prob_o = os.path.join(objdir, "harm_problem.o")
    io_o = os.path.join(objdir, "harm_io.o")
    kern_o = os.path.join(objdir, "harm_kernels.o")
    _run(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, "harm_problem.c"), "-o", prob_o])
    _run(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, "harm_io.c"), "-o", io_o])
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "harm_kernels.hip"), "-o", kern_o])
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", kern_o, prob_o, io_o, "-o", tmp, "-lm"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-j4"])


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_lib()
    if "--all" in argv:
        build_oracle()


if __name__ == "__main__":
    main()
