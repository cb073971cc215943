"""Build recipe for libewfe.so (gfx950). Called by __graft_entry__.build() and usable directly:

    python -m edelweissfe_amd.build
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libewfe.so")
SOURCES = ["ewfe_kernels.hip", "ewfe_solver.hip", "ewfe_amg.hip", "ewfe_plan.cpp", "ewfe_amg_host.cpp"]
ARCH = os.environ.get("EWFE_OFFLOAD_ARCH", "gfx950")


def hipcc_cmd(out):
    return [
        "hipcc", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        # no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        *[os.path.join(CSRC, s) for s in SOURCES],
        "-o", out, "-lpthread",
    ]


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = hipcc_cmd(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
