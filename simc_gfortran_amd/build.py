"""Build libsimc_gpu.so for gfx950 (hipcc), in-tree.

    python -m simc_gfortran_amd.build

Steps: generate the COSY polynomial code from data/maps (gen_cosy.py), then
compile the HIP kernels and the host map reader into one shared library.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
GEN = os.path.join(CSRC, "_gen")
LIB = os.path.join(PKG, "libsimc_gpu.so")
MAP_DIR = os.path.join(REPO, "data", "maps")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SIMC_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: only the explicit FMAs of the generated polynomials fuse,
# so every other expression rounds exactly as the reference's does.
# -packed-fp32-ops off: no v_pk_*_f32 generation.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def build(verbose: bool = True) -> str:
    sys.path.insert(0, CSRC)
    import gen_cosy  # noqa: E402
    os.makedirs(GEN, exist_ok=True)
    summary = gen_cosy.generate(MAP_DIR, os.path.join(GEN, "cosy_gen"))
    if verbose:
        print("gen_cosy:", summary)
    srcs = [os.path.join(CSRC, "simc_kernels.hip"), os.path.join(CSRC, "simc_maps.cpp")]
    cmd = [HIPCC, *HIP_FLAGS, "-I", GEN, "-I", CSRC, "-I", os.path.join(REPO, "include"),
           *srcs, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build()
