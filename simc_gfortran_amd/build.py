"""Build libsimc_gpu.so for gfx950 (hipcc), in-tree.

    python -m simc_gfortran_amd.build

Steps: generate the COSY polynomial code from data/maps (gen_cosy.py), then
compile the HIP kernels, the output formats and the host map reader into one
shared library;
build the per-event Fortran-ABI drop-in libsimc_dropin.so on top of it; and,
when a Fortran compiler is present, compile the Fortran module simc_gpu and
the Fortran test programs of tests/fortran against both libraries.
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
DROPIN = os.path.join(PKG, "libsimc_dropin.so")
FORTRAN = os.environ.get("SIMC_FC", "/opt/rocm/lib/llvm/bin/flang")
FBUILD = os.path.join(PKG, "fortran", "_build")
FBIN = os.path.join(REPO, "tests", "fortran", "_bin")
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
    gen_hut(verbose)
    srcs = [os.path.join(CSRC, "simc_kernels.hip"), os.path.join(CSRC, "simc_io.hip"),
            os.path.join(CSRC, "simc_maps.cpp")]
    cmd = [HIPCC, *HIP_FLAGS, "-I", GEN, "-I", CSRC, "-I", os.path.join(REPO, "include"),
           *srcs, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    build_dropin(verbose)
    build_fortran(verbose)
    return LIB


def gen_hut(verbose: bool = True) -> str:
    """_gen/hut_gen.inc: the detector huts of the step programs as
    straight-line device code (gen_hut.cpp, built and run on the host)."""
    exe = os.path.join(GEN, "gen_hut")
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-I", os.path.join(REPO, "include"),
           os.path.join(CSRC, "gen_hut.cpp"), "-o", exe]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    out = os.path.join(GEN, "hut_gen.inc")
    subprocess.run([exe, out], check=True)
    return out


def build_dropin(verbose: bool = True) -> str:
    """libsimc_dropin.so: mc_hms_/mc_shms_/mc_*_recon_ on top of libsimc_gpu.so."""
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", os.path.join(CSRC, "simc_dropin.cpp"),
           "-I", os.path.join(REPO, "include"), "-L", PKG, "-lsimc_gpu", "-Wl,-rpath,$ORIGIN",
           "-o", DROPIN + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(DROPIN + ".tmp", DROPIN)
    return DROPIN


def build_fortran(verbose: bool = True) -> bool:
    """The module simc_gpu and the Fortran test programs (skipped without a
    Fortran compiler)."""
    if not os.path.exists(FORTRAN):
        if verbose:
            print("no Fortran compiler at", FORTRAN, "- Fortran module and programs not built")
        return False
    os.makedirs(FBUILD, exist_ok=True)
    os.makedirs(FBIN, exist_ok=True)
    mod_src = os.path.join(PKG, "fortran", "simc_gpu.f90")
    mod_obj = os.path.join(FBUILD, "simc_gpu.o")
    cmds = [[FORTRAN, "-O2", "-c", mod_src, "-module-dir", FBUILD, "-o", mod_obj]]
    tdir = os.path.join(REPO, "tests", "fortran")
    rpath = "-Wl,-rpath,$ORIGIN/../../../simc_gfortran_amd"
    cmds.append([FORTRAN, "-O2", "-I", FBUILD, os.path.join(tdir, "batch_host.f90"), mod_obj,
                 "-L", PKG, "-lsimc_gpu", rpath, "-o", os.path.join(FBIN, "batch_host")])
    cmds.append([FORTRAN, "-O2", "-I", FBUILD, os.path.join(tdir, "gen_events.f90"), mod_obj,
                 "-L", PKG, "-lsimc_gpu", rpath, "-o", os.path.join(FBIN, "gen_events")])
    cmds.append([FORTRAN, "-O2", os.path.join(tdir, "dropin_events.f90"),
                 "-L", PKG, "-lsimc_dropin", "-lsimc_gpu", rpath, "-o", os.path.join(FBIN, "dropin_events")])
    cmds.append([FORTRAN, "-O2", os.path.join(tdir, "field_events.f90"),
                 "-L", PKG, "-lsimc_dropin", "-lsimc_gpu", rpath, "-o", os.path.join(FBIN, "field_events")])
    for c in cmds:
        if verbose:
            print(" ".join(c))
        subprocess.run(c, check=True)
    return True


if __name__ == "__main__":
    build()
