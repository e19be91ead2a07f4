"""The Fortran side of the boundary, without a GPU: the module simc_gpu binds
every entry point of include/simc_gpu.h and compiles; the per-event drop-in
library exports the reference's subroutine symbols (gfortran/flang ABI)."""
import ctypes
import os
import re
import subprocess

import pytest

from simc_gfortran_amd import build as B
from tests.test_library import declared_symbols

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODULE = os.path.join(REPO, "simc_gfortran_amd", "fortran", "simc_gpu.f90")


def test_module_binds_every_entry_point():
    src = open(MODULE).read()
    bound = set(re.findall(r'bind\(C,\s*name="(simc_\w+)"\)', src))
    assert bound == set(declared_symbols())


def test_module_compiles(tmp_path):
    if not os.path.exists(B.FORTRAN):
        pytest.skip("no Fortran compiler")
    subprocess.run([B.FORTRAN, "-c", MODULE, "-module-dir", str(tmp_path), "-o", str(tmp_path / "m.o")],
                   check=True)


def test_dropin_exports_reference_symbols():
    lib = ctypes.CDLL(B.DROPIN)
    for name in ("mc_hms_", "mc_shms_", "mc_sos_", "mc_hrsr_", "mc_hrsl_", "mc_hms_recon_", "mc_shms_recon_",
                 "mc_sos_recon_", "mc_hrsr_recon_", "mc_hrsl_recon_", "trginit_", "track_from_tgt_",
                 "track_to_tgt_", "simc_dropin_seed_", "simc_dropin_close_"):
        assert hasattr(lib, name), name
