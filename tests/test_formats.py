"""Output formats on the CPU: the .bin ntuple written by the library equals,
byte for byte, the file the reference's own NtupleInit / NtupleClose (and
records written like results_ntu_write) produce for the same mode and values
(oracle/_ref), for every analysis mode; the reader inverts the writer."""
import os
import shutil
import tempfile

import numpy as np
import pytest

from simc_gfortran_amd import api

CASES = [("elastic", False, 46), ("pion", False, 53), ("pion", True, 61), ("semi", False, 56),
         ("semi", True, 64)]


@pytest.fixture
def short_dir():
    d = tempfile.mkdtemp(prefix="nt", dir="/tmp")  # NtupleInit takes a CHARACTER*80 name
    yield d
    shutil.rmtree(d)


@pytest.mark.parametrize("mode,tgt,ncols", CASES, ids=[f"{m}{'_tgtfield' if t else ''}" for m, t, _ in CASES])
def test_ntuple_matches_reference_writer(oracle, short_dir, mode, tgt, ncols):
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built")
    tags = api.ntuple_tags(mode, tgt_field=tgt)
    assert len(tags) == ncols  # the column numbers NtupleInit.f comments give
    rng = np.random.default_rng(ncols)
    rows = rng.normal(size=(7, ncols)) * 10.0 ** rng.integers(-5, 5, size=(7, ncols))
    rows[0, :3] = [0.0, -0.0, np.inf]
    ref = os.path.join(short_dir, "r.bin")
    oracle.ref_ntuple(ref, mode, tgt, rows)
    ours = os.path.join(short_dir, "o.bin")
    with api.Ntuple(ours, tags) as nt:
        nt.write(rows[:3])
        nt.write(rows[3:])
    assert open(ours, "rb").read() == open(ref, "rb").read()
    t, v = api.read_ntuple(ours)
    assert t == tags
    np.testing.assert_array_equal(v, rows)


def test_ntuple_rejects_bad_shapes(short_dir):
    with pytest.raises(api.SimcError):
        api.Ntuple(os.path.join(short_dir, "x.bin"), ["a"] * 81)
    with api.Ntuple(os.path.join(short_dir, "y.bin"), ["a", "b"]) as nt:
        with pytest.raises(ValueError):
            nt.write(np.zeros((3, 3)))


@pytest.mark.parametrize("case", ["elastic", "pion_tgtfield"])
def test_ntuple_golden_files(short_dir, case):
    """The committed reference-written files (tests/golden/ntuple/) are what
    the host writer produces for the same selected rows."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "ntuple", case + ".npz"))
    tags = api.ntuple_tags(str(g["mode"]), tgt_field=bool(g["tgt_field"]))
    ours = os.path.join(short_dir, "o.bin")
    with api.Ntuple(ours, tags) as nt:
        nt.write(g["rows"][g["mask"] == 1])
    assert open(ours, "rb").read() == g["ref_bin"].tobytes()
