"""The counter-based stream with Box-Muller normals against the reference's
RANLUX stream with polar-method normals, on the CPU oracle: two independent
samples of the same input distribution, histograms of every focal-plane and
target variable statistically equal.  A planted change of the wire
resolution must be detected, which shows the test's power."""
import numpy as np
import pytest

from simc_gfortran_amd import workloads as W
from tests._stats import compare

N = 60000


@pytest.mark.parametrize("wl", [W.HMS_C0, W.SHMS_C1], ids=["hms", "shms"])
def test_philox_box_muller_matches_ranlux_polar(oracle, wl):
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, wire_sigma=wl.wire_sigma, fit_f32=False)
    a = oracle.run(wl.arm, W.generate(wl, N, seed=31), rng="philox", seed=7, **kw)
    b = oracle.run(wl.arm, W.generate(wl, N, seed=131), rng="ranlux", seed=314159, nthreads=1, **kw)
    ok, rep = compare(a, b, N)
    print(wl.name, rep)
    assert ok, rep


def test_statistics_detect_a_planted_resolution_change(oracle):
    wl = W.SHMS_C1
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, fit_f32=False)
    a = oracle.run(wl.arm, W.generate(wl, N, seed=32), rng="philox", seed=8, wire_sigma=0.030, **kw)
    b = oracle.run(wl.arm, W.generate(wl, N, seed=132), rng="ranlux", seed=271828, nthreads=1,
                   wire_sigma=0.060, **kw)
    ok, rep = compare(a, b, N)
    print(rep)
    assert not ok, rep
