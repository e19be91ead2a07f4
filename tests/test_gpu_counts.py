"""The algorithmic work counts the measurement divides by (simc_stop_flops*,
simc_stop_flops_hut), restated here from the map files and the counting
rules, not read back from the library.

Polynomial work of an event (per forward segment it passes and the
reconstruction of an accepted event): one multiply per monomial of degree 2
or more that the monomial tree needs (every monomial is its parent times one
variable; parents of needed monomials are needed too) plus one FMA (2
operations) per non-zero coefficient.  Segments are the classes transp()
evaluates (shared/transp.f:1-281 via hms/mc_hms.f, shms/mc_shms.f).
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAPS = os.path.join(REPO, "data", "maps")
FILES = {1: ("hms", "hms/forward_cosy.dat", "hms/recon_cosy.dat"),
         5: ("shms", "shms/shms_forward.dat", "shms/shms_recon.dat")}


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simc_gfortran_amd import api
    c = api.Spectrometers()
    yield c
    c.close()


def _poly_count(terms, nout):
    nodes, nnz = set(), 0
    for c, e in terms:
        k = sum(1 for x in c[:nout] if x != 0.0)
        if not k:
            continue
        nnz += k
        e = tuple(e)
        while sum(e) >= 2 and e not in nodes:
            nodes.add(e)
            j = max(i for i in range(len(e)) if e[i])  # the last variable present
            e = e[:j] + (e[j] - 1,) + e[j + 1:]
    return len(nodes) + 2 * nnz


@pytest.mark.parametrize("arm", [1, 5])
def test_polynomial_counts(ctx, arm):
    sys.path.insert(0, os.path.join(REPO, "simc_gfortran_amd", "csrc"))
    import gen_cosy
    name, fwd, rec = FILES[arm]
    classes = gen_cosy.parse_forward(os.path.join(MAPS, fwd))
    forward = sum(_poly_count(classes[c - 1], 5) for c in gen_cosy.TRANSP_CLASSES[name])
    recon = _poly_count(gen_cosy.parse_recon(os.path.join(MAPS, rec)), 4)
    assert (forward, recon) == {1: (10397, 4144), 5: (14722, 4144)}[arm]
    whole = np.array(ctx.stop_flops(arm, 0))
    stages = [np.array(ctx.stop_flops(arm, k)) for k in (1, 2, 3)]
    assert whole[0] == forward + recon
    assert stages[2][0] == recon and not stages[2][1:].any()  # the reconstruction: accepted events only
    np.testing.assert_array_equal(stages[0] + stages[1] + stages[2], whole)


@pytest.mark.parametrize("arm", [1, 5])
def test_hut_counts(ctx, arm):
    full = np.array(ctx.hut_flops(arm, ms=True, wcs=True))
    no_wcs = np.array(ctx.hut_flops(arm, ms=True, wcs=False))
    bare = np.array(ctx.hut_flops(arm, ms=False, wcs=False))
    # stop codes of the target-to-hut stages hold no hut work
    poly = np.array(ctx.stop_flops(arm, 1)) + np.array(ctx.stop_flops(arm, 2))
    before_hut = (poly > 0) & (poly < poly[0])
    assert before_hut.any() and not full[before_hut].any()
    # the accepted event does the most hut work; 12 wire planes, 2 deviates each
    assert full[0] == full.max() > bare[0] > 0
    assert full[0] - no_wcs[0] == 12 * 2 * 4
    # the fit: 11 per plane + 18; every other step without deviates is a
    # drift (11), a rotated box (13) or the fiducial cut (4): what is left
    # after the fit and the planes' 3 each is made of those
    rest = bare[0] - (11 * 12 + 18) - 12 * 3
    assert rest > 0 and any((rest - 4 * f - 13 * b) % 11 == 0 for f in (0, 1) for b in (0, 1, 2))
