"""Full-stream runs: the GPU (counter-based Philox stream, Box-Muller normals)
against the oracle driven by the reference's RANLUX stream and polar-method
normals, on independent samples of the same input distribution: histograms
of the accepted events' focal-plane and reconstructed target variables agree
in chi2/ndf and the accepted fractions agree within binomial error
(5 sigma)."""
import numpy as np
import pytest

from simc_gfortran_amd import workloads as W
from tests._stats import compare

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N = 100000


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simc_gfortran_amd import api
    c = api.Spectrometers()
    yield c
    c.close()


@pytest.mark.parametrize("wl", [W.HMS_C0, W.SHMS_C1, W.SOS_1p5, W.HRSR_2, W.HRSL_3],
                         ids=["hms", "shms", "sos", "hrsr", "hrsl"])
def test_gpu_full_stream_histograms_match_reference_stream(ctx, oracle, wl):
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, wire_sigma=wl.wire_sigma)
    ref = oracle.run(wl.arm, W.generate(wl, N, seed=41), rng="ranlux", seed=1234567, nthreads=1,
                     fit_f32=False, **kw)
    ev = W.generate(wl, N, seed=141)
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in ev.items()}
    out = ctx.run(wl.arm, d, rng="philox", seed=99, fit="fp64", **kw)
    torch.cuda.synchronize()
    g = {k: v.cpu().numpy() for k, v in out.items() if v is not None}
    ok, rep = compare(g, ref, N)
    print(wl.name, rep)
    assert ok, rep
