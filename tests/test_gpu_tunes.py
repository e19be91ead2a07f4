"""The reference's other SHMS tunes on the GPU (shared/transp.f:287-478 and
shms/mc_shms_recon.f:63-128 read whichever map files shms_forward.dat /
shms_recon.dat point to).  Their term structures differ from the compiled
polynomials', so the library evaluates them from run-time tables; the
oracle reads the same files.  Flags and stop codes identical, coordinates
within 1e-12, counters and uniforms drawn equal.
"""
import numpy as np
import pytest

from simc_gfortran_amd import workloads as W
from tests._compare import parity_report

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TUNES = ["shms-2017-26cm-monte_quads_p18", "shms-2011-26cm-monte_ideal_6ord"]


def _dev(ev):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in ev.items()}


def _host(out):
    return {k: v.cpu().numpy() for k, v in out.items() if v is not None}


@pytest.fixture(scope="module")
def tuned():
    """One context per tune with the tune loaded for the SHMS."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from oracle import oracle as O
    from simc_gfortran_amd import api
    ctxs = {}
    for tune in TUNES:
        c = api.Spectrometers()
        c.load_arm(5, *O.tune_files(tune))
        ctxs[tune] = c
    yield ctxs
    for c in ctxs.values():
        c.close()


@pytest.mark.parametrize("tune", TUNES)
@pytest.mark.parametrize("rng", ["philox", "tape"])
def test_tune_parity(tuned, oracle, tmp_path_factory, tune, rng):
    wl = W.SHMS_C1
    n = 40000 if rng == "philox" else 4000
    ev = W.generate(wl, n, seed=91)
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, wire_sigma=wl.wire_sigma)
    d = oracle.tune_map_dir(str(tmp_path_factory.getbasetemp()), tune)
    ctx = tuned[tune]
    if rng == "tape":
        stride = 1024  # the SHMS hut's polar draws take about 520 uniforms
        tape = W.uniform_tape(n, stride, 17)
        ref = oracle.run(5, ev, rng="tape", tape=tape, stride=stride, fit_f32=False, map_dir=d, **kw)
        ctx.reset_counters(5)
        out = ctx.run(5, _dev(ev), rng="tape", tape=torch.from_numpy(tape).cuda(), stride=stride, fit="fp64",
                      **kw)
    else:
        ref = oracle.run(5, ev, rng="philox", seed=5, event0=300, fit_f32=False, map_dir=d, **kw)
        ctx.reset_counters(5)
        out = ctx.run(5, _dev(ev), rng="philox", seed=5, event0=300, fit="fp64", **kw)
    torch.cuda.synchronize()
    g = _host(out)
    print("tune", tune, rng, n, flush=True)
    rep = parity_report(g, ref, 1e-12)
    print(rep)
    assert rep["flag_mismatch"] == 0
    assert all(v == 0 for v in rep["bad"].values()), rep
    np.testing.assert_array_equal(g["stop"], ref["stop"])
    np.testing.assert_array_equal(g["used"], ref["used"])
    cnt = ctx.counters(5)
    for i, k in enumerate(cnt):
        if k != "stop_id":
            assert cnt[k] == ref["counters"][i], (k, cnt[k], ref["counters"][i])
    assert 0.2 < g["ok"].mean() < 0.7 and not (g["stop"] == 250).any()


@pytest.mark.parametrize("tune", TUNES)
def test_tune_recon_only(tuned, oracle, tmp_path_factory, tune):
    rng = np.random.default_rng(8)
    n = 10000
    x, dx = rng.uniform(-30, 30, n), rng.uniform(-0.05, 0.05, n)
    y, dy = rng.uniform(-20, 20, n), rng.uniform(-0.04, 0.04, n)
    fry = rng.normal(0, 0.2, n)
    fry[:10] = 0.0
    d = oracle.tune_map_dir(str(tmp_path_factory.getbasetemp()), tune)
    ref = oracle.recon(5, x, dx, y, dy, fry, map_dir=d)
    default = oracle.recon(5, x, dx, y, dy, fry)
    got = [t.cpu().numpy() for t in tuned[tune].recon(5, *[torch.from_numpy(v).cuda() for v in (x, dx, y, dy, fry)])]
    for a, b, s in zip(got, ref, (1.0, 1e-2, 1e-2, 1.0)):
        assert np.all(np.abs(a - b) <= 1e-12 * (np.abs(b) + s))
    assert np.abs(ref[0] - default[0]).max() > 1e-6  # a different tune indeed


def test_tune_replaced_by_default_maps(oracle):
    """Loading the default pair again returns the arm to the compiled
    polynomials: the same bits as a fresh context."""
    from oracle import oracle as O
    from simc_gfortran_amd import api
    wl = W.SHMS_C1
    ev = _dev(W.generate(wl, 20000, seed=4))
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, rng="philox", seed=9, fit="fp64")
    a = api.Spectrometers()
    a.load_arm(5, *O.tune_files(TUNES[0]))
    a.load_arm(5, O.MAP_DIR + "/shms/shms_forward.dat", O.MAP_DIR + "/shms/shms_recon.dat")
    b = api.Spectrometers()
    ra = _host(a.run(5, ev, **kw))
    rb = _host(b.run(5, ev, **kw))
    torch.cuda.synchronize()
    for k in ("ok", "dpp", "x_fp", "dy_fp"):
        np.testing.assert_array_equal(ra[k], rb[k])
    a.close()
    b.close()


def test_tune_adjoint_refused(tuned):
    from simc_gfortran_amd import api
    wl = W.SHMS_C1
    ctx = tuned[TUNES[0]]
    n = 256
    ev = _dev(W.generate(wl, n, seed=2))
    mask = torch.ones(n, dtype=torch.uint8, device="cuda")
    coef = torch.zeros((461, 4), dtype=torch.float64, device="cuda")
    with pytest.raises(api.SimcError) as e:
        ctx.adjoint(5, ev, mask, coef, [0.0] * 6, p_spec=wl.p_spec, m2=wl.m2)
    assert e.value.code == -8
