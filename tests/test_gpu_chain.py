"""The H(e,e'p) coincidence chain on the GPU against the oracle's chain:
simc_gen_hep -> simc_target_prep for the electron (SHMS windows) and the
proton (HMS windows) -> SHMS and HMS transport + reconstruction, every step
on its counter-based stream, with the generator's columns passed on as the
integration notes describe.

Two comparisons: (1) the whole oracle chain from its own generator: flags,
the coincidence set and the summed weight must agree (the generator's
last-bit differences, amplified near p_yptar = 0 by spectrometer_angles'
conditioning, reach the proton arm's slopes at the 1e-11 level, so
coordinates are not compared here); (2) the oracle's target step and arms
run on the GPU generator's own columns: flags identical and coordinates
within 1e-12, as for the arms alone.
"""
import numpy as np
import pytest

from simc_gfortran_amd import workloads as W
from tests._compare import parity_report

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simc_gfortran_amd import api
    c = api.Spectrometers()
    yield c
    c.close()


def _vertices(g, particle):
    if particle == "e":
        return {"x": g["x"], "y": g["y"], "z": g["z"], "E": g["e_E"], "p": g["e_E"], "theta": g["e_theta"],
                "xptar": g["e_xptar"], "yptar": g["e_yptar"]}
    return {"x": g["x"], "y": g["y"], "z": g["z"], "E": g["p_E"], "p": g["p_P"], "theta": g["p_theta"],
            "delta": g["p_delta"], "xptar": g["p_xptar"], "yptar": g["p_yptar"]}


def test_coincidence_chain_matches_oracle(ctx, oracle):
    from simc_gfortran_amd import api
    n = 20000
    H = W.HEP_ELASTIC
    arms = {"e": (5, 2, H["e_P"], H["e_theta"], H["e_phi"], W.ME, 0.03),
            "p": (1, 3, H["p_P"], H["p_theta"], H["p_phi"], W.MP, 0.0)}
    To = oracle.make_target(**W.LH2_CELL)
    T = api.Target(**W.LH2_CELL)
    print("chain", n, flush=True)
    gd = ctx.gen_hep(T, api.hep_settings(**H), n, seed=41, event0=100)
    go, rc = oracle.gen_hep(To, oracle.make_hep(**H), n, rng="philox", seed=41, event0=100)
    assert rc == 0
    both_g = np.ones(n, bool)
    both_r = np.ones(n, bool)
    for name, (arm, narm, p0, th0, phi0, mass, wire) in arms.items():
        Po = oracle.make_prep(narm, arm, True, True, p_spec=p0, th_spec=th0, phi=phi0, mass=mass,
                              beam_energy=H["Ebeam"])
        P = api.PrepSettings(narm, arm, 1, 1, p0, th0, phi0, mass, 0.0, H["Ebeam"], (api.C.c_double * 5)())
        seed = 7 if name == "e" else 8
        kw = dict(p_spec=p0, m2=mass * mass, ms=True, wcs=True, wire_sigma=wire, rng="philox", seed=seed,
                  event0=100)

        def oracle_chain(vertices):
            ev_r, rc = oracle.target_prep(To, Po, vertices, rng="philox", seed=seed, event0=100)
            assert rc == 0
            r = oracle.run(arm, ev_r, fit_f32=False, **kw)
            assert r["rc"] == 0
            return r

        ev = ctx.target_prep(T, P, _vertices(gd, name), rng="philox", seed=seed, event0=100)
        out = ctx.run(arm, {k: ev[k] for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")}, fit="fp64", **kw)
        torch.cuda.synchronize()
        g = {k: t.cpu().numpy() for k, t in out.items() if t is not None}
        # (2) identical inputs: the GPU generator's columns
        same = oracle_chain({k: v.cpu().numpy() for k, v in _vertices(gd, name).items()})
        rep = parity_report(g, same, 1e-12)
        print(name, rep)
        assert rep["flag_mismatch"] == 0
        assert all(b == 0 for b in rep["bad"].values()), rep
        # (1) the oracle's own chain from its own generator
        ref = oracle_chain(_vertices(go, name))
        assert (g["stop"] == ref["stop"]).all()
        both_g &= g["stop"] == 0
        both_r &= ref["stop"] == 0
    assert (both_g == both_r).all()
    # a real coincidence sample: both arms accept a tenth or more of the events
    assert both_g.mean() > 0.1
    w = gd["weight"].cpu().numpy()
    np.testing.assert_allclose(w[both_g].sum(), go["weight"][both_r].sum(), rtol=1e-12)


def test_two_arms_on_two_streams_match_sequential_runs(ctx):
    """One context, SHMS and HMS issued on two streams at once: outputs and
    STOP counters bit-identical to the same runs one after the other (the
    survivor queues and counters are per arm)."""
    from simc_gfortran_amd import api  # noqa: F401
    n = 2_000_000
    wls = (W.SHMS_C1, W.HMS_C0)
    evs = [{k: torch.from_numpy(v).cuda() for k, v in W.generate(wl, n, seed=50 + i).items()}
           for i, wl in enumerate(wls)]
    kws = [dict(p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, wire_sigma=wl.wire_sigma, rng="philox", seed=60,
                fit="fp64") for wl in wls]
    print("two streams", n, flush=True)
    seq, seq_cnt = [], []
    for wl, ev, kw in zip(wls, evs, kws):
        ctx.reset_counters(wl.arm)
        seq.append(ctx.run(wl.arm, ev, **kw))
        torch.cuda.synchronize()
        seq_cnt.append(list(ctx.counters(wl.arm).values()))
    for wl in wls:
        ctx.reset_counters(wl.arm)
    main = torch.cuda.current_stream()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    par = []
    for s, wl, ev, kw in zip(streams, wls, evs, kws):
        s.wait_stream(main)
        with torch.cuda.stream(s):
            par.append(ctx.run(wl.arm, ev, **kw))
    for s in streams:
        main.wait_stream(s)
    torch.cuda.synchronize()
    for i, wl in enumerate(wls):
        for k, t in seq[i].items():
            if t is not None:
                assert torch.equal(t, par[i][k]), (wl.name, k)
        assert list(ctx.counters(wl.arm).values()) == seq_cnt[i]
