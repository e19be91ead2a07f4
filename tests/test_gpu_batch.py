"""The batched form: many kinematic settings in one run.  Every setting's
events must equal a plain run of that setting alone (same Philox streams:
bit-identical), its counters must equal the plain run's, the launch sequence
must replay correctly from a captured HIP graph, and the autotuner must leave
results unchanged."""
import numpy as np
import pytest

from simc_gfortran_amd import workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

OUT = ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen", "ok", "stop")
MPI = 139.57018


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simc_gfortran_amd import api
    c = api.Spectrometers()
    yield c
    c.close()


def mixed(counts, gap=17, seeds=None):
    """Settings alternating arms with assorted momenta, particles and flags, laid
    out with gaps between their event ranges; returns (settings, events).
    seeds: per-setting Philox keys (default: one key, distinct event ids)."""
    specs = [
        dict(arm=1, p_spec=3000.0, m2=W.MP ** 2),
        dict(arm=5, p_spec=6000.0, m2=W.ME ** 2, wire_sigma=0.030, fit="fp64"),
        dict(arm=1, p_spec=1500.0, m2=W.MP ** 2, ms=False),
        dict(arm=5, p_spec=2500.0, m2=MPI ** 2, coll=True),
        dict(arm=1, p_spec=8000.0, m2=W.ME ** 2, fit="fp64"),
        dict(arm=5, p_spec=4200.0, m2=W.MP ** 2),
        dict(arm=1, p_spec=2000.0, m2=MPI ** 2, decay=True, ctau=780.45),
        dict(arm=2, p_spec=1200.0, m2=W.MP ** 2, fit="fp64"),
        dict(arm=3, p_spec=2000.0, m2=W.MP ** 2, fit="fp64"),
        dict(arm=4, p_spec=3000.0, m2=W.ME ** 2),
    ]
    settings, parts, off = [], [], 0
    for i, cnt in enumerate(counts):
        d = dict(specs[i % len(specs)])
        wl = {1: W.HMS_C0, 2: W.SOS_1p5, 3: W.HRSR_2, 4: W.HRSL_3, 5: W.SHMS_C1}[d["arm"]]
        if seeds is None:
            d.update(offset=off, count=cnt, seed=11, event0=(i + 1) << 32)
        else:
            d.update(offset=off, count=cnt, seed=seeds[i], event0=0)
        settings.append(d)
        ev = W.generate(wl, cnt, seed=100 + i)
        parts.append((off, ev))
        off += cnt + gap
    n = off
    events = {k: np.zeros(n) for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")}
    for o, ev in parts:
        for k in events:
            events[k][o:o + len(ev["dpp"])] = ev[k]
    return settings, {k: torch.from_numpy(v).cuda() for k, v in events.items()}


def plain(ctx, d, events):
    sl = slice(d["offset"], d["offset"] + d["count"])
    ev = {k: v[sl].contiguous() for k, v in events.items()}
    kw = {k: d[k] for k in ("p_spec", "m2", "ms", "wcs", "decay", "ctau", "coll", "fit", "wire_sigma")
          if k in d}
    ctx.reset_counters(d["arm"])
    out = ctx.run(d["arm"], ev, rng="philox", seed=d["seed"], event0=d["event0"], **kw)
    torch.cuda.synchronize()
    return out, ctx.counters(d["arm"])


def test_batch_equals_plain_runs(ctx):
    from simc_gfortran_amd import api
    counts = [20000, 3000, 5, 0, 45000, 1, 12000, 9000, 7000, 6000]
    settings, events = mixed(counts)
    b = api.Batch(ctx, settings)
    out = b.run(events)
    torch.cuda.synchronize()
    for i, d in enumerate(settings):
        if d["count"] == 0:
            continue
        ref, cnt = plain(ctx, d, events)
        sl = slice(d["offset"], d["offset"] + d["count"])
        for k in OUT:
            np.testing.assert_array_equal(out[k][sl].cpu().numpy(), ref[k].cpu().numpy(), err_msg=f"{i} {k}")
        got = b.counters(i)
        for k, v in cnt.items():
            if k != "stop_id":
                assert got[k] == v, (i, k, got[k], v)
    b.close()


def test_batch_per_setting_seeds_equal_plain_runs(ctx):
    """Settings with their own Philox keys (a SIMC run seeds per run,
    simc.f:107) and overlapping event ids share passes: the lanes key each
    event's stream from its own setting, and every setting still equals its
    plain run bit for bit."""
    from simc_gfortran_amd import api
    counts = [20000, 3000, 5, 0, 45000, 1, 12000, 9000, 7000, 6000, 15000, 4000]
    seeds = [11 * 1000 + i for i in range(len(counts))]
    settings, events = mixed(counts, seeds=seeds)
    b = api.Batch(ctx, settings)
    n_passes, n_lane = b.passes()
    assert n_lane >= 2 and n_passes < len(counts) - 1, (n_passes, n_lane)
    out = b.run(events)
    torch.cuda.synchronize()
    for i, d in enumerate(settings):
        if d["count"] == 0:
            continue
        ref, cnt = plain(ctx, d, events)
        sl = slice(d["offset"], d["offset"] + d["count"])
        for k in OUT:
            np.testing.assert_array_equal(out[k][sl].cpu().numpy(), ref[k].cpu().numpy(), err_msg=f"{i} {k}")
        got = b.counters(i)
        for k, v in cnt.items():
            if k != "stop_id":
                assert got[k] == v, (i, k, got[k], v)
    b.close()


def test_batch_c3_configuration(ctx, oracle):
    """The batched configuration as specified: 64 settings (HMS protons and
    SHMS electrons, p0 U(1, 9) GeV/c, counts log-uniform in [1e4, 1e7],
    1.5e8 events), each with its own seed 2026*1000 + s and events numbered
    from 0.  Per setting: a sample of events against the oracle (stop codes
    identical, coordinates within 1e-12), the whole range bit-identical to
    the setting's plain run, its counters equal to the plain run's and its
    trials equal to its event count."""
    from simc_gfortran_amd import api
    from tests._compare import parity_report
    dev = torch.device("cuda:0")
    sets = W.mixed_settings(64)
    total = sum(w.n for w in sets)
    assert 1.4e8 < total < 1.6e8
    ev = {k: torch.empty(total, dtype=torch.float64, device=dev) for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")}
    settings, off = [], 0
    for w in sets:
        W.generate_device(w, w.n, seed=w.seed, device=dev, out=ev, at=off)
        settings.append(dict(arm=w.arm, p_spec=w.p_spec, m2=w.m2, wire_sigma=w.wire_sigma, fit="fp64",
                             seed=w.seed, event0=0, offset=off, count=w.n))
        off += w.n
    b = api.Batch(ctx, settings)
    n_passes, n_lane = b.passes()
    print("C3:", total, "events,", n_passes, "passes,", n_lane, "keyed per setting", flush=True)
    assert n_passes == 2 and n_lane == 2
    out = b.run(ev)
    torch.cuda.synchronize()
    outs = b.new_outputs(ev)
    sample = 2000
    for i, d in enumerate(settings):
        sl = slice(d["offset"], d["offset"] + d["count"])
        kw = {k: d[k] for k in ("p_spec", "m2", "wire_sigma", "fit")}
        ctx.reset_counters(d["arm"])
        ctx.run(d["arm"], {k: v[sl] for k, v in ev.items()}, rng="philox", seed=d["seed"], event0=0,
                out={k: v[sl] for k, v in outs.items()}, **kw)
        torch.cuda.synchronize()
        for k in OUT:
            assert torch.equal(out[k][sl], outs[k][sl]), (i, k)
        cnt = ctx.counters(d["arm"])
        got = b.counters(i)
        for k, v in cnt.items():
            if k != "stop_id":
                assert got[k] == v, (i, k, got[k], v)
        assert got["trials"] == d["count"]
        m = min(sample, d["count"])
        ev_h = {k: v[d["offset"]:d["offset"] + m].cpu().numpy() for k, v in ev.items()}
        ref = oracle.run(d["arm"], ev_h, p_spec=d["p_spec"], m2=d["m2"], wire_sigma=d["wire_sigma"],
                         rng="philox", seed=d["seed"], event0=0, fit_f32=False)
        g = {k: v[d["offset"]:d["offset"] + m].cpu().numpy() for k, v in out.items() if v is not None}
        rep = parity_report(g, ref, 1e-12)
        assert rep["flag_mismatch"] == 0, (i, rep)
        assert all(v == 0 for v in rep["bad"].values()), (i, rep)
    b.close()


def test_batch_queues_survive_context_runs(ctx):
    """A batch keeps its own survivor queues: a larger plain run of the same
    arm afterwards (which grows the context's queues) leaves a captured batch
    run replaying correctly."""
    from simc_gfortran_amd import api
    settings, events = mixed([30000, 20000, 10000, 4000])
    b = api.Batch(ctx, settings)
    ref = b.run(events)
    torch.cuda.synchronize()
    ref = {k: v.clone() for k, v in ref.items()}
    out = b.new_outputs(events)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        b.run(events, out=out)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        b.run(events, out=out)
    big = {k: torch.from_numpy(v).cuda() for k, v in W.generate(W.HMS_C0, 300000, seed=5).items()}
    for arm, wl in ((1, W.HMS_C0), (5, W.SHMS_C1)):
        ctx.run(arm, big, p_spec=wl.p_spec, m2=wl.m2, rng="philox", seed=3)
    torch.cuda.synchronize()
    for k in ("dpp", "y", "dxdz", "dydz"):
        out[k].copy_(events[k])  # INOUT fields back to their inputs
    for k in ("x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen"):
        out[k].zero_()           # pathlen accumulates
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    for k in OUT:
        assert torch.equal(out[k], ref[k]), k
    b.close()


def test_batch_rejects_bad_outputs(ctx):
    from simc_gfortran_amd import api
    settings, events = mixed([1000, 500])
    b = api.Batch(ctx, settings)
    out = b.new_outputs(events)
    bad = dict(out, x_fp=out["x_fp"].float())
    with pytest.raises(TypeError):
        b.run(events, out=bad)
    bad = dict(out, ok=out["ok"][:10])
    with pytest.raises(ValueError):
        b.run(events, out=bad)
    bad = dict(out, pathlen=out["pathlen"].cpu())
    with pytest.raises(TypeError):
        b.run(events, out=bad)
    with pytest.raises(ValueError):
        b.run(events, out={k: v for k, v in out.items() if k != "stop"})
    b.close()


def test_batch_graph_replays(ctx):
    from simc_gfortran_amd import api
    settings, events = mixed([30000, 20000, 10000, 4000])
    b = api.Batch(ctx, settings)
    ref = b.run(events)
    torch.cuda.synchronize()
    ref = {k: v.clone() for k, v in ref.items()}
    out = b.new_outputs(events)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        b.run(events, out=out)  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        b.run(events, out=out)
    for rep in range(3):
        for k in ("dpp", "y", "dxdz", "dydz"):
            out[k].copy_(events[k])  # INOUT fields back to their inputs (eager work between replays)
        for k in ("x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen"):
            out[k].zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k in OUT:
            assert torch.equal(out[k], ref[k]), (rep, k)
    b.close()


def test_batch_tuner_keeps_results(ctx):
    from simc_gfortran_amd import api
    settings, events = mixed([200000, 50000, 90000])
    b = api.Batch(ctx, settings)
    ref = b.run(events)
    torch.cuda.synchronize()
    ref = {k: v.clone() for k, v in ref.items()}
    chunk, ms = b.tune(events, b.new_outputs(events))
    assert chunk in [1 << k for k in range(18, 25)] and ms > 0
    out = b.run(events)
    torch.cuda.synchronize()
    for k in OUT:
        assert torch.equal(out[k], ref[k]), k
    b.close()


def test_batch_rejects_bad_layouts(ctx):
    from simc_gfortran_amd import api
    base = dict(arm=1, p_spec=3000.0, m2=W.MP ** 2)
    with pytest.raises(api.SimcError):
        api.Batch(ctx, [dict(base, offset=0, count=10), dict(base, offset=5, count=10)])
    with pytest.raises(api.SimcError):
        api.Batch(ctx, [dict(base, offset=(1 << 29) - 5, count=10)])
    with pytest.raises(api.SimcError):
        api.Batch(ctx, [dict(base, offset=0, count=10, decay=True)])
