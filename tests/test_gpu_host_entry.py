"""The host-array entry point (simc_arm_run_host): chunked through pinned
staging on two streams, it must give the device entry point's results bit for
bit -- over several chunks, on tapes, with the INOUT columns starting from the
caller's values, and with the optional outputs left out."""
import numpy as np
import pytest

from simc_gfortran_amd import workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FIELDS = ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen", "resmult", "m2", "ok", "stop",
          "used")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simc_gfortran_amd import api
    c = api.Spectrometers()
    yield c
    c.close()


def _device(ctx, wl, ev, **kw):
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in ev.items()}
    out = ctx.run(wl.arm, d, p_spec=wl.p_spec, m2=wl.m2, wire_sigma=wl.wire_sigma, fit="fp64", **kw)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items() if v is not None}


@pytest.mark.parametrize("wl", [W.SHMS_C1, W.HMS_C0], ids=["shms", "hms"])
def test_host_entry_matches_device_over_chunks(ctx, wl):
    n = 2_500_000  # three chunks of at most 2^20
    ev = W.generate(wl, n, seed=12)
    print("host entry", wl.name, n, flush=True)
    ref = _device(ctx, wl, ev, rng="philox", seed=4, event0=77)
    got = ctx.run_host(wl.arm, ev, p_spec=wl.p_spec, m2=wl.m2, wire_sigma=wl.wire_sigma, fit="fp64", seed=4,
                       event0=77)
    for k in FIELDS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


def test_host_entry_tape_and_inout(ctx):
    wl = W.SHMS_C1
    n, stride = 3000, 1024
    ev = W.generate(wl, n, seed=13)
    tape = W.uniform_tape(n, stride, 99)
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in ev.items()}
    # INOUT: the fitted track keeps the caller's values where the fit is not
    # reached, pathlen accumulates onto them, resmult is kept by events that
    # stop before the hut
    init = {k: np.full(n, 7.25) for k in ("x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen", "resmult")}
    dev_out = {k: torch.from_numpy(v.copy()).cuda() for k, v in init.items()}
    for k in ("dpp", "y", "dxdz", "dydz", "m2"):
        dev_out[k] = torch.empty(n, dtype=torch.float64, device="cuda")
    dev_out["ok"] = torch.empty(n, dtype=torch.uint8, device="cuda")
    dev_out["stop"] = torch.empty(n, dtype=torch.uint8, device="cuda")
    dev_out["used"] = torch.empty(n, dtype=torch.int32, device="cuda")
    kw = dict(p_spec=wl.p_spec, m2=wl.m2, wire_sigma=wl.wire_sigma, fit="fp64", rng="tape", stride=stride)
    ref = ctx.run(wl.arm, d, tape=torch.from_numpy(tape).cuda(), out=dev_out, **kw)
    torch.cuda.synchronize()
    ref = {k: v.cpu().numpy() for k, v in ref.items() if v is not None}
    out = {k: v.copy() for k, v in init.items()}
    for k in ("dpp", "y", "dxdz", "dydz", "m2"):
        out[k] = np.zeros(n)
    out["ok"] = np.zeros(n, np.uint8)
    out["stop"] = np.zeros(n, np.uint8)
    out["used"] = np.zeros(n, np.int32)
    got = ctx.run_host(wl.arm, ev, tape=tape, out=out, **kw)
    for k in FIELDS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    before_hut = (got["stop"] >= 1) & (got["stop"] <= 70)  # SHMS stops up to the dipole exit
    assert before_hut.sum() > 100 and (got["x_fp"][before_hut] == 7.25).all()  # kept: no fit reached
    in_hut = (got["stop"] == 0) | ((got["stop"] > 70) & (got["stop"] < 250))
    assert (got["resmult"][before_hut] == 7.25).all() and np.isin(got["resmult"][in_hut], (1.0, 2.0)).all()
    assert (got["pathlen"] > 7.25).all()


def test_host_entry_single_event_calls(ctx):
    """One event per call, as the per-event drop-in calls it: the same
    results as one batched call over the same event ids."""
    wl = W.HMS_C0
    n = 64
    ev = W.generate(wl, n, seed=14)
    ref = ctx.run_host(wl.arm, ev, p_spec=wl.p_spec, m2=wl.m2, fit="fp64", seed=8, event0=500)
    for i in range(n):
        one = {k: v[i:i + 1] for k, v in ev.items()}
        got = ctx.run_host(wl.arm, one, p_spec=wl.p_spec, m2=wl.m2, fit="fp64", seed=8, event0=500 + i)
        for k in FIELDS:
            assert got[k][0] == ref[k][i] or (np.isnan(got[k][0]) and np.isnan(ref[k][i])), (i, k)
