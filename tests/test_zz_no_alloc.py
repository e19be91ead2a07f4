"""The allocation-free entry point simc_arm_run_noalloc, called through the C
ABI: its outputs and STOP counters equal simc_arm_run's on the same events,
on two calls in a row (the second smaller, over the scratch the first one
left), for the Philox stream and a tape; sizes beyond the scratch capacity
are refused with SIMC_E_ARG."""
import ctypes as C

import numpy as np
import pytest

from simc_gfortran_amd import workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MAX_EVENTS = 524288
COLS = ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen", "resmult", "m2", "ok", "stop",
        "used")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simc_gfortran_amd import api
    c = api.Spectrometers()
    yield c
    c.close()


def outputs(ev):
    n = ev["dpp"].numel()
    dev = ev["dpp"].device
    out = {k: ev[k].clone() for k in ("dpp", "y", "dxdz", "dydz")}
    for k in ("x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen", "resmult", "m2"):
        out[k] = torch.zeros(n, dtype=torch.float64, device=dev)
    out["ok"] = torch.zeros(n, dtype=torch.uint8, device=dev)
    out["stop"] = torch.zeros(n, dtype=torch.uint8, device=dev)
    out["used"] = torch.zeros(n, dtype=torch.int32, device=dev)
    return out


def call(ctx, entry, arm, wl, ev, n, rng):
    """One call of `entry` (simc_arm_run or simc_arm_run_noalloc) through the
    C ABI; returns (rc, outputs, counters)."""
    from simc_gfortran_amd import api
    L = api.library()
    S = api.Settings(wl.p_spec, wl.m2, 0.0, 1, 1, 0, 0, api.FIT_FP64, 0, wl.wire_sigma)
    out = outputs(ev)
    ein = api.EventsIn(*[api._ptr(ev.get(k)) for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")])
    eout = api.EventsOut(*[api._ptr(out.get(k)) for k in api.OUT_FIELDS])
    ctx.reset_counters(arm)
    st = torch.cuda.current_stream().cuda_stream
    rc = getattr(L, entry)(ctx._h, arm, C.byref(S), C.byref(ein), C.byref(eout), n, C.byref(rng), C.c_void_p(st))
    torch.cuda.synchronize()
    return rc, out, ctx.counters(arm)


def events(wl, n, seed):
    return {k: torch.from_numpy(v).cuda() for k, v in W.generate(wl, n, seed=seed).items()}


def philox(seed, event0):
    from simc_gfortran_amd import api
    return api.Rng(api.RNG_PHILOX, 0, seed, event0, None, 0)


@pytest.mark.parametrize("first,second", [((5, 400000), (1, 150000)), ((1, MAX_EVENTS), (5, 3000))])
def test_noalloc_equals_arm_run_twice(ctx, first, second):
    for k, (arm, n) in enumerate((first, second)):
        wl = W.HMS_C0 if arm == 1 else W.SHMS_C1
        ev = events(wl, n, 40 + k)
        print("noalloc call", k, "arm", arm, "n", n, flush=True)
        rc_a, ref, cnt_a = call(ctx, "simc_arm_run", arm, wl, ev, n, philox(9 + k, 1000 * k))
        rc_b, got, cnt_b = call(ctx, "simc_arm_run_noalloc", arm, wl, ev, n, philox(9 + k, 1000 * k))
        assert rc_a == 0 and rc_b == 0
        for c in COLS:
            assert torch.equal(got[c], ref[c]), (k, c)
        assert cnt_a == cnt_b
        assert cnt_b["trials"] == n
        assert 0.2 * n < int(got["ok"].sum()) < 0.8 * n


def test_noalloc_tape_equals_arm_run(ctx):
    from simc_gfortran_amd import api
    n, stride = 20000, 1024
    wl = W.SHMS_C1
    ev = events(wl, n, 3)
    tape = torch.from_numpy(W.uniform_tape(n, stride, 17)).cuda()
    R = api.Rng(api.RNG_TAPE, 0, 0, 0, api._ptr(tape), stride)
    rc_a, ref, cnt_a = call(ctx, "simc_arm_run", 5, wl, ev, n, R)
    rc_b, got, cnt_b = call(ctx, "simc_arm_run_noalloc", 5, wl, ev, n, R)
    assert rc_a == 0 and rc_b == 0
    for c in COLS:
        assert torch.equal(got[c], ref[c]), c
    assert cnt_a == cnt_b


def test_noalloc_refuses_beyond_capacity(ctx):
    from simc_gfortran_amd import api
    n = MAX_EVENTS + 1
    ev = events(W.HMS_C0, n, 5)
    rc, _, _ = call(ctx, "simc_arm_run_noalloc", 1, W.HMS_C0, ev, n, philox(1, 0))
    assert rc == -6
    assert b"scratch capacity" in api.library().simc_last_error()
