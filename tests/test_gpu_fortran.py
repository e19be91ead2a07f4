"""Fortran callers on the GPU: the batch driver through the module simc_gpu
(simc_arm_run_host) and the per-event drop-in mc_hms / mc_sos / mc_hrs? / mc_shms called like
montecarlo() does, both against the library's own batched run on the same
events and Philox streams (bit-identical), including the STOP counters the
drop-in adds into the program's common blocks.  The programs are built by
build() (tests/fortran/_bin)."""
import os
import subprocess

import numpy as np
import pytest

from simc_gfortran_amd import workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "tests", "fortran", "_bin")
MAPS = os.path.join(REPO, "data", "maps")
OUT9 = ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(os.path.join(BIN, "batch_host")):
        pytest.skip("Fortran programs not built")
    from simc_gfortran_amd import api
    c = api.Spectrometers()
    yield c
    c.close()


def write_events(path, ev):
    np.concatenate([ev[k] for k in ("dpp", "x", "y", "z", "dxdz", "dydz")]).astype(np.float64).tofile(path)


def read_outputs(path, n, ok_dtype):
    raw = np.fromfile(path, dtype=np.uint8)
    o = raw[:9 * n * 8].view(np.float64).reshape(9, n)
    ok = raw[9 * n * 8:].view(ok_dtype)[:n]
    return {k: o[i] for i, k in enumerate(OUT9)}, ok


def counters_line(stdout):
    for line in stdout.splitlines():
        if line.startswith("counters"):
            return [int(v) for v in line.split()[1:]]
    raise AssertionError(stdout)


@pytest.mark.parametrize("wl", [W.HMS_C0, W.SHMS_C1, W.SOS_1p5, W.HRSR_2, W.HRSL_3],
                         ids=["hms", "shms", "sos", "hrsr", "hrsl"])
def test_fortran_batch_driver_matches_library(ctx, tmp_path, wl):
    n, seed = 5000, 77
    ev = W.generate(wl, n, seed=3)
    write_events(tmp_path / "in.bin", ev)
    print("batch_host", wl.name, n, flush=True)
    r = subprocess.run([os.path.join(BIN, "batch_host"), MAPS, str(wl.arm), str(n), repr(wl.p_spec),
                        repr(wl.m2), str(seed), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got, ok = read_outputs(tmp_path / "out.bin", n, np.uint8)
    ctx.reset_counters(wl.arm)
    ref = ctx.run_host(wl.arm, ev, p_spec=wl.p_spec, m2=wl.m2, fit="fp64", seed=seed)
    assert (ok == ref["ok"]).all()
    for k in OUT9:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    cnt = list(ctx.counters(wl.arm).values())
    assert counters_line(r.stdout) == cnt


@pytest.mark.parametrize("wl", [W.HMS_C0, W.SHMS_C1, W.SOS_1p5, W.HRSR_2, W.HRSL_3],
                         ids=["hms", "shms", "sos", "hrsr", "hrsl"])
def test_fortran_per_event_dropin_matches_batched(ctx, tmp_path, wl):
    n, seed = 400, 5
    ev = W.generate(wl, n, seed=4)
    write_events(tmp_path / "in.bin", ev)
    env = dict(os.environ, SIMC_MAP_DIR=MAPS)
    print("dropin_events", wl.name, n, flush=True)
    r = subprocess.run([os.path.join(BIN, "dropin_events"), str(wl.arm), str(n), repr(wl.p_spec),
                        repr(wl.m2), str(seed), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    got, ok = read_outputs(tmp_path / "out.bin", n, np.int32)
    # the drop-in keeps the reference's REAL*4 track fit and the spectrometer's wire table
    ctx.reset_counters(wl.arm)
    ref = ctx.run_host(wl.arm, ev, p_spec=wl.p_spec, m2=wl.m2, fit="ref_f32", seed=seed)
    assert (ok == ref["ok"]).all()
    for k in OUT9:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    cnt = list(ctx.counters(wl.arm).values())
    dropin = counters_line(r.stdout)
    names = list(ctx.counters(wl.arm))
    for i, k in enumerate(names):
        if k != "stop_id":
            assert dropin[i] == cnt[i], (k, dropin[i], cnt[i])
    worst = float([l for l in r.stdout.splitlines() if l.startswith("recon_worst")][0].split()[1])
    # the HRS drop-in restores the fitted y from y_fp + VDC offset (one rounding)
    assert worst == 0.0 if wl.arm not in (3, 4) else worst < 1e-12
