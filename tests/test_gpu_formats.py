"""Output formats from device results: the ntuple's GPU-side compaction and
record formatting equals the host writer on the same selected rows, byte for
byte; the STOP-counter report carries the context's counters in the
reference's line layout (simc.f:462-530)."""
import os

import numpy as np
import pytest

from simc_gfortran_amd import workloads as W

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


def test_device_ntuple_equals_host_ntuple(ctx, tmp_path):
    from simc_gfortran_amd import api
    wl = W.HMS_C0
    n = 300_001
    ev = W.generate_device(wl, n, seed=9, device=torch.device("cuda:0"))
    inputs = {k: v.clone() for k, v in ev.items()}
    out = ctx.run(wl.arm, ev, p_spec=wl.p_spec, m2=wl.m2, rng="philox", seed=4, fit="fp64")
    torch.cuda.synchronize()
    # hadron-arm columns of an elastic ntuple (results_write.f:90-101): recon, focal plane, generated
    cols = [out["dpp"], out["dydz"], out["dxdz"], out["y"], out["x_fp"], out["dx_fp"], out["y_fp"],
            out["dy_fp"], inputs["dpp"], inputs["dydz"], inputs["dxdz"], inputs["y"]]
    tags = api.ntuple_tags("elastic")[:12]
    print("ntuple", n, flush=True)
    with api.Ntuple(str(tmp_path / "d.bin"), tags) as nt:
        written = nt.write_device(cols, out["ok"])
    sel = out["ok"].cpu().numpy() == 1
    assert written == int(sel.sum()) > 0.2 * n
    rows = np.stack([c.cpu().numpy()[sel] for c in cols], axis=1)
    with api.Ntuple(str(tmp_path / "h.bin"), tags) as nt:
        nt.write(rows)
    assert open(tmp_path / "d.bin", "rb").read() == open(tmp_path / "h.bin", "rb").read()
    with api.Ntuple(str(tmp_path / "all.bin"), tags) as nt:  # no selection: every event
        assert nt.write_device(cols) == n
    t, v = api.read_ntuple(str(tmp_path / "all.bin"))
    assert t == tags and v.shape == (n, 12)
    np.testing.assert_array_equal(v[:, 4], out["x_fp"].cpu().numpy())


def test_stop_report_layout_and_counts(ctx, tmp_path):
    from simc_gfortran_amd import api
    for wl in (W.HMS_C0, W.SHMS_C1, W.HRSR_2):
        ctx.reset_counters(wl.arm)
        ev = W.generate_device(wl, 50_000, seed=3, device=torch.device("cuda:0"))
        ctx.run(wl.arm, ev, p_spec=wl.p_spec, m2=wl.m2, rng="philox", seed=8, fit="fp64")
    torch.cuda.synchronize()
    path = str(tmp_path / "run.geni")
    api.stop_report(ctx, path, 5, 1)  # SHMS electrons, HMS hadrons
    lines = open(path).read().split("\n")
    h = ctx.counters(1)
    assert lines[0] == " HMS Trials:           " + "%12d" % h["trials"]
    assert lines[1] == " Slit hor/vert/corners " + "".join("%12d" % h[k] for k in ("slit_hor", "slit_vert",
                                                                                     "slit_oct"))
    assert lines[8] == " Successes             " + "%12d" % h["successes"] and lines[9] == ""
    s = ctx.counters(5)
    assert lines[10] == " SHMS Trials:          " + "%12d" % s["trials"]
    assert lines[15].split()[-5:] == [str(s[k]) for k in ("Q3_in", "Q3_men", "Q2_mid", "Q3_mex", "Q3_out")]
    assert lines[22] == " Successes             " + "%12d" % s["successes"]
    api.stop_report(ctx, path, 3, 3)
    lines = open(path).read().split("\n")
    r = ctx.counters(3)
    assert lines[0] == " HRSr Trials:          " + "%12d" % r["trials"]
    assert lines[1].split()[-2:] == [str(r["slit_vert"]), str(r["slit_hor"])]
    assert lines[8] == " S1, S2, Cal\t    " + "".join("%12d" % r[k] for k in ("s1", "s2", "cal"))


def test_device_ntuple_empty_and_no_selection(ctx, tmp_path):
    from simc_gfortran_amd import api
    z = torch.zeros(0, dtype=torch.float64, device="cuda")
    with api.Ntuple(str(tmp_path / "e.bin"), ["a", "b"]) as nt:
        assert nt.write_device([z, z]) == 0
        none = torch.zeros(5, dtype=torch.uint8, device="cuda")
        c = torch.arange(5, dtype=torch.float64, device="cuda")
        assert nt.write_device([c, c], none) == 0
    t, v = api.read_ntuple(str(tmp_path / "e.bin"))
    assert t == ["a", "b"] and v.shape == (0, 2)


@pytest.mark.parametrize("case", ["elastic", "pion_tgtfield"])
def test_device_ntuple_equals_reference_file(ctx, tmp_path, case):
    """The GPU writer against the file the reference's own NtupleInit /
    NtupleClose wrote for the selected rows (tests/golden/ntuple/, made by
    tests/golden/make_golden.py ntuple through oracle/_ref): byte for byte."""
    from simc_gfortran_amd import api
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "ntuple", case + ".npz"))
    tags = api.ntuple_tags(str(g["mode"]), tgt_field=bool(g["tgt_field"]))
    rows = g["rows"]
    cols = [torch.from_numpy(np.ascontiguousarray(rows[:, k])).cuda() for k in range(rows.shape[1])]
    mask = torch.from_numpy(g["mask"]).cuda()
    path = tmp_path / "d.bin"
    with api.Ntuple(str(path), tags) as nt:
        assert nt.write_device(cols, mask) == int(g["mask"].sum())
    assert open(path, "rb").read() == g["ref_bin"].tobytes()
