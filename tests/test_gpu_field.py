"""Target-field tracking on the GPU (simc_track_from_tgt, simc_track_to_tgt)
against the reference's outputs (tests/golden/field) and the oracle.

track_from_tgt performs the reference's operations in the same order with
FP contraction off, so the GPU reproduces it bit for bit.  track_to_tgt
evaluates the arm's reconstruction polynomial in the generated (trie) order
instead of the map file's term order, so its coordinates agree within
1e-12 relative (floor 1e-12 absolute) and its flags are identical.
"""
import glob
import os

import numpy as np
import pytest

from simc_gfortran_amd import workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = sorted(glob.glob(os.path.join(HERE, "golden", "field", "*.npz")))
TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simc_gfortran_amd import api
    c = api.Spectrometers()
    yield c
    c.close()


def dev(v):
    return torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).cuda()


def field_of(api, mp, te, tp):
    return api.TargetField(api.FIELD_MAP if mp == "map" else (None if mp == "None" else mp), te, tp)


def close(a, b, scale=1.0):
    return np.abs(a - b) <= TOL * (np.abs(b) + scale)


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_gpu_matches_reference_field_golden(ctx, path):
    from simc_gfortran_amd import api
    g = np.load(path)
    F = field_of(api, str(g["map"]), float(g["theta_e"]), float(g["theta_p"]))
    print(os.path.basename(path), flush=True)
    if str(g["kind"]) == "from":
        ev = {k: dev(g["in_" + k]) for k in ("x", "y", "z", "dx", "dy")}
        ok = F.track_from_tgt(ev, dev(g["mom"]), float(g["mass"]), int(g["spect"]))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(ok.cpu().numpy(), g["out_ok"])
        for k in ("x", "y", "z", "dx", "dy"):
            np.testing.assert_array_equal(ev[k].cpu().numpy(), g["out_" + k], err_msg=k)
    else:
        trk = {k: dev(g["trk_" + k]) for k in ("x_fp", "dx_fp", "y_fp", "dy_fp")}
        rec = {k: dev(g["in_" + k]) for k in ("delta", "y", "dx", "dy")}
        ok = F.track_to_tgt(ctx, int(g["arm"]), trk, rec, dev(g["frx"]), dev(g["fry"]), dev(g["mom"]),
                            float(g["mass"]), float(g["ctheta"]), float(g["stheta"]), int(g["spect"]))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(ok.cpu().numpy(), g["out_ok"])
        for k, sc in (("delta", 1.0), ("y", 1.0), ("dx", 1e-2), ("dy", 1e-2)):
            assert close(rec[k].cpu().numpy(), g["out_" + k], sc).all(), k
    F.close()


def test_from_tgt_parity_large(ctx, oracle):
    """1e5 vertices through the real map at a tilted field: bit for bit."""
    from simc_gfortran_amd import api
    n = 100_000
    rng = np.random.default_rng(11)
    ev = dict(x=rng.uniform(-1, 1, n), y=rng.uniform(-1, 1, n), z=rng.uniform(-3, 3, n),
              dx=rng.uniform(-0.1, 0.1, n), dy=rng.uniform(-0.08, 0.08, n))
    mom = np.where(rng.random(n) < 0.5, -1.0, 1.0) * rng.uniform(300.0, 5000.0, n)
    Fo = oracle.field(oracle.FIELD_MAP, -7.0, 12.0)
    for spect, mass in ((1, 938.272), (-1, 0.511)):
        ref = oracle.track_from_tgt(Fo, ev, mom, mass, spect)
        F = api.TargetField(api.FIELD_MAP, -7.0, 12.0)
        d = {k: dev(v) for k, v in ev.items()}
        ok = F.track_from_tgt(d, dev(mom), mass, spect)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(ok.cpu().numpy(), ref["ok"])
        for k in ("x", "y", "z", "dx", "dy"):
            np.testing.assert_array_equal(d[k].cpu().numpy(), ref[k], err_msg=k)
        F.close()


@pytest.mark.parametrize("arm", [1, 2, 3, 4, 5])
def test_to_tgt_parity_every_arm(ctx, oracle, arm):
    """The accepted events of an oracle run of each arm, reconstructed
    through the field: flags identical, coordinates within 1e-12."""
    from simc_gfortran_amd import api
    wl = {1: W.HMS_C0, 2: W.SOS_1p5, 3: W.HRSR_2, 4: W.HRSL_3, 5: W.SHMS_C1}[arm]
    electron = wl.mass < 1.0
    spect, sgn = (-1, -1.0) if electron else (1, 1.0)
    ev = W.generate(wl, 8000, seed=21)
    res = oracle.run(arm, ev, p_spec=wl.p_spec, m2=wl.m2, rng="philox", seed=5, fit_f32=False)
    sel = res["ok"] == 1
    trk = {k: res[k][sel] for k in ("x_fp", "dx_fp", "y_fp", "dy_fp")}
    if arm in (3, 4):
        trk["y_fp"] = trk["y_fp"] + (0.48 if arm == 3 else 0.78)
    first = dict(delta=res["dpp"][sel], y=res["y"][sel], dx=res["dxdz"][sel], dy=res["dydz"][sel])
    m = int(sel.sum())
    rng = np.random.default_rng(arm)
    frx, fry = rng.uniform(-0.3, 0.3, m), rng.uniform(-0.3, 0.3, m)
    mom = sgn * wl.p_spec * (1 + first["delta"] / 100)
    th = np.radians(wl.theta_deg)
    st = np.sin(th) if arm in (1, 3) else -np.sin(th)
    Fo = oracle.field(oracle.FIELD_MAP, 9.0, 9.0)
    ref = oracle.track_to_tgt(Fo, arm, trk, first, frx, fry, mom, wl.mass, np.cos(th), st, spect)
    F = api.TargetField(api.FIELD_MAP, 9.0, 9.0)
    rec = {k: dev(v) for k, v in first.items()}
    print("to_tgt arm", arm, m, flush=True)
    ok = F.track_to_tgt(ctx, arm, {k: dev(v) for k, v in trk.items()}, rec, dev(frx), dev(fry), dev(mom),
                        wl.mass, np.cos(th), st, spect)
    torch.cuda.synchronize()
    okg = ok.cpu().numpy()
    assert (okg == ref["ok"]).all()
    assert ref["ok"].mean() > 0.9
    for k, sc in (("delta", 1.0), ("y", 1.0), ("dx", 1e-2), ("dy", 1e-2)):
        bad = ~close(rec[k].cpu().numpy(), ref[k], sc)
        assert bad.sum() == 0, (k, int(bad.sum()))
    F.close()


def test_host_entry_point_matches_device(ctx):
    from simc_gfortran_amd import api
    n = 3000
    rng = np.random.default_rng(5)
    ev = dict(x=rng.uniform(-1, 1, n), y=rng.uniform(-1, 1, n), z=rng.uniform(-2, 2, n),
              dx=rng.uniform(-0.05, 0.05, n), dy=rng.uniform(-0.05, 0.05, n))
    mom = rng.uniform(500.0, 3000.0, n)
    F = api.TargetField(None, 0.0, 20.0)  # the uniform test field
    h = F.track_from_tgt_host(ev, mom, 938.272, 1)
    d = {k: dev(v) for k, v in ev.items()}
    ok = F.track_from_tgt(d, dev(mom), 938.272, 1)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(h["ok"], ok.cpu().numpy())
    for k in ("x", "y", "z", "dx", "dy"):
        np.testing.assert_array_equal(h[k], d[k].cpu().numpy(), err_msg=k)
    F.close()


def test_field_errors(ctx):
    from simc_gfortran_amd import api
    with pytest.raises(api.SimcError):
        api.TargetField("/nonexistent/field.map")
    F = api.TargetField("0")
    ev = {k: dev(np.zeros(4)) for k in ("x", "y", "z", "dx", "dy")}
    with pytest.raises(api.SimcError) as e:
        F.track_from_tgt(ev, dev(np.ones(4)), 1.0, 0)
    assert e.value.code == -6
    F.close()


def test_empty_inputs(ctx):
    from simc_gfortran_amd import api
    F = api.TargetField(api.FIELD_MAP, 0.0, 0.0)
    e = {k: torch.zeros(0, dtype=torch.float64, device="cuda") for k in ("x", "y", "z", "dx", "dy")}
    assert F.track_from_tgt(e, torch.zeros(0, dtype=torch.float64, device="cuda"), 938.0, 1).numel() == 0
    z = torch.zeros(0, dtype=torch.float64, device="cuda")
    trk = {k: z for k in ("x_fp", "dx_fp", "y_fp", "dy_fp")}
    rec = {k: z.clone() for k in ("delta", "y", "dx", "dy")}
    assert F.track_to_tgt(ctx, 1, trk, rec, z, z, z, 938.0, 1.0, 0.0, 1).numel() == 0
    F.close()
