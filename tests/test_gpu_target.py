"""Target step on the GPU (simc_target_prep) against the oracle and the
reference's golden vectors.

* tape: the reference's own trip_thru_target / target_musc results
  (tests/golden/target) and the oracle's full preparation, same tapes;
* philox: the oracle's preparation on the same counter-based stream;
* the C0 chain vertex -> target step -> HMS against the oracle's chain;
* the host-pointer entry point, tape exhaustion and argument errors.
Tolerance: 1e-12 relative (floor: the coordinate's scale); the device's
libm differs from the host's in the last bit, nothing else.  Uniforms drawn
agree exactly.
"""
import glob
import math
import os

import numpy as np
import pytest

from simc_gfortran_amd import workloads as W
from tests._compare import parity_report
from tests.test_target_oracle import TARGET_GOLDEN, golden_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MP = 938.27231
ME = 0.51099906
SCALE = {"dpp": 1.0, "x": 1.0, "y": 1.0, "z": 1.0, "dxdz": 1e-2, "dydz": 1e-2, "fry": 1.0, "eloss": 1.0,
         "teff": 1e-2, "dang1": 1e-3, "dang2": 1e-3}


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simc_gfortran_amd import api
    c = api.Spectrometers()
    yield c
    c.close()


def to_dev(d):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in d.items() if v is not None}


def api_target(T):
    from simc_gfortran_amd import api
    return api.Target(T.A, T.Z, T.rho, T.X0, T.length, T.angle, T.zoffset, T.can, 0)


def api_prep(P):
    from simc_gfortran_amd import api
    return api.PrepSettings(P.narm, P.spectrometer, P.using_eloss, P.mc_smear, P.p_spec, P.th_spec, P.phi,
                            P.mass, P.coulomb, P.beam_energy, (api.C.c_double * 5)(*P.offset))


def assert_close(g, r, keys, rtol=1e-12):
    for k in keys:
        a, b = np.asarray(g[k]), np.asarray(r[k])
        tol = rtol * (np.maximum(np.abs(a), np.abs(b)) + SCALE[k])
        bad = ~(np.abs(a - b) <= tol)
        assert not bad.any(), (k, int(bad.sum()), float(np.abs(a - b).max()))


@pytest.mark.parametrize("path", TARGET_GOLDEN, ids=[os.path.basename(p)[:-4] for p in TARGET_GOLDEN])
def test_target_step_against_reference_golden(ctx, oracle, path):
    g = np.load(path)
    T, P, v, tape = golden_case(oracle, g)
    stride = int(g["stride"])
    out = ctx.target_prep(api_target(T), api_prep(P), to_dev(v), rng="tape",
                          tape=torch.from_numpy(tape).cuda(), stride=stride)
    torch.cuda.synchronize()
    h = {k: t.cpu().numpy() for k, t in out.items()}
    np.testing.assert_array_equal(h["used"], g["used"])
    assert_close(h, g, ("eloss", "teff", "dang1", "dang2"))
    ref, rc = oracle.target_prep(T, P, v, rng="tape", tape=tape, stride=stride)
    assert rc == 0
    keys = ("eloss", "teff", "dang1", "dang2") if P.narm == 1 else tuple(SCALE)
    assert_close(h, ref, keys)


PREP_CASES = [
    # narm, spectrometer, mass, p_spec, th_spec, phi, using_eloss, mc_smear, can
    ("hadron_hms", 3, 1, MP, 3000.0, 0.35, 1.5 * math.pi, True, True, 3),
    ("hadron_hms_no_loss", 3, 1, MP, 3000.0, 0.35, 1.5 * math.pi, False, True, 1),
    ("electron_shms", 2, 5, ME, 6000.0, 0.22, 0.5 * math.pi, True, True, 2),
    ("electron_shms_no_smear", 2, 5, ME, 6000.0, 0.22, 0.5 * math.pi, True, False, 3),
    ("beam", 1, 0, ME, 0.0, 0.0, 0.0, True, True, 3),
]


def _vertex(n, seed, mass, p0, th0, phi):
    rng = np.random.default_rng(seed)
    delta = rng.uniform(-10, 20, n)
    p = p0 * (1 + delta / 100) if p0 > 0 else np.full(n, 10600.0)
    xptar = rng.uniform(-0.06, 0.06, n)
    yptar = rng.uniform(-0.04, 0.04, n)
    return {"x": rng.normal(0, 0.05, n), "y": rng.normal(0, 0.05, n), "z": rng.uniform(-5, 5, n),
            "E": np.sqrt(p * p + mass * mass), "p": p, "delta": delta,
            "theta": W.physics_theta(th0, phi, xptar, yptar) if p0 > 0 else np.zeros(n),
            "xptar": xptar, "yptar": yptar, "dang_yp": rng.normal(0, 1e-4, n), "dang_xp": rng.normal(0, 1e-4, n)}


@pytest.mark.parametrize("case", PREP_CASES, ids=[c[0] for c in PREP_CASES])
@pytest.mark.parametrize("rng_kind", ["tape", "philox"])
def test_target_step_parity(ctx, oracle, case, rng_kind):
    name, narm, spec, mass, p0, th0, phi, eloss, smear, can = case
    n, stride = 20000, 64
    v = _vertex(n, 31, mass, p0, th0, phi)
    T = oracle.make_target(can=can, zoffset=0.25)
    v["z"] = v["z"] + 0.25
    P = oracle.make_prep(narm, spec, eloss, smear, p_spec=p0, th_spec=th0, phi=phi, mass=mass, coulomb=0.0,
                         beam_energy=10600.0, offset=(0.05, -0.03, 0.2, 1e-3, -2e-3))
    tape = W.uniform_tape(n, stride, 77)
    kw = dict(rng="tape", tape=tape, stride=stride) if rng_kind == "tape" else dict(rng="philox", seed=99,
                                                                                    event0=12345)
    ref, rc = oracle.target_prep(T, P, v, **kw)
    assert rc == 0
    if rng_kind == "tape":
        kw["tape"] = torch.from_numpy(tape).cuda()
    out = ctx.target_prep(api_target(T), api_prep(P), to_dev(v), **kw)
    torch.cuda.synchronize()
    h = {k: t.cpu().numpy() for k, t in out.items()}
    np.testing.assert_array_equal(h["used"], ref["used"])
    keys = ("eloss", "teff", "dang1", "dang2") if narm == 1 else tuple(SCALE)
    assert_close(h, ref, keys)


def test_target_step_host_entry_point(ctx):
    from simc_gfortran_amd import api
    n = 5000
    v = _vertex(n, 5, MP, 3000.0, 0.35, 1.5 * math.pi)
    T = api.Target(1.0, 1.0, 0.0723, 61.28, 10.0, 0.0, 0.0, 3, 0)
    P = api.PrepSettings(3, 1, 1, 1, 3000.0, 0.35, 1.5 * math.pi, MP, 0.0, 0.0, (api.C.c_double * 5)())
    d = ctx.target_prep(T, P, to_dev(v), rng="philox", seed=7, event0=3)
    torch.cuda.synchronize()
    h = ctx.target_prep_host(T, P, v, rng="philox", seed=7, event0=3)
    for k in d:
        np.testing.assert_array_equal(d[k].cpu().numpy(), h[k], err_msg=k)


def test_target_step_tape_exhaustion(ctx):
    """A tape slot too short for an event's draws: used = -1 and NaN outputs
    on the device entry point, SIMC_E_TAPE from the host entry point."""
    from simc_gfortran_amd import api
    n = 64
    v = _vertex(n, 6, MP, 3000.0, 0.35, 1.5 * math.pi)
    T = api.Target(1.0, 1.0, 0.0723, 61.28, 10.0, 0.0, 0.0, 3, 0)
    P = api.PrepSettings(3, 1, 1, 1, 3000.0, 0.35, 1.5 * math.pi, MP, 0.0, 0.0, (api.C.c_double * 5)())
    tape = W.uniform_tape(n, 3, 5)
    out = ctx.target_prep(T, P, to_dev(v), rng="tape", tape=torch.from_numpy(tape).cuda(), stride=3)
    torch.cuda.synchronize()
    used = out["used"].cpu().numpy()
    assert (used == -1).all()
    assert np.isnan(out["dpp"].cpu().numpy()).all()
    with pytest.raises(api.SimcError) as e:
        ctx.target_prep_host(T, P, v, rng="tape", tape=tape, stride=3)
    assert e.value.code == -7


def test_target_step_argument_errors(ctx):
    from simc_gfortran_amd import api
    v = to_dev(_vertex(16, 7, MP, 3000.0, 0.35, 1.5 * math.pi))
    T = api.Target(1.0, 1.0, 0.0723, 61.28, 10.0, 0.0, 0.0, 3, 0)
    bad = [api.PrepSettings(4, 1, 1, 1, 3000.0, 0.3, 0.0, MP, 0.0, 0.0, (api.C.c_double * 5)()),
           api.PrepSettings(3, 9, 1, 1, 3000.0, 0.3, 0.0, MP, 0.0, 0.0, (api.C.c_double * 5)()),
           api.PrepSettings(3, 1, 1, 1, 3000.0, 0.3, 0.0, 0.0, 0.0, 0.0, (api.C.c_double * 5)())]
    for P in bad:
        with pytest.raises(api.SimcError) as e:
            ctx.target_prep(T, P, v)
        assert e.value.code == -6
    P = api.PrepSettings(3, 1, 0, 1, 3000.0, 0.3, 0.0, MP, 0.0, 0.0, (api.C.c_double * 5)())
    with pytest.raises(api.SimcError):  # a hadron arm without loss reads delta
        ctx.target_prep(T, P, {k: t for k, t in v.items() if k != "delta"})
    with pytest.raises(api.SimcError):
        ctx.target_prep(api.Target(1.0, 1.0, 0.0, 61.28, 10.0, 0.0, 0.0, 3, 0), P, v)


def test_c0_chain_vertex_to_focal_plane(ctx, oracle):
    """HMS 3 GeV protons from a 10 cm LH2 cell: target step then mc_hms, both
    on the counter-based streams, against the oracle's chain."""
    wl = W.HMS_C0
    n = 20000
    v = W.generate_vertex(wl, n, seed=2222)
    th0, phi = math.radians(wl.theta_deg), 1.5 * math.pi
    T = oracle.make_target(**W.LH2_CELL)
    P = oracle.make_prep(3, 1, True, True, p_spec=wl.p_spec, th_spec=th0, phi=phi, mass=wl.mass)
    ev_ref, rc = oracle.target_prep(T, P, v, rng="philox", seed=1001, event0=0)
    assert rc == 0
    ref = oracle.run(1, ev_ref, p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, rng="philox", seed=1001,
                     event0=0, fit_f32=False)
    assert ref["rc"] == 0
    ev = ctx.target_prep(api_target(T), api_prep(P), to_dev(v), rng="philox", seed=1001, event0=0)
    arm_in = {k: ev[k] for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")}
    out = ctx.run(1, arm_in, p_spec=wl.p_spec, m2=wl.m2, ms=True, wcs=True, rng="philox", seed=1001,
                  event0=0, fit="fp64")
    torch.cuda.synchronize()
    g = {k: t.cpu().numpy() for k, t in out.items() if t is not None}
    rep = parity_report(g, ref, 1e-12)
    print(rep)
    assert rep["flag_mismatch"] == 0
    assert all(b == 0 for b in rep["bad"].values()), rep
    assert rep["n_both_ok"] > 0.2 * n
