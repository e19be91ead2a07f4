"""The oracle's target step (trip_thru_target, enerloss_new, target_musc and
montecarlo()'s vertex-to-arm transform) on the CPU:

* golden vectors written by the reference's own target.f (compiled from its
  sources into oracle/_ref, tests/golden/make_golden.py target): energy loss,
  radiation lengths, scattering angles and uniforms drawn, bit for bit;
* radiation lengths of init.f's Tsai formula against the tabulated values;
* closed forms of the transform and of delta, and the width of the sampled
  scattering angles (Lynch-Dahl) on the counter-based stream.
"""
import glob
import hashlib
import math
import os

import numpy as np
import pytest

from simc_gfortran_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
TARGET_GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "target", "*.npz")))
ME = 0.51099906
MP = 938.27231


def golden_case(oracle, g):
    """Target, prep settings, vertex dict and tape of one golden file."""
    A, Z, rho, X0, length, angle, zoffset = g["target"]
    T = oracle.make_target(A=A, Z=Z, rho=rho, X0=X0, length=length, angle=angle, zoffset=zoffset,
                           can=int(g["can"]))
    narm = int(g["narm"])
    n = len(g["zpos"])
    P = oracle.make_prep(narm, int(g["spectrometer"]), True, True, p_spec=3000.0, th_spec=0.3,
                         phi=1.5 * math.pi, mass=float(g["mass"]), beam_energy=float(g["energy"][0]))
    v = {"z": g["zpos"] + zoffset, "E": g["e_musc"], "p": g["p_musc"], "E_vertex": g["energy"],
         "theta": g["theta"], "x": np.zeros(n), "y": np.zeros(n), "xptar": np.zeros(n),
         "yptar": np.zeros(n)}
    tape = W.uniform_tape(n, int(g["stride"]), int(g["tape_seed"]))
    assert hashlib.sha256(tape.tobytes()).hexdigest() == str(g["tape_sha256"])
    return T, P, v, tape


@pytest.mark.parametrize("path", TARGET_GOLDEN, ids=[os.path.basename(p)[:-4] for p in TARGET_GOLDEN])
def test_oracle_target_step_matches_reference_golden(oracle, path):
    g = np.load(path)
    T, P, v, tape = golden_case(oracle, g)
    out, rc = oracle.target_prep(T, P, v, rng="tape", tape=tape, stride=int(g["stride"]))
    assert rc == 0
    for k in ("eloss", "teff", "dang1", "dang2", "used"):
        np.testing.assert_array_equal(out[k], g[k], err_msg=k)


def test_target_golden_cover_every_branch():
    """The fixtures reach every can shape, both end-cap regions of the 2017
    cell, every window table and the beam."""
    seen = set()
    for p in TARGET_GOLDEN:
        g = np.load(p)
        seen.add((int(g["narm"]), int(g["spectrometer"]) if int(g["narm"]) > 1 else 0, int(g["can"])))
        if int(g["can"]) == 3 and int(g["narm"]) > 1:
            length = g["target"][4]
            ecir = 1.315 * 2.54
            side = (g["zpos"] + length / 2 + ecir / np.tan(g["theta"])) < length - ecir
            assert side.any() and (~side).any()
    assert {c for _, _, c in seen} >= {0, 1, 2, 3}
    assert {s for a, s, _ in seen if a > 1} >= {1, 2, 3, 4, 5}
    assert {a for a, _, _ in seen} == {1, 2, 3}


@pytest.mark.parametrize("A,Z,x0_table", [
    (12.0107, 6.0, 42.70), (26.9815, 13.0, 24.01), (63.546, 29.0, 12.86), (207.2, 82.0, 6.37)])
def test_tsai_radiation_length_matches_tables(oracle, A, Z, x0_table):
    """init.f's Tsai formula (with the Coulomb correction f(Z)) reproduces the
    tabulated radiation lengths to 0.3%; H, D and 4He take the table."""
    assert abs(oracle.target_X0(A, Z) / x0_table - 1.0) < 3e-3
    assert oracle.target_X0(1.0, 1.0) == 61.28
    assert oracle.target_X0(2.0, 1.0) == 122.4
    assert oracle.target_X0(4.0, 2.0) == 94.32


def _vertex(n, seed, mass=MP, p0=3000.0):
    rng = np.random.default_rng(seed)
    delta = rng.uniform(-10, 10, n)
    p = p0 * (1 + delta / 100)
    return {"x": rng.normal(0, 0.05, n), "y": rng.normal(0, 0.05, n), "z": rng.uniform(-5, 5, n),
            "E": np.sqrt(p * p + mass * mass), "p": p, "theta": rng.uniform(0.25, 0.5, n), "delta": delta,
            "xptar": rng.uniform(-0.05, 0.05, n), "yptar": rng.uniform(-0.04, 0.04, n)}


def test_transform_without_smearing_or_loss(oracle):
    """No loss, no scattering: delta passes through; a vertex on the beam line
    (x_tgt = 0) seen from angle th lies at y = -z sin(th) sin(phi), z_arm =
    z cos(th); after the drift the track passes through that point."""
    n = 500
    v = _vertex(n, 3)
    v["x"] = np.zeros(n)
    th, phi = 0.35, 1.5 * math.pi
    T = oracle.make_target()
    P = oracle.make_prep(3, 1, False, False, p_spec=3000.0, th_spec=th, phi=phi, mass=MP)
    out, rc = oracle.target_prep(T, P, v, rng="philox", seed=1)
    assert rc == 0
    np.testing.assert_array_equal(out["dpp"], v["delta"])
    assert (out["dang1"] == 0).all() and (out["dang2"] == 0).all()
    np.testing.assert_array_equal(out["dxdz"], v["xptar"])
    np.testing.assert_array_equal(out["dydz"], v["yptar"])
    x_arm = -v["y"]
    y_arm = -v["z"] * math.sin(th) * math.sin(phi)
    z_arm = v["z"] * math.cos(th)
    np.testing.assert_allclose(out["x"] + z_arm * out["dxdz"], x_arm, rtol=0, atol=1e-13)
    np.testing.assert_allclose(out["y"] + z_arm * out["dydz"], y_arm, rtol=0, atol=1e-13)
    np.testing.assert_array_equal(out["fry"], out["x"])
    assert (out["z"] == 0).all()
    # the loss is still sampled (and zeroed): the draws of the window materials
    assert (out["eloss"] == 0).all() and (out["used"] > 0).all()


def test_offsets_and_beam_angles(oracle):
    """Offsets are subtracted before the drift; the beam's angles add to
    yptar and, scaled by cos(th), to xptar (simc.f:1398-1399, 1427-1431)."""
    n = 300
    v = _vertex(n, 4)
    v["dang_yp"] = np.full(n, 1e-3)
    v["dang_xp"] = np.full(n, -2e-3)
    th = 0.4
    off = (0.1, -0.2, 0.3, 0.004, -0.005)
    T = oracle.make_target()
    base, _ = oracle.target_prep(T, oracle.make_prep(3, 1, False, False, p_spec=3000., th_spec=th,
                                                     phi=1.5 * math.pi, mass=MP),
                                 {k: v[k] for k in v if not k.startswith("dang")}, rng="philox", seed=2)
    P = oracle.make_prep(3, 1, False, False, p_spec=3000., th_spec=th, phi=1.5 * math.pi, mass=MP, offset=off)
    out, rc = oracle.target_prep(T, P, v, rng="philox", seed=2)
    assert rc == 0
    np.testing.assert_allclose(out["dydz"], v["yptar"] + 1e-3 - off[4], rtol=0, atol=1e-15)
    np.testing.assert_allclose(out["dxdz"], v["xptar"] - 2e-3 * math.cos(th) - off[3], rtol=0, atol=1e-15)
    # undo the drift of both runs: positions differ by the offsets
    sinp = math.sin(1.5 * math.pi)
    z_arm = v["z"] * math.cos(th) + v["x"] * math.sin(th) * sinp
    xb = base["x"] + z_arm * base["dxdz"]
    xo = out["x"] + (z_arm - off[2]) * out["dxdz"]
    np.testing.assert_allclose(xo, xb - off[0], rtol=0, atol=1e-12)


def test_delta_after_loss(oracle):
    """Hadron: delta from the momentum after the sampled loss; electron: from
    the energy after loss and Coulomb energy (simc.f:1379-1384, 1653-1654)."""
    n = 400
    T = oracle.make_target()
    v = _vertex(n, 5)
    out, rc = oracle.target_prep(T, oracle.make_prep(3, 1, True, True, p_spec=3000., th_spec=0.3,
                                                     phi=1.5 * math.pi, mass=MP), v, rng="philox", seed=3)
    assert rc == 0 and (out["eloss"] > 0).all()
    p_after = np.sqrt((v["E"] - out["eloss"]) ** 2 - MP * MP)
    np.testing.assert_allclose(out["dpp"], (p_after - 3000.) / 3000. * 100, rtol=0, atol=1e-11)
    ve = _vertex(n, 6, mass=ME, p0=6000.)
    out, rc = oracle.target_prep(T, oracle.make_prep(2, 5, True, True, p_spec=6000., th_spec=0.2,
                                                     phi=0.5 * math.pi, mass=ME, coulomb=0.7), ve,
                                 rng="philox", seed=3)
    assert rc == 0
    np.testing.assert_allclose(out["dpp"], (ve["E"] - out["eloss"] - 0.7 - 6000.) / 6000. * 100, rtol=0,
                               atol=1e-11)


def test_scattering_angles_follow_lynch_dahl(oracle):
    """dang / theta_sigma(p, beta, teff) is a standard normal truncated at 3.5
    (target.f:565-570), on the counter-based stream."""
    from scipy import stats
    n = 40000
    v = _vertex(n, 7)
    T = oracle.make_target()
    out, rc = oracle.target_prep(T, oracle.make_prep(3, 1, True, True, p_spec=3000., th_spec=0.3,
                                                     phi=1.5 * math.pi, mass=MP), v, rng="philox", seed=11)
    assert rc == 0
    beta = v["p"] / v["E"]
    sig = 13.6 / (v["p"] * beta) * np.sqrt(out["teff"]) * (1 + 0.088 * np.log10(out["teff"] / beta ** 2))
    for k in ("dang1", "dang2"):
        u = out[k] / sig
        assert np.abs(u).max() <= 3.5
        tn = stats.truncnorm(-3.5, 3.5)
        assert stats.kstest(u, tn.cdf).pvalue > 1e-3, k


def test_unknown_spectrometer_is_rejected(oracle):
    v = _vertex(4, 8)
    out, rc = oracle.target_prep(oracle.make_target(), oracle.make_prep(3, 7, True, True, p_spec=3000.,
                                                                          mass=MP), v, rng="philox")
    assert rc == -1
