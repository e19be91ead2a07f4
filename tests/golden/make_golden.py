"""Write the golden vectors of tests/golden/*.npz from the reference itself.

The reference's Fortran hot path (mc_hms, mc_shms, mc_sos, mc_hrsr, mc_hrsl and
everything they call)
is compiled from its own sources by `make -C oracle ref` into
oracle/_ref/libsimcref.so; this script drives it over small event samples on
per-event uniform tapes and stores inputs, tape recipe and every output.  The
tapes are regenerated from their seed by the tests (sha256 recorded here).

    python tests/golden/make_golden.py [prefix] # arm goldens (*.npz), optionally
                                                # only the cases named prefix*
    python tests/golden/make_golden.py target   # target step (target/*.npz):
        the reference's trip_thru_target and target_musc (target.f) over
        vertices on per-event tapes
    python tests/golden/make_golden.py field    # target-field tracking (field/*.npz):
        the reference's trgInit, track_from_tgt and track_to_tgt (trg_track.f)
    python tests/golden/make_golden.py rad      # radiative tails (gen/rad_*.npz): the
        reference's generate with using_rad (generate_rad, brem, bremos,
        radc_init_ev) + complete_main on per-event tapes, one process per case
    python tests/golden/make_golden.py lfit     # track fit (lfit/ref_lfit.npz): the
        reference's LFIT (cern/lfit.f) over seeded REAL*4 point sets
    python tests/golden/make_golden.py sieve    # mock sieve slit (sieve/*.npz): the
        reference built with use_sieve / use_2sieve / both on (make -C oracle sieve)
    python tests/golden/make_golden.py ntuple   # .bin ntuple (ntuple/*.npz): the file the
        reference's NtupleInit / NtupleClose write for the selected rows of
        seeded tables (elastic; pion with the target-field columns)
    python tests/golden/make_golden.py gen      # generator (gen/*.npz):
        the reference's physics_angles, spectrometer_angles (event.f), sigep and
        fofa_best_fit (physics_proton.f) over sampled arguments, and its
        generate + complete_main for H(e,e'p) elastic on per-event tapes
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import oracle as O  # noqa: E402
from simc_gfortran_amd import workloads as W  # noqa: E402

MPI = 139.57018

# name, arm, workload (event distribution), overrides
CASES = [
    ("hms_proton_ms_wcs", 1, W.HMS_C0, dict(p_spec=3000.0, m2=W.MP ** 2)),
    ("shms_electron_ms_wcs", 5, W.SHMS_C1, dict(p_spec=6000.0, m2=W.ME ** 2)),
    ("hms_proton_no_ms_no_wcs", 1, W.HMS_C0, dict(p_spec=3000.0, m2=W.MP ** 2, ms=False, wcs=False)),
    ("shms_electron_no_ms_no_wcs", 5, W.SHMS_C1, dict(p_spec=6000.0, m2=W.ME ** 2, ms=False, wcs=False)),
    ("hms_pion_coll_decay", 1, W.HMS_C0, dict(p_spec=2000.0, m2=MPI ** 2, coll=True, decay=True, ctau=780.45)),
    ("shms_pion_coll_decay", 5, W.SHMS_C1, dict(p_spec=2000.0, m2=MPI ** 2, coll=True, decay=True, ctau=780.45)),
    ("sos_proton_ms_wcs", 2, W.SOS_1p5, dict(p_spec=1500.0, m2=W.MP ** 2)),
    ("sos_electron_no_ms_no_wcs", 2, W.SOS_1p5, dict(p_spec=1200.0, m2=W.ME ** 2, ms=False, wcs=False)),
    ("sos_pion_decay", 2, W.SOS_1p5, dict(p_spec=1000.0, m2=MPI ** 2, decay=True, ctau=780.45)),
    ("hrsr_proton_ms_wcs", 3, W.HRSR_2, dict(p_spec=2000.0, m2=W.MP ** 2)),
    ("hrsl_electron_ms_wcs", 4, W.HRSL_3, dict(p_spec=3000.0, m2=W.ME ** 2)),
    ("hrsl_electron_no_ms_no_wcs", 4, W.HRSL_3, dict(p_spec=3000.0, m2=W.ME ** 2, ms=False, wcs=False)),
    ("hrsr_pion_decay", 3, W.HRSR_2, dict(p_spec=1500.0, m2=MPI ** 2, decay=True, ctau=780.45)),
]
N = 1500
STRIDE = 2048


def tape_for(i):
    return W.uniform_tape(N, STRIDE, 900 + i)


def main(only=None):
    for i, (name, arm, wl, kw) in enumerate(CASES):
        if only and not name.startswith(only):
            continue
        ev = W.generate(wl, N, seed=wl.seed + 17 * i)
        tape = tape_for(i)
        args = dict(ms=True, wcs=True)
        args.update(kw)
        out = O.ref_run(arm, ev, tape=tape, stride=STRIDE, **args)
        assert out["rc"] == 0, (name, out["rc"])
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"),
            arm=arm, tape_seed=900 + i, stride=STRIDE,
            tape_sha256=hashlib.sha256(tape.tobytes()).hexdigest(),
            settings=np.array([args.get("p_spec"), args.get("m2"), float(args.get("ms", True)),
                               float(args.get("wcs", True)), float(args.get("decay", False)),
                               float(args.get("coll", False)), args.get("ctau", 0.0)]),
            **{"in_" + k: ev[k] for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")},
            **{"out_" + k: out[k] for k in ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp",
                                             "dy_fp", "pathlen", "resmult", "m2", "ok", "used")},
            counters=out["counters"], stop_mask=out["stop_mask"])
        print(name, "accepted", int(out["ok"].sum()), "of", N)


# The mock sieve slit: the reference built with use_sieve, use_2sieve or both
# on (oracle/_ref/libsimcref_sieve1.so, _sieve2.so, _sieve3.so,
# `make -C oracle sieve`).
SIEVE_CASES = [
    ("hms_proton_sieve", 1, W.HMS_C0, dict(p_spec=3000.0, m2=W.MP ** 2, sieve=1)),
    ("shms_electron_sieve", 5, W.SHMS_C1, dict(p_spec=6000.0, m2=W.ME ** 2, sieve=1)),
    ("shms_electron_sieve2", 5, W.SHMS_C1, dict(p_spec=6000.0, m2=W.ME ** 2, sieve=2)),
    ("sos_proton_sieve", 2, W.SOS_1p5, dict(p_spec=1500.0, m2=W.MP ** 2, sieve=1)),
    ("hrsr_proton_sieve", 3, W.HRSR_2, dict(p_spec=2000.0, m2=W.MP ** 2, sieve=1)),
    ("hrsl_electron_sieve", 4, W.HRSL_3, dict(p_spec=3000.0, m2=W.ME ** 2, sieve=1)),
    ("shms_electron_sieve3", 5, W.SHMS_C1, dict(p_spec=6000.0, m2=W.ME ** 2, sieve=3)),
]


def main_sieve():
    os.makedirs(os.path.join(HERE, "sieve"), exist_ok=True)
    for i, (name, arm, wl, kw) in enumerate(SIEVE_CASES):
        ev = W.generate(wl, N, seed=wl.seed + 31 * i)
        tape = W.uniform_tape(N, STRIDE, 950 + i)
        args = dict(ms=True, wcs=True)
        args.update(kw)
        out = O.ref_run(arm, ev, tape=tape, stride=STRIDE, **args)
        assert out["rc"] == 0, (name, out["rc"])
        np.savez_compressed(
            os.path.join(HERE, "sieve", name + ".npz"),
            arm=arm, tape_seed=950 + i, stride=STRIDE, sieve=args["sieve"],
            tape_sha256=hashlib.sha256(tape.tobytes()).hexdigest(),
            settings=np.array([args.get("p_spec"), args.get("m2"), 1.0, 1.0, 0.0, 0.0, 0.0]),
            **{"in_" + k: ev[k] for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")},
            **{"out_" + k: out[k] for k in ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp",
                                             "dy_fp", "pathlen", "resmult", "m2", "ok", "used")},
            counters=out["counters"], stop_mask=out["stop_mask"])
        print(name, "accepted", int(out["ok"].sum()), "of", N)


# Target step (trip_thru_target + target_musc of the reference's target.f):
# name, narm, spectrometer, mass, target (A, Z, rho, length, angle, zoffset,
# can), energy range (MeV), angle range (rad)
ME_ = 0.51099906
MP_ = 938.27231
TARGET_CASES = [
    ("hadron_hms_lh2_cryo2017", 3, 1, MP_, (1.0, 1.0, 0.0723, 10.0, 0.0, 0.0, 3), (1500., 5000.), (0.2, 0.9)),
    ("hadron_hms_lh2_beercan", 3, 1, MP_, (1.0, 1.0, 0.0723, 4.0, 0.0, 0.0, 1), (1500., 5000.), (0.2, 0.9)),
    ("electron_shms_lh2_pudding", 2, 5, ME_, (1.0, 1.0, 0.0723, 4.0, 0.0, 0.0, 2), (2000., 9000.), (0.1, 0.6)),
    ("electron_shms_ld2_cryo2017", 2, 5, ME_, (2.0, 1.0, 0.167, 10.0, 0.0, 0.5, 3), (2000., 9000.), (0.1, 0.6)),
    ("electron_sos_lh2_beercan", 2, 2, ME_, (1.0, 1.0, 0.0723, 4.0, 0.0, 0.0, 1), (500., 1700.), (0.3, 1.2)),
    ("hadron_hrs_lh2_cryo2017", 3, 3, MP_, (1.0, 1.0, 0.0723, 10.0, 0.0, 0.0, 3), (1500., 5000.), (0.2, 0.9)),
    ("electron_hrsl_lh2_pudding", 2, 4, ME_, (1.0, 1.0, 0.0723, 4.0, 0.0, 0.0, 2), (2000., 9000.), (0.1, 0.6)),
    ("hadron_shms_carbon_tilted", 3, 5, MP_, (12.0, 6.0, 2.26, 0.3, 0.35, 0.0, 0), (1500., 5000.), (0.2, 0.9)),
    ("beam_lh2_cryo2017", 1, 1, ME_, (1.0, 1.0, 0.0723, 10.0, 0.0, 0.0, 3), (10600., 10600.), (0.0, 0.0)),
    ("beam_gold_foil", 1, 1, ME_, (197.0, 79.0, 19.32, 0.05, 0.0, 0.0, 0), (6400., 6400.), (0.0, 0.0)),
]
TARGET_N = 1000
TARGET_STRIDE = 64


def target_tape_for(i):
    return W.uniform_tape(TARGET_N, TARGET_STRIDE, 700 + i)


def main_target():
    out_dir = os.path.join(HERE, "target")
    os.makedirs(out_dir, exist_ok=True)
    for i, (name, narm, spec, mass, tg, erange, trange) in enumerate(TARGET_CASES):
        A, Z, rho, length, angle, zoffset, can = tg
        T = O.make_target(A=A, Z=Z, rho=rho, length=length, angle=angle, zoffset=zoffset, can=can)
        rng = np.random.default_rng(4000 + i)
        zpos = rng.uniform(-0.5, 0.5, TARGET_N) * length
        energy = rng.uniform(*erange, TARGET_N)
        theta = rng.uniform(*trange, TARGET_N)
        e_musc = rng.uniform(*erange, TARGET_N)  # energy after radiation (orig)
        # the beam's scattering uses Ein itself as the momentum (simc.f:1365)
        p_musc = e_musc.copy() if narm == 1 else np.sqrt(e_musc ** 2 - mass ** 2)
        beta = p_musc / e_musc if narm == 3 else np.ones(TARGET_N)
        tape = target_tape_for(i)
        r = O.ref_target_step(T, narm, spec, mass, zpos, energy, theta, tape, TARGET_STRIDE, p_musc=p_musc,
                              beta_musc=beta)
        assert (r["used"] >= 0).all(), name
        np.savez_compressed(
            os.path.join(out_dir, name + ".npz"),
            narm=narm, spectrometer=spec, mass=mass,
            target=np.array([A, Z, rho, T.X0, length, angle, zoffset]), can=can,
            tape_seed=700 + i, stride=TARGET_STRIDE,
            tape_sha256=hashlib.sha256(tape.tobytes()).hexdigest(),
            zpos=zpos, energy=energy, theta=theta, e_musc=e_musc, p_musc=p_musc,
            **{k: r[k] for k in ("eloss", "teff", "dang1", "dang2", "used")})
        print(name, "mean eloss", float(r["eloss"].mean()), "mean teff", float(r["teff"].mean()))


# Target-field tracking (trg_track.f): track_from_tgt over vertices, and
# track_to_tgt over the accepted events of an oracle run of the arm.
# name, map (None = uniform test field, "0" = zero field), theta_e, theta_p (deg)
FIELD_FROM_CASES = [
    ("from_map_proton", "map", 0.0, 10.0, +1, MP_, (800., 3000.)),
    ("from_map_electron", "map", -6.0, 0.0, -1, ME_, (800., 4000.)),
    ("from_uniform_proton", None, 0.0, 25.0, +1, MP_, (600., 2000.)),
    ("from_zero_field", "0", 0.0, 0.0, +1, MP_, (1000., 2000.)),
]
# name, arm, workload, spect, mass, charge sign, arm angle (deg), map angles (deg)
FIELD_TO_CASES = [
    ("to_hms_proton", 1, W.HMS_C0, +1, MP_, +1.0, 20.0, (10.0, 10.0)),
    ("to_shms_electron", 5, W.SHMS_C1, -1, ME_, -1.0, 12.5, (-8.0, 0.0)),
    ("to_hrsl_electron", 4, W.HRSL_3, -1, ME_, -1.0, 15.0, (5.0, 0.0)),
]
NF = 1000


def main_field():
    out_dir = os.path.join(HERE, "field")
    os.makedirs(out_dir, exist_ok=True)
    for i, (name, mp, te, tp, spect, mass, prange) in enumerate(FIELD_FROM_CASES):
        rng = np.random.default_rng(700 + i)
        ev = dict(x=rng.uniform(-0.5, 0.5, NF), y=rng.uniform(-0.5, 0.5, NF), z=rng.uniform(-1.5, 1.5, NF),
                  dx=rng.uniform(-0.06, 0.06, NF), dy=rng.uniform(-0.04, 0.04, NF))
        mom = spect * rng.uniform(*prange, NF)
        path = O.FIELD_MAP if mp == "map" else mp
        r = O.ref_track_from_tgt(path, te, tp, ev, mom, mass, spect)
        np.savez_compressed(os.path.join(out_dir, name + ".npz"), kind="from", map=str(mp), theta_e=te,
                            theta_p=tp, spect=spect, mass=mass, mom=mom,
                            **{"in_" + k: ev[k] for k in ev}, **{"out_" + k: r[k] for k in r})
        print(name, "ok", int(r["ok"].sum()), "of", NF)
    for i, (name, arm, wl, spect, mass, sgn, th_deg, (te, tp)) in enumerate(FIELD_TO_CASES):
        rng = np.random.default_rng(800 + i)
        ev = W.generate(wl, 4 * NF, seed=900 + i)
        res = O.run(arm, ev, p_spec=wl.p_spec, m2=mass * mass, rng="philox", seed=31 + i, fit_f32=False)
        sel = np.flatnonzero(res["ok"] == 1)[:NF]
        trk = {k: res[k][sel] for k in ("x_fp", "dx_fp", "y_fp", "dy_fp")}
        if arm in (3, 4):  # /track/ holds the fitted y, y_fp carries the VDC offset
            trk["y_fp"] = trk["y_fp"] + (0.48 if arm == 3 else 0.78)
        first = dict(delta=res["dpp"][sel], y=res["y"][sel], dx=res["dxdz"][sel], dy=res["dydz"][sel])
        m = len(sel)
        frx = rng.uniform(-0.2, 0.2, m)
        fry = rng.uniform(-0.2, 0.2, m)
        mom = sgn * wl.p_spec * (1 + first["delta"] / 100)
        th = np.radians(th_deg)
        st = np.sin(th) if arm in (1, 3) else -np.sin(th)  # simc.f:1573-1584
        r = O.ref_track_to_tgt(O.FIELD_MAP, te, tp, arm, trk, first, frx, -fry, mom, mass, np.cos(th), st, spect)
        np.savez_compressed(os.path.join(out_dir, name + ".npz"), kind="to", map="map", theta_e=te, theta_p=tp,
                            arm=arm, spect=spect, mass=mass, mom=mom, frx=frx, fry=-fry, ctheta=np.cos(th),
                            stheta=st, **{"trk_" + k: trk[k] for k in trk},
                            **{"in_" + k: first[k] for k in first}, **{"out_" + k: r[k] for k in r})
        print(name, "ok", int(r["ok"].sum()), "of", m)


NG = 4000


def main_gen():
    out_dir = os.path.join(HERE, "gen")
    os.makedirs(out_dir, exist_ok=True)
    rng = np.random.default_rng(1200)
    # central angles of an arm on either side of the beam (phi0 pi/2 or 3 pi/2),
    # slopes inside and beyond the spectrometers' acceptances, and dx = 0
    th0 = rng.uniform(0.05, 1.4, NG)
    ph0 = np.where(rng.random(NG) < 0.5, np.pi / 2, 1.5 * np.pi)
    dx = rng.uniform(-0.15, 0.15, NG)
    dx[::97] = 0.0
    dy = rng.uniform(-0.1, 0.1, NG)
    th = np.zeros(NG)
    ph = np.zeros(NG)
    sdx = np.zeros(NG)
    sdy = np.zeros(NG)
    for i in range(NG):
        th[i], ph[i] = O.ref_physics_angles(th0[i], ph0[i], dx[i], dy[i])
        sdx[i], sdy[i] = O.ref_spectrometer_angles(th0[i], ph0[i], th[i], ph[i])
    np.savez_compressed(os.path.join(out_dir, "angles.npz"), theta0=th0, phi0=ph0, dx=dx, dy=dy, theta=th,
                        phi=ph, back_dx=sdx, back_dy=sdy)
    # elastic kinematics over beam energies 0.5-11 GeV and electron angles 3-80 deg
    ein = rng.uniform(500.0, 11000.0, NG)
    the = rng.uniform(0.05, 1.4, NG)
    mp = 938.27231
    e = ein * mp / (mp + ein * (1 - np.cos(the)))
    q2 = 2 * ein * e * (1 - np.cos(the))
    sig = np.array([O.ref_sigep(q2[i], e[i], the[i], ein[i]) for i in range(NG)])
    qsq = -rng.uniform(0.0, 25.0, NG) * 1e6 / 197.327053 ** 2
    qsq[::101] = 0.0
    ff = np.array([O.ref_fofa(v) for v in qsq])
    np.savez_compressed(os.path.join(out_dir, "sigep.npz"), Q2=q2, E=e, theta=the, Ein=ein, sigep=sig, qsquar=qsq,
                        GE=ff[:, 0], GM=ff[:, 1])
    print("gen goldens", NG, "angles,", NG, "sigep")
    # the reference's generate + complete_main (H(e,e'p) elastic, radiation
    # off) on per-event tapes, one case per raster pattern
    tgt = O.make_target(**W.LH2_CELL)
    for pat in (0, 1, 2, 3):
        kw = dict(W.HEP_ELASTIC)
        kw.update(fr_pattern=pat, fr1=0.1, fr2=0.2, using_coulomb=1, coulomb=0.7, using_eloss=int(pat != 1))
        tape = W.uniform_tape(GEN_N, GEN_STRIDE, GEN_SEED + pat)
        r, rc = O.ref_gen_hep(tgt, O.make_hep(**kw), GEN_N, tape, GEN_STRIDE)
        assert rc == 0
        np.savez_compressed(os.path.join(out_dir, f"hep_pattern{pat}.npz"), tape_seed=GEN_SEED + pat,
                            stride=GEN_STRIDE, tape_sha256=hashlib.sha256(tape.tobytes()).hexdigest(),
                            settings=np.array(repr(kw)), **{"out_" + k: v for k, v in r.items()})
        print("hep_pattern", pat, "ok", int(r["ok"].sum()), "of", GEN_N)
    # degenerate settings: xptar range of zero width (physics_angles' dx = 0
    # branch), no beam spread or spot, an annulus of zero width, losses off
    kw = dict(W.HEP_ELASTIC)
    kw.update(fr_pattern=2, fr1=0.1, fr2=0.1, xwid=0.0, ywid=0.0, dEbeam=0.0, e_xptar=(0.0, 0.0), using_eloss=0)
    tape = W.uniform_tape(GEN_N, GEN_STRIDE, GEN_SEED + 10)
    r, rc = O.ref_gen_hep(tgt, O.make_hep(**kw), GEN_N, tape, GEN_STRIDE)
    assert rc == 0
    np.savez_compressed(os.path.join(out_dir, "hep_pattern2_edges.npz"), tape_seed=GEN_SEED + 10, stride=GEN_STRIDE,
                        tape_sha256=hashlib.sha256(tape.tobytes()).hexdigest(), settings=np.array(repr(kw)),
                        **{"out_" + k: v for k, v in r.items()})
    print("hep_pattern2_edges ok", int(r["ok"].sum()), "of", GEN_N)


GEN_N = 2000
GEN_STRIDE = 96
GEN_SEED = 1300

# Radiative tails (generate with using_rad: generate_rad of radc.f, brem.f,
# radc_init_ev of init.f): name, radiative settings over W.HEP_RAD, generator
# settings over W.HEP_ELASTIC.  Each case runs in a fresh process: with
# intcor_mode 0 the reference's radc_init_ev reads a saved local
# (dsoft_prime) that only brem/bremos set, so an earlier case would leak
# into it; in a fresh process it is 0.
RAD_CASES = [
    ("rad_default", {}, {}),
    ("rad_onshell_brem", dict(use_offshell_rad=0), {}),
    ("rad_schwinger", dict(intcor_mode=0), {}),
    ("rad_flag1", dict(rad_flag=1), {}),
    ("rad_flag2_friedrich", dict(rad_flag=2, extrad_flag=3), {}),
    ("rad_flag3", dict(rad_flag=3), {}),
    ("rad_extrad3", dict(extrad_flag=3), {}),
    ("rad_expon", dict(use_expon=1), {}),
    ("rad_hardwired", dict(Egamma_gen_max=50.0), {}),
    ("rad_no_proton_tail", dict(doing_tail=(1, 1, 0)), {}),
    ("rad_edges_as_built", dict(W.hep_rad_limits(W.HEP_ELASTIC, Em_edge_from_cuts=False, dE_edge_test=5.0)), {}),
    ("rad_raster1_no_eloss", {}, dict(fr_pattern=1, using_eloss=0)),
]
RAD_N = 800
RAD_STRIDE = 160
RAD_SEED = 1600


def rad_settings(i):
    name, rkw, hkw = RAD_CASES[i]
    rad = dict(W.HEP_RAD)
    rad.update(rkw)
    rad.update(etatzai=O.rad_etatzai(W.LH2_CELL["Z"]), targ_A=W.LH2_CELL["A"])
    hep = dict(W.HEP_ELASTIC)
    hep.update(hkw)
    return name, rad, hep


def main_rad_case(i):
    name, rad, hep = rad_settings(i)
    out_dir = os.path.join(HERE, "gen")
    tgt = O.make_target(**W.LH2_CELL)
    tape = W.uniform_tape(RAD_N, RAD_STRIDE, RAD_SEED + i)
    r, rc = O.ref_gen_hep_rad(tgt, O.make_hep(**hep), O.make_rad(**rad), RAD_N, tape, RAD_STRIDE)
    assert rc == 0, (name, rc)
    np.savez_compressed(os.path.join(out_dir, name + ".npz"), case=i, tape_seed=RAD_SEED + i, stride=RAD_STRIDE,
                        tape_sha256=hashlib.sha256(tape.tobytes()).hexdigest(),
                        settings=np.array(repr((rad, hep))), **{"out_" + k: v for k, v in r.items()})
    print(name, "ok", int(r["ok"].sum()), "of", RAD_N, "tails", np.bincount(r["ntail"][r["ok"] == 1], minlength=4))


def main_rad():
    import subprocess
    for i in range(len(RAD_CASES)):
        subprocess.run([sys.executable, __file__, "rad_case", str(i)], check=True)


# LFIT (cern/lfit.f, compiled into oracle/_ref as gfortran sees it): the
# reference's A, B, E over seeded REAL*4 point sets (W.lfit_point_sets).
LFIT_N = 12000
LFIT_SEED = 1500


def main_lfit():
    X, Y = W.lfit_point_sets(LFIT_N, LFIT_SEED)
    A, B, E = O.ref_lfit(X, Y)
    np.savez_compressed(os.path.join(HERE, "lfit", "ref_lfit.npz"), seed=LFIT_SEED, n=LFIT_N,
                        points_sha256=hashlib.sha256(X.tobytes() + Y.tobytes()).hexdigest(), A=A, B=B, E=E)
    print("lfit", LFIT_N, "sets; A = B = 0 for", int(((A == 0) & (B == 0)).sum()))


# ntuple: the reference's NtupleInit / record writes / NtupleClose over the
# selected rows of seeded tables, for the GPU writer to reproduce byte for byte.
NTUPLE_CASES = [("elastic", False, 500, 61), ("pion", True, 400, 62)]


def main_ntuple():
    import tempfile
    from simc_gfortran_amd import api
    os.makedirs(os.path.join(HERE, "ntuple"), exist_ok=True)
    for mode, tgt, n, seed in NTUPLE_CASES:
        ncols = len(api.ntuple_tags(mode, tgt_field=tgt))
        rng = np.random.default_rng(seed)
        rows = rng.normal(size=(n, ncols)) * 10.0 ** rng.integers(-6, 6, size=(n, ncols))
        rows[0, :4] = [0.0, -0.0, 1e-50, -3.5e38]
        mask = (rng.random(n) < 0.6).astype(np.uint8)
        with tempfile.TemporaryDirectory(prefix="nt", dir="/tmp") as d:
            path = os.path.join(d, "r.bin")
            O.ref_ntuple(path, mode, tgt, rows[mask == 1])
            data = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
        name = f"{mode}{'_tgtfield' if tgt else ''}"
        np.savez_compressed(os.path.join(HERE, "ntuple", name + ".npz"), mode=mode, tgt_field=tgt, rows=rows,
                            mask=mask, ref_bin=data)
        print("ntuple", name, n, "rows,", int(mask.sum()), "written,", len(data), "bytes")


if __name__ == "__main__":
    if sys.argv[1:] == ["sieve"]:
        main_sieve()
    elif sys.argv[1:] == ["ntuple"]:
        main_ntuple()
    elif sys.argv[1:] == ["gen"]:
        main_gen()
    elif sys.argv[1:] == ["target"]:
        main_target()
    elif sys.argv[1:] == ["field"]:
        main_field()
    elif sys.argv[1:] == ["lfit"]:
        main_lfit()
    elif sys.argv[1:] == ["rad"]:
        main_rad()
    elif sys.argv[1:2] == ["rad_case"]:
        main_rad_case(int(sys.argv[2]))
    else:
        main(sys.argv[1] if sys.argv[1:] else None)
