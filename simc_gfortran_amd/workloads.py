"""Synthetic event workloads for the spectrometer path (the inputs the caller
of mc_hms / mc_shms prepares in simc.f:1403-1443).

Each workload draws target-frame vertices and angles with numpy from a fixed
seed and converts them to spectrometer TRANSPORT coordinates exactly as
montecarlo() does before calling the arm (simc.f:1411-1441):

    x_arm = -y_tgt
    y_arm = -x_tgt cos(th) - z_tgt sin(th) sin(phi)
    z_arm =  z_tgt cos(th) + x_tgt sin(th) sin(phi)
    (x_arm, y_arm) -= z_arm (dx/dz, dy/dz)        # drift back to z = 0
    fry = x_arm                                    # xtar_init, simc.f:1441

with phi = 3 pi/2 for a left-side spectrometer (HMS) and pi/2 for the SOS and
the SHMS (dbase.f sets spec%p%phi / spec%e%phi this way).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

ME = 0.51099906        # constants.inc:17
MP = 938.27231         # constants.inc:19
MPI = 139.57018        # constants.inc:23


@dataclass
class Workload:
    name: str
    arm: int                 # 1 = HMS, 2 = SOS, 3 = HRS-R, 4 = HRS-L, 5 = SHMS
    n: int
    p_spec: float            # MeV/c
    theta_deg: float
    mass: float              # MeV
    dpp: tuple               # uniform range, %
    dxdz: tuple              # uniform range, rad
    dydz: tuple              # uniform range, rad
    z: tuple = (-5.0, 5.0)   # cm, 10 cm LH2 cell
    sigma_xy: float = 0.01   # cm
    wire_sigma: float = 0.0  # cm; 0 keeps the spectrometer's own value
    seed: int = 1001
    extra: dict = field(default_factory=dict)

    @property
    def m2(self):
        return self.mass * self.mass


# BASELINE configurations (sizes as quoted there)
HMS_C0 = Workload("hms_proton_3GeV", 1, 100_000, 3000.0, 20.0, MP, (-12.0, 12.0),
                  (-0.08, 0.08), (-0.05, 0.05), seed=1001)
SHMS_C1 = Workload("shms_electron_6GeV", 5, 10_000_000, 6000.0, 12.5, ME, (-15.0, 25.0),
                   (-0.06, 0.06), (-0.04, 0.04), wire_sigma=0.030, seed=1002)
# SOS (arm 2): 1.5 GeV/c protons at 30 degrees, ranges a little beyond the
# collimator (h 7.2, v 4.7 cm at 126 cm, sos/mc_sos.f:62-65)
SOS_1p5 = Workload("sos_proton_1.5GeV", 2, 1_000_000, 1500.0, 30.0, MP, (-15.0, 15.0),
                   (-0.045, 0.045), (-0.07, 0.07), seed=1003)
# HRS-R (arm 3) protons and HRS-L (arm 4) electrons: ranges a little beyond the
# slit (h 3.1, v 6.1 cm at 110-111 cm, hrsl/mc_hrsl.f:54-57) and the +-4.5 %
# momentum bite
HRSR_2 = Workload("hrsr_proton_2GeV", 3, 1_000_000, 2000.0, 20.0, MP, (-6.0, 6.0),
                  (-0.07, 0.07), (-0.04, 0.04), seed=1004)
HRSL_3 = Workload("hrsl_electron_3GeV", 4, 1_000_000, 3000.0, 15.0, ME, (-6.0, 6.0),
                  (-0.07, 0.07), (-0.04, 0.04), seed=1005)


def arm_phi(arm: int) -> float:
    """Spectrometer phi: 3 pi/2 for HMS and HRS-R (arms 1, 3), pi/2 for SOS,
    HRS-L and SHMS (2, 4, 5, 6), dbase.f:246-262."""
    return 1.5 * math.pi if arm in (1, 3) else 0.5 * math.pi


def generate(w: Workload, n: int | None = None, seed: int | None = None) -> dict:
    """Spectrometer-frame SoA inputs (float64) for workload w."""
    n = w.n if n is None else n
    rng = np.random.default_rng(w.seed if seed is None else seed)
    xt = rng.normal(0.0, w.sigma_xy, n)
    yt = rng.normal(0.0, w.sigma_xy, n)
    zt = rng.uniform(w.z[0], w.z[1], n)
    dpp = rng.uniform(w.dpp[0], w.dpp[1], n)
    dxdz = rng.uniform(w.dxdz[0], w.dxdz[1], n)
    dydz = rng.uniform(w.dydz[0], w.dydz[1], n)
    th = math.radians(w.theta_deg)
    phi = arm_phi(w.arm)
    cth, sth, sphi = math.cos(th), math.sin(th), math.sin(phi)
    x = -yt
    y = -xt * cth - zt * sth * sphi
    z = zt * cth + xt * sth * sphi
    x = x - z * dxdz
    y = y - z * dydz
    z = np.zeros(n)
    return {"dpp": dpp, "x": x, "y": y, "z": z, "dxdz": dxdz, "dydz": dydz, "fry": x.copy()}


def generate_device(w: Workload, n: int, seed: int, device, out: dict | None = None, at: int = 0) -> dict:
    """The same distributions and transform as generate(), drawn with torch on
    `device` (for large batches); writes into out[k][at:at+n] when given."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    f64 = dict(dtype=torch.float64, device=device)

    def uni(lo, hi):
        return torch.rand(n, generator=g, **f64) * (hi - lo) + lo

    xt = torch.randn(n, generator=g, **f64) * w.sigma_xy
    yt = torch.randn(n, generator=g, **f64) * w.sigma_xy
    zt = uni(*w.z)
    dpp = uni(*w.dpp)
    dxdz = uni(*w.dxdz)
    dydz = uni(*w.dydz)
    th = math.radians(w.theta_deg)
    phi = arm_phi(w.arm)
    cth, sth, sphi = math.cos(th), math.sin(th), math.sin(phi)
    x = -yt
    y = -xt * cth - zt * sth * sphi
    z = zt * cth + xt * sth * sphi
    x = x - z * dxdz
    y = y - z * dydz
    ev = {"dpp": dpp, "x": x, "y": y, "z": torch.zeros_like(x), "dxdz": dxdz, "dydz": dydz, "fry": x.clone()}
    if out is None:
        return ev
    for k, v in ev.items():
        out[k][at:at + n] = v
    return out


# The C0 target: liquid hydrogen in the 2017 10 cm cell (init.f:33 X0 value)
LH2_CELL = dict(A=1.0, Z=1.0, rho=0.0723, X0=61.28, length=10.0, angle=0.0, zoffset=0.0, can=3)


def physics_theta(th0: float, phi0: float, dxdz, dydz):
    """Angle from the beam line of a track with spectrometer angles (dx/dz,
    dy/dz) (physics_angles, event.f:1572-1620, for phi0 = pi/2 or 3 pi/2)."""
    r = np.sqrt(1.0 + dxdz ** 2 + dydz ** 2)
    return np.arccos((math.cos(th0) - dydz * math.sin(th0) * math.sin(phi0)) / r)


def generate_vertex(w: Workload, n: int | None = None, seed: int | None = None) -> dict:
    """Target-frame vertices for the target step (the caller's side of
    simc.f:1376-1441): position x, y ~ N(0, sigma_xy), z ~ U(z); delta,
    xptar, yptar uniform as in generate(); p = p_spec (1 + delta/100),
    E = sqrt(p^2 + m^2) (no radiation: E_vertex = E), theta from
    physics_theta.  float64 numpy arrays."""
    n = w.n if n is None else n
    rng = np.random.default_rng(w.seed if seed is None else seed)
    xt = rng.normal(0.0, w.sigma_xy, n)
    yt = rng.normal(0.0, w.sigma_xy, n)
    zt = rng.uniform(w.z[0], w.z[1], n)
    delta = rng.uniform(w.dpp[0], w.dpp[1], n)
    xptar = rng.uniform(w.dxdz[0], w.dxdz[1], n)
    yptar = rng.uniform(w.dydz[0], w.dydz[1], n)
    phi = arm_phi(w.arm)
    p = w.p_spec * (1.0 + delta / 100.0)
    E = np.sqrt(p * p + w.mass * w.mass)
    theta = physics_theta(math.radians(w.theta_deg), phi, xptar, yptar)
    return {"x": xt, "y": yt, "z": zt, "E": E, "p": p, "theta": theta, "delta": delta,
            "xptar": xptar, "yptar": yptar}


def mixed_settings(n_settings: int = 64, seed: int = 2026, lo: float = 1e4, hi: float = 1e7,
                   total: float | None = 1.5e8):
    """The batched configuration: n_settings kinematic settings alternating
    HMS (protons) and SHMS (electrons), central momenta U(1, 9) GeV/c, angles
    U(10, 40) degrees, beam energy from {6.4, 8.5, 10.6} GeV (carried along;
    the spectrometer model does not use it), event counts log-uniform in
    [lo, hi], then scaled by one common factor so that they sum to `total`
    (64 log-uniform draws in [1e4, 1e7] average 9.3e7 events).  Returns a
    list of Workload objects with n = the setting's count."""
    rng = np.random.default_rng(seed)
    draws = []
    for i in range(n_settings):
        p = float(rng.uniform(1000.0, 9000.0))
        th = float(rng.uniform(10.0, 40.0))
        beam = float(rng.choice([6400.0, 8500.0, 10600.0]))
        draws.append((p, th, beam, math.exp(rng.uniform(math.log(lo), math.log(hi)))))
    scale = 1.0 if total is None else total / sum(d[3] for d in draws)
    out = []
    for i, (p, th, beam, c) in enumerate(draws):
        hms = i % 2 == 0
        cnt = int(round(c * scale))
        base = HMS_C0 if hms else SHMS_C1
        out.append(Workload(f"setting{i:02d}_{'hms' if hms else 'shms'}", base.arm, cnt, p, th, base.mass,
                            base.dpp, base.dxdz, base.dydz, wire_sigma=base.wire_sigma,
                            seed=seed * 1000 + i, extra={"beam_MeV": beam}))
    return out


def uniform_tape(n: int, stride: int, seed: int) -> np.ndarray:
    """Per-event uniform tapes in (0,1): event i draws tape[i*stride + k]."""
    rng = np.random.default_rng(seed)
    t = rng.random(n * stride)
    t[t == 0.0] = 0.5
    return t


def lfit_point_sets(n: int, seed: int, npts: int = 12):
    """Point sets for the straight-line fit LFIT (cern/lfit.f:11-60) with
    KEY = 0: X are plane positions, Y hits, both REAL*4 (float32).

    Most sets look like a wire chamber's: 12 planes in two chambers of six
    (z within +-60 cm, irregular spacing), hits on a line with 0.02-0.3 cm
    smearing, none of the values exactly representable in few bits; some
    planes read 0.0 (a y plane in an x fit, or a hit exactly 0), which LFIT
    skips.  The rest are edge cases: no, one or two non-zero hits, all planes
    at one position (SUMXX = 0), large and tiny magnitudes, negative zero.
    """
    rng = np.random.default_rng(seed)
    z0 = rng.uniform(-60.0, -40.0, (n, 1)) + np.sort(rng.uniform(0.0, 12.0, (n, npts // 2)), axis=1)
    z1 = rng.uniform(30.0, 45.0, (n, 1)) + np.sort(rng.uniform(0.0, 12.0, (n, npts - npts // 2)), axis=1)
    X = np.concatenate([z0, z1], axis=1)
    slope = rng.uniform(-0.1, 0.1, (n, 1))
    icpt = rng.uniform(-40.0, 40.0, (n, 1))
    Y = slope * X + icpt + rng.normal(0.0, 1.0, (n, 1)) * rng.uniform(0.02, 0.3, (n, 1)) * \
        rng.standard_normal((n, npts))
    # 0-4 planes read exactly zero
    nz = rng.integers(0, 5, n)
    for i in np.nonzero(nz)[0]:
        Y[i, rng.choice(npts, nz[i], replace=False)] = 0.0
    k = n // 20  # edge cases in the last 5 % of sets, 1/8 of them each
    e = n - k
    kinds = np.arange(k) % 8
    for j, kind in enumerate(kinds):
        i = e + j
        if kind == 0:
            Y[i] = 0.0                                     # no points
        elif kind == 1:
            Y[i] = 0.0
            Y[i, rng.integers(npts)] = rng.uniform(-5, 5)  # one point
        elif kind == 2:
            keep = rng.choice(npts, 2, replace=False)      # two points: E = 0 branch
            m = np.zeros(npts, bool)
            m[keep] = True
            Y[i, ~m] = 0.0
        elif kind == 3:
            X[i] = X[i, 0]                                 # all planes at one z: SUMXX = 0
        elif kind == 4:
            Y[i] *= 1e4                                    # large hits
        elif kind == 5:
            Y[i] *= 1e-6                                   # tiny hits
        elif kind == 6:
            Y[i, ::3] = -0.0                               # negative zeros are skipped too
        else:
            Y[i] = rng.uniform(-1.0, 1.0) + 1e-3 * rng.standard_normal(npts)  # nearly flat
    return X.astype(np.float32), Y.astype(np.float32)


def _elastic_central(ebeam: float, th_e_deg: float):
    """E' and the proton's momentum and angle of H(e,e'p) at the central
    electron angle (two-body elastic kinematics)."""
    th = math.radians(th_e_deg)
    e1 = ebeam * MP / (MP + ebeam * (1.0 - math.cos(th)))
    px, pz = -e1 * math.sin(th), ebeam - e1 * math.cos(th)
    return e1, math.hypot(px, pz), math.atan2(-px, pz)


# H(e,e'p) elastic coincidence in Hall C (the coincidence configuration):
# 10.6 GeV beam on the 10 cm LH2 cell, electron in the SHMS at 12.5 deg,
# proton in the HMS at the conjugate angle, flat square raster +-1 mm.  The
# reference spreads the beam energy uniformly over dEbeam (event.f:204); a
# Gaussian of 5e-4 relative sigma is replaced by the uniform width with the
# same RMS, dEbeam = sqrt(12) * 5e-4 * Ebeam.
_E1, _PP, _THP = _elastic_central(10600.0, 12.5)
HEP_ELASTIC = dict(Ebeam=10600.0, dEbeam=math.sqrt(12.0) * 5e-4 * 10600.0, xwid=0.01, ywid=0.01, xoffset=0.0,
                   yoffset=0.0, fr1=0.1, fr2=0.1, coulomb=0.0, e_yptar=(-0.035, 0.035), e_xptar=(-0.05, 0.05),
                   e_theta=math.radians(12.5), e_phi=arm_phi(5), e_P=_E1, p_theta=_THP, p_phi=arm_phi(1), p_P=_PP,
                   Mh=MP, fr_pattern=3, using_eloss=1, using_coulomb=0)


def hep_rad_limits(hep: dict, *, e_delta=(-10.0, 22.0), p_delta=(-10.0, 10.0), Em_cut=(0.0, 200.0),
                   eloss_max=(4.0, 4.0, 8.0), slop_delta=(0.0, 0.0), dE_edge_test=0.0,
                   Em_edge_from_cuts=True) -> dict:
    """Edges of an H(e,e'p) run that generate_rad reads, formed with
    limits_init's formulas (init.f:206-226 true edges, :282-302 Em slop,
    :405-421 Egamma_tot_max) from the spectrometer acceptances (SPedge delta,
    %), the Em cut window (MeV), the largest target losses of the beam, the
    electron and the proton (MeV; the smallest are taken as 0) and the
    reconstruction slops of delta (%).  No Coulomb correction.

    Em_edge_from_cuts: the reference's limits_init forms edge%Em from the cut
    window only for A >= 3 (init.f:297-302) and leaves it 0 for hydrogen, which
    limits tails 1 and 2 to photons below dE_edge_test; True applies the
    A >= 3 formula (the evident intent), False reproduces the zeros."""
    e_P, p_P, Mh = hep["e_P"], hep["p_P"], hep["Mh"]
    sd_e, sd_p = slop_delta
    e_min = (1. + (e_delta[0] - sd_e) / 100.) * e_P - dE_edge_test
    e_max = (1. + (e_delta[1] + sd_e) / 100.) * e_P + dE_edge_test
    pp_min = max(0.001, (1. + (p_delta[0] - sd_p) / 100.) * p_P - dE_edge_test)
    pp_max = (1. + (p_delta[1] + sd_p) / 100.) * p_P + dE_edge_test
    pE_min = math.sqrt(pp_min ** 2 + Mh * Mh)
    pE_max = math.sqrt(pp_max ** 2 + Mh * Mh)
    e_max += eloss_max[1]
    pE_max += eloss_max[2]
    ebeam_max = hep["Ebeam"] + hep["dEbeam"] / 2.
    if Em_edge_from_cuts:
        slop_ebeam = hep["dEbeam"] / 2. + eloss_max[0]
        slop_ee = sd_e / 100. * e_P + eloss_max[1]
        r = math.sqrt(pE_max ** 2 - Mh * Mh)
        slop_ep = math.sqrt((r + sd_p / 100. * p_P) ** 2 + Mh * Mh) - pE_max + eloss_max[2]
        slop = slop_ebeam + slop_ee + slop_ep + dE_edge_test
        Em = (max(0.0, Em_cut[0] - slop), Em_cut[1] + slop)
    else:
        Em = (0.0, 0.0)
    tot = ebeam_max + Mh - e_min - pE_min
    return dict(edge_e_E=(e_min, e_max), edge_p_E=(pE_min, pE_max), edge_Em=Em, vertex_Trec=(0.0, 0.0),
                Egamma_tot_max=tot, dE_edge_test=dE_edge_test)


# The radiative settings of the reference's H(e,e'p) example input
# (infiles/extra_hydrogen.inp: rad_flag 0, extrad_flag 2, intcor_mode 1,
# use_offshell_rad 1, one_tail 0, use_expon 0) for the coincidence
# configuration, electron in the SHMS (5) and proton in the HMS (1).
HEP_RAD = dict(rad_flag=0, extrad_flag=2, intcor_mode=1, use_expon=0, use_offshell_rad=1,
               doing_tail=(1, 1, 1), e_arm=5, p_arm=1, Egamma_gen_max=0.0,
               **hep_rad_limits(HEP_ELASTIC))
