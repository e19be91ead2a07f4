// Event generator on the device: H(e,e'p) elastic scattering without
// radiation, one event per lane (FP64).
//
//   generate          event.f:157-212   beam position (gauss1(3) widths),
//                                       raster, vertex z, beam energy loss
//                                       (trip_thru_target narm 1, sampled),
//                                       Coulomb energy, Ein
//                     event.f:275-277   electron xptar, yptar, uniform
//   physics_angles    event.f:1572-1616
//   complete_ev       event.f:497-563   elastic E', nu, Q2, q, p = q
//                     event.f:1013-1014 jacobian 1 / cos^3
//   spectrometer_angles event.f:1618-1647
//   sigep             physics_proton.f:1-23 (fofa_best_fit :137-171,
//                                       sigMott :176-190)
//   complete_main     event.f:1397-1398, :1450-1452, :1558 weight
//
// Operation order follows the reference so that the oracle and the device
// agree to the last bits of the elementary functions (sin, cos, acos, atan,
// atan2, tan, pow, log).
#pragma once

#include "simc_target.h"

namespace simcdev {

constexpr uint32_t kGenTag = 0x48454E47u;  // "GNEH"

// constants.inc:17-44 (kPi: simc_device.h)
constexpr double kMe = 0.51099906;
constexpr double kMp = 938.27231;
constexpr double kHbarc = 197.327053;
constexpr double kAlpha = 1. / 137.0359895;

struct HepDev {
    double Ebeam, dEbeam, xwid, ywid, xoffset, yoffset, fr1, fr2, coulomb;
    double e_yp0, e_ypw, e_xp0, e_xpw;  // min and (max - min) of the slopes
    double e_cos, e_sin, e_sinph;       // cos, sin of the electron arm's theta, sin of its phi
    double e_phi;                       // phi0 of the electron arm
    double e_P;
    double p_x0, p_y0, p_z0;            // the hadron arm's central unit vector
    double p_P;
    double Mh, Mh2;
    int fr_pattern, using_eloss;
};

// physics_angles, event.f:1572-1616 (phi0 = pi/2 or 3 pi/2: its cos term
// is not used)
__device__ __forceinline__ void physics_angles(double costh, double sinth, double sinph, double phi0, double dx,
                                               double dy, double& theta, double& phi)
{
    const double r = sqrt(1. + dx * dx + dy * dy);
    theta = acos((costh - dy * sinth * sinph) / r);
    if (dx != 0.0) {
        double p = atan((dy * costh + sinth * sinph) / dx);
        if (p <= 0) p = p + kPi;
        if (sinph < 0.) p = p + kPi;
        phi = p;
    } else {
        phi = phi0;
    }
}

// spectrometer_angles, event.f:1618-1647, with the central unit vector
// (x0, y0, z0) formed on the host
__device__ __forceinline__ void spectrometer_angles(double x0, double y0, double z0, double theta, double phi,
                                                    double& dx, double& dy)
{
    const double st = sin(theta);
    const double x = st * cos(phi), y = st * sin(phi), z = cos(theta);
    const double cos_dtheta = x * x0 + y * y0 + z * z0;
    dx = x / cos_dtheta;
    double d = sqrt(1 / (cos_dtheta * cos_dtheta) - 1. - dx * dx);
    const double y_event = y / cos_dtheta;
    if (y_event < y0) d = -d;
    dy = d;
}

// sigep with fofa_best_fit and sigMott (physics_proton.f).  The reference's
// compiler forms Q**2. and Q**3. as products and Q**4., Q**5. with pow.
__device__ __forceinline__ double sigep(double Q2, double E, double theta, double Ein)
{
    const double qsquar = -Q2 / (kHbarc * kHbarc);
    const double Qsq = -qsquar * (kHbarc * kHbarc) * 1.e-6;
    const double Q = sqrt(fmax(Qsq, 0.0));
    const double Q3 = Q * Q * Q, Q4 = pow(Q, 4.), Q5 = pow(Q, 5.);
    const double GE = 1. / (1. + 0.62 * Q + 0.68 * Qsq + 2.8 * Q3 + 0.83 * Q4);
    const double GM = 2.793 / (1. + 0.35 * Q + 2.44 * Qsq + 0.5 * Q3 + 1.04 * Q4 + 0.34 * Q5);
    const double qmu4mp = Q2 / 4. / (kMp * kMp);
    const double W1p = GM * GM * qmu4mp;
    const double W2p = (GE * GE + GM * GM * qmu4mp) / (1.0 + qmu4mp);
    const double t = tan(theta / 2.);
    const double Wp = W2p + 2. * W1p * (t * t);
    const double m = 2. * kAlpha * kHbarc * E * cos(theta / 2.) / Q2;
    const double mott = (m * m) * 1.e4;
    return mott * E / Ein * Wp;
}

struct HepEvent {
    double x, y, z, rasterx, rastery, Ein, eloss, teff;
    double e_xptar, e_yptar, e_theta, e_phi, e_E, e_delta;
    double p_xptar, p_yptar, p_theta, p_phi, p_P, p_E, p_delta;
    double Q2, nu, q, jacobian, sigcc, weight;
    bool ok;
};

// One event: the uniforms in the reference's order (beam x, y; raster;
// z; the beam's losses; Ein; electron yptar, xptar).
template <class R>
__device__ __forceinline__ void gen_hep(R& g, const HepDev& H, const PrepDev& P, HepEvent& v)
{
    v.x = gauss1(g, 3.0) * H.xwid + H.xoffset;
    v.y = gauss1(g, 3.0) * H.ywid + H.yoffset;
    double t5 = 0.0, t6 = 0.0;
    if (H.fr_pattern == 1) {
        const double t3 = g.next() * kPi;
        const double t4 = g.next() * kPi;
        t5 = cos(t3) * H.fr1;
        t6 = cos(t4) * H.fr2;
    } else if (H.fr_pattern == 2) {
        const double t3 = g.next() * 2. * kPi;
        const double t4 = sqrt(g.next()) * (H.fr2 - H.fr1) + H.fr1;
        t5 = cos(t3) * t4;
        t6 = sin(t3) * t4;
    } else if (H.fr_pattern == 3) {
        const double t3 = 2. * g.next() - 1.0;
        const double t4 = 2. * g.next() - 1.0;
        t5 = H.fr1 * t3;
        t6 = H.fr2 * t4;
    }
    v.x = v.x + t5;
    v.y = v.y + t6;
    v.rasterx = t5;
    v.rastery = t6;
    v.z = (0.5 - g.next()) * P.length + P.zoffset;
    v.eloss = trip_beam(g, P, v.z - P.zoffset, v.teff);
    if (!H.using_eloss) v.eloss = 0.0;
    v.Ein = H.Ebeam + (g.next() - 0.5) * H.dEbeam + H.coulomb - v.eloss;
    v.e_yptar = H.e_yp0 + g.next() * H.e_ypw;
    v.e_xptar = H.e_xp0 + g.next() * H.e_xpw;
    physics_angles(H.e_cos, H.e_sin, H.e_sinph, H.e_phi, v.e_xptar, v.e_yptar, v.e_theta, v.e_phi);
    // complete_ev, hydrogen elastic
    const double ste = sin(v.e_theta);
    const double uex = ste * cos(v.e_phi), uey = ste * sin(v.e_phi), uez = cos(v.e_theta);
    const double Ein = v.Ein;
    const double E = Ein * H.Mh / (H.Mh + Ein * (1. - uez));
    v.ok = !(E > Ein);
    v.e_E = E;
    v.e_delta = (E - H.e_P) * 100. / H.e_P;
    v.nu = Ein - E;
    v.Q2 = 2 * Ein * E * (1. - uez);
    v.q = sqrt(v.Q2 + v.nu * v.nu);
    const double uqx = -(E * uex / v.q), uqy = -(E * uey / v.q), uqz = (Ein - E * uez) / v.q;
    v.p_P = v.q;
    v.p_theta = acos(uqz);
    double pphi = atan2(uqy, uqx);
    if (pphi < 0.) pphi = pphi + 2. * kPi;
    v.p_phi = pphi;
    spectrometer_angles(H.p_x0, H.p_y0, H.p_z0, v.p_theta, v.p_phi, v.p_xptar, v.p_yptar);
    v.p_E = sqrt(v.q * v.q + H.Mh2);
    v.p_delta = (v.q - H.p_P) * 100. / H.p_P;
    const double r = sqrt(1. + v.e_yptar * v.e_yptar + v.e_xptar * v.e_xptar);
    v.jacobian = 1.0 / (r * r * r);
    v.sigcc = sigep(v.Q2, E, v.e_theta, Ein);
    v.weight = v.jacobian * v.sigcc;
}

}  // namespace simcdev
