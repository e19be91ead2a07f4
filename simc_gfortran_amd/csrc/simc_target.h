// Target step on the device: the per-event preparation that turns a vertex
// in the target frame into the inputs of one arm (one event per lane, FP64).
//
//   trip_thru_target  target.f:1-246   path lengths through target and windows,
//                                      radiation lengths, sampled energy loss
//   enerloss_new      enerloss_new.f   one material's loss (typeflag 1)
//   target_musc       target.f:548-574 two planar scattering angles
//   montecarlo        simc.f:1376-1441 (hadron arm), :1649-1706 (electron arm),
//                     :1364-1369 (beam): delta, angles, TRANSPORT transform,
//                     spectrometer offsets, drift to z = 0
//
// Constants that depend only on the setting (window thicknesses of the
// spectrometer, material terms of the Bethe formula, cos/sin of the angles)
// are formed once on the host with the same expressions and libm calls; the
// per-event arithmetic keeps the reference's operation order.
#pragma once

#include "simc_device.h"

namespace simcdev {

// One material of trip_thru_target with the setting-constant terms of
// enerloss_new (enerloss_new.f:23-31,51-55):
//   k_mp  = 0.1536e-03 * zeff / aeff          (first factors of Eloss_mp_new)
//   l0    = log(me / I**2) + 1.063
//   k_log = 0.1536 * zeff / aeff               (argument factor of the 4th log)
//   k_chi = 0.307075 / 2. * zeff / aeff        (chsi)
//   co    = log(hnup) - log(I) + 0.5
struct MaterialDev {
    double rho, k_mp, l0, k_log, k_chi, co, co27;  // co27 = |co / 27|
};

enum { MAT_TARGET = 0, MAT_AL, MAT_AIR, MAT_KEVLAR, MAT_MYLAR, N_MAT };

struct PrepDev {
    MaterialDev mat[N_MAT];
    double x0_cm[N_MAT];          // radiation lengths in cm (target.inc:10,16,22,28; init.f:43)
    double length, cos_angle, angle, zoffset;
    double s_al, s_air, s_kevlar, s_mylar;  // window thicknesses of the spectrometer (target.f:71-101)
    double angle_sign;            // forward path uses cos(theta + angle_sign * angle)
    double p_spec, cos_th, sin_th, sin_phi, mass, coulomb, beam_energy;
    double off[5];                // x, y, z, xptar, yptar
    double ln10;                  // log(10.)
    int narm, can, liquid, using_eloss, mc_smear;
};

// The particle-dependent part of enerloss_new (enerloss_new.f:20-22,29):
// gamma, beta, beta**2 and log(beta*gamma), shared by every material.
struct ElossKin {
    double epart, mpart, gamma, beta, b2, lbg, log10bg;
};

__device__ __forceinline__ ElossKin eloss_kin(double epart, double mpart, double ln10)
{
    ElossKin k;
    k.epart = epart;
    k.mpart = mpart;
    k.gamma = epart / mpart;
    k.beta = sqrt(1.0 - 1.0 / (k.gamma * k.gamma));
    k.b2 = k.beta * k.beta;
    k.lbg = log(k.beta * k.gamma);
    k.log10bg = k.lbg / ln10;
    return k;
}

// enerloss_new(len, dens, zeff, aeff, epart, mpart, typeflag=1), enerloss_new.f:19-82
template <class R>
__device__ __forceinline__ double eloss_material(R& g, const MaterialDev& m, const ElossKin& k, double len,
                                                 double ln10)
{
    const double thick = len * m.rho;
    double E = 0.0;
    if (thick > 0.0) {
        double denscorr;
        if (k.log10bg < 0.0) denscorr = 0.0;
        else if (k.log10bg < 3.0) {
            const double t3 = 3.0 - k.log10bg;
            denscorr = m.co + ln10 * k.log10bg + m.co27 * (t3 * t3 * t3);
        } else if (k.log10bg < 4.7) denscorr = m.co + ln10 * k.log10bg;
        else denscorr = m.co + ln10 * 4.7;
        const double mp_new =
            m.k_mp * thick / k.b2 * (m.l0 + 2.0 * k.lbg + log(m.k_log * thick / k.b2) - k.b2 - denscorr);
        const double mp = mp_new * 1000.0;
        const double chsi = m.k_chi * thick / k.b2;
        const double x = fabs(gauss1(g, 10.0));
        const double lambda = x > 0.0 ? -2.0 * log(x) : 100000.0;
        E = lambda * chsi + mp;
    }
    if (E > (k.epart - k.mpart)) E = (k.epart - k.mpart) - 0.0000001;
    return E;
}

// trip_thru_target for narm 2/3 (target.f:71-246); returns the loss, sets
// radlen.  Inlined: the stream's key must stay provably uniform.
template <class R>
__device__ __forceinline__ double trip_scattered(R& g, const PrepDev& P, double zpos, double energy, double theta,
                                                 double& radlen)
{
    const double forward_path = (P.length / 2.0 - zpos) / fabs(cos(theta + P.angle_sign * P.angle));
    double s_target = forward_path;
    double s_al = P.s_al;
    if (P.liquid) {
        if (P.can == 1) {  // beer can, target.f:105-113
            const double side_path = 1.325 * 2.54 / fabs(sin(theta));
            if (forward_path < side_path) {
                s_al = s_al + 0.005 * 2.54 / fabs(cos(theta));
            } else {
                s_target = side_path;
                s_al = s_al + 0.005 * 2.54 / fabs(sin(theta));
            }
        } else if (P.can == 2) {  // pudding can, target.f:114-139
            const double tt = tan(theta);
            const double t = tt * tt;
            const double atmp = 1.0 + t;
            const double btmp = -2.0 * zpos * t;
            const double half = P.length / 2.0;
            const double ctmp = zpos * zpos * t - half * half;
            const double z_can = (-btmp + sqrt(btmp * btmp - 4.0 * atmp * ctmp)) / 2.0 / atmp;
            s_target = (z_can - zpos) / fabs(cos(theta));
            const double costmp = z_can / half;
            // |costmp| slightly above 1 (rounding) gives 0, as target.f:131-133
            const double th_can = fabs(costmp) <= 1.0 ? acos(z_can / half) : 0.0;
            s_al = s_al + 0.0050 * 2.54 / fabs(sin(3.14159265358979 / 2.0 - (theta - th_can)));
        } else if (P.can == 3) {  // 2017 cryo cell, target.f:140-165
            const double ecir = 1.315 * 2.54;
            const double ecor = (1.315 + 0.0071) * 2.54;
            const double entec = P.length - ecir;
            const double twall = ecor - ecir;
            const double tcm = zpos + P.length / 2.0;
            double tliquid, tal;
            if ((tcm + ecir / tan(theta)) < entec) {
                tliquid = ecir / sin(theta);
                tal = twall / sin(theta);
            } else {
                const double d = P.length - ecir - tcm;
                const double ds = d * sin(theta);
                const double si = sqrt(ecir * ecir - ds * ds);
                tliquid = si + d * cos(theta);
                tal = (sqrt(ecor * ecor - ds * ds) - si) * twall / (ecor - ecir);
            }
            s_al = s_al + tal;
            s_target = tliquid;
        }
    }
    radlen = s_target / P.x0_cm[MAT_TARGET] + s_al / P.x0_cm[MAT_AL] + P.s_air / P.x0_cm[MAT_AIR] +
             P.s_kevlar / P.x0_cm[MAT_KEVLAR] + P.s_mylar / P.x0_cm[MAT_MYLAR];
    const ElossKin k = eloss_kin(energy, P.mass, P.ln10);
    const double e_t = eloss_material(g, P.mat[MAT_TARGET], k, s_target, P.ln10);
    const double e_al = eloss_material(g, P.mat[MAT_AL], k, s_al, P.ln10);
    const double e_air = eloss_material(g, P.mat[MAT_AIR], k, P.s_air, P.ln10);
    const double e_kev = eloss_material(g, P.mat[MAT_KEVLAR], k, P.s_kevlar, P.ln10);
    const double e_myl = eloss_material(g, P.mat[MAT_MYLAR], k, P.s_mylar, P.ln10);
    return e_t + e_al + e_air + e_kev + e_myl;
}

// trip_thru_target for the incoming beam (narm 1, target.f:33-50)
template <class R>
__device__ __forceinline__ double trip_beam(R& g, const PrepDev& P, double zpos, double& radlen)
{
    const double s_target = (P.length / 2.0 + zpos) / P.cos_angle;  // cos_angle = |cos(angle)|
    radlen = s_target / P.x0_cm[MAT_TARGET] + P.s_al / P.x0_cm[MAT_AL];
    const ElossKin k = eloss_kin(P.beam_energy, P.mass, P.ln10);
    const double e_t = eloss_material(g, P.mat[MAT_TARGET], k, s_target, P.ln10);
    const double e_al = eloss_material(g, P.mat[MAT_AL], k, P.s_al, P.ln10);
    return e_t + e_al;
}

// target_musc, target.f:548-574
template <class R>
__device__ __forceinline__ void target_musc(R& g, double p, double beta, double teff, double& d1, double& d2)
{
    const double Es = 13.6, epsilon = 0.088, nsig_max = 3.5;
    const double theta_sigma = Es / p / beta * sqrt(teff) * (1.0 + epsilon * log10(teff / (beta * beta)));
    d1 = theta_sigma * gauss1(g, nsig_max);
    d2 = theta_sigma * gauss1(g, nsig_max);
}

}  // namespace simcdev
