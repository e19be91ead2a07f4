// Radiative tails of H(e,e'p) elastic on the device, one event per lane
// (FP64): SIMC's generate with using_rad for doing_hyd_elast.
//
//   generate_rad       radc.f:120-517   tail choice, photon energies per tail
//                                       in their kinematic windows, orig
//                                       kinematics, gen_weight
//   basicrad           radc.f:3-91      Egamma ~ g Egamma^(g-1) (one uniform)
//   peaked_rad_weight  radc.f:523-650   weight in the peaked basis (brem or
//                                       bremos soft deltas at the window ends)
//   extrad_phi         radc.f:668-718   external correction phi
//   brem, inter        brem.f:6-248     soft/hard photon deltas, on shell
//   bremos             brem.f:344-578   the same with off-shell protons
//   schwinger, spen    radc.f:721-761
//   lambda_dave        radc.f:763-812
//   radc_init_ev       init.f:655-720   bt, lambda, hardcorfac, g(4)
//   basicrad_init_ev   init.f:732-804   g, c constants of the tails
//   complete_ev        event.f:497-563, :1013-1014, :1031-1045 (hydrogen
//                                       elastic, the scattered particles'
//                                       target losses, radc_init_ev)
//
// The reference's callers of gamma() reach the Fortran intrinsic GAMMA
// (tgamma), and its compiled x**0.5 is a square root.  Operation order
// follows the reference so that the oracle and the device agree to the
// last bits of the elementary functions.
#pragma once

#include "simc_gen.h"

namespace simcdev {

// Run constants of radc_init (init.f:576-651) and limits_init (init.f:91-470)
// that generate_rad reads.
struct RadDev {
    int rad_flag, extrad_flag, intcor_mode, use_expon, use_offshell_rad;
    int doing_tail[3];
    int hardwired;                       // Egamma_gen_max > 0.01 (dbase.f:514-518)
    double dE_edge_test, Egamma_gen_max;
    double edge_e_E[2], edge_p_E[2], edge_Em[2], vertex_Trec[2];
    double Egamma_tot_max, Egamma1_max;  // init.f:409-424
    double etatzai, targ_A;
};

constexpr double kAmu = 931.49432;               // constants.inc:38
constexpr double kAlpi = kAlpha / 3.141592653589793;
constexpr double kEuler = 0.577215665;

// spence, brem.f:287-322 (the reference's approximate form)
__device__ __forceinline__ double rad_spence(double ax)
{
    const double bx = fabs(ax);
    if (bx <= 1) return 0.0;
    const double l = log(bx);
    return -0.5 * (l * l);
}

// inter, brem.f:218-248 (calculate_spence = .true.).  Kept out of line:
// brem/bremos reach it up to 28 times per event, and inlined copies of its
// eight logs push the kernel to 256 VGPRs with scratch spills.  No call
// site holds the stream's key live across it (the key must stay provably
// uniform for philox4x32_10).
__device__ __noinline__ double rad_inter(double alpha, double ar1, double ar2, double e1, double e2, double de)
{
    const double de2 = e1 - e2;
    const double amult = -1. / (alpha * (ar1 - ar2));
    double r = log(fabs((e2 / de) + ar1 * (de2 / de))) * log(fabs((ar1 - 1.) / ar1)) -
               log(fabs((e2 / de) + ar2 * (de2 / de))) * log(fabs((ar2 - 1.) / ar2));
    const double arg1 = (de2 / (e2 + ar1 * de2)) * (ar1 - 1.);
    const double arg2 = (de2 / (e2 + ar1 * de2)) * (ar1);
    const double arg3 = (de2 / (e2 + ar2 * de2)) * (ar2 - 1.);
    const double arg4 = (de2 / (e2 + ar2 * de2)) * (ar2);
    r = (((r - rad_spence(arg1)) + rad_spence(arg2)) + rad_spence(arg3)) - rad_spence(arg4);
    return r * amult / kPi;
}

// inter_prime, brem.f:581-598
__device__ __forceinline__ double rad_inter_prime(double alpha, double ar1, double ar2, double de)
{
    const double amult = -1. / (alpha * (ar1 - ar2));
    return (-1. / de) * amult / kPi * (log(fabs((ar1 - 1.) / ar1)) - log(fabs((ar2 - 1.) / ar2)));
}

// brem, brem.f:6-216 (include_hard, calculate_spence; exponentiate =
// use_expon).  MeV in; bsoft, bhard, dbsoft (1/MeV) out.
__device__ __forceinline__ double rad_brem(int use_expon, double ein, double eout, double egamma, bool radiate_proton,
                                        double& bsoft, double& bhard, double& dbsoft)
{
    const double pi = 3.141592653589793, am = .93827231, ame = .00051099906, e2 = 1. / 137.0359895;
    const double ame2 = ame * ame, ame4 = ame2 * ame2, am2 = am * am;
    const double ak = ein / 1000., akp = eout / 1000., de = egamma / 1000.;
    const double eang = 2. * asin(sqrt(am / (2. * ak) * (ak / akp - 1.)));
    const double se = sin(eang / 2.);
    const double q2 = 4. * ak * akp * (se * se);
    const double ape = am + ak - akp;
    const double ap = sqrt(ape * ape - am2);
    const double pang = acos((ak - akp * cos(eang)) / ap);
    double aprod = 1.e0;
    const double bei = aprod * (-1. / (2. * pi)) * log(ak / de);
    const double dbei = aprod * (1. / (2. * pi * de));
    const double bef = aprod * (-1. / (2. * pi)) * log(akp / de);
    const double dbef = aprod * (1. / (2. * pi * de));
    aprod = -1.e0;
    double adot = ak * akp * (1. - cos(eang));
    double alpha = 2. * ame2 - 2. * adot;
    double ar1 = 0.5 + sqrt(adot * adot - ame4) / alpha;
    double ar2 = 0.5 - sqrt(adot * adot - ame4) / alpha;
    const double bee = aprod * adot * rad_inter(alpha, ar1, ar2, ak, akp, de);
    const double dbee = aprod * adot / (pi * alpha * (ar1 - ar2) * de) * (log((ar1 - 1) / ar1) - log((ar2 - 1) / ar2));
    double bz = 0.0, bzz = 0.0, dbz = 0.0, dbzz = 0.0;
    if (radiate_proton) {
        const double ame_am2 = (ame * am) * (ame * am);
        aprod = 1.e0;
        const double bpi = aprod * (-1. / (2. * pi)) * log(am / de);
        const double dbpi = aprod * (1. / (2. * pi * de));
        const double bpf = aprod * (-1. / (2. * pi)) * log(ape / de);
        const double dbpf = aprod * (1 / (2. * pi * de));
        aprod = -1.e0;
        adot = am * ape;
        alpha = 2. * am2 - 2. * adot;
        ar1 = 0.5 + sqrt(adot * adot - am2 * am2) / alpha;
        ar2 = 0.5 - sqrt(adot * adot - am2 * am2) / alpha;
        const double bpp = aprod * adot * rad_inter(alpha, ar1, ar2, am, ape, de);
        const double dbpp =
            aprod * adot / (pi * alpha * (ar1 - ar2) * de) * (log((ar1 - 1) / ar1) - log((ar2 - 1) / ar2));
        // the four electron-proton interference terms: ei-pi, ef-pf, ei-pf, ef-pi
        double bb[4], dbb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double e1 = (k == 0 || k == 2) ? ak : akp;
            const double e2v = (k == 1 || k == 2) ? ape : am;
            aprod = k < 2 ? -1.e0 : 1.e0;
            adot = k == 0 ? ak * am : k == 1 ? akp * ape - akp * ap * cos(eang + pang)
                 : k == 2 ? ak * ape - ak * ap * cos(pang) : akp * am;
            alpha = am2 + ame2 - 2. * adot;
            ar1 = (am2 - adot + sqrt(adot * adot - ame_am2)) / alpha;
            ar2 = (am2 - adot - sqrt(adot * adot - ame_am2)) / alpha;
            bb[k] = aprod * adot * rad_inter(alpha, ar1, ar2, e1, e2v, de);
            dbb[k] = aprod * adot / (pi * alpha * (ar1 - ar2) * de) * (log((ar1 - 1) / ar1) - log((ar2 - 1) / ar2));
        }
        bzz = 2. * e2 * (bpi + bpf + bpp);
        bz = 2. * e2 * (bb[0] + bb[1] + bb[2] + bb[3]);
        dbzz = 2. * e2 * (dbpi + dbpf + dbpp);
        dbz = 2. * e2 * (dbb[0] + dbb[1] + dbb[2] + dbb[3]);
    }
    const double b = 2. * e2 * (bei + bef + bee);
    bsoft = b + bz + bzz;
    bhard = -1. * (e2 / pi) * (-28 / 9. + 13. / 6. * log(q2 / ame2));
    const double db = 2. * e2 * (dbei + dbef + dbee);
    dbsoft = db + dbz + dbzz;
    dbsoft = dbsoft / 1000.;
    const double r = use_expon == 1 ? -dbsoft / exp(bsoft) : 1. - dbsoft;
    return r * (1. - bhard);
}

// bremos, brem.f:344-578 (MeV four-vectors in; the initial proton at rest
// with the proton mass, its momentum 0 at every call site)
__device__ __forceinline__ double rad_bremos(int use_expon, double egamma, double k_iz, double k_fx, double k_fy,
                                          double k_fz, double p_fx, double p_fy, double p_fz, double p_fe,
                                          bool radiate_proton, double& bsoft, double& bhard, double& dbsoft)
{
    const double pi = 3.141592653589793, twopi = 2. * 3.141592653589793, ame = .00051099906;
    const double e2 = 1. / 137.0359895, mp = .93827231;
    const double ame2 = ame * ame, ame4 = ame2 * ame2;
    const double de = egamma / 1000.;
    const double kix = 0.0 / 1000., kiy = 0.0 / 1000., kiz = k_iz / 1000.;
    const double kfx = k_fx / 1000., kfy = k_fy / 1000., kfz = k_fz / 1000.;
    const double pix = 0.0 / 1000., piy = 0.0 / 1000., piz = 0.0 / 1000.;
    const double pfe = p_fe / 1000., pfx = p_fx / 1000., pfy = p_fy / 1000., pfz = p_fz / 1000.;
    const double kie = sqrt(kix * kix + kiy * kiy + kiz * kiz + ame2);
    const double kfe = sqrt(kfx * kfx + kfy * kfy + kfz * kfz + ame2);
    const double pie = mp;
    const double dke = kfe - kie, dkx = kfx - kix, dky = kfy - kiy, dkz = kfz - kiz;
    const double q2 = -1. * (dke * dke - dkx * dkx - dky * dky - dkz * dkz);
    const double ami = mp;
    const double amf = sqrt(pfe * pfe - pfx * pfx - pfy * pfy - pfz * pfz);
    double aprod = 1.e0;
    const double bei = aprod * (-1. / twopi) * log(kie / de);
    const double dbei = aprod * (-1. / twopi) * (-1. / de);
    const double bef = aprod * (-1. / twopi) * log(kfe / de);
    const double dbef = aprod * (-1. / twopi) * (-1. / de);
    aprod = -1.e0;
    double adot = kie * kfe - kix * kfx - kiy * kfy - kiz * kfz;
    double alpha = 2. * ame2 - 2. * adot;
    double ar1 = 0.5 + sqrt(4. * (adot * adot) - 4. * ame4) / (2. * alpha);
    double ar2 = 0.5 - sqrt(4. * (adot * adot) - 4. * ame4) / (2. * alpha);
    const double bee = aprod * adot * rad_inter(alpha, ar1, ar2, kie, kfe, de);
    const double dbee = aprod * adot * rad_inter_prime(alpha, ar1, ar2, de);
    double bz = 0.0, bzz = 0.0, dbz = 0.0, dbzz = 0.0;
    if (radiate_proton) {
        aprod = 1.e0;
        const double bpi = aprod * (-1. / twopi) * log(pie / de);
        const double dbpi = aprod * (-1. / twopi) * (-1. / de);
        const double bpf = aprod * (-1. / twopi) * log(pfe / de);
        const double dbpf = aprod * (-1. / twopi) * (-1. / de);
        aprod = -1.e0;
        adot = pie * pfe - pix * pfx - piy * pfy - piz * pfz;
        alpha = ami * ami + amf * amf - 2. * adot;
        {
            const double m = ami * amf;
            ar1 = (2. * (amf * amf) - 2. * adot + sqrt(4. * (adot * adot) - 4. * (m * m))) / (2. * alpha);
            ar2 = (2. * (amf * amf) - 2. * adot - sqrt(4. * (adot * adot) - 4. * (m * m))) / (2. * alpha);
        }
        const double bpp = aprod * adot * rad_inter(alpha, ar1, ar2, pie, pfe, de);
        const double dbpp = aprod * adot * rad_inter_prime(alpha, ar1, ar2, de);
        // interference terms ei-pi, ef-pf, ei-pf, ef-pi
        double bb[4], dbb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool ki = k == 0 || k == 2, pf = k == 1 || k == 2;
            const double ee = ki ? kie : kfe, ex = ki ? kix : kfx, ey = ki ? kiy : kfy, ez = ki ? kiz : kfz;
            const double pe = pf ? pfe : pie, px = pf ? pfx : pix, py = pf ? pfy : piy, pz = pf ? pfz : piz;
            const double M = pf ? amf : ami, m = ame * M;
            aprod = k < 2 ? -1.e0 : 1.e0;
            adot = ee * pe - ex * px - ey * py - ez * pz;
            alpha = M * M + ame2 - 2. * adot;
            ar1 = (2. * (M * M) - 2. * adot + sqrt(4. * (adot * adot) - 4. * (m * m))) / (2. * alpha);
            ar2 = (2. * (M * M) - 2. * adot - sqrt(4. * (adot * adot) - 4. * (m * m))) / (2. * alpha);
            bb[k] = aprod * adot * rad_inter(alpha, ar1, ar2, ee, pe, de);
            dbb[k] = aprod * adot * rad_inter_prime(alpha, ar1, ar2, de);
        }
        bzz = 2. * e2 * (bpi + bpf + bpp);
        bz = 2. * e2 * (bb[0] + bb[1] + bb[2] + bb[3]);
        dbzz = 2. * e2 * (dbpi + dbpf + dbpp);
        dbz = 2. * e2 * (dbb[0] + dbb[1] + dbb[2] + dbb[3]);
    }
    const double b = 2. * e2 * (bei + bef + bee);
    bsoft = b + bz + bzz;
    bhard = -1. * (e2 / pi) * (-28 / 9. + 13. / 6. * log(q2 / ame2));
    const double db = 2. * e2 * (dbei + dbef + dbee);
    dbsoft = db + dbz + dbzz;
    dbsoft = dbsoft / 1000.;
    const double r = use_expon == 1 ? -dbsoft / exp(bsoft) : 1. - dbsoft;
    return r * (1. - bhard);
}

// lambda_dave, radc.f:763-812 (plus_flag = 1)
__device__ __forceinline__ double rad_lambda_dave(int itail, bool doing_proton, double e1, double e2, double e3,
                                                  double p3, double th)
{
    double plus_term = 0.0;
    if (itail < 3) {
        plus_term = log((1. - cos(th)) / 2.);
        if (doing_proton) plus_term = plus_term + 2. * log(e1 / e2);
    }
    if (itail == 1) return kAlpi * ((2. * log(2. * e1 / kMe) - 1.) + plus_term);
    if (itail == 2) return kAlpi * ((2. * log(2. * e2 / kMe) - 1.) + plus_term);
    if (!doing_proton) return 0.0;
    const double lam = kAlpi * ((e3 / p3) * log((e3 + p3) / (e3 - p3)) - 2.);
    return lam < 0 ? 0.0 : lam;
}

// The event fields generate_rad and its helpers read and change.
struct RadEvent {
    double Ein, e_E, e_P, e_delta, e_theta, e_phi, e_xptar, e_yptar;
    double uex, uey, uez;
    double p_P, p_E, p_delta, p_theta, p_phi, p_xptar, p_yptar;
    double upx, upy, upz;
    double Q2, nu, q;
    double jacobian, z;
    double eloss[3], teff[3];
};

// /radccom/ per event (radc.inc:5-14)
struct RadState {
    double bt[2], lambda[3], g[5], c[5], c_int[4], c_ext[4], g_int, g_ext, frac[3];
    double hardcorfac;
    bool rad_proton_this_ev;
};

// basicrad_init_ev, init.f:732-804
__device__ __forceinline__ void rad_basicrad_init_ev(RadState& S, double e1, double e2, double e3)
{
    const double one = 1.;
    const double* lam = S.lambda;
    const double* bt = S.bt;
    double* g = S.g;
    double* c = S.c;
    double* ci = S.c_int;
    double* ce = S.c_ext;
    g[1] = lam[0] + bt[0];
    g[2] = lam[1] + bt[1];
    g[3] = lam[2];
    g[0] = g[1] + g[2] + g[3];
    ci[1] = lam[0] / pow(e1 * e2, lam[0] / 2.);
    ci[2] = lam[1] / pow(e1 * e2, lam[1] / 2.);
    ci[3] = lam[2] / pow(kMp * e3, lam[2] / 2.);
    for (int i = 1; i <= 3; ++i) ci[i] = ci[i] * exp(-kEuler * lam[i - 1]) / tgamma(one + lam[i - 1]);
    S.g_int = lam[0] + lam[1] + lam[2];
    ci[0] = ci[1] * ci[2] * S.g_int / lam[0] / lam[1];
    if (lam[2] > 0) ci[0] = ci[0] * ci[3] / lam[2];
    ci[0] = ci[0] * tgamma(one + lam[0]) * tgamma(one + lam[1]) * tgamma(one + lam[2]) / tgamma(one + S.g_int);
    ce[1] = bt[0] / pow(e1, bt[0]) / tgamma(one + bt[0]);
    ce[2] = bt[1] / pow(e2, bt[1]) / tgamma(one + bt[1]);
    ce[3] = 0.0;
    S.g_ext = bt[0] + bt[1];
    ce[0] = ce[1] * ce[2] * S.g_ext / bt[0] / bt[1];
    ce[0] = ce[0] * tgamma(one + bt[0]) * tgamma(one + bt[1]) / tgamma(one + S.g_ext);
    for (int i = 1; i <= 2; ++i)
        c[i] = ci[i] * ce[i] * g[i] / lam[i - 1] / bt[i - 1] * tgamma(one + lam[i - 1]) * tgamma(one + bt[i - 1]) /
               tgamma(one + g[i]);
    c[3] = ci[3];
    c[0] = c[1] * c[2] * g[0] / g[1] / g[2];
    if (g[3] > 0) c[0] = c[0] * c[3] / g[3];
    c[0] = c[0] * tgamma(one + g[1]) * tgamma(one + g[2]) * tgamma(one + g[3]) / tgamma(one + g[0]);
    c[4] = g[4] / pow(e1 * e2, g[4]) / tgamma(one + g[4]);
    if (g[3] > 0) c[4] = c[4] / pow(e3, g[4]);
}

// schwinger, radc.f:721-745, and spen, radc.f:747-761 (dsoft, dhard)
__device__ __forceinline__ void rad_schwinger(const RadDev& R, const RadEvent& v, double Ecutoff, double& dsoft,
                                              double& dhard)
{
    const double lq = log(v.Q2 / (kMe * kMe)) - 1.0;
    const double sn = sin(v.e_theta / 2.);
    const double s2 = sn * sn;
    const double b = 1. + 2. * v.nu * s2 / (R.targ_A * kAmu);
    const double x = 1. - s2;
    double y = 1.0, s = 0.0;
    int i = 0;
    while (i <= 100 && fabs(y) > fabs(s) * 1.e-4) {
        i = i + 1;
        y = x * y;
        s = s + y / (double)(i * i);
    }
    const double spence = s - 2.5893784;
    const double etta = 1.0;
    dsoft = kAlpi * lq * log(v.Ein / (etta * etta) * v.e_E * b / (Ecutoff * Ecutoff));
    const double l = log(v.Ein / v.e_E);
    dhard = -kAlpi * (2.166666 * lq + spence - (l * l) / 2.0);
}

// radc_init_ev, init.f:655-720.  In schwinger mode the reference leaves
// dsoft_prime (a saved local that only brem/bremos set) at its initial 0.
__device__ __forceinline__ void rad_init_ev(const RadDev& R, RadState& S, const RadEvent& v)
{
    for (int i = 0; i < 2; ++i) S.bt[i] = R.etatzai * v.teff[i];
    for (int i = 1; i <= 3; ++i)
        S.lambda[i - 1] = rad_lambda_dave(i, R.doing_tail[2] != 0, v.Ein, v.e_E, v.p_E, v.p_P, v.e_theta);
    S.rad_proton_this_ev = S.lambda[2] > 0;
    const double Ecutoff = 450.;
    double dsoft, dhard, dsoft_prime = 0.0;
    if (R.intcor_mode == 0)
        rad_schwinger(R, v, Ecutoff, dsoft, dhard);
    else if (!R.use_offshell_rad)
        rad_brem(R.use_expon, v.Ein, v.e_E, Ecutoff, S.rad_proton_this_ev, dsoft, dhard, dsoft_prime);
    else
        rad_bremos(R.use_expon, Ecutoff, v.Ein, v.e_P * v.uex, v.e_P * v.uey, v.e_P * v.uez, v.p_P * v.upx,
                   v.p_P * v.upy, v.p_P * v.upz, v.p_E, S.rad_proton_this_ev, dsoft, dhard, dsoft_prime);
    S.hardcorfac = 1. / (1. - dhard);
    S.g[4] = -dsoft_prime * Ecutoff + S.bt[0] + S.bt[1];
    rad_basicrad_init_ev(S, v.Ein, v.e_E, v.p_E);
    for (int i = 1; i <= 3; ++i) S.frac[i - 1] = S.g[i] / S.g[0];
}

// basicrad, radc.f:3-91 (itail 0 is the combined tail 4)
template <class R>
__device__ __forceinline__ void rad_basicrad(R& gr, const RadState& S, int itail, double lo, double hi,
                                             double& Egamma, double& weight)
{
    Egamma = 0.0;
    weight = 0.0;
    if (itail == 0) itail = 4;
    const double g = S.g[itail];
    if (g <= 0) {
        weight = 1.0;
        return;
    }
    if (hi <= lo || hi <= 0) return;
    const double power_hi = pow(hi, g);
    double power_lo = 0.0;
    if (lo > 0) power_lo = pow(lo, g);
    const double ymin = power_lo / power_hi;
    const double y = ymin + gr.next() * (1. - ymin);
    const double x = pow(y, 1. / g);
    Egamma = x * hi;
    weight = S.c[itail] / g * (power_hi - power_lo);
}

// extrad_phi, radc.f:668-718
__device__ __forceinline__ double rad_extrad_phi(const RadDev& R, const RadState& S, int itail, double E1,
                                                 double E2, double Egamma)
{
    double phi = 1.0;
    if (R.extrad_flag == 2) {
        if (itail == 0)
            phi = 1. - (S.bt[0] / E1 + S.bt[1] / E2) / (S.g[1] + S.g[2]) * Egamma;
        else if (itail == 1 || itail == 2)
            phi = 1. - S.bt[itail - 1] / (itail == 1 ? E1 : E2) / S.g[itail] * Egamma;
    } else if (R.extrad_flag == 3 && (itail == 1 || itail == 2)) {
        const double x = Egamma / (itail == 1 ? E1 : E2);
        const double t = S.bt[itail - 1] / R.etatzai;
        phi = phi * (1. - x + (x * x) / R.etatzai) * exp(t * ((R.etatzai - 0.5) - R.etatzai * x + (x * x) / 2.)) *
              tgamma(1. + S.bt[itail - 1]);
    }
    return phi;
}

// peaked_rad_weight, radc.f:523-650 (rad_flag 0 or 1)
__device__ __forceinline__ double rad_peaked_weight(const RadDev& R, const RadState& S, const RadEvent& v, double Egamma,
                                                 double emin, double emax, double basicrad_weight)
{
    const double ein = v.Ein, eout = v.e_E, eul = 0.577215665;
    if (R.extrad_flag <= 2 && R.rad_flag == 1) return basicrad_weight * rad_extrad_phi(R, S, 0, ein, eout, Egamma);
    double dsoft_intmin = 1.0, dsoft_intmax, dhard, dp;
    if (!R.use_offshell_rad) {
        if (emin > 0) rad_brem(R.use_expon, ein, eout, emin, S.rad_proton_this_ev, dsoft_intmin, dhard, dp);
        rad_brem(R.use_expon, ein, eout, emax, S.rad_proton_this_ev, dsoft_intmax, dhard, dp);
    } else {
        const double kx = v.e_E * v.uex, ky = v.e_E * v.uey, kz = v.e_E * v.uez;
        const double px = v.p_P * v.upx, py = v.p_P * v.upy, pz = v.p_P * v.upz;
        if (emin > 0)
            rad_bremos(R.use_expon, emin, ein, kx, ky, kz, px, py, pz, v.p_E, S.rad_proton_this_ev, dsoft_intmin,
                       dhard, dp);
        rad_bremos(R.use_expon, emax, ein, kx, ky, kz, px, py, pz, v.p_E, S.rad_proton_this_ev, dsoft_intmax, dhard,
                   dp);
    }
    double w;
    if (emin > 0)
        w = S.c_ext[0] / S.g_ext * (exp(-dsoft_intmax) * pow(emax, S.g_ext) - exp(-dsoft_intmin) * pow(emin, S.g_ext));
    else
        w = S.c_ext[0] / S.g_ext * (exp(-dsoft_intmax) * pow(emax, S.g_ext));
    const double* g = S.g;
    const double* bt = S.bt;
    w = w * exp(-eul * g[4]) / tgamma(1. + g[4]) * tgamma(1. + g[4] - bt[0] - bt[1]) * tgamma(1. + bt[0]) *
        tgamma(1. + bt[1]) / tgamma(1. + g[4]);
    if (w < 0) w = 0;
    return w;
}

// complete_ev for hydrogen elastic with the scattered particles' target
// losses and radc_init_ev; false when E' > Ein (complete_ev's failure).
template <class R>
__device__ __forceinline__ bool rad_complete_ev(R& gr, const HepDev& H, const PrepDev& Pe, const PrepDev& Pp,
                                                const RadDev& Rd, RadState& S, RadEvent& v)
{
    v.jacobian = 1.0;
    const double ste = sin(v.e_theta);
    v.uex = ste * cos(v.e_phi);
    v.uey = ste * sin(v.e_phi);
    v.uez = cos(v.e_theta);
    v.e_E = v.Ein * H.Mh / (H.Mh + v.Ein * (1. - v.uez));
    if (v.e_E > v.Ein) return false;
    v.e_P = v.e_E;
    v.e_delta = (v.e_P - H.e_P) * 100. / H.e_P;
    v.nu = v.Ein - v.e_E;
    v.Q2 = 2 * v.Ein * v.e_E * (1. - v.uez);
    v.q = sqrt(v.Q2 + v.nu * v.nu);
    v.upx = -v.e_P * v.uex / v.q;
    v.upy = -v.e_P * v.uey / v.q;
    v.upz = (v.Ein - v.e_P * v.uez) / v.q;
    v.p_P = v.q;
    v.p_theta = acos(v.upz);
    double pphi = atan2(v.upy, v.upx);
    if (pphi < 0.) pphi = pphi + 2. * kPi;
    v.p_phi = pphi;
    spectrometer_angles(H.p_x0, H.p_y0, H.p_z0, v.p_theta, v.p_phi, v.p_xptar, v.p_yptar);
    v.p_E = sqrt(v.p_P * v.p_P + H.Mh2);
    v.p_delta = (v.p_P - H.p_P) * 100. / H.p_P;
    const double r = sqrt(1. + v.e_yptar * v.e_yptar + v.e_xptar * v.e_xptar);
    v.jacobian = v.jacobian / (r * r * r);
    v.eloss[1] = trip_scattered(gr, Pe, v.z - Pe.zoffset, v.e_E, v.e_theta, v.teff[1]);
    v.eloss[2] = trip_scattered(gr, Pp, v.z - Pp.zoffset, v.p_E, v.p_theta, v.teff[2]);
    if (!H.using_eloss) {
        v.eloss[1] = 0.0;
        v.eloss[2] = 0.0;
    }
    rad_init_ev(Rd, S, v);
    return true;
}

struct RadOut {
    double Eg[3], gen_weight, orig_Ein, orig_e_E, orig_e_delta, orig_p_E, orig_p_P, orig_p_delta;
    int ntail;
};


// One tail's photon (generate_rad, e.g. radc.f:299-311): the window
// [emin, emax] widened by dE_edge_test (or emax hardwired), then basicrad;
// false when basicrad finds no window (the event fails).
template <class R>
__device__ __forceinline__ bool rad_draw(R& gr, const RadDev& Rd, const RadState& S, int itail_basis, double& emin,
                                         double& emax, double& Egamma, double& bw)
{
    emin = emin - Rd.dE_edge_test;
    emax = emax + Rd.dE_edge_test;
    if (Rd.hardwired) emax = Rd.Egamma_gen_max;
    rad_basicrad(gr, S, itail_basis, emin, emax, Egamma, bw);
    return bw > 0;
}

// The event weight after one tail (radc.f:319-325): peaked_rad_weight in the
// peaked basis (it replaces the product), else the product with extrad_phi.
__device__ __forceinline__ double rad_weigh(const RadDev& Rd, const RadState& S, const RadEvent& v, int itail,
                                            double rad_weight, double Egamma, double emin, double emax, double bw)
{
    rad_weight = rad_weight * bw;
    if (Rd.rad_flag <= 1) return rad_peaked_weight(Rd, S, v, Egamma, emin, emax, bw);
    return rad_weight * rad_extrad_phi(Rd, S, itail, v.Ein, v.e_E, Egamma);
}

// One event of generate with using_rad: the uniforms in the reference's
// order (beam x, y; raster; z; the beam's losses; Ein; electron yptar,
// xptar; complete_ev's losses of the electron and the proton; the tail
// choice; tail 1's photon and complete_ev's losses again; tails 2, 3).
// o receives the VERTEX kinematics, ro the photons and the ORIG ones.
template <class R>
__device__ __forceinline__ void gen_hep_rad(R& gr, const HepDev& H, const PrepDev& Pb, const PrepDev& Pe, const PrepDev& Pp,
                            const RadDev& Rd, HepEvent& o, RadOut& ro, RadEvent& v, double& hardcorfac)
{
    RadState S;
    o.x = gauss1(gr, 3.0) * H.xwid + H.xoffset;
    o.y = gauss1(gr, 3.0) * H.ywid + H.yoffset;
    double t5 = 0.0, t6 = 0.0;
    if (H.fr_pattern == 1) {
        const double t3 = gr.next() * kPi;
        const double t4 = gr.next() * kPi;
        t5 = cos(t3) * H.fr1;
        t6 = cos(t4) * H.fr2;
    } else if (H.fr_pattern == 2) {
        const double t3 = gr.next() * 2. * kPi;
        const double t4 = sqrt(gr.next()) * (H.fr2 - H.fr1) + H.fr1;
        t5 = cos(t3) * t4;
        t6 = sin(t3) * t4;
    } else if (H.fr_pattern == 3) {
        const double t3 = 2. * gr.next() - 1.0;
        const double t4 = 2. * gr.next() - 1.0;
        t5 = H.fr1 * t3;
        t6 = H.fr2 * t4;
    }
    o.x = o.x + t5;
    o.y = o.y + t6;
    o.rasterx = t5;
    o.rastery = t6;
    v.z = (0.5 - gr.next()) * Pb.length + Pb.zoffset;
    v.eloss[0] = trip_beam(gr, Pb, v.z - Pb.zoffset, v.teff[0]);
    if (!H.using_eloss) v.eloss[0] = 0.0;
    v.Ein = H.Ebeam + (gr.next() - 0.5) * H.dEbeam + H.coulomb - v.eloss[0];
    v.e_yptar = H.e_yp0 + gr.next() * H.e_ypw;
    v.e_xptar = H.e_xp0 + gr.next() * H.e_xpw;
    physics_angles(H.e_cos, H.e_sin, H.e_sinph, H.e_phi, v.e_xptar, v.e_yptar, v.e_theta, v.e_phi);
    bool ok = rad_complete_ev(gr, H, Pe, Pp, Rd, S, v);
    double gen_weight = 1.0;
    ro.Eg[0] = ro.Eg[1] = ro.Eg[2] = 0.0;
    ro.orig_Ein = ro.orig_e_E = ro.orig_e_delta = ro.orig_p_E = ro.orig_p_P = ro.orig_p_delta = 0.0;
    ro.ntail = 0;
    if (ok) {
        // generate_rad
        ok = false;
        double rad_weight = 1;
        int peaked_basis_flag = 1, ntail;
        if (Rd.rad_flag <= 1) {
            peaked_basis_flag = 0;
            const double u = gr.next();
            ntail = u >= S.frac[0] + S.frac[1] ? 3 : u >= S.frac[0] ? 2 : 1;
        } else if (Rd.rad_flag == 2) {
            ntail = (int)(gr.next() * 3.) + 1;
            if (ntail == 4) ntail = 3;
        } else {
            ntail = 0;
        }
        ro.ntail = ntail;
        // hydrogen elastic: vertex Trec = 0 and Em = 0 (event.f:547,938)
        const double max_delta_Trec = fmax(0.0 - Rd.vertex_Trec[0], Rd.vertex_Trec[1] - 0.0);
        bool fail = false;
        double bw;
        if (Rd.doing_tail[0] && (ntail == 0 || ntail == 1)) {
            // tail 1: the beam energy after radiation must send the electron into
            // the electron arm's energy edges (radc.f:265-274)
            const double Mp = kMp;
            double ebeam_max = Mp * Rd.edge_e_E[1] / (Mp - Rd.edge_e_E[1] * (1. - v.uez));
            if (ebeam_max < 0) ebeam_max = 1.e10;
            const double ebeam_min = Mp * Rd.edge_e_E[0] / (Mp - Rd.edge_e_E[0] * (1. - v.uez));
            double emin = v.Ein - ebeam_max;
            double emax = v.Ein - ebeam_min;
            emax = fmin(emax, Rd.edge_Em[1]);
            emax = fmin(emax, Rd.Egamma1_max);
            fail = !rad_draw(gr, Rd, S, 1 * peaked_basis_flag, emin, emax, ro.Eg[0], bw);
            if (!fail) {
                v.Ein = v.Ein - ro.Eg[0];
                fail = !rad_complete_ev(gr, H, Pe, Pp, Rd, S, v);
                if (!fail) rad_weight = rad_weigh(Rd, S, v, 1, rad_weight, ro.Eg[0], emin, emax, bw);
            }
        }
        if (!fail && Rd.doing_tail[1] && (ntail == 0 || ntail == 2)) {
            // tail 2: the electron arm's edges and the measured Em (radc.f:341-362)
            double emin = v.e_E - Rd.edge_e_E[1];
            double emax = v.e_E - Rd.edge_e_E[0];
            emax = fmin(emax, (Rd.edge_Em[1] - 0.0) - ro.Eg[0] + max_delta_Trec);
            if (ntail != 0 || !S.rad_proton_this_ev)
                emin = fmax(emin, (Rd.edge_Em[0] - 0.0) - ro.Eg[0] - max_delta_Trec);
            emax = fmin(emax, Rd.Egamma_tot_max - ro.Eg[0]);
            fail = !rad_draw(gr, Rd, S, 2 * peaked_basis_flag, emin, emax, ro.Eg[1], bw);
            if (!fail) rad_weight = rad_weigh(Rd, S, v, 2, rad_weight, ro.Eg[1], emin, emax, bw);
        }
        if (!fail && S.rad_proton_this_ev && (ntail == 0 || ntail == 3)) {
            // tail 3: the hadron arm's edges and the measured Em (radc.f:396-416)
            double emin = v.p_E - Rd.edge_p_E[1];
            double emax = v.p_E - Rd.edge_p_E[0];
            emax = fmin(emax, (Rd.edge_Em[1] - 0.0) - ro.Eg[0] - ro.Eg[1] + max_delta_Trec);
            emin = fmax(emin, (Rd.edge_Em[0] - 0.0) - ro.Eg[0] - ro.Eg[1] - max_delta_Trec);
            emax = fmin(emax, Rd.Egamma_tot_max - ro.Eg[0] - ro.Eg[1]);
            fail = !rad_draw(gr, Rd, S, 3 * peaked_basis_flag, emin, emax, ro.Eg[2], bw);
            if (!fail) rad_weight = rad_weigh(Rd, S, v, 3, rad_weight, ro.Eg[2], emin, emax, bw);
        }
        if (!fail) {
            // orig = vertex with the photons (radc.f:474-503)
            ro.orig_Ein = v.Ein + ro.Eg[0];
            ro.orig_e_E = v.e_E - ro.Eg[1];
            if (ro.orig_e_E > 0e0) {
                ro.orig_e_delta = (ro.orig_e_E - H.e_P) / H.e_P * 100.;
                ro.orig_p_E = v.p_E - ro.Eg[2];
                if (ro.orig_p_E > H.Mh) {
                    ro.orig_p_P = sqrt(ro.orig_p_E * ro.orig_p_E - H.Mh2);
                    ro.orig_p_delta = (ro.orig_p_P - H.p_P) / H.p_P * 100.;
                    gen_weight = gen_weight * rad_weight / S.hardcorfac;
                    ok = true;
                }
            }
        }
    }
    hardcorfac = S.hardcorfac;
    ro.gen_weight = gen_weight;
    o.ok = ok;
    o.z = v.z;
    o.Ein = v.Ein;
    o.eloss = v.eloss[0];
    o.teff = v.teff[0];
    o.e_xptar = v.e_xptar;
    o.e_yptar = v.e_yptar;
    o.e_theta = v.e_theta;
    o.e_phi = v.e_phi;
    o.e_E = v.e_E;
    o.e_delta = v.e_delta;
    o.p_xptar = v.p_xptar;
    o.p_yptar = v.p_yptar;
    o.p_theta = v.p_theta;
    o.p_phi = v.p_phi;
    o.p_P = v.p_P;
    o.p_E = v.p_E;
    o.p_delta = v.p_delta;
    o.Q2 = v.Q2;
    o.nu = v.nu;
    o.q = v.q;
    o.jacobian = v.jacobian;
    o.sigcc = sigep(v.Q2, v.e_E, v.e_theta, v.Ein);
    o.weight = v.jacobian * gen_weight * o.sigcc;
}

}  // namespace simcdev
