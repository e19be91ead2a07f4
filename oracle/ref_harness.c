/*
 * ref_harness.c -- drives the reference Fortran hot path (compiled by
 * oracle/Makefile from the reference's own sources into oracle/_ref/) event
 * by event.  TEST INFRASTRUCTURE ONLY.
 *
 * The harness supplies one routine of the path:
 *   grnd_()  -- the uniform source, read from a per-event tape (the reference
 *               draws from RANLUX through call_ranlux.f:58-73).
 * Everything else (mc_hms, mc_hms_hut, mc_hms_recon, mc_sos, mc_sos_hut,
 * mc_sos_recon, mc_hrsr*, mc_hrsl*, mc_shms, mc_shms_hut, mc_shms_recon,
 * transp/transp_init, project, musc, musc_ext, rotate_*, LFIT (cern/lfit.f,
 * compiled with its CERNLIB '#' lines removed as gfortran ignores them),
 * gauss1, enerloss_new, loren, the collimator routines) is the reference's
 * own code.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "simc_oracle.h"

/* Fortran entry points (gfortran/flang ABI: lower case, trailing underscore,
 * all arguments by reference, LOGICAL = 4-byte integer). */
extern void mc_hms_(double *p_spec, double *th_spec, double *dpp, double *x, double *y, double *z,
                    double *dxdz, double *dydz, double *x_fp, double *dx_fp, double *y_fp,
                    double *dy_fp, double *m2, int32_t *ms_flag, int32_t *wcs_flag,
                    int32_t *decay_flag, double *resmult, double *fry, int32_t *ok_spec,
                    double *pathlen, int32_t *using_coll);
extern void mc_shms_(double *p_spec, double *th_spec, double *dpp, double *x, double *y,
                     double *z, double *dxdz, double *dydz, double *x_fp, double *dx_fp,
                     double *y_fp, double *dy_fp, double *m2, int32_t *ms_flag, int32_t *wcs_flag,
                     int32_t *decay_flag, double *resmult, double *fry, int32_t *ok_spec,
                     double *pathlen, int32_t *spectr, int32_t *using_coll);

/* Common blocks (hms/struct_hms.inc:34-42, shms/struct_shms.inc:26-44,
 * simulate.inc:153 /decd/ whose first member is ctau). */
extern void mc_sos_(double *p_spec, double *th_spec, double *dpp, double *x, double *y, double *z,
                    double *dxdz, double *dydz, double *x_fp, double *dx_fp, double *y_fp,
                    double *dy_fp, double *m2, int32_t *ms_flag, int32_t *wcs_flag,
                    int32_t *decay_flag, double *resmult, double *fry, int32_t *ok_spec,
                    double *pathlen);
/* mc_hrsr / mc_hrsl share mc_sos's argument list (hrsr/mc_hrsr.f:1-3, hrsl/mc_hrsl.f:1-3) */
extern void mc_hrsr_(double *p_spec, double *th_spec, double *dpp, double *x, double *y, double *z,
                     double *dxdz, double *dydz, double *x_fp, double *dx_fp, double *y_fp,
                     double *dy_fp, double *m2, int32_t *ms_flag, int32_t *wcs_flag,
                     int32_t *decay_flag, double *resmult, double *fry, int32_t *ok_spec,
                     double *pathlen);
extern void mc_hrsl_(double *p_spec, double *th_spec, double *dpp, double *x, double *y, double *z,
                     double *dxdz, double *dydz, double *x_fp, double *dx_fp, double *y_fp,
                     double *dy_fp, double *m2, int32_t *ms_flag, int32_t *wcs_flag,
                     int32_t *decay_flag, double *resmult, double *fry, int32_t *ok_spec,
                     double *pathlen);
extern struct { int32_t v[H_NCOUNT]; } spectrometer1_;
extern struct { int32_t v[R_NCOUNT]; } spectrometer3_;  /* hrsr/struct_hrsr.inc:24-40 */
extern struct { int32_t v[R_NCOUNT]; } spectrometer4_;  /* hrsl/struct_hrsl.inc:24-40 */
extern struct { int32_t v[SO_NCOUNT]; } spectrometer2_;  /* sos/struct_sos.inc:9-17 */
extern struct { int32_t v[S_NCOUNT]; } spectrometer_shms_;
extern struct { double ctau; } decd_;

static const double *g_tape;
static int64_t g_stride, g_pos, g_used;
static int g_overflow;

double grnd_(void)
{
    g_used++;
    if (g_pos >= g_stride) {
        g_overflow = 1;
        return 0.5;
    }
    return g_tape[g_pos++];
}

/* The reference's LFIT (cern/lfit.f:11-60) over n point sets of L points
 * each (rows of X and Y), KEY = 0: A, B, E per set. */
extern void lfit_(float *X, float *Y, int32_t *L, int32_t *KEY, float *A, float *B, float *E);
void ref_lfit(int64_t n, int32_t L, const float *X, const float *Y, float *A, float *B, float *E)
{
    for (int64_t i = 0; i < n; i++) {
        int32_t l = L, key = 0;
        lfit_((float *)(X + i * L), (float *)(Y + i * L), &l, &key, A + i, B + i, E + i);
    }
}

/* Runs n events through mc_hms (arm 1), mc_sos (2), mc_hrsr (3), mc_hrsl (4) or mc_shms (5).  Same SoA
 * conventions as or_run; counters receive the change of the reference's
 * STOP counters over the call.  mapdir must hold hms/ and shms/ map files
 * (the reference opens them relative to the working directory). */
int ref_run(int arm, double p_spec, double m2_in, int ms_flag, int wcs_flag, int decay_flag,
            double ctau, int using_coll, int64_t n, const double *dpp_in, const double *x_in,
            const double *y_in, const double *z_in, const double *dxdz_in,
            const double *dydz_in, const double *fry_in, double *dpp, double *y, double *dxdz,
            double *dydz, double *x_fp, double *dx_fp, double *y_fp, double *dy_fp,
            double *pathlen, double *resmult, double *m2_out, uint8_t *ok, int32_t *used,
            const double *tape, int64_t stride, const char *mapdir, int64_t *counters,
            uint64_t *stop_mask)
{
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return -1;
    if (chdir(mapdir) != 0) return -2;
    int ncount = arm == 1 ? H_NCOUNT : arm == 2 ? SO_NCOUNT : arm == 5 ? S_NCOUNT : R_NCOUNT;
    const int32_t *common = arm == 1 ? spectrometer1_.v : arm == 2 ? spectrometer2_.v
                          : arm == 3 ? spectrometer3_.v : arm == 4 ? spectrometer4_.v
                                     : spectrometer_shms_.v;
    int32_t before[S_NCOUNT];
    memcpy(before, common, ncount * sizeof(int32_t));
    decd_.ctau = ctau;
    int rc = 0;
    int32_t prev[S_NCOUNT];
    for (int64_t i = 0; i < n; i++) {
        memcpy(prev, common, ncount * sizeof(int32_t));
        g_tape = tape + i * stride;
        g_stride = stride;
        g_pos = 0;
        g_used = 0;
        g_overflow = 0;
        double ps = p_spec, th = 0.0, d = dpp_in[i], xx = x_in[i], yy = y_in[i];
        double zz = z_in ? z_in[i] : 0.0, dx = dxdz_in[i], dy = dydz_in[i];
        double fry = fry_in ? fry_in[i] : x_in[i];
        double xf = x_fp[i], dxf = dx_fp[i], yf = y_fp[i], dyf = dy_fp[i];
        double m2 = m2_in, rm = resmult ? resmult[i] : 0.0, pl = pathlen[i];
        int32_t msf = ms_flag, wcf = wcs_flag, dcf = decay_flag, okv = 0, uc = using_coll;
        int32_t spectr = 5;
        if (arm == 1)
            mc_hms_(&ps, &th, &d, &xx, &yy, &zz, &dx, &dy, &xf, &dxf, &yf, &dyf, &m2, &msf, &wcf,
                    &dcf, &rm, &fry, &okv, &pl, &uc);
        else if (arm == 2)
            mc_sos_(&ps, &th, &d, &xx, &yy, &zz, &dx, &dy, &xf, &dxf, &yf, &dyf, &m2, &msf, &wcf,
                    &dcf, &rm, &fry, &okv, &pl);
        else if (arm == 3)
            mc_hrsr_(&ps, &th, &d, &xx, &yy, &zz, &dx, &dy, &xf, &dxf, &yf, &dyf, &m2, &msf, &wcf,
                     &dcf, &rm, &fry, &okv, &pl);
        else if (arm == 4)
            mc_hrsl_(&ps, &th, &d, &xx, &yy, &zz, &dx, &dy, &xf, &dxf, &yf, &dyf, &m2, &msf, &wcf,
                     &dcf, &rm, &fry, &okv, &pl);
        else
            mc_shms_(&ps, &th, &d, &xx, &yy, &zz, &dx, &dy, &xf, &dxf, &yf, &dyf, &m2, &msf,
                     &wcf, &dcf, &rm, &fry, &okv, &pl, &spectr, &uc);
        dpp[i] = d;
        y[i] = yy;
        dxdz[i] = dx;
        dydz[i] = dy;
        x_fp[i] = xf;
        dx_fp[i] = dxf;
        y_fp[i] = yf;
        dy_fp[i] = dyf;
        pathlen[i] = pl;
        if (resmult) resmult[i] = rm;
        if (m2_out) m2_out[i] = m2;
        ok[i] = (uint8_t)(okv != 0);
        if (used) used[i] = (int32_t)g_used;
        if (g_overflow) rc = OR_E_TAPE;
        /* which of the reference's STOP counters this event moved (bit c for
         * counter c of the common block): its exit point as the reference
         * records it */
        if (stop_mask) {
            uint64_t m = 0;
            for (int c = 0; c < ncount && c < 64; c++)
                if (common[c] != prev[c]) m |= (uint64_t)1 << c;
            stop_mask[i] = m;
        }
    }
    const int32_t *after = common;
    for (int c = 0; c < ncount; c++) counters[c] += (int64_t)(after[c] - before[c]);
    if (chdir(cwd) != 0) return -3;
    return rc;
}

/* ---- target step: the reference's trip_thru_target and target_musc ---- */
extern void ref_set_target_(double *a, double *z, double *rho, double *x0, double *tlen,
                            double *tang, double *zoff, int32_t *ican, int32_t *iearm,
                            int32_t *iharm);
extern void trip_thru_target_(int32_t *narm, double *zpos, double *energy, double *theta,
                              double *eloss, double *radlen, double *mass, int32_t *typeflag);
extern void target_musc_(double *p, double *beta, double *teff, double *dangles);

/* n events of trip_thru_target(narm, zpos[i], energy[i], theta[i], ..., mass,
 * typeflag) followed, when p_musc is not NULL, by target_musc(p_musc[i],
 * beta_musc[i], teff, dangles), each event on its own tape slot.  tgt =
 * {A, Z, rho, X0 (g/cm^2), length, angle, zoffset}; spec is the arm's
 * spectrometer number (electron_arm and hadron_arm both set to it).
 * used[i] = uniforms drawn, -1 on tape overflow. */
int ref_target_step(const double *tgt, int can, int narm, int spec, double mass, int typeflag,
                    int64_t n, const double *zpos, const double *energy, const double *theta,
                    const double *p_musc, const double *beta_musc, double *eloss, double *teff,
                    double *dang1, double *dang2, int32_t *used, const double *tape,
                    int64_t stride)
{
    double a = tgt[0], z = tgt[1], rho = tgt[2], x0 = tgt[3], len = tgt[4], ang = tgt[5],
           zoff = tgt[6];
    int32_t ican = can, arm = spec;
    ref_set_target_(&a, &z, &rho, &x0, &len, &ang, &zoff, &ican, &arm, &arm);
    for (int64_t i = 0; i < n; i++) {
        g_tape = tape + i * stride;
        g_stride = stride;
        g_pos = 0;
        g_used = 0;
        g_overflow = 0;
        int32_t na = narm, tf = typeflag;
        double zp = zpos[i], en = energy[i], th = theta[i], el = 0.0, rl = 0.0, m = mass;
        trip_thru_target_(&na, &zp, &en, &th, &el, &rl, &m, &tf);
        double d[2] = {0.0, 0.0};
        if (p_musc) {
            double p = p_musc[i], b = beta_musc[i];
            target_musc_(&p, &b, &rl, d);
        }
        eloss[i] = el;
        teff[i] = rl;
        dang1[i] = d[0];
        dang2[i] = d[1];
        used[i] = g_overflow ? -1 : (int32_t)g_used;
    }
    return 0;
}

/* ---- target-field tracking: the reference's trg_track.f ---- */
extern void trginit_(const char *map, double *theta_e, double *phi_e, double *theta_p,
                     double *phi_p, size_t map_len);
extern void track_from_tgt_(double *x, double *y, double *z, double *dx, double *dy,
                            double *mom, double *mass, int32_t *spect, int32_t *ok);
extern void track_to_tgt_(double *delta, double *y, double *dx, double *dy, double *frx,
                          double *fry, double *mom, double *mass, double *ctheta,
                          double *stheta, int32_t *spect, int32_t *ok, int32_t *arm);
extern struct { double xs, ys, zs, dpps, dxdzs, dydzs; } track_;  /* spectrometers.inc:47 */

/* trgInit(map, ...): map NULL = the uniform test field (' '), "0" = zero field. */
void ref_trg_init(const char *map, double theta_e, double phi_e, double theta_p, double phi_p)
{
    const char *m = map ? map : " ";
    trginit_(m, &theta_e, &phi_e, &theta_p, &phi_p, strlen(m));
}

void ref_track_from_tgt(int64_t n, double *x, double *y, double *z, double *dx, double *dy,
                        const double *mom, double mass, int spect, uint8_t *ok)
{
    for (int64_t i = 0; i < n; i++) {
        double m = mom[i], ms = mass;
        int32_t sp = spect, o = 0;
        track_from_tgt_(&x[i], &y[i], &z[i], &dx[i], &dy[i], &m, &ms, &sp, &o);
        ok[i] = (uint8_t)(o != 0);
    }
}

/* track_to_tgt per event, with /track/ set to the event's focal-plane track;
 * mapdir holds the recon maps the reference opens on first use. */
int ref_track_to_tgt(int arm, int64_t n, const double *x_fp, const double *dx_fp,
                     const double *y_fp, const double *dy_fp, double *delta, double *y,
                     double *dx, double *dy, const double *frx, const double *fry,
                     const double *mom, double mass, double ctheta, double stheta, int spect,
                     uint8_t *ok, const char *mapdir)
{
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return -1;
    if (chdir(mapdir) != 0) return -2;
    for (int64_t i = 0; i < n; i++) {
        track_.xs = x_fp[i];
        track_.ys = y_fp[i];
        track_.dxdzs = dx_fp[i];
        track_.dydzs = dy_fp[i];
        double fx = frx[i], fy = fry[i], m = mom[i], ms = mass, ct = ctheta, st = stheta;
        int32_t sp = spect, o = 1, a = arm;
        track_to_tgt_(&delta[i], &y[i], &dx[i], &dy[i], &fx, &fy, &m, &ms, &ct, &st, &sp, &o, &a);
        ok[i] = (uint8_t)(o != 0);
    }
    if (chdir(cwd) != 0) return -3;
    return 0;
}

/* ---- generator: the reference's generate + complete_main (event.f) ---- */
void ref_gen_one_(const double *par, const int32_t *ipar, double *out, int32_t *ok);

/* n events of H(e,e'p) elastic generation without radiation, one tape slot
 * each; target = A, Z, rho, X0, length, angle, zoffset; can; the electron
 * and hadron arms' spectrometer codes (their windows enter complete_ev's
 * losses, drawn after the generator's uniforms).  out: 27 values per event. */
int ref_gen_hep(const double *target, int32_t can, int32_t earm, int32_t harm, const double *par,
                const int32_t *ipar, int64_t n, const double *tape, int64_t stride, double *out,
                int32_t *ok, int32_t *used)
{
    double a = target[0], z = target[1], rho = target[2], x0 = target[3], len = target[4];
    double ang = target[5], zoff = target[6];
    ref_set_target_(&a, &z, &rho, &x0, &len, &ang, &zoff, &can, &earm, &harm);
    int rc = 0;
    for (int64_t i = 0; i < n; i++) {
        g_tape = tape + i * stride;
        g_stride = stride;
        g_pos = 0;
        g_used = 0;
        g_overflow = 0;
        ref_gen_one_(par, ipar, out + 27 * i, ok + i);
        used[i] = (int32_t)g_used;
        if (g_overflow) rc = -2;
    }
    return rc;
}

/* ---- radiative tails: the reference's generate with using_rad ---- */
void ref_gen_rad_one_(const double *par, const int32_t *ipar, const double *rpar, const int32_t *irpar,
                      double *out, int32_t *ok);

/* n events of H(e,e'p) elastic generation with radiation (generate_rad),
 * one tape slot each; out: 44 values per event (ref_gen.f). */
int ref_gen_hep_rad(const double *target, int32_t can, int32_t earm, int32_t harm, const double *par,
                    const int32_t *ipar, const double *rpar, const int32_t *irpar, int64_t n,
                    const double *tape, int64_t stride, double *out, int32_t *ok, int32_t *used)
{
    double a = target[0], z = target[1], rho = target[2], x0 = target[3], len = target[4];
    double ang = target[5], zoff = target[6];
    ref_set_target_(&a, &z, &rho, &x0, &len, &ang, &zoff, &can, &earm, &harm);
    int rc = 0;
    for (int64_t i = 0; i < n; i++) {
        g_tape = tape + i * stride;
        g_stride = stride;
        g_pos = 0;
        g_used = 0;
        g_overflow = 0;
        ref_gen_rad_one_(par, ipar, rpar, irpar, out + 44 * i, ok + i);
        used[i] = (int32_t)g_used;
        if (g_overflow) rc = -2;
    }
    return rc;
}

/* the reference's brem, bremos, gamma and lambda_dave on given arguments */
extern void ref_rad_setflags_(int32_t *iexpon);
extern double brem_(double *ein, double *eout, double *egamma, int32_t *radiate_proton, double *bsoft,
                    double *bhard, double *dbsoft);
extern double bremos_(double *egamma, double *kix, double *kiy, double *kiz, double *kfx, double *kfy,
                      double *kfz, double *pix, double *piy, double *piz, double *pfx, double *pfy,
                      double *pfz, double *pfe, int32_t *radiate_proton, double *bsoft, double *bhard,
                      double *dbsoft);
extern double gamma_(double *x);
extern double lambda_dave_(int32_t *itail, int32_t *plus_flag, int32_t *doing_proton, double *e1,
                           double *e2, double *e3, double *p3, double *th);

/* out: r, bsoft, bhard, dbsoft per call (4 values) */
void ref_brem(int64_t n, int32_t use_expon, const double *ein, const double *eout, const double *egamma,
              const int32_t *radiate_proton, double *out)
{
    ref_rad_setflags_(&use_expon);
    for (int64_t i = 0; i < n; i++) {
        double a = ein[i], b = eout[i], c = egamma[i];
        int32_t rp = radiate_proton[i];
        out[4 * i] = brem_(&a, &b, &c, &rp, &out[4 * i + 1], &out[4 * i + 2], &out[4 * i + 3]);
    }
}

/* args: 14 doubles per call (egamma, k_i, k_f, p_i, p_f, p_fe); out as ref_brem */
void ref_bremos(int64_t n, int32_t use_expon, const double *args, const int32_t *radiate_proton, double *out)
{
    ref_rad_setflags_(&use_expon);
    for (int64_t i = 0; i < n; i++) {
        double v[14];
        memcpy(v, args + 14 * i, sizeof v);
        int32_t rp = radiate_proton[i];
        out[4 * i] = bremos_(&v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9], &v[10],
                             &v[11], &v[12], &v[13], &rp, &out[4 * i + 1], &out[4 * i + 2], &out[4 * i + 3]);
    }
}

void ref_rad_gamma(int64_t n, const double *x, double *out)
{
    for (int64_t i = 0; i < n; i++) {
        double v = x[i];
        out[i] = gamma_(&v);
    }
}

/* args: itail, plus_flag, doing_proton as ints; e1, e2, e3, p3, th as doubles (5 per call) */
void ref_lambda_dave(int64_t n, const int32_t *iargs, const double *dargs, double *out)
{
    for (int64_t i = 0; i < n; i++) {
        int32_t it = iargs[3 * i], pf = iargs[3 * i + 1], dp = iargs[3 * i + 2];
        double d[5];
        memcpy(d, dargs + 5 * i, sizeof d);
        out[i] = lambda_dave_(&it, &pf, &dp, &d[0], &d[1], &d[2], &d[3], &d[4]);
    }
}
