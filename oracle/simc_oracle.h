/*
 * simc_oracle.h -- FP64 CPU restatement of SIMC's spectrometer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in simc_gfortran_amd/ and the timed CPU baseline of
 * bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every routine restates one reference routine (file:line cited at its
 * definition in simc_oracle.c) with the same operation order, constants and
 * random-number consumption order, so that the reference compiled from its
 * own sources (oracle/_ref) and this restatement agree to the last bit on the
 * same uniform tape, up to libm ulp differences.
 */
#ifndef SIMC_ORACLE_H
#define SIMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_CLASS 41     /* spectrometers.inc:5 */
#define OR_MAX_TERMS 1000   /* shared/transp.f:81, hms/mc_hms_recon.f:43 */

/* error codes (negative) */
#define OR_OK 0
#define OR_E_IO (-1)
#define OR_E_MAPFMT (-2)
#define OR_E_NCLASS (-3)
#define OR_E_TOF (-4)
#define OR_E_TOOMANY (-5)
#define OR_E_ARG (-6)
#define OR_E_TAPE (-7)

typedef struct {
    int n_classes;
    int n_terms[OR_MAX_CLASS];
    double length[OR_MAX_CLASS];     /* cm, from !LENGTH: comments (0 if none) */
    int adrift[OR_MAX_CLASS];
    double driftdist[OR_MAX_CLASS];  /* cm */
    double coeff[OR_MAX_CLASS][OR_MAX_TERMS][5];
    short expon[OR_MAX_CLASS][OR_MAX_TERMS][5]; /* x, x', y, y', delta */
} or_fwd_map;

typedef struct {
    int n_terms;
    int max_order;
    double coeff[OR_MAX_TERMS][4];
    short expon[OR_MAX_TERMS][5];
} or_recon_map;

/* Uniform source.  kind: 0 = tape, 1 = philox (per event), 2 = ranlux (global). */
typedef struct {
    int kind;
    /* tape */
    const double *tape;
    int64_t stride;
    int64_t pos;
    /* philox */
    uint32_t key[2];
    uint64_t event;
    uint32_t buf[4];
    uint32_t tag;       /* fourth counter word; 0 = the arms' stream (OR_PHILOX_TAG) */
    /* ranlux */
    struct or_ranlux *rl;
    /* bookkeeping */
    int64_t used;
    int overflow;
    /* philox normal deviates: Box-Muller second value pending */
    int have;
    double pending;
} or_rng;

typedef struct or_ranlux {
    double seeds[25];           /* 1-based as in cern/ranlux.f */
    int next[25];
    int i24, j24, in24, nskip, luxlev;
    double carry, twom24, twom12;
    int64_t kount;
    /* grnd() buffer, call_ranlux.f:58-73 */
    double rvec[1001];
    int latest;
} or_ranlux;

/* Per-event transport state: the reference's /track/ common plus the decay
 * bookkeeping of /ntuple_variables/ and /gnrl/. */
typedef struct {
    double xs, ys, zs, dxdzs, dydzs, dpps;
    double decdist, Mh2_final;
    double xs_fp, ys_fp, dxdzs_fp, dydzs_fp;
} or_track;

/* HMS stop counters, same order as common /spectrometer1/ (hms/struct_hms.inc:34-42) */
enum {
    H_TRIALS, H_SLIT_HOR, H_SLIT_VERT, H_SLIT_OCT, H_Q1_IN, H_Q1_MID, H_Q1_OUT,
    H_Q2_IN, H_Q2_MID, H_Q2_OUT, H_Q3_IN, H_Q3_MID, H_Q3_OUT, H_D1_IN, H_D1_OUT,
    H_HUT, H_DC1, H_DC2, H_SCIN, H_CAL, H_SUCCESSES, H_COLL, H_NCOUNT
};
/* SHMS stop counters, same order as common /spectrometer_shms/ (shms/struct_shms.inc:26-44) */
enum {
    S_TRIALS, S_SUCCESSES, S_TARG_HOR, S_TARG_VERT, S_TARG_OCT, S_SLIT_HOR, S_SLIT_VERT,
    S_SLIT_OCT, S_HB_IN, S_HB_MEN, S_HB_MEX, S_HB_OUT, S_Q1_IN, S_Q1_MEN, S_Q1_MID,
    S_Q1_MEX, S_Q1_OUT, S_Q2_IN, S_Q2_MEN, S_Q2_MID, S_Q2_MEX, S_Q2_OUT, S_Q3_IN,
    S_Q3_MEN, S_Q3_MID, S_Q3_MEX, S_Q3_OUT, S_D1_IN, S_D1_FLR, S_D1_MEN, S_D1_MID1,
    S_D1_MID2, S_D1_MID3, S_D1_MID4, S_D1_MID5, S_D1_MID6, S_D1_MID7, S_D1_MEX,
    S_D1_OUT, S_BP_IN, S_BP_OUT, S_HUT, S_DC1, S_DC2, S_CER1, S_CER2, S_S1, S_S2,
    S_S3, S_CAL, S_STOP_ID, S_COLL, S_NCOUNT
};

/* SOS stop counters, same order as common /spectrometer2/ (sos/struct_sos.inc:1-17) */
enum {
    SO_TRIALS, SO_SLIT_HOR, SO_SLIT_VERT, SO_SLIT_OCT, SO_QUAD_IN, SO_QUAD_MID, SO_QUAD_OUT,
    SO_BM01_IN, SO_BM01_OUT, SO_BM02_IN, SO_BM02_OUT, SO_EXIT, SO_HUT, SO_DC1, SO_DC2,
    SO_SCIN, SO_CAL, SO_SUCCESSES, SO_NCOUNT
};

/* HRS stop counters, same order as common /spectrometer3/ (HRS-R, hrsr/struct_hrsr.inc:24-40)
 * and /spectrometer4/ (HRS-L, hrsl/struct_hrsl.inc:24-40) */
enum {
    R_TRIALS, R_SLIT_HOR, R_SLIT_VERT, R_Q1_IN, R_Q1_MID, R_Q1_OUT, R_Q2_IN, R_Q2_MID,
    R_Q2_OUT, R_Q3_IN, R_Q3_MID, R_Q3_OUT, R_D1_IN, R_D1_OUT, R_HUT, R_DC1, R_DC2, R_S1,
    R_S2, R_CAL, R_SUCCESSES, R_NCOUNT
};

/* Stop codes: one unambiguous code per exit point of the path.  The batch
 * drivers write one per event; 0 = accepted. */
enum {
    STOP_OK = 0,
    /* shared front end */
    STOP_SLIT_HOR = 1, STOP_SLIT_VERT, STOP_SLIT_OCT, STOP_COLL,
    /* HMS */
    STOP_H_Q1_IN = 10, STOP_H_Q1_MID, STOP_H_Q1_OUT, STOP_H_Q2_IN, STOP_H_Q2_MID,
    STOP_H_Q2_OUT, STOP_H_Q3_IN, STOP_H_Q3_MID, STOP_H_Q3_OUT, STOP_H_D1_IN,
    STOP_H_D1_OUT, STOP_H_PIPE0, STOP_H_PIPE1, STOP_H_PIPE2, STOP_H_PIPE3,
    STOP_H_DC1, STOP_H_DC2, STOP_H_SCIN,
    /* SHMS */
    STOP_S_HB_IN = 40, STOP_S_HB_MEN, STOP_S_HB_MEX, STOP_S_HB_OUT,
    STOP_S_Q1_IN, STOP_S_Q1_MEN, STOP_S_Q1_MID, STOP_S_Q1_MEX, STOP_S_Q1_OUT,
    STOP_S_Q2_IN, STOP_S_Q2_MEN, STOP_S_Q2_MID, STOP_S_Q2_MEX, STOP_S_Q2_OUT,
    STOP_S_Q3_IN, STOP_S_Q3_MEN, STOP_S_Q3_MID, STOP_S_Q3_MEX, STOP_S_Q3_OUT,
    STOP_S_D1_IN, STOP_S_D1_FLR, STOP_S_D1_MEN, STOP_S_D1_MID1, STOP_S_D1_MID2,
    STOP_S_D1_MID3, STOP_S_D1_MID4, STOP_S_D1_MID5, STOP_S_D1_MID6, STOP_S_D1_MID7,
    STOP_S_D1_MEX, STOP_S_D1_OUT, STOP_S_DC1, STOP_S_DC2, STOP_S_S1X, STOP_S_S1Y,
    STOP_S_S2X, STOP_S_S2Y, STOP_S_CAL, STOP_S_CALFID,
    /* SOS */
    STOP_SO_QUAD_IN = 100, STOP_SO_QUAD_MID, STOP_SO_QUAD_OUT, STOP_SO_BM01_IN,
    STOP_SO_BM01_OUT, STOP_SO_BM02_IN, STOP_SO_BM02_OUT, STOP_SO_EXIT_FLANGE,
    STOP_SO_EXIT_APERTURE, STOP_SO_EXIT_WINDOW, STOP_SO_DC1, STOP_SO_DC2, STOP_SO_SCIN,
    /* HRS (both arms) */
    STOP_R_PIPE_65 = 120, STOP_R_PIPE_80, STOP_R_Q1_PRE, STOP_R_Q1_IN, STOP_R_Q1_MID,
    STOP_R_Q1_OUT, STOP_R_Q1_POST, STOP_R_Q2_PRE, STOP_R_Q2_IN, STOP_R_Q2_MID, STOP_R_Q2_OUT,
    STOP_R_Q2_POST1, STOP_R_Q2_POST2, STOP_R_D1_PRE, STOP_R_D1_IN, STOP_R_D1_OUT,
    STOP_R_D1_POST_R, STOP_R_D1_POST_BOX, STOP_R_Q3_PRE, STOP_R_Q3_IN, STOP_R_Q3_MID,
    STOP_R_Q3_OUT, STOP_R_Q3_POST, STOP_R_WINDOW, STOP_R_DC1, STOP_R_DC2, STOP_R_S1, STOP_R_S2,
    /* bookkeeping */
    STOP_TAPE_OVERFLOW = 250
};

typedef struct {
    double p_spec;      /* MeV/c */
    double m2;          /* MeV^2, initial particle mass squared */
    double ctau;        /* cm, decay length constant (simulate.inc:145) */
    int ms_flag, wcs_flag, decay_flag, using_coll;
    int fit_f32;        /* 1: REAL*4 track fit as the reference; 0: FP64 fit */
    double wire_sigma;  /* cm; <= 0 keeps the reference table */
    double mom_offset;  /* s_p of the adjoint form's momentum scale: the HMS and SHMS maps see
                           delta' = 100((1 + delta/100)/(1 + s_p) - 1), scattering the true
                           momentum; 0 = off (the reference) */
    int sieve;          /* mock sieve slit, compile-time off in the reference: 0 off, 1 the arm's
                           use_sieve, 2 the SHMS's use_2sieve (in front of the bender) */
} or_settings;

typedef struct {
    or_fwd_map *hms_fwd, *shms_fwd;
    or_recon_map *hms_rec, *shms_rec;
    or_fwd_map *sos_fwd;        /* NULL when dir has no sos/ maps */
    or_recon_map *sos_rec;
    or_fwd_map *hrsr_fwd, *hrsl_fwd;   /* NULL when dir has no hrsr/ (hrsl/) maps */
    or_recon_map *hrsr_rec, *hrsl_rec;
} or_maps;

/* ---- map files ---- */
int or_read_forward(const char *path, or_fwd_map *m);
int or_read_recon(const char *path, or_recon_map *m);
or_maps *or_maps_load(const char *dir);   /* dir holds hms/ and shms/ (+ optional sos/, hrsr/, hrsl/) */
void or_maps_free(or_maps *m);

/* ---- random numbers ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_philox_uniform(uint32_t key0, uint32_t key1, uint64_t event, uint64_t k);
or_ranlux *or_ranlux_new(int lux, int seed);
void or_ranlux_fill(or_ranlux *r, double *rvec, int len);
void or_ranlux_save(const or_ranlux *r, int32_t isdext[25]);
void or_ranlux_restore(or_ranlux *r, const int32_t isdext[25]);
double or_grnd(or_rng *g);

/* ---- single routines (exported for unit tests) ---- */
double or_powi(double x, int m);
double or_gauss1(or_rng *g, double nsigmax);
double or_bm_log(double u);
void or_bm_sincos2pi(double u, double *sn, double *cs);
/* n normal deviates of the philox stream of one event (unit tests);
 * returns the uniforms consumed. */
int64_t or_philox_normals(uint64_t seed, uint64_t event, int n, double nsigmax, double *out);
void or_lfit_f32(const float *X, const float *Y, int L, int KEY, float *A, float *B, float *E);
void or_lfit_f64(const double *X, const double *Y, int L, int KEY, double *A, double *B, double *E);
void or_loren(double gam, double bx, double by, double bz, double e, double x, double y,
              double z, double *ef1, double *pxf, double *pyf, double *pzf, double *pf1);
void or_enerloss_new(or_rng *g, double len, double dens, double zeff, double aeff,
                     double epart, double mpart, int typeflag, double *eloss);
double or_pion_coll_absorb(double ppi, double thick);
void or_rotate_haxis(const or_track *t, double rotang, double *xp0, double *yp0);
void or_rotate_vaxis(const or_track *t, double rotang, double *xp0, double *yp0);
void or_transp_eval(const or_fwd_map *m, int cls, const double ray[5], double sum[5]);
void or_recon_eval(const or_recon_map *m, const double hut[5], double sum[4]);

/* ---- per-event arms (argument lists of hms/mc_hms.f:1-4, shms/mc_shms.f:1-4) ---- */
void or_mc_hms(const or_maps *M, const or_settings *S, or_rng *g, or_track *t, int64_t *cnt,
               double p_spec, double *dpp, double *x, double *y, double *z, double *dxdz,
               double *dydz, double *x_fp, double *dx_fp, double *y_fp, double *dy_fp,
               double *m2, double *resmult, double fry, int *ok_spec, double *pathlen,
               int *stop_code, int *fit_done);
void or_mc_sos(const or_maps *M, const or_settings *S, or_rng *g, or_track *t, int64_t *cnt,
               double p_spec, double *dpp, double *x, double *y, double *z, double *dxdz,
               double *dydz, double *x_fp, double *dx_fp, double *y_fp, double *dy_fp,
               double *m2, double *resmult, double fry, int *ok_spec, double *pathlen,
               int *stop_code, int *fit_done);
/* mc_hrsr (arm 3) / mc_hrsl (arm 4), hrsr/mc_hrsr.f:1-3, hrsl/mc_hrsl.f:1-3 */
void or_mc_hrs(const or_maps *M, int arm, const or_settings *S, or_rng *g, or_track *t,
               int64_t *cnt, double p_spec, double *dpp, double *x, double *y, double *z,
               double *dxdz, double *dydz, double *x_fp, double *dx_fp, double *y_fp,
               double *dy_fp, double *m2, double *resmult, double fry, int *ok_spec,
               double *pathlen, int *stop_code, int *fit_done);
void or_mc_shms(const or_maps *M, const or_settings *S, or_rng *g, or_track *t, int64_t *cnt,
                double p_spec, double *dpp, double *x, double *y, double *z, double *dxdz,
                double *dydz, double *x_fp, double *dx_fp, double *y_fp, double *dy_fp,
                double *m2, double *resmult, double fry, int *ok_spec, double *pathlen,
                int *stop_code, int *fit_done);

/* ---- batch driver ----
 * arm: 1 = HMS, 2 = SOS, 3 = HRS-R, 4 = HRS-L, 5 = SHMS.  rng_kind: 0 tape (tape[i*stride + k]), 1 philox
 * (key = seed, stream = global event id = event0 + i), 2 ranlux (one global
 * stream in event order; forces a single thread).
 * Inputs are SoA; outputs follow the reference's INOUT conventions:
 *   dpp/y/dxdz/dydz: recon values on success, input values otherwise;
 *   x_fp..dy_fp: the track fit when it was reached, else left untouched;
 *   pathlen: input + accumulated path.
 * counters: the arm's H_/SO_/R_/S_NCOUNT int64, accumulated (not cleared). */
typedef struct {
    const double *dpp, *x, *y, *z, *dxdz, *dydz, *fry;
} or_events_in;
typedef struct {
    double *dpp, *y, *dxdz, *dydz;        /* recon (INOUT semantics) */
    double *x_fp, *dx_fp, *y_fp, *dy_fp;  /* focal plane */
    double *pathlen;                      /* INOUT */
    double *resmult, *m2;                 /* optional (may be NULL) */
    uint8_t *ok, *stop;                   /* stop code per event */
    int32_t *used;                        /* uniforms consumed (optional) */
} or_events_out;

int or_run(const or_maps *M, int arm, const or_settings *S, int64_t n, const or_events_in *in,
           or_events_out *out, int rng_kind, const double *tape, int64_t stride,
           uint64_t seed, uint64_t event0, int nthreads, int64_t *counters);

double or_gauss1_tape(const double *tape, int64_t len, double nsigmax, int64_t *used);
void or_ranlux_sequence(int lux, int seed, int32_t isdext[25], double *rvec, int n);
int or_recon_run(const or_maps *M, int arm, int64_t n, const double *x_fp, const double *dx_fp,
                 const double *y_fp, const double *dy_fp, const double *fry, double *delta,
                 double *dxdz, double *dydz, double *ytar);

/* ---- target step (event preparation before an arm) ----
 * Target of trip_thru_target (target.inc:31-53 fields used there); X0 in
 * g/cm^2 (or_target_X0 gives init.f's value). */
typedef struct {
    double A, Z, rho, X0, length, angle, zoffset;
    int can;
} or_target;
/* One particle's preparation: narm 1 beam, 2 electron, 3 hadron; spectrometer
 * = electron_arm / hadron_arm code; angles in rad, energies in MeV. */
typedef struct {
    int narm, spectrometer, using_eloss, mc_smear;
    double p_spec, th_spec, phi, mass, coulomb, beam_energy;
    double offset[5]; /* x, y, z (cm), xptar, yptar (rad) */
} or_prep;
typedef struct {
    /* E, p: after radiation (orig); E_vertex: before (vertex), NULL = E */
    const double *x, *y, *z, *E, *p, *E_vertex, *theta, *delta, *xptar, *yptar, *dang_yp, *dang_xp;
} or_vertex_in;
typedef struct {
    double *dpp, *x, *y, *z, *dxdz, *dydz, *fry, *eloss, *teff, *dang1, *dang2;
} or_prep_out;
#define OR_TARGET_TAG 0x54475453u   /* philox stream of the target step */
double or_target_X0(double A, double Z);
int or_trip_thru_target(or_rng *g, const or_target *T, int narm, int spec, double zpos,
                        double energy, double theta, double mass, int typeflag, double *eloss,
                        double *radlen);
void or_target_musc(or_rng *g, double p, double beta, double teff, double dangles[2]);
/* rng_kind 0 tape, 1 philox (key seed, event event0 + i, tag OR_TARGET_TAG).
 * Returns 0, -1 for an unknown spectrometer code, -2 on tape overflow. */
int or_target_prep(const or_target *T, const or_prep *P, int64_t n, const or_vertex_in *in,
                   or_prep_out *out, int rng_kind, const double *tape, int64_t stride,
                   uint64_t seed, uint64_t event0, int nthreads, int32_t *used);

/* ---- target-field tracking (trg_track.f) ---- */
#define OR_TRG_N 51
typedef struct {
    double bz[OR_TRG_N][OR_TRG_N];   /* B_field_z(iz, ir) at bz[ir][iz] (T) */
    double br[OR_TRG_N][OR_TRG_N];
    double zz[OR_TRG_N], rr[OR_TRG_N];
    double stht[2], ctht[2];         /* [0]: spect -1 (electron arm), [1]: spect +1 */
} or_field;
/* trgInit: map = NULL for the uniform 5 T test field, "0" for zero field,
 * else the 2601-line map file; angles in degrees.  Returns 0, -1 when the
 * file cannot be read. */
int or_trg_init(or_field *F, const char *map, double theta_e, double phi_e, double theta_p,
                double phi_p);
void or_trg_field(const or_field *F, const double x[3], double B[3], int spect);
void or_trg_rk4(const or_field *F, const double u0[9], double u1[9], double h, int spect,
                double factor);
void or_trg_track_to_plane(const or_field *F, double u[9], double E, double dl, double a,
                           double b, double c, double d, int *ok, int spect);
void or_track_from_tgt(const or_field *F, double *x, double *y, double *z, double *dx,
                       double *dy, double mom, double mass, int spect, int *ok);
/* batch forms (OpenMP over events): per-event mom (signed), one mass and
 * spect per call; ok as bytes */
void or_track_from_tgt_run(const or_field *F, int64_t n, double *x, double *y, double *z,
                           double *dx, double *dy, const double *mom, double mass, int spect,
                           uint8_t *ok, int nthreads);
void or_track_to_tgt_run(const or_field *F, const or_maps *M, int arm, int64_t n,
                         const double *x_fp, const double *dx_fp, const double *y_fp,
                         const double *dy_fp, double *delta, double *y, double *dx, double *dy,
                         const double *frx, const double *fry, const double *mom, double mass,
                         double ctheta, double stheta, int spect, uint8_t *ok, int nthreads);
/* track_to_tgt with the arm's reconstruction of the focal-plane track
 * (x_fp, dx_fp, y_fp, dy_fp: the /track/ values the recon reads). */
void or_track_to_tgt(const or_field *F, const or_maps *M, int arm, const double track[4],
                     double *delta, double *y, double *dx, double *dy, double frx, double fry,
                     double mom, double mass, double ctheta, double stheta, int spect, int *ok);

/* ---- event generator: H(e,e'p) elastic scattering, radiation off ----
 * generate (event.f:126-428) for doing_hyd_elast without using_rad,
 * complete_ev's hydrogen-elastic branch (event.f:432-1052) up to the
 * scattered particles' target losses (the target step does those), and
 * complete_main's weight (event.f:1393-1560) with sigep. */
typedef struct {
    double Ebeam, dEbeam;            /* MeV */
    double xwid, ywid;               /* gen%xwid, gen%ywid: beam sigma, cm */
    double xoffset, yoffset;         /* targ%xoffset, targ%yoffset, cm */
    double fr1, fr2;                 /* raster half-widths or radii, cm */
    double coulomb;                  /* targ%Coulomb_constant, MeV */
    double e_yptar[2], e_xptar[2];   /* gen%e%yptar, gen%e%xptar (min, max), rad */
    double e_theta, e_phi, e_P;      /* spec%e: central angles (rad), momentum (MeV/c) */
    double p_theta, p_phi, p_P;      /* spec%p */
    double Mh;                       /* hadron mass, MeV (Mp) */
    int fr_pattern, using_eloss, using_coulomb;
} or_hep;
typedef struct {
    double *x, *y, *z, *rasterx, *rastery, *Ein, *eloss, *teff;
    double *e_xptar, *e_yptar, *e_theta, *e_phi, *e_E, *e_delta;
    double *p_xptar, *p_yptar, *p_theta, *p_phi, *p_P, *p_E, *p_delta;
    double *Q2, *nu, *q, *jacobian, *sigcc, *weight;
    uint8_t *ok;
} or_hep_out;
#define OR_GEN_TAG 0x48454E47u   /* philox stream of the generator, "GNEH" */
#define OR_HUT_TAG 0x48555453u   /* philox stream of the detector huts, "HUTS" */
void or_physics_angles(double theta0, double phi0, double dx, double dy, double *theta,
                       double *phi);
void or_spectrometer_angles(double theta0, double phi0, double *dx, double *dy, double theta,
                            double phi);
void or_fofa_best_fit(double qsquar, double *GE, double *GM);
double or_sigmott(double e0, double theta, double Q2);
/* sigep of a vertex with these Q2, e%E, e%theta, Ein */
double or_sigep(double Q2, double E, double theta, double Ein);
/* rng_kind 0 tape, 1 philox (key seed, event event0 + i, tag OR_GEN_TAG).
 * Returns 0, -2 on tape overflow. */
int or_gen_hep(const or_target *T, const or_hep *H, int64_t n, or_hep_out *out, int rng_kind,
               const double *tape, int64_t stride, uint64_t seed, uint64_t event0, int nthreads,
               int32_t *used);

/* ---- radiative tails of H(e,e'p) elastic (radc.f, brem.f, init.f) ----
 * Run constants of radc_init (init.f:576-651) and limits_init
 * (init.f:91-470) that generate_rad reads, as explicit inputs. */
typedef struct {
    int rad_flag;          /* 0..3 (radc.f:141-159) */
    int extrad_flag;       /* 1..3; 0 = radc_init's default (init.f:621-629) */
    int intcor_mode;       /* 0 schwinger, 1 brem / bremos (init.f:690-702) */
    int use_expon;         /* brem's exponentiate (init.f:641-645) */
    int use_offshell_rad;  /* bremos instead of brem */
    int doing_tail[3];     /* dbase.f:494-511 from one_tail */
    int e_arm, p_arm;      /* spectrometers of complete_ev's window losses */
    double dE_edge_test;
    double Egamma_gen_max; /* hardwired limit when > 0.01 (dbase.f:514-518) */
    double edge_e_E[2], edge_p_E[2], edge_Em[2], vertex_Trec[2];  /* min, max */
    double Egamma_tot_max; /* init.f:409-421 */
    double etatzai;        /* init.f:634 */
    double targ_A;         /* schwinger's recoil factor (radc.f:729) */
} or_rad;
typedef struct {
    double *Egamma1, *Egamma2, *Egamma3;   /* Egamma_used(1..3) */
    double *gen_weight;                    /* main%gen_weight = rad_weight / hardcorfac */
    double *orig_Ein, *orig_e_E, *orig_e_delta, *orig_p_E, *orig_p_P, *orig_p_delta;
    double *eloss2, *teff2, *eloss3, *teff3;  /* complete_ev's losses (last call) */
    double *hardcorfac;
    int32_t *ntail;                        /* ntup%radarm */
} or_rad_out;
/* radc_init's etatzai for a target of charge Z (target_init's L1, L2). */
double or_rad_etatzai(double Z);
/* generate with using_rad for doing_hyd_elast (event.f:126-428 calling
 * generate_rad, radc.f:120-517) and complete_main's weight; hep outputs
 * hold the VERTEX kinematics (after tail 1), rad outputs the photons, the
 * ORIG kinematics and complete_ev's losses.  Returns 0, -2 on tape
 * overflow, -3 for an unsupported setting. */
int or_gen_hep_rad(const or_target *T, const or_hep *H, const or_rad *R, int64_t n, or_hep_out *out,
                   or_rad_out *rout, int rng_kind, const double *tape, int64_t stride, uint64_t seed,
                   uint64_t event0, int nthreads, int32_t *used);
/* pieces, for pins (or_rad_gamma: the GAMMA the reference's callers reach) */
double or_rad_gamma(double x);
double or_brem(const or_rad *R, double ein, double eout, double egamma, int radiate_proton, double *bsoft,
               double *bhard, double *dbsoft);
double or_bremos(const or_rad *R, double egamma, double k_ix, double k_iy, double k_iz, double k_fx,
                 double k_fy, double k_fz, double p_ix, double p_iy, double p_iz, double p_fx, double p_fy,
                 double p_fz, double p_fe, int radiate_proton, double *bsoft, double *bhard, double *dbsoft);
double or_lambda_dave(int itail, int plus_flag, int doing_proton, double e1, double e2, double e3, double p3,
                      double th);

#ifdef __cplusplus
}
#endif
#endif
