/*
 * simc_gpu.h -- C ABI of the MI355X spectrometer transport + reconstruction
 * library (libsimc_gpu.so).
 *
 * It replaces, for SIMC's per-event spectrometer hot path, the Fortran
 * subroutines
 *   mc_hms(p_spec, th_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp,
 *          dy_fp, m2, ms_flag, wcs_flag, decay_flag, resmult, fry, ok_spec,
 *          pathlen, using_HMScoll)                       hms/mc_hms.f:1-4
 *   mc_shms(... same ..., pathlen, spectr, using_SHMScoll) shms/mc_shms.f:1-4
 *   mc_hms_recon(delta_p, delta_t, delta_phi, y_tgt, fry) hms/mc_hms_recon.f:1
 *   mc_shms_recon(..., fry, spectr)                       shms/mc_shms_recon.f:1-2
 *   mc_sos(... as mc_hms without using_HMScoll)            sos/mc_sos.f:1-3
 *   mc_sos_recon(delta_p, delta_t, delta_phi, y_tgt, fry)  sos/mc_sos_recon.f:1
 *   transp_init(spectr, n_classes)                       shared/transp.f:294
 * called once per event per arm from montecarlo() (simc.f:1464-1488,
 * 1730-1754).  Here one call transports a whole batch of events held in
 * device memory as structure-of-arrays.
 *
 * Conventions: every entry point returns 0 or a negative SIMC_E_* code and
 * never exits the process; simc_last_error() describes the last failure of
 * the calling thread.  Pointers in simc_events_in / simc_events_out / tapes
 * are DEVICE pointers.  A context owns the device copies of the maps and the
 * stop counters; use one context per host thread.
 */
#ifndef SIMC_GPU_H
#define SIMC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SIMC_OK 0
#define SIMC_E_IO (-1)        /* map file cannot be opened (transp.f:318-321 stops)  */
#define SIMC_E_MAPFMT (-2)    /* malformed map line (Fortran read error)             */
#define SIMC_E_NCLASS (-3)    /* wrong class count (hms/mc_hms.f:188, shms/mc_shms.f:344) */
#define SIMC_E_TOF (-4)       /* non-zero TOF term (transp.f:385-389)                */
#define SIMC_E_TOOMANY (-5)   /* > 1000 terms or > 41 classes (transp.f:365,371)     */
#define SIMC_E_ARG (-6)       /* invalid argument                                    */
#define SIMC_E_TAPE (-7)      /* an event needed more uniforms than its tape slot    */
#define SIMC_E_MAPSTRUCT (-8) /* map terms differ from the compiled polynomial structure (adjoint form), or exceed order 8 */
#define SIMC_E_HIP (-9)       /* HIP runtime error                                   */
#define SIMC_E_NOMEM (-10)
#define SIMC_E_INTERNAL (-11) /* inconsistent step program (library defect)          */

#define SIMC_ARM_HMS 1        /* spectrometer numbers of transp_init, transp.f:300-313 */
#define SIMC_ARM_SOS 2
#define SIMC_ARM_HRSR 3       /* HRS-R (the reference's former hadron arm)               */
#define SIMC_ARM_HRSL 4       /* HRS-L (the reference's former electron arm)             */
#define SIMC_ARM_SHMS 5

#define SIMC_RNG_PHILOX 1     /* counter-based: uniform k of event e = Philox4x32-10(key=seed, ctr=(k/4, e, tag)) */
#define SIMC_RNG_TAPE 0       /* uniform k of event i = tape[i*stride + k] (parity runs)               */

#define SIMC_FIT_REF_F32 1    /* REAL*4 hits and fit parameters, as cern/lfit.f:18 */
#define SIMC_FIT_FP64 0       /* the same fit carried out in FP64                 */

#define SIMC_HMS_NCOUNTERS 22  /* common /spectrometer1/,      hms/struct_hms.inc:34-42   */
#define SIMC_SHMS_NCOUNTERS 52 /* common /spectrometer_shms/, shms/struct_shms.inc:26-44 */
#define SIMC_SOS_NCOUNTERS 18  /* common /spectrometer2/,     sos/struct_sos.inc:9-17    */
#define SIMC_HRS_NCOUNTERS 21  /* common /spectrometer3/ (HRS-R), /spectrometer4/ (HRS-L),
                                  hrsr/struct_hrsr.inc:24-40, hrsl/struct_hrsl.inc:24-40  */

typedef struct simc_ctx simc_ctx;

/* Per-call (per kinematic setting) scalars: the by-value arguments of mc_hms. */
typedef struct {
    double p_spec;        /* central momentum, MeV/c                          */
    double m2;            /* particle mass squared, MeV^2                     */
    double ctau;          /* decay length c*tau, cm (simulate.inc:145)        */
    int32_t ms_flag;      /* multiple scattering                              */
    int32_t wcs_flag;     /* wire-chamber smearing                            */
    int32_t decay_flag;   /* in-flight decay                                  */
    int32_t using_coll;   /* step pions/muons through the collimator          */
    int32_t fit_mode;     /* SIMC_FIT_REF_F32 or SIMC_FIT_FP64                */
    int32_t sieve;        /* mock sieve slit, compile-time off in the reference
                             (hms/mc_hms.f:155-168, shms/mc_shms.f:276-310,
                             sos/mc_sos.f:122-137, hrs?/mc_hrs?.f:122-132):
                             a bit mask: 0 off, 1 the arm's use_sieve,
                             2 the SHMS's use_2sieve (in front of the
                             bender), 3 both (SHMS)                       */
    double wire_sigma;    /* cm; <= 0 keeps the spectrometer's own table      */
} simc_settings;

/* Event inputs (device SoA, n doubles each).  z is accepted for interface
 * parity (the arms store it but never use it); fry may be NULL (= x). */
typedef struct {
    const double *dpp, *x, *y, *z, *dxdz, *dydz, *fry;
} simc_events_in;

/* Event outputs (device SoA).  dpp, y, dxdz, dydz, pathlen are INOUT with
 * mc_hms semantics and may alias the inputs: on success they receive the
 * reconstructed delta (%), y_tar (cm), x'_tar, y'_tar; on failure the input
 * values (hms/mc_hms.f:427-440).  pathlen accumulates.  x_fp..dy_fp are
 * written when the track fit was reached, else left untouched.  resmult is
 * set in the hut (hms/mc_hms_hut.f:292-297) and left untouched by events
 * that stop before it.  resmult, m2, used may be NULL.  stop receives the
 * exit point (SIMC_STOP_*), ok 1/0. */
typedef struct {
    double *dpp, *y, *dxdz, *dydz;
    double *x_fp, *dx_fp, *y_fp, *dy_fp;
    double *pathlen;
    double *resmult, *m2;
    uint8_t *ok, *stop;
    int32_t *used;
} simc_events_out;

typedef struct {
    int32_t kind;          /* SIMC_RNG_PHILOX or SIMC_RNG_TAPE */
    int32_t reserved;
    uint64_t seed;         /* Philox key                        */
    uint64_t event0;       /* global id of event 0 of this call */
    const double *tape;    /* device tape (SIMC_RNG_TAPE)       */
    int64_t stride;        /* uniforms per event slot           */
} simc_rng;

/* transp_init + the recon readers for the arms.  map_dir holds
 * hms/forward_cosy.dat, hms/recon_cosy.dat, shms/shms_forward.dat,
 * shms/shms_recon.dat and, optionally, sos/forward_cosy.dat,
 * sos/recon_cosy.dat, hrsr/hrs_forward_cosy.dat, hrsr/hrs_recon_cosy.dat,
 * hrsl/hrs_forward_cosy.dat and hrsl/hrs_recon_cosy.dat (the reference's
 * relative paths, transp.f:300-309, hms/mc_hms_recon.f:73,
 * shms/mc_shms_recon.f:74, sos/mc_sos_recon.f:73, hrsr/mc_hrsr_recon.f:74,
 * hrsl/mc_hrsl_recon.f:74); the optional arms are loaded when present. */
int simc_ctx_create(const char *map_dir, simc_ctx **out);
int simc_ctx_create_files(const char *hms_forward, const char *hms_recon,
                          const char *shms_forward, const char *shms_recon, simc_ctx **out);
/* Load (or replace) one arm's maps: transp_init(spectr) and the recon reader
 * of that arm (arm = SIMC_ARM_HMS, _SOS, _HRSR, _HRSL or _SHMS).  Maps with
 * the term structure the library was compiled for run its generated
 * polynomials; any other tune or fit (other exponents or non-zero pattern,
 * terms up to order 8) is evaluated from run-time tables built here, at a
 * lower rate.  The adjoint form needs the compiled structure
 * (SIMC_E_MAPSTRUCT otherwise). */
int simc_ctx_load_arm(simc_ctx *ctx, int arm, const char *forward, const char *recon);
void simc_ctx_destroy(simc_ctx *ctx);

/* Transport + reconstruct n events through one arm (mc_hms / mc_shms).
 * stream: a hipStream_t (NULL = default stream).  Asynchronous: returns after
 * the launch.  The context keeps one survivor queue per arm (92 bytes
 * per event of a launch, at most 2^24 events, allocated on first use): runs of
 * the same arm on one context must be ordered (same stream, or synchronised
 * streams). */
int simc_hms_run(simc_ctx *ctx, const simc_settings *s, const simc_events_in *in,
                 simc_events_out *out, int64_t n, const simc_rng *rng, void *stream);
int simc_shms_run(simc_ctx *ctx, const simc_settings *s, const simc_events_in *in,
                  simc_events_out *out, int64_t n, const simc_rng *rng, void *stream);
/* mc_sos(p_spec, th_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp,
 *        m2, ms_flag, wcs_flag, decay_flag, resmult, fry, ok_spec, pathlen)
 * sos/mc_sos.f:1-3 (no collimator stepping: using_coll is ignored). */
int simc_sos_run(simc_ctx *ctx, const simc_settings *s, const simc_events_in *in,
                 simc_events_out *out, int64_t n, const simc_rng *rng, void *stream);
/* mc_hrsr / mc_hrsl (same argument list as mc_sos), hrsr/mc_hrsr.f:1-3,
 * hrsl/mc_hrsl.f:1-3 (no collimator stepping: using_coll is ignored).  y_fp of
 * an accepted event is the fitted value minus the VDC offset (0.48 cm HRS-R,
 * 0.78 cm HRS-L, hrsr/mc_hrsr.f:524, hrsl/mc_hrsl.f:525); the reconstruction
 * uses the unshifted track, as the reference does. */
int simc_hrsr_run(simc_ctx *ctx, const simc_settings *s, const simc_events_in *in,
                  simc_events_out *out, int64_t n, const simc_rng *rng, void *stream);
int simc_hrsl_run(simc_ctx *ctx, const simc_settings *s, const simc_events_in *in,
                  simc_events_out *out, int64_t n, const simc_rng *rng, void *stream);
int simc_arm_run(simc_ctx *ctx, int arm, const simc_settings *s, const simc_events_in *in,
                 simc_events_out *out, int64_t n, const simc_rng *rng, void *stream);

/* simc_arm_run without any allocation (no counterpart in the reference: for
 * callers that cannot allocate device memory, such as code under HIP graph
 * capture or a tight serving loop).  The survivor queues live in a static
 * device array of the library (about 92 MiB per device, held until the
 * process ends, its address set up by simc_ctx_create); the call allocates
 * nothing and only launches kernels on `stream`.  At most
 * SIMC_NOALLOC_MAX_EVENTS events per call (larger n: SIMC_E_ARG).  Every call
 * on one device shares that array: calls that overlap on one device (other
 * streams, other contexts, other host threads) race and must be ordered by
 * the caller.  Results and counters equal simc_arm_run's. */
#define SIMC_NOALLOC_MAX_EVENTS 524288
int simc_arm_run_noalloc(simc_ctx *ctx, int arm, const simc_settings *s, const simc_events_in *in,
                         simc_events_out *out, int64_t n, const simc_rng *rng, void *stream);

/* simc_arm_run on HOST arrays (the same structs, host pointers; the tape too),
 * synchronous: every passed output is written before the call returns; ok
 * and stop may be NULL here.  The events go through in chunks of up to 2^20
 * on two library-owned streams with pinned staging buffers, so that the
 * copies into and out of the caller's arrays (split over host threads), the
 * transfers and the three stages of neighbouring chunks overlap.  Only the
 * columns the kernels read are sent (the inputs, and x_fp..dy_fp and
 * pathlen, which keep or accumulate the caller's values).  For callers
 * without device memory management, e.g. a Fortran batch driver of
 * montecarlo() (simc.f:1464-1488).  Calls on one context are serialised. */
int simc_arm_run_host(simc_ctx *ctx, int arm, const simc_settings *s, const simc_events_in *in,
                      simc_events_out *out, int64_t n, const simc_rng *rng);

/* simc_recon_run on host arrays (synchronous, as simc_arm_run_host). */
int simc_recon_run_host(simc_ctx *ctx, int arm, int64_t n, const double *x_fp, const double *dx_fp,
                        const double *y_fp, const double *dy_fp, const double *fry, double *delta,
                        double *dxdz_tar, double *dydz_tar, double *y_tar);

/* simc_arm_run with stage timing (no counterpart in the reference: a
 * measurement hook).  stage_events[0..4] are hipEvent_t handles created by
 * the caller; they are recorded on `stream` before stage 1 (target to
 * collimator), after stage 1, after stage 2 (the magnets), after stage 3
 * (detector hut and fit) and after the reconstruction of the accepted
 * events, in the first launch chunk (2^24 events, fewer for long tapes). */
int simc_arm_run_timed(simc_ctx *ctx, int arm, const simc_settings *s, const simc_events_in *in,
                       simc_events_out *out, int64_t n, const simc_rng *rng, void *stream,
                       void *const *stage_events);

/* ---- Batched form: many kinematic settings in one run -------------------
 *
 * Each setting is one SIMC run's worth of events for one arm (the by-value
 * arguments of mc_hms / mc_shms, its own Philox key and event numbering) over
 * a range [offset, offset + count) of the event arrays.  Settings may mix
 * arms and particle types; their ranges must not overlap and must lie in
 * [0, 2^29).  simc_batch_create groups the settings by arm and kernel
 * variant, uploads their table and reserves all device memory, so that
 * simc_batch_run only launches kernels on `stream` (it may be captured into
 * a HIP graph).  Counters are kept per setting.  Philox streams only:
 * event j of a setting draws from stream (seed, event0 + j), so settings
 * sharing a seed need disjoint event-id ranges.  Settings are launched
 * together (one pass) when they share arm, kernel variant and flags; when
 * their seeds differ, each event's lane keys its own setting's stream, so a
 * SIMC run's per-setting seed (simc.f:107, sgrnd) splits no pass.
 * The batch owns its survivor queues (sized by the launch tile): runs of the
 * context do not disturb a captured batch run.  simc_batch_set_chunk and
 * simc_batch_tune may replace them, and simc_ctx_load_arm replaces the arm's
 * tables: graphs captured before either call must be captured again. */
typedef struct {
    int32_t arm;            /* SIMC_ARM_HMS, _SOS, _HRSR, _HRSL or _SHMS        */
    int32_t reserved;
    simc_settings s;        /* this setting's scalars                           */
    uint64_t seed;          /* Philox key of this setting's streams             */
    uint64_t event0;        /* global id of its first event                     */
    int64_t offset;         /* index of its first event in the event arrays     */
    int64_t count;          /* its number of events                             */
} simc_setting;

typedef struct simc_batch simc_batch;

int simc_batch_create(simc_ctx *ctx, const simc_setting *settings, int n_settings, simc_batch **out);
/* in / out: device arrays covering every setting's range (the plain run's
 * semantics per event). */
int simc_batch_run(simc_batch *b, const simc_events_in *in, simc_events_out *out, void *stream);
/* STOP counters of setting i (input order) from the last run, common-block order. */
int simc_batch_counters(simc_batch *b, int setting, int64_t *c, int ncounters);
/* Launch tile: events per launch sequence (256 .. 2^24; reserves queues). */
int simc_batch_set_chunk(simc_batch *b, int64_t chunk);
/* Autotuner: times the run (warm-up + 3 repetitions) for launch tiles of
 * 2^18 .. 2^24 events, keeps the fastest and returns it and its time.  The
 * runs write `out` (use outputs that do not alias the inputs). */
int simc_batch_tune(simc_batch *b, const simc_events_in *in, simc_events_out *out, void *stream,
                    int64_t *best_chunk, double *best_ms);
void simc_batch_destroy(simc_batch *b);
/* How simc_batch_create grouped the settings: number of passes, and how many
 * of them key their lanes per setting (no counterpart in the reference). */
int simc_batch_passes(simc_batch *b, int32_t *n_passes, int32_t *n_lane_key);

/* ---- Adjoint form: gradients for optics fits ---------------------------
 *
 * Loss and gradients over the events with mask[i] != 0 (typically `ok` of a
 * plain run), their random draws replayed from the same stream (Philox: the
 * same seed and event0; tape: the same tape) and the mask held fixed:
 *   delta' = 100 ((1 + delta/100) / (1 + s_p) - 1), dy/dz' = dy/dz - dth,
 *   transport (FP64 track fit), X = x_fp + ox, Y = y_fp + oy, X' = dx_fp +
 *   ox', Y' = dy_fp + oy', reconstruction with recon_coef,
 *   L = 1/N sum [(delta_r - delta)^2 + (x'_r - dx/dz)^2 + (y'_r - dy/dz)^2 + (y_r - y)^2]
 * offsets = {ox, oy (cm), ox', oy' (rad), s_p, dth (rad)} (host array).
 * recon_coef, grad_coef: device [n_terms][4] (map file order: x'_tar, y_tar,
 * y'_tar, delta columns); out9: device, receives L, dL/d offsets (6), the
 * events excluded (a wire hit exactly 0, which the track fit would skip)
 * and N.  Asynchronous on `stream`.  No decay or collimator stepping. */
int simc_adjoint_run(simc_ctx *ctx, int arm, const simc_settings *s, const simc_events_in *in,
                     const uint8_t *mask, int64_t n, const simc_rng *rng, const double *recon_coef,
                     const double *offsets, double *grad_coef, double *out9, void *stream);
/* The loaded reconstruction map: n_terms, and its coefficients [n_terms][4]
 * when coef is not NULL (host array). */
int simc_recon_coefficients(simc_ctx *ctx, int arm, double *coef, int32_t *n_terms);

/* Reconstruction only (mc_hms_recon / mc_shms_recon) of focal-plane tracks:
 * x_fp, y_fp in cm, slopes in rad, fry in cm.  Outputs delta (%), x'_tar,
 * y'_tar (rad), y_tar (cm). */
int simc_recon_run(simc_ctx *ctx, int arm, int64_t n, const double *x_fp, const double *dx_fp,
                   const double *y_fp, const double *dy_fp, const double *fry, double *delta,
                   double *dxdz_tar, double *dydz_tar, double *y_tar, void *stream);

/* ---- Target step: event preparation before an arm ----------------------
 *
 * Per event, the part of SIMC's event loop between the vertex and the arm:
 *   trip_thru_target (target.f:1-246) with sampled losses (enerloss_new,
 *     typeflag 1) at the vertex energy, as complete_ev does (event.f:1031-1040;
 *     event.f:194 for the beam), the loss zeroed without using_Eloss;
 *   target_musc (target.f:548-574) when mc_smear;
 *   montecarlo()'s delta, angles, TRANSPORT transform, spectrometer offsets
 *     and drift to z = 0 (simc.f:1376-1441 hadron arm, :1649-1706 electron
 *     arm; :1364-1369 the beam's angles dang_in).
 * The outputs of an electron or hadron preparation are the inputs of
 * simc_arm_run (dpp, x, y, z = 0, dxdz, dydz, fry = x).  Draws come from a
 * stream of their own: Philox (seed, event0 + i) with the counter word
 * "TGTS", independent of the arm's stream of the same event; on a tape, the
 * reference's order (losses: target, Al, air, kevlar, mylar, each drawing one
 * gauss1(10) when its thickness is > 0; then the two gauss1(3.5) angles). */
#define SIMC_TARGET_BEAM 1      /* narm of trip_thru_target, target.f:27 */
#define SIMC_TARGET_ELECTRON 2
#define SIMC_TARGET_HADRON 3

typedef struct {
    double A, Z;          /* target nucleus (targ%A, targ%Z)                          */
    double rho;           /* g/cm^3                                                   */
    double X0;            /* radiation length, g/cm^2; <= 0: init.f:18-42 from A, Z  */
    double length;        /* cm                                                       */
    double angle;         /* target angle, rad                                        */
    double zoffset;       /* cm                                                       */
    int32_t can;          /* 1 beer can, 2 pudding can, 3 2017 10 cm cell (liquid: Z < 2.4) */
    int32_t reserved;
} simc_target;

typedef struct {
    int32_t narm;          /* SIMC_TARGET_*                                             */
    int32_t spectrometer;  /* electron_arm / hadron_arm: 1 HMS, 2 SOS, 3 HRS-R, 4 HRS-L,
                              5 or 6 SHMS (window materials, target.f:63-97)            */
    int32_t using_eloss;   /* using_Eloss                                               */
    int32_t mc_smear;      /* target multiple scattering                                */
    double p_spec;         /* central momentum, MeV/c                                   */
    double th_spec;        /* central angle, rad                                        */
    double phi;            /* spectrometer phi, rad: 3 pi/2 beam left, pi/2 beam right  */
    double mass;           /* particle mass, MeV (Me for the electron and the beam)     */
    double coulomb;        /* Coulomb energy (main%target%Coulomb), MeV; electron arm   */
    double beam_energy;    /* Ebeam, MeV: energy of the beam's loss sampling            */
    double offset[5];      /* spectrometer offsets x, y, z (cm), xptar, yptar (rad)     */
} simc_prep_settings;

/* Vertex quantities (device SoA).  x, y, z: target frame, cm, z including
 * the target's zoffset.  E, p: energy and momentum after radiation (orig),
 * MeV; E_vertex: before radiation (vertex), NULL = E.  theta: the particle's
 * physics scattering angle, rad.  delta: orig delta (%), read only by a
 * hadron arm without using_eloss (NULL otherwise).  xptar, yptar: its
 * spectrometer angles, rad.  dang_yp, dang_xp: the beam's angles (dang_in),
 * NULL = 0.  The beam reads z and E (= Ein) only. */
typedef struct {
    const double *x, *y, *z, *E, *p, *E_vertex, *theta, *delta, *xptar, *yptar, *dang_yp, *dang_xp;
} simc_vertex_in;

/* Outputs (device SoA; every pointer may be NULL).  dpp..fry: the arm's
 * inputs; eloss (MeV) and teff (radiation lengths) of trip_thru_target;
 * dang1, dang2: the target_musc angles (the beam's dang_in(1), dang_in(2));
 * used: uniforms drawn, -1 when a tape slot ran out (that event's outputs
 * are NaN). */
typedef struct {
    double *dpp, *x, *y, *z, *dxdz, *dydz, *fry;
    double *eloss, *teff, *dang1, *dang2;
    int32_t *used;
} simc_prep_out;

/* Asynchronous on `stream`. */
int simc_target_prep(simc_ctx *ctx, const simc_target *t, const simc_prep_settings *s,
                     const simc_vertex_in *in, simc_prep_out *out, int64_t n, const simc_rng *rng,
                     void *stream);
/* The same on host arrays (synchronous); returns SIMC_E_TAPE when an event's
 * tape slot ran out. */
int simc_target_prep_host(simc_ctx *ctx, const simc_target *t, const simc_prep_settings *s,
                          const simc_vertex_in *in, simc_prep_out *out, int64_t n, const simc_rng *rng);

/* ---- Event generator: H(e,e'p) elastic scattering, radiation off -------
 *
 * Per event, SIMC's generate (event.f:126-428) for doing_hyd_elast without
 * using_rad: beam position (gauss1(3) sigma widths, event.f:157-158), raster
 * (fr_pattern 1 bedpost, 2 circle, 3 flat square; :165-183), vertex z
 * (:187), the beam's sampled loss (trip_thru_target narm 1, :194-196), the
 * Coulomb energy (:197-203), Ein (:204-205), electron yptar, xptar uniform
 * in the generation limits (:275-277), physics_angles; then complete_ev's
 * hydrogen-elastic branch (event.f:497-563: E' = Ein Mh / (Mh + Ein (1 -
 * cos theta_e)), nu, Q2, q, the proton along q, spectrometer_angles,
 * :1013-1014 jacobian) and complete_main's weight (event.f:1558: jacobian
 * x sigep, physics_proton.f:1-23).  The scattered particles' losses of
 * complete_ev (event.f:1031-1040) belong to simc_target_prep (narm 2, 3),
 * which takes these outputs as its vertex.  Draws come from a stream of their
 * own: Philox (seed, event0 + i) with the counter word "GNEH"; on a tape,
 * the reference's order (gauss1 x, gauss1 y, raster 2, z 1, the beam's
 * losses, Ein 1, yptar 1, xptar 1). */
typedef struct {
    double Ebeam, dEbeam;           /* MeV: beam energy, full width of its flat spread   */
    double xwid, ywid;              /* gen%xwid, gen%ywid: beam sigma, cm                 */
    double xoffset, yoffset;        /* targ%xoffset, targ%yoffset, cm                     */
    double fr1, fr2;                /* targ%fr1, targ%fr2: raster half-widths or radii    */
    double coulomb;                 /* targ%Coulomb_constant, MeV                         */
    double e_yptar[2], e_xptar[2];  /* gen%e%yptar, gen%e%xptar (min, max), rad           */
    double e_theta, e_phi, e_P;     /* spec%e: central angle, phi (pi/2 or 3 pi/2), MeV/c */
    double p_theta, p_phi, p_P;     /* spec%p                                             */
    double Mh;                      /* hadron mass, MeV (Mp for hydrogen)                 */
    int32_t fr_pattern;             /* 0 none, 1, 2, 3                                    */
    int32_t using_eloss, using_coulomb, reserved;
} simc_hep_settings;

/* Outputs (device SoA, every pointer may be NULL): main%target x, y, z (cm,
 * z with zoffset), rasterx, rastery; Ein; the beam's eloss (MeV) and teff
 * (radiation lengths); the electron's and the proton's xptar, yptar, theta,
 * phi (rad), E and P (MeV), delta (%); Q2 (MeV^2), nu, q; jacobian, sigcc
 * (microbarn/sr), weight; ok: complete_ev's success; used: uniforms drawn,
 * -1 when a tape slot ran out (that event's outputs are NaN). */
typedef struct {
    double *x, *y, *z, *rasterx, *rastery, *Ein, *eloss, *teff;
    double *e_xptar, *e_yptar, *e_theta, *e_phi, *e_E, *e_delta;
    double *p_xptar, *p_yptar, *p_theta, *p_phi, *p_P, *p_E, *p_delta;
    double *Q2, *nu, *q, *jacobian, *sigcc, *weight;
    uint8_t *ok;
    int32_t *used;
} simc_hep_out;

/* Asynchronous on `stream`; the target description gives the beam's path
 * (length, angle, zoffset, can, material). */
int simc_gen_hep(simc_ctx *ctx, const simc_target *t, const simc_hep_settings *h, simc_hep_out *out,
                 int64_t n, const simc_rng *rng, void *stream);
/* The same on host arrays (synchronous); returns SIMC_E_TAPE when an event's
 * tape slot ran out. */
int simc_gen_hep_host(simc_ctx *ctx, const simc_target *t, const simc_hep_settings *h, simc_hep_out *out,
                      int64_t n, const simc_rng *rng);

/* ---- Radiative tails of H(e,e'p) elastic (radc.f, brem.f, init.f) ----
 * simc_gen_hep_rad is generate with using_rad for doing_hyd_elast
 * (event.f:126-428 calling generate_rad, radc.f:120-517): simc_gen_hep's
 * event, then complete_ev's target losses of the scattered electron and
 * proton (event.f:1031-1040, in the arms' windows), radc_init_ev
 * (init.f:655-720), the tail choice and the photons of generate_rad with
 * basicrad, peaked_rad_weight / extrad_phi and brem / bremos / schwinger,
 * and complete_main's weight = jacobian x gen_weight x sigep.  The run
 * constants that radc_init (init.f:576-651) and limits_init (init.f:91-470)
 * give the reference are settings here. */
typedef struct {
    int32_t rad_flag;         /* 0..3 (radc.f:141-159)                          */
    int32_t extrad_flag;      /* 1..3; 0 = radc_init's default (init.f:621-629) */
    int32_t intcor_mode;      /* 0 schwinger, 1 brem / bremos                   */
    int32_t use_expon;        /* brem's exponentiate (init.f:641-645)           */
    int32_t use_offshell_rad; /* bremos instead of brem                         */
    int32_t doing_tail[3];    /* from one_tail, dbase.f:494-511                 */
    int32_t e_arm, p_arm;     /* spectrometers whose windows the electron / proton cross */
    double dE_edge_test;      /* MeV                                             */
    double Egamma_gen_max;    /* hardwired photon limit when > 0.01 (dbase.f:514) */
    double edge_e_E[2];       /* edge%e%E min, max (MeV)                         */
    double edge_p_E[2];       /* edge%p%E                                        */
    double edge_Em[2];        /* edge%Em (limits_init leaves 0 for hydrogen)     */
    double vertex_Trec[2];    /* VERTEXedge%Trec                                 */
    double Egamma_tot_max;    /* init.f:409-421                                  */
    double etatzai;           /* radc_init, init.f:634; <= 0: from the target's Z */
} simc_rad_settings;

/* Radiative outputs (device, each may be NULL): Egamma_used(1..3) (MeV),
 * gen_weight = rad_weight / hardcorfac, the ORIG kinematics the arms see
 * (Ein, e E and delta, p E, P and delta), complete_ev's last losses and
 * radiation lengths of the electron (2) and the proton (3), hardcorfac,
 * and the radiating tail (ntup%radarm: 1..3, 0 = all tails possible). */
typedef struct {
    double *Egamma1, *Egamma2, *Egamma3, *gen_weight;
    double *orig_Ein, *orig_e_E, *orig_e_delta, *orig_p_E, *orig_p_P, *orig_p_delta;
    double *eloss2, *teff2, *eloss3, *teff3, *hardcorfac;
    int32_t *ntail;
} simc_rad_out;

/* out holds the VERTEX kinematics (after tail 1's photon: Ein is the beam
 * energy at the vertex); ok = generate's success.  On a tape, the uniforms
 * are consumed in the reference's order. */
int simc_gen_hep_rad(simc_ctx *ctx, const simc_target *t, const simc_hep_settings *h,
                     const simc_rad_settings *r, simc_hep_out *out, simc_rad_out *rout, int64_t n,
                     const simc_rng *rng, void *stream);

/* ---- Target-field tracking (polarized-target field, trg_track.f) ----
 * simc_field_create = trgInit(map, theta_e, phi_e, theta_p, phi_p)
 * (trg_track.f:243-325): map = NULL or blank for the uniform 5 T test field,
 * "0" for no field, else the 51 x 51 field-map file (e.g. trg_field_map.dat); angles
 * in degrees between the field and the electron (e) / hadron (p) arm, phi
 * ignored as in the reference.  The map is copied to the current device. */
typedef struct simc_field simc_field;
int simc_field_create(const char *map, double theta_e, double phi_e, double theta_p, double phi_p,
                      simc_field **out);
void simc_field_destroy(simc_field *f);
/* track_from_tgt(x, y, z, dx, dy, mom, mass, spect, ok), trg_track.f:591-658,
 * for n events (device arrays, asynchronous on `stream`): x..dy INOUT, from
 * the vertex to the field-free plane z = 100 cm; mom signed (MeV/c, < 0 for
 * negative particles), spect -1 (electron-arm field angle) or +1 (hadron
 * arm); ok may be NULL. */
int simc_track_from_tgt(const simc_field *f, int spect, int64_t n, double *x, double *y, double *z,
                        double *dx, double *dy, const double *mom, double mass, uint8_t *ok,
                        void *stream);
/* track_to_tgt(delta, y, dx, dy, frx, fry, mom, mass, ctheta, stheta, spect,
 * ok, arm), trg_track.f:738-876, for n events: delta, y, dx, dy INOUT (first
 * guess from the arm's reconstruction in, field-corrected reconstruction
 * out); x_fp..dy_fp the focal-plane track the arm's reconstruction reads
 * (/track/: for the HRS the fitted y, i.e. y_fp + VDC offset); fry as the
 * reference receives it (simc.f passes -fry); ok INOUT (NULL: all true in,
 * not written). */
int simc_track_to_tgt(simc_ctx *ctx, const simc_field *f, int arm, int spect, int64_t n,
                      const double *x_fp, const double *dx_fp, const double *y_fp,
                      const double *dy_fp, double *delta, double *y, double *dx, double *dy,
                      const double *frx, const double *fry, const double *mom, double mass,
                      double ctheta, double stheta, uint8_t *ok, void *stream);
/* The same two on host arrays (synchronous; calls on one field are serialised). */
int simc_track_from_tgt_host(simc_field *f, int spect, int64_t n, double *x, double *y, double *z,
                             double *dx, double *dy, const double *mom, double mass, uint8_t *ok);
int simc_track_to_tgt_host(simc_ctx *ctx, simc_field *f, int arm, int spect, int64_t n,
                           const double *x_fp, const double *dx_fp, const double *y_fp,
                           const double *dy_fp, double *delta, double *y, double *dx, double *dy,
                           const double *frx, const double *fry, const double *mom, double mass,
                           double ctheta, double stheta, uint8_t *ok);

/* ---- Output formats ----
 * The .bin ntuple (simc.f:71-74): a Fortran unformatted sequential file with
 * 4-byte record lengths around every record -- the column count (INTEGER*4)
 * and the 16-character column names (NtupleInit.f:351-355), then one REAL*8
 * record per column of every event written (results_write.f:262-264).
 * tags: ncols (<= 80) names as NtupleInit.f gives them for the run's mode. */
typedef struct simc_ntuple simc_ntuple;
int simc_ntuple_open(const char *path, int ncols, const char *const *tags, simc_ntuple **out);
/* nrows events from a host array, row-major (nrows x ncols). */
int simc_ntuple_write(simc_ntuple *nt, const double *rows, int64_t nrows);
/* The events i with select[i] != 0 (select NULL: all n) of ncols DEVICE
 * columns cols[c][i], in event order: compacted and formatted on the GPU,
 * appended to the file before returning (the stream is synchronised);
 * *written (may be NULL) receives the number of events written. */
int simc_ntuple_write_device(simc_ntuple *nt, const double *const *cols, const uint8_t *select, int64_t n,
                             int64_t *written, void *stream);
int simc_ntuple_close(simc_ntuple *nt);
/* The STOP-counter report of simc.f:462-530 (the .geni file) from the
 * context's counters, for the arms electron_arm and hadron_arm (codes as in
 * the reference, 6 = SHMS too), in gfortran's list-directed layout. */
int simc_stop_report(simc_ctx *ctx, const char *path, int electron_arm, int hadron_arm);

/* STOP counters accumulated by the context since creation / reset, in the
 * order of the reference's common blocks.  Synchronises the device. */
int simc_counters_get(simc_ctx *ctx, int arm, int64_t *counters, int ncounters);
int simc_counters_reset(simc_ctx *ctx, int arm);

/* Central-ray geometry of the loaded forward map: drift length of a class
 * (cm) and its pure-drift flag (/driftdistances/, spectrometers.inc:9-12). */
int simc_map_class_info(simc_ctx *ctx, int arm, int cls, double *driftdist, int32_t *adrift,
                        int32_t *n_terms);

/* Algorithmic FP64 operations of the polynomial part for one event that
 * reaches every segment of the arm (monomial multiplies + 2 per coefficient). */
int64_t simc_poly_flops_full(simc_ctx *ctx, int arm);

/* The same count for an event that ends with stop code c, written to
 * flops[c] for c = 0..255 (nflops >= 256): the segments it passes through. */
int simc_stop_flops(simc_ctx *ctx, int arm, int64_t *flops, int nflops);
/* The same restricted to the segments of one kernel stage (1: target to
 * collimator, 2: the magnets, 3: the reconstruction of accepted events;
 * 0: all). */
int simc_stop_flops_stage(simc_ctx *ctx, int arm, int stage, int64_t *flops, int nflops);

/* FP64 operations of the detector hut for an event that ends with stop code
 * c, written to flops[c] (no counterpart in the reference: a measurement
 * hook).  The hut holds no COSY polynomial, so the counts above hold none of
 * its work; this one counts the arithmetic of the reference's hut routines
 * for the steps the event executes: a drift 11 (shared/project.f:39-41), a
 * scattering 18 (shared/musc.f:48-56) or 36 (shared/musc_ext.f:36-51) plus
 * its normal deviates, a wire plane 3 plus its two deviates
 * (shms/mc_shms_hut.f:183-192), the two track fits 11 per plane + 18
 * (cern/lfit.f:11-59), a rotated box 13 (shared/rotate_haxis.f:56-61), the
 * calorimeter fiducial cut 4; +, -, *, /, sqrt, log and sin/cos one each,
 * 4 per normal deviate (one Box-Muller pair: ln, *-2, sqrt, 2 pi u, sin,
 * cos, two products).  Scattering and wire deviates count as ms_flag and
 * wcs_flag say.  Stop codes before the hut: 0.  nflops >= 256. */
int simc_stop_flops_hut(simc_ctx *ctx, int arm, int ms_flag, int wcs_flag, int64_t *flops, int nflops);

const char *simc_last_error(void);
const char *simc_version(void);

/* Stop codes written to simc_events_out.stop */
#define SIMC_STOP_OK 0
#define SIMC_STOP_SLIT_HOR 1
#define SIMC_STOP_SLIT_VERT 2
#define SIMC_STOP_SLIT_OCT 3
#define SIMC_STOP_COLL 4
#define SIMC_STOP_TAPE_OVERFLOW 250
#define SIMC_STOP_INTERNAL 251    /* a step outside its stage (cannot occur: checked at load) */

#ifdef __cplusplus
}
#endif
#endif
