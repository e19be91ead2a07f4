// Per-event drop-in replacements of the reference's Fortran subroutines
// mc_hms, mc_shms, mc_sos, mc_hrsr, mc_hrsl and their recon routines (gfortran/flang ABI:
// lower-case names with a trailing underscore, every argument by reference,
// LOGICAL as 4-byte integers), built as libsimc_dropin.so on top of
// libsimc_gpu.so.  Linking a SIMC executable against this library instead of
// the objects of hms/mc_hms*.f, shms/mc_shms*.f, sos/mc_sos*.f and hrs?/mc_hrs?*.f keeps every call site of
// simc.f unchanged; each call transports ONE event on the GPU (synchronous),
// so this path is for compatibility -- throughput needs the batched entry
// points of simc_gpu.h.
//
// State shared with the caller's program through its common blocks, when
// they exist (weak references; a program without them still links):
//   /spectrometer1/, /spectrometer_shms/, /spectrometer2/, /spectrometer3/,
//   /spectrometer4/  STOP counters, incremented per event (hms/struct_hms.inc:34-42,
//            shms/struct_shms.inc:26-44, sos/struct_sos.inc:22-35,
//            hrsr/struct_hrsr.inc:24-40, hrsl/struct_hrsl.inc:24-40);
//   /track/  xs, ys, zs, dpps, dxdzs, dydzs (spectrometers.inc:47-60): set to
//            the fitted focal-plane track of an accepted event, as
//            hms/mc_hms.f:419-423 does, and read by the recon entry points
//            (hms/mc_hms_recon.f:113-118).  The HRS arms return y_fp minus the
//            VDC offset while /track/ keeps the fitted y (hrsl/mc_hrsl.f:512-525);
//            the drop-in adds the offset back, so /track/ ys may differ from the
//            fitted value by one rounding;
//   /decd/   ctau, its first member (simulate.inc:153).
// trgInit / track_from_tgt / track_to_tgt (trg_track.f) are replaced too: the
// target field lives on the GPU, one synchronous launch per call.
//
// Random numbers: a counter-based Philox4x32-10 stream per event, keyed by
// (seed, call number); seed from SIMC_SEED (default 0) or simc_dropin_seed.
// Maps are read from SIMC_MAP_DIR, default the current directory, where the
// reference opens hms/forward_cosy.dat etc. (shared/transp.f:300-309).
// The mock sieve slit, a compile-time DATA value in the reference
// (hms/mc_hms.f use_sieve, shms/mc_shms.f use_sieve / use_2sieve, ...), is
// chosen once per process by SIMC_SIEVE: unset or 0 off, 1 use_sieve,
// 2 the SHMS's use_2sieve, 3 both (bit 1 applies to the SHMS only).
// Errors are printed to stderr and leave ok_spec = .false.; they never end
// the process.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>

#include "../../include/simc_gpu.h"

extern "C" {
extern int32_t spectrometer1_[SIMC_HMS_NCOUNTERS] __attribute__((weak));
extern int32_t spectrometer_shms_[SIMC_SHMS_NCOUNTERS] __attribute__((weak));
extern int32_t spectrometer2_[SIMC_SOS_NCOUNTERS] __attribute__((weak));
extern int32_t spectrometer3_[SIMC_HRS_NCOUNTERS] __attribute__((weak));
extern int32_t spectrometer4_[SIMC_HRS_NCOUNTERS] __attribute__((weak));
extern double track_[6] __attribute__((weak));
extern double decd_[1] __attribute__((weak));
}

namespace {

std::mutex g_mu;
simc_ctx* g_ctx = nullptr;
bool g_failed = false;
uint64_t g_seed = 0;
uint64_t g_event = 0;
int32_t g_sieve = 0;
bool g_seed_set = false;
simc_field* g_field = nullptr;

simc_ctx* context()
{
    if (g_ctx || g_failed) return g_ctx;
    const char* dir = std::getenv("SIMC_MAP_DIR");
    if (!g_seed_set) {
        const char* sd = std::getenv("SIMC_SEED");
        g_seed = sd ? std::strtoull(sd, nullptr, 10) : 0;
    }
    const char* sv = std::getenv("SIMC_SIEVE");
    g_sieve = sv ? (int32_t)std::strtol(sv, nullptr, 10) : 0;
    if (simc_ctx_create(dir ? dir : ".", &g_ctx) != SIMC_OK) {
        std::fprintf(stderr, "simc_dropin: %s\n", simc_last_error());
        g_failed = true;
        g_ctx = nullptr;
    }
    return g_ctx;
}

void add_counters(int arm)
{
    const int nc = arm == SIMC_ARM_HMS ? SIMC_HMS_NCOUNTERS
                 : arm == SIMC_ARM_SOS ? SIMC_SOS_NCOUNTERS
                 : arm == SIMC_ARM_SHMS ? SIMC_SHMS_NCOUNTERS
                                        : SIMC_HRS_NCOUNTERS;
    int32_t* common = arm == SIMC_ARM_HMS ? spectrometer1_
                    : arm == SIMC_ARM_SOS ? spectrometer2_
                    : arm == SIMC_ARM_HRSR ? spectrometer3_
                    : arm == SIMC_ARM_HRSL ? spectrometer4_
                                           : spectrometer_shms_;
    if (!common) return;
    int64_t c[SIMC_SHMS_NCOUNTERS];
    if (simc_counters_get(g_ctx, arm, c, nc) != SIMC_OK) {
        std::fprintf(stderr, "simc_dropin: %s\n", simc_last_error());
        return;
    }
    for (int i = 0; i < nc; ++i) {
        // stop_id (SHMS slot 50) is a last-value register in the reference
        if (arm == SIMC_ARM_SHMS && i == 50) continue;
        common[i] += (int32_t)c[i];
    }
}

void arm_event(int arm, double* p_spec, double* dpp, double* x, double* y, double* z, double* dxdz,
               double* dydz, double* x_fp, double* dx_fp, double* y_fp, double* dy_fp, double* m2,
               int32_t* ms_flag, int32_t* wcs_flag, int32_t* decay_flag, double* resmult, double* fry,
               int32_t* ok_spec, double* pathlen, int32_t* using_coll)
{
    std::lock_guard<std::mutex> lock(g_mu);
    *ok_spec = 0;
    if (!context()) return;
    simc_settings s = {};
    s.p_spec = *p_spec;
    s.m2 = *m2;
    s.ctau = decd_ ? decd_[0] : 0.0;
    s.ms_flag = *ms_flag != 0;
    s.wcs_flag = *wcs_flag != 0;
    s.decay_flag = *decay_flag != 0;
    s.using_coll = *using_coll != 0;
    s.fit_mode = SIMC_FIT_REF_F32;
    s.sieve = arm == SIMC_ARM_SHMS ? g_sieve : (g_sieve & 1);
    s.wire_sigma = 0.0;
    simc_events_in in = {dpp, x, y, z, dxdz, dydz, fry};
    uint8_t ok = 0, stop = 0;
    simc_events_out out = {dpp, y, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, pathlen, resmult, m2, &ok, &stop,
                           nullptr};
    simc_rng rng = {};
    rng.kind = SIMC_RNG_PHILOX;
    rng.seed = g_seed;
    rng.event0 = g_event++;
    if (simc_counters_reset(g_ctx, arm) != SIMC_OK ||
        simc_arm_run_host(g_ctx, arm, &s, &in, &out, 1, &rng) != SIMC_OK) {
        std::fprintf(stderr, "simc_dropin: %s\n", simc_last_error());
        return;
    }
    add_counters(arm);
    *ok_spec = ok ? 1 : 0;
    if (ok && track_) {  // hms/mc_hms.f:419-423
        track_[0] = *x_fp;
        track_[1] = arm == SIMC_ARM_HRSR ? *y_fp + 0.48 : arm == SIMC_ARM_HRSL ? *y_fp + 0.78 : *y_fp;
        track_[4] = *dx_fp;
        track_[5] = *dy_fp;
    }
}

void recon_event(int arm, double* delta_p, double* delta_t, double* delta_phi, double* y_tgt, double* fry)
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (!track_) {
        std::fprintf(stderr, "simc_dropin: common /track/ is not present in this program\n");
        return;
    }
    if (!context()) return;
    double xs = track_[0], ys = track_[1], dxdzs = track_[4], dydzs = track_[5];
    double delta, dxdz, dydz, yt;
    if (simc_recon_run_host(g_ctx, arm, 1, &xs, &dxdzs, &ys, &dydzs, fry, &delta, &dxdz, &dydz, &yt) !=
        SIMC_OK) {
        std::fprintf(stderr, "simc_dropin: %s\n", simc_last_error());
        return;
    }
    // hms/mc_hms_recon.f:134-137: delta_phi = x'_tar, y_tgt, delta_t = y'_tar, delta_p
    *delta_p = delta;
    *delta_t = dydz;
    *delta_phi = dxdz;
    *y_tgt = yt;
}

}  // namespace

extern "C" {

// mc_hms(p_spec, th_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp,
//        m2, ms_flag, wcs_flag, decay_flag, resmult, fry, ok_spec, pathlen,
//        using_HMScoll), hms/mc_hms.f:1-4.  th_spec is unused (hms/mc_hms.f:153).
void mc_hms_(double* p_spec, double* th_spec, double* dpp, double* x, double* y, double* z, double* dxdz,
             double* dydz, double* x_fp, double* dx_fp, double* y_fp, double* dy_fp, double* m2, int32_t* ms_flag,
             int32_t* wcs_flag, int32_t* decay_flag, double* resmult, double* fry, int32_t* ok_spec,
             double* pathlen, int32_t* using_coll)
{
    (void)th_spec;
    arm_event(SIMC_ARM_HMS, p_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, m2, ms_flag, wcs_flag,
              decay_flag, resmult, fry, ok_spec, pathlen, using_coll);
}

// mc_shms(... same ..., pathlen, spectr, using_SHMScoll), shms/mc_shms.f:1-4;
// spectr must be 5 (shms/mc_shms.f:336-341).
void mc_shms_(double* p_spec, double* th_spec, double* dpp, double* x, double* y, double* z, double* dxdz,
              double* dydz, double* x_fp, double* dx_fp, double* y_fp, double* dy_fp, double* m2,
              int32_t* ms_flag, int32_t* wcs_flag, int32_t* decay_flag, double* resmult, double* fry,
              int32_t* ok_spec, double* pathlen, int32_t* spectr, int32_t* using_coll)
{
    (void)th_spec;
    if (*spectr != SIMC_ARM_SHMS) {
        std::fprintf(stderr, "simc_dropin: mc_shms called with spectr = %d (must be 5)\n", (int)*spectr);
        *ok_spec = 0;
        return;
    }
    arm_event(SIMC_ARM_SHMS, p_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, m2, ms_flag,
              wcs_flag, decay_flag, resmult, fry, ok_spec, pathlen, using_coll);
}

// mc_sos(p_spec, th_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp,
//        m2, ms_flag, wcs_flag, decay_flag, resmult, fry, ok_spec, pathlen),
// sos/mc_sos.f:1-3 (no collimator switch: the octagons always apply).
void mc_sos_(double* p_spec, double* th_spec, double* dpp, double* x, double* y, double* z, double* dxdz,
             double* dydz, double* x_fp, double* dx_fp, double* y_fp, double* dy_fp, double* m2, int32_t* ms_flag,
             int32_t* wcs_flag, int32_t* decay_flag, double* resmult, double* fry, int32_t* ok_spec,
             double* pathlen)
{
    (void)th_spec;
    int32_t no_coll = 0;
    arm_event(SIMC_ARM_SOS, p_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, m2, ms_flag, wcs_flag,
              decay_flag, resmult, fry, ok_spec, pathlen, &no_coll);
}

// mc_sos_recon(delta_p, delta_t, delta_phi, y_tgt, fry), sos/mc_sos_recon.f:1
void mc_sos_recon_(double* delta_p, double* delta_t, double* delta_phi, double* y_tgt, double* fry)
{
    recon_event(SIMC_ARM_SOS, delta_p, delta_t, delta_phi, y_tgt, fry);
}

// mc_hrsr / mc_hrsl (same arguments as mc_sos), hrsr/mc_hrsr.f:1-3, hrsl/mc_hrsl.f:1-3
void mc_hrsr_(double* p_spec, double* th_spec, double* dpp, double* x, double* y, double* z, double* dxdz,
              double* dydz, double* x_fp, double* dx_fp, double* y_fp, double* dy_fp, double* m2,
              int32_t* ms_flag, int32_t* wcs_flag, int32_t* decay_flag, double* resmult, double* fry,
              int32_t* ok_spec, double* pathlen)
{
    (void)th_spec;
    int32_t no_coll = 0;
    arm_event(SIMC_ARM_HRSR, p_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, m2, ms_flag,
              wcs_flag, decay_flag, resmult, fry, ok_spec, pathlen, &no_coll);
}

void mc_hrsl_(double* p_spec, double* th_spec, double* dpp, double* x, double* y, double* z, double* dxdz,
              double* dydz, double* x_fp, double* dx_fp, double* y_fp, double* dy_fp, double* m2,
              int32_t* ms_flag, int32_t* wcs_flag, int32_t* decay_flag, double* resmult, double* fry,
              int32_t* ok_spec, double* pathlen)
{
    (void)th_spec;
    int32_t no_coll = 0;
    arm_event(SIMC_ARM_HRSL, p_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, m2, ms_flag,
              wcs_flag, decay_flag, resmult, fry, ok_spec, pathlen, &no_coll);
}

// mc_hrsr_recon / mc_hrsl_recon(delta_p, delta_t, delta_phi, y_tgt, fry),
// hrsr/mc_hrsr_recon.f:1, hrsl/mc_hrsl_recon.f:1
void mc_hrsr_recon_(double* delta_p, double* delta_t, double* delta_phi, double* y_tgt, double* fry)
{
    recon_event(SIMC_ARM_HRSR, delta_p, delta_t, delta_phi, y_tgt, fry);
}

void mc_hrsl_recon_(double* delta_p, double* delta_t, double* delta_phi, double* y_tgt, double* fry)
{
    recon_event(SIMC_ARM_HRSL, delta_p, delta_t, delta_phi, y_tgt, fry);
}

// mc_hms_recon(delta_p, delta_t, delta_phi, y_tgt, fry), hms/mc_hms_recon.f:1
void mc_hms_recon_(double* delta_p, double* delta_t, double* delta_phi, double* y_tgt, double* fry)
{
    recon_event(SIMC_ARM_HMS, delta_p, delta_t, delta_phi, y_tgt, fry);
}

// mc_shms_recon(delta_p, delta_t, delta_phi, y_tgt, fry, spectr), shms/mc_shms_recon.f:1-2
void mc_shms_recon_(double* delta_p, double* delta_t, double* delta_phi, double* y_tgt, double* fry,
                    int32_t* spectr)
{
    (void)spectr;
    recon_event(SIMC_ARM_SHMS, delta_p, delta_t, delta_phi, y_tgt, fry);
}

// ---- target-field tracking (trg_track.f) ----
// trgInit(map, theta_e, phi_e, theta_p, phi_p), trg_track.f:243: map is a
// blank-padded CHARACTER (hidden length argument): all blanks = the uniform
// test field, '0' = no field, else the map file (opened relative to the
// working directory, like the reference).
void trginit_(const char* map, double* theta_e, double* phi_e, double* theta_p, double* phi_p, size_t map_len)
{
    std::lock_guard<std::mutex> lock(g_mu);
    std::string m(map, map_len);
    while (!m.empty() && (m.back() == ' ' || m.back() == '\0')) m.pop_back();
    if (g_field) simc_field_destroy(g_field);
    g_field = nullptr;
    if (simc_field_create(m.empty() ? nullptr : m.c_str(), *theta_e, *phi_e, *theta_p, *phi_p, &g_field) != SIMC_OK)
        std::fprintf(stderr, "simc_dropin: %s\n", simc_last_error());
}

// track_from_tgt(x, y, z, dx, dy, mom, mass, spect, ok), trg_track.f:591
void track_from_tgt_(double* x, double* y, double* z, double* dx, double* dy, double* mom, double* mass,
                     int32_t* spect, int32_t* ok)
{
    std::lock_guard<std::mutex> lock(g_mu);
    *ok = 0;
    if (!g_field) {
        std::fprintf(stderr, "simc_dropin: track_from_tgt before trgInit\n");
        return;
    }
    uint8_t o = 0;
    if (simc_track_from_tgt_host(g_field, *spect, 1, x, y, z, dx, dy, mom, *mass, &o) != SIMC_OK) {
        std::fprintf(stderr, "simc_dropin: %s\n", simc_last_error());
        return;
    }
    *ok = o ? 1 : 0;
}

// track_to_tgt(delta, y, dx, dy, frx, fry, mom, mass, ctheta, stheta, spect,
// ok, arm), trg_track.f:738: the arm's reconstruction reads the focal-plane
// track from /track/ (arm 6 is the SHMS, :845).
void track_to_tgt_(double* delta, double* y, double* dx, double* dy, double* frx, double* fry, double* mom,
                   double* mass, double* ctheta, double* stheta, int32_t* spect, int32_t* ok, int32_t* arm)
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_field || !track_ || !context()) {
        std::fprintf(stderr, "simc_dropin: track_to_tgt needs trgInit, /track/ and the maps\n");
        *ok = 0;
        return;
    }
    const int a = *arm == 6 ? SIMC_ARM_SHMS : *arm;
    uint8_t o = *ok ? 1 : 0;
    if (simc_track_to_tgt_host(g_ctx, g_field, a, *spect, 1, &track_[0], &track_[4], &track_[1], &track_[5], delta,
                               y, dx, dy, frx, fry, mom, *mass, *ctheta, *stheta, &o) != SIMC_OK) {
        std::fprintf(stderr, "simc_dropin: %s\n", simc_last_error());
        *ok = 0;
        return;
    }
    *ok = o ? 1 : 0;
}

// Seed of the per-event streams (Fortran: call simc_dropin_seed(seed) with an
// integer*8 seed); restarts the event numbering.
void simc_dropin_seed_(int64_t* seed)
{
    std::lock_guard<std::mutex> lock(g_mu);
    g_seed = (uint64_t)*seed;
    g_seed_set = true;
    g_event = 0;
}

// Release the GPU context (Fortran: call simc_dropin_close()).
void simc_dropin_close_()
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ctx) simc_ctx_destroy(g_ctx);
    g_ctx = nullptr;
    if (g_field) simc_field_destroy(g_field);
    g_field = nullptr;
    g_failed = false;
}

}  // extern "C"
