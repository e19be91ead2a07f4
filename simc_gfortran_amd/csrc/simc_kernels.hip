// Spectrometer transport + reconstruction kernels (HMS, SHMS) and the C ABI
// of libsimc_gpu.so (include/simc_gpu.h).
//
// Layout: events are structure-of-arrays in HBM, one event per lane; a lane
// carries its event from the target through the magnets, the detector hut
// and the track fit to the reconstructed target quantities.  The COSY
// polynomials are generated straight-line code (cosy_gen.dev.inc) whose
// coefficients come from a packed per-arm table read with uniform (scalar)
// loads.  Stop codes are histogrammed per workgroup in LDS and folded into
// the context's device counters with one atomic per non-empty bin.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/simc_gpu.h"
#include "simc_device.h"
#include "simc_maps.h"
#include "simc_ops.h"
#include "simc_programs.h"
#include "simc_target.h"
#include "simc_field.h"
#include "simc_gen.h"
#include "simc_rad.h"
#include "cosy_gen.dev.inc"
#include "hut_gen.inc"

using namespace simcdev;

namespace {

constexpr int kHistBins = 256;
constexpr int kBlock = 256;  // threads per workgroup (one event each)
constexpr int kHistSlots = kHistBins + 3;  // + per-step collimator slit hits (hor, vert, oct)

// Per-arm tables owned by the context (device memory, read through the
// constant address space).
struct ArmTables {
    CollGeom coll;     // collimator stepping geometry
    double zdc[12];    // wire-plane z positions, planes 1..12
    // straight-line fit as a linear map of the hits (all planes of the
    // plane's coordinate present): slope = sum fw_a h, intercept = sum fw_b h
    double fw_a[12], fw_b[12];
    // subtracted from y_fp of an accepted event after the reconstruction
    // (the HRS VDC offset, hrsl/mc_hrsl.f:525, hrsr/mc_hrsr.f:524)
    double yfp_shift;
};

// One kinematic setting: the by-value arguments of mc_hms / mc_shms and the
// setting's place in the launch.  Events of a launch are numbered by a pass
// index P; the setting s covers P in [start, start + count) and its event j
// = P - start is element offset + j of the event arrays and event id
// event0 + j of the random streams.
struct SetDev {
    double p_spec, m2, ctau, sigma;
    uint64_t start, seed, event0, offset;
    int32_t flags;   // KF_* bits (the same for every setting of a launch)
    uint32_t hist;   // index of the setting's stop-code histogram
};

// Per-launch arguments.  The event arrays are reached through buffer
// resources (n_arr elements from the base pointers, one 32-bit byte offset
// per lane), so no 64-bit per-lane addresses are kept.  A plain run has one
// setting, passed by value (set == nullptr); a batched run reads its table.
struct KArgs {
    SetDev set0;
    const SetDev* set;    // n_set settings sorted by start, or nullptr
    uint32_t n_set;
    int32_t flags;        // KF_* bits of every setting of the launch (wave-uniform control flow)
    uint64_t seed;        // Philox key of every setting of the launch (uniform round keys); a
                          // launch of settings with different keys reads each setting's own
    uint32_t count;       // events in this launch: pass indices pass_base + [0, count)
    uint64_t pass_base;
    uint64_t ix_base;     // element of the event arrays that the base pointers address
    uint32_t n_arr;       // elements addressable from the base pointers
    const ArmTables* tab;
    const double* coef;
    const uint32_t* gen;  // run-time polynomial tables of the arm's maps, or nullptr (compiled)
    const Op* ops;
    const double *dpp, *x, *y, *dxdz, *dydz, *fry;
    double *o_dpp, *o_y, *o_dxdz, *o_dydz, *x_fp, *dx_fp, *y_fp, *dy_fp, *pathlen, *resmult, *m2o;
    uint8_t *ok, *stop;
    int32_t* used;
    const double* tape;   // tape slot of launch event li at li * stride_bytes (plain runs)
    uint32_t stride_bytes;
    unsigned long long* hist;   // per-setting histograms of kHistSlots bins
};
enum : int32_t { KF_MS = 1, KF_WCS = 2, KF_DECAY = 4, KF_COLL = 8, KF_FIT_F32 = 16, KF_SIEVE = 32, KF_SIEVE2 = 64 };

typedef const double __attribute__((address_space(4)))* cdptr;
typedef const Op __attribute__((address_space(4)))* coptr;
typedef const ArmTables __attribute__((address_space(4)))* ctptr;
typedef const uint32_t __attribute__((address_space(4)))* cuptr;

// Results of one event besides the track.
struct EvOut {
    double x_fp, dx_fp, y_fp, dy_fp;
    double dpp, dxdz, dydz, y;
    double resmult;
    bool fit_done;
};

// COSY forward class `cls` of the arm (generated straight-line code).
template <int ARM>
__device__ __forceinline__ void poly_fwd(int cls, cdptr c, double r0, double r1, double r2, double r3,
                                         double r4, double& o0, double& o1, double& o2, double& o3,
                                         double& o4)
{
    if (ARM == SIMC_ARM_HMS) {
        switch (cls) {
            case 2: cosy_hms_fwd_c2(c + COSY_HMS_C2_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
            case 3: cosy_hms_fwd_c3(c + COSY_HMS_C3_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
            case 5: cosy_hms_fwd_c5(c + COSY_HMS_C5_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
            case 6: cosy_hms_fwd_c6(c + COSY_HMS_C6_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
            case 8: cosy_hms_fwd_c8(c + COSY_HMS_C8_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
            case 9: cosy_hms_fwd_c9(c + COSY_HMS_C9_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
            default: cosy_hms_fwd_c11(c + COSY_HMS_C11_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
        }
    } else if (ARM == SIMC_ARM_SOS) {
#define OFWD(N) case N: cosy_sos_fwd_c##N(c + COSY_SOS_C##N##_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
        switch (cls) {
            OFWD(2) OFWD(3) OFWD(4) OFWD(5) OFWD(6) OFWD(7) OFWD(8)
            default: cosy_sos_fwd_c9(c + COSY_SOS_C9_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
        }
#undef OFWD
    } else if (ARM == SIMC_ARM_HRSR) {  // both HRS arms (one map structure)
#define RFWD(N) case N: cosy_hrs_fwd_c##N(c + COSY_HRS_C##N##_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
        switch (cls) {
            RFWD(2) RFWD(3) RFWD(5) RFWD(6) RFWD(8) RFWD(10)
            default: cosy_hrs_fwd_c11(c + COSY_HRS_C11_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
        }
#undef RFWD
    } else {
#define SFWD(N) case N: cosy_shms_fwd_c##N(c + COSY_SHMS_C##N##_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
        switch (cls) {
            SFWD(3) SFWD(7) SFWD(8) SFWD(12) SFWD(13) SFWD(17) SFWD(18) SFWD(23) SFWD(24) SFWD(25)
            SFWD(26) SFWD(27) SFWD(28) SFWD(29)
            default: cosy_shms_fwd_c30(c + COSY_SHMS_C30_OFF, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4); break;
        }
#undef SFWD
    }
}

// Run-time polynomial of class cls (0 = recon) from simc::build_generic's
// tables, for maps whose term structure differs from the compiled one:
// out_k = sum over the class's trie nodes of c_k * monomial, each node's
// monomial one multiply from its parent's (kept per depth), one FMA per
// non-zero coefficient (shared/transp.f:205-218, hms/mc_hms_recon.f:122-130).
// Every index is wavefront-uniform: scalar loads and scalar branches.
template <int NOUT>
__device__ __forceinline__ void poly_generic(cuptr g, int cls, double r0, double r1, double r2, double r3, double r4,
                                             double (&out)[NOUT])
{
    const uint32_t first = g[cls], n = g[33 + cls];
    const cuptr rec = g + g[66];
    const cdptr coef = (cdptr)g + g[67];
#pragma unroll
    for (int o = 0; o < NOUT; ++o) out[o] = 0.0;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0, s6 = 0.0, s7 = 0.0, s8 = 0.0;
    for (uint32_t k = first; k < first + n; ++k) {
        const uint32_t h = rec[2 * k];
        uint32_t c = rec[2 * k + 1];
        double x;
        switch ((h >> 4) & 15u) {
        case 0: x = r0; break;
        case 1: x = r1; break;
        case 2: x = r2; break;
        case 3: x = r3; break;
        default: x = r4; break;
        }
        double m;
        switch (h & 15u) {
        case 0: m = 1.0; break;
        case 1: s1 = x; m = s1; break;
        case 2: s2 = s1 * x; m = s2; break;
        case 3: s3 = s2 * x; m = s3; break;
        case 4: s4 = s3 * x; m = s4; break;
        case 5: s5 = s4 * x; m = s5; break;
        case 6: s6 = s5 * x; m = s6; break;
        case 7: s7 = s6 * x; m = s7; break;
        default: s8 = s7 * x; m = s8; break;
        }
#pragma unroll
        for (int o = 0; o < NOUT; ++o)
            if (h & (256u << o)) {
                out[o] = __builtin_fma(coef[c], m, out[o]);
                ++c;
            }
    }
}

template <int ARM>
__device__ __forceinline__ void poly_recon(cdptr c, double h0, double h1, double h2, double h3, double h4,
                                           double& s0, double& s1, double& s2, double& s3)
{
    if (ARM == SIMC_ARM_HMS) cosy_hms_recon(c + COSY_HMS_RECON_OFF, h0, h1, h2, h3, h4, s0, s1, s2, s3);
    else if (ARM == SIMC_ARM_SOS) cosy_sos_recon(c + COSY_SOS_RECON_OFF, h0, h1, h2, h3, h4, s0, s1, s2, s3);
    else if (ARM == SIMC_ARM_HRSR) cosy_hrs_recon(c + COSY_HRS_RECON_OFF, h0, h1, h2, h3, h4, s0, s1, s2, s3);
    else cosy_shms_recon(c + COSY_SHMS_RECON_OFF, h0, h1, h2, h3, h4, s0, s1, s2, s3);
}

// COSY forward class `cls` with two tangent directions (adjoint form).
template <int ARM>
__device__ __forceinline__ void poly_fwd_tan(int cls, cdptr c, const double (&r)[5], const double (&ra)[5],
                                             const double (&rb)[5], double (&o)[5], double (&oa)[5],
                                             double (&ob)[5])
{
#define TAN_ARGS r[0], r[1], r[2], r[3], r[4], ra[0], ra[1], ra[2], ra[3], ra[4], rb[0], rb[1], rb[2], rb[3], rb[4], \
                 o[0], o[1], o[2], o[3], o[4], oa[0], oa[1], oa[2], oa[3], oa[4], ob[0], ob[1], ob[2], ob[3], ob[4]
    if (ARM == SIMC_ARM_HMS) {
#define HT(N) case N: cosy_hms_fwd_c##N##_tan(c + COSY_HMS_C##N##_OFF, TAN_ARGS); break;
        switch (cls) {
            HT(2) HT(3) HT(5) HT(6) HT(8) HT(9)
            default: cosy_hms_fwd_c11_tan(c + COSY_HMS_C11_OFF, TAN_ARGS); break;
        }
#undef HT
    } else {
#define ST(N) case N: cosy_shms_fwd_c##N##_tan(c + COSY_SHMS_C##N##_OFF, TAN_ARGS); break;
        switch (cls) {
            ST(3) ST(7) ST(8) ST(12) ST(13) ST(17) ST(18) ST(23) ST(24) ST(25) ST(26) ST(27) ST(28) ST(29)
            default: cosy_shms_fwd_c30_tan(c + COSY_SHMS_C30_OFF, TAN_ARGS); break;
        }
#undef ST
    }
#undef TAN_ARGS
}

// Reconstruction (hms/mc_hms_recon.f:113-137, shms/mc_shms_recon.f:131-160,
// hrsl/mc_hrsl_recon.f:113-137, sos/mc_sos_recon.f:113-137: the SOS keeps
// every input away from 0, the others only x_tar)
template <int ARM>
__device__ __forceinline__ void recon(cdptr c, cuptr gen, const EvOut& o, double fry, EvOut& r)
{
    double h[5] = {o.x_fp / 100.0, o.dx_fp, o.y_fp / 100.0, o.dy_fp, fry / 100.0};
#pragma unroll
    for (int k = ARM == SIMC_ARM_SOS ? 0 : 4; k < 5; ++k)
        if (fabs(h[k]) <= 1.e-30) h[k] = 1.e-30;
    double s0, s1, s2, s3;
    if (gen) {
        double s[4];
        poly_generic<4>(gen, 0, h[0], h[1], h[2], h[3], h[4], s);
        s0 = s[0]; s1 = s[1]; s2 = s[2]; s3 = s[3];
    } else {
        poly_recon<ARM>(c, h[0], h[1], h[2], h[3], h[4], s0, s1, s2, s3);
    }
    r.dxdz = s0;         // delta_phi
    r.y = s1 * 100.0;    // y_tgt
    r.dydz = s2;         // delta_t
    r.dpp = s3 * 100.0;  // delta_p
}

// hit_dipole, hms/mc_hms.f:445-492 with hms/apertures_hms.inc:20-32
__device__ __forceinline__ bool hms_hit_dipole(double x, double y)
{
    const double x_d1 = 34.29, y_d1 = 12.07, x_d2 = 27.94, y_d2 = 18.42, x_d3 = 13.97,
                 y_d3 = 18.95, x_d4 = 1.956, y_d4 = 20.32, x_d5 = 27.94, y_d5 = 12.065,
                 r_d5 = 6.35, a_d6 = -0.114, b_d6 = 20.54;
    const double xl = fabs(x), yl = fabs(y);
    const bool c1 = (xl <= x_d1) && (yl <= y_d1);
    const bool c2 = (xl <= x_d2) && (yl <= y_d2);
    const bool c3 = (xl <= x_d3) && (yl <= y_d3);
    const bool c4 = (xl <= x_d4) && (yl <= y_d4);
    const bool c5 = ((xl - x_d5) * (xl - x_d5) + (yl - y_d5) * (yl - y_d5)) <= r_d5 * r_d5;
    const bool c6 = (xl >= x_d4) && (xl <= x_d3) && ((yl - a_d6 * xl - b_d6) <= 0.0);
    return !(c1 || c2 || c3 || c4 || c5 || c6);
}

// Track fit of the 12 chamber planes (hms/mc_hms_hut.f:439-455,
// shms/mc_shms_hut.f:410-426, sos/mc_sos_hut.f:409-414) with LFIT
// (cern/lfit.f:22-59, KEY = 0).  Each plane measures one coordinate; the
// reference stores the other coordinate as zero, which LFIT's zero test
// skips, so each fit runs over its own planes in increasing order.  Hits are
// re-read from LDS in the second pass instead of being held in registers.
// T is float for the reference's REAL*4 hits and results (sums in FP64).
template <class T, int NP>
__device__ __forceinline__ void lfit_planes(const int (&pl)[NP], const double* lds, const double (&zdc)[12],
                                            T& A, T& B)
{
    double count = 0.0, sumx = 0.0, sumy = 0.0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const T y = (T)lds[pl[j] * kBlock + threadIdx.x];
        const T x = (T)zdc[pl[j]];
        if (y != (T)0) {
            sumx = sumx + (double)x;
            sumy = sumy + (double)y;
            count = count + 1.0;
        }
    }
    if (count <= 1.0) { A = (T)0; B = (T)0; return; }
    const double ymed = sumy / count;
    const double xmed = sumx / count;
    double sumxy = 0.0, sumxx = 0.0;
    __asm__ volatile("" ::: "memory");  // second pass reloads the hits
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const T y = (T)lds[pl[j] * kBlock + threadIdx.x];
        const T x = (T)zdc[pl[j]];
        if (y != (T)0) {
            const double sx = (double)x - xmed;
            const double sy = (double)y - ymed;
            sumxy = sumxy + sx * sy;
            sumxx = sumxx + sx * sx;
        }
    }
    if (sumxx == 0.0) { A = (T)0; B = (T)0; return; }
    A = (T)(sumxy / sumxx);
    B = (T)(ymed - (double)A * xmed);
}

template <class T, int NX, int NY>
__device__ __forceinline__ void fit_planes(const int (&xpl)[NX], const int (&ypl)[NY], const double* lds,
                                           const double (&zdc)[12], double& x_fp, double& dx_fp, double& y_fp,
                                           double& dy_fp)
{
    T a, b;
    lfit_planes<T, NX>(xpl, lds, zdc, a, b);
    dx_fp = (double)a; x_fp = (double)b;
    lfit_planes<T, NY>(ypl, lds, zdc, a, b);
    dy_fp = (double)a; y_fp = (double)b;
}

// The SOS and the HRS measure y on planes 1, 3, 5 of each chamber
// (sos/mc_sos_hut.f:324-329, hrsl/mc_hrsl_hut.f:275-279), the HMS and SHMS on
// planes 2 and 5.
template <int ARM = SIMC_ARM_HMS>
__device__ __forceinline__ void fit_track(const Flags& F, const double* lds, const double (&zdc)[12],
                                          double& x_fp, double& dx_fp, double& y_fp, double& dy_fp)
{
    if constexpr (ARM == SIMC_ARM_SOS || ARM == SIMC_ARM_HRSR) {
        const int xpl[6] = {1, 3, 5, 7, 9, 11};
        const int ypl[6] = {0, 2, 4, 6, 8, 10};
        if (F.fit_f32) fit_planes<float>(xpl, ypl, lds, zdc, x_fp, dx_fp, y_fp, dy_fp);
        else fit_planes<double>(xpl, ypl, lds, zdc, x_fp, dx_fp, y_fp, dy_fp);
    } else {
        const int xpl[8] = {0, 2, 3, 5, 6, 8, 9, 11};
        const int ypl[4] = {1, 4, 7, 10};
        if (F.fit_f32) fit_planes<float>(xpl, ypl, lds, zdc, x_fp, dx_fp, y_fp, dy_fp);
        else fit_planes<double>(xpl, ypl, lds, zdc, x_fp, dx_fp, y_fp, dy_fp);
    }
}

// ---------------------------------------------------------------------------
// The step-program interpreter of the magnet stages: one event of arm ARM,
// steps [pc_begin, pc_end) (drifts, COSY segments, apertures and the
// collimator).  Only the track, the random stream and the tilted-plane point
// (xt, yt) are carried from step to step.  GEN: the arm's maps run from
// run-time polynomial tables (a copy of its own, so that the compiled
// polynomials' code is not shaped by the other path).
// ---------------------------------------------------------------------------
template <int ARM, class R, bool RARE, bool GEN>
__device__ __forceinline__ int run_program(R& g, Track& t, const KArgs& a, const Flags& FA, uint32_t off,
                                           unsigned* coll_cnt, int pc_begin, int pc_end)
{
    // The in-flight decay and the collimator stepping of pions/muons are
    // compiled only into the RARE variant, so the common variant does not
    // pay their register cost.
    Flags F = FA;
    if (!RARE) {
        F.decay = 0;
        F.coll = 0;
    }
    const cdptr coef = (cdptr)a.coef;
    const coptr ops = (coptr)a.ops;
    const ctptr tab = (ctptr)a.tab;
    double xt = 0.0, yt = 0.0;
    int st = ST_OK;
    for (int pc = pc_begin; pc < pc_end; ++pc) {
        const int code = ops[pc].code;
        if (code == OP_END) break;
        const double p0 = ops[pc].p[0], p1 = ops[pc].p[1];
        switch (code) {
        case OP_PROJECT:
            project(g, t, F, p0);
            break;
        case OP_TRANSP: {
            const int cls = ops[pc].arg;
            transp(g, t, F, p0, [&](double r0, double r1, double r2, double r3, double r4, double& o0,
                                    double& o1, double& o2, double& o3, double& o4) {
                if constexpr (GEN) {
                    double o[5];
                    poly_generic<5>((cuptr)a.gen, cls, r0, r1, r2, r3, r4, o);
                    o0 = o[0]; o1 = o[1]; o2 = o[2]; o3 = o[3]; o4 = o[4];
                } else {
                    poly_fwd<ARM>(cls, coef, r0, r1, r2, r3, r4, o0, o1, o2, o3, o4);
                }
            });
            break;
        }
        case OP_CHECK_R2:
            if ((t.xs * t.xs + t.ys * t.ys) > p0) st = ops[pc].stop;
            break;
        case OP_ROT_H:
        case OP_ROT_V: {
            const Rot r{p0, p1, ops[pc].p[2]};
            xt = t.xs;
            yt = t.ys;
            if (code == OP_ROT_H) rotate_haxis(t, r, xt, yt);
            else rotate_vaxis(t, r, xt, yt);
            xt = xt + ops[pc].p[3];
            yt = yt + ops[pc].p[4];
            break;
        }
        case OP_CHECK_T_R2:
            if ((xt * xt + yt * yt) > p0) st = ops[pc].stop;
            break;
        case OP_CHECK_T_HB:
            if ((xt * xt > p0) || (yt > p1) || (yt < ops[pc].p[2])) st = ops[pc].stop;
            break;
        case OP_HIT_DIPOLE_T:
            if (hms_hit_dipole(xt, yt)) st = ops[pc].stop;
            break;
        case OP_PIPE_T: {
            const double dx = xt - p1, dy = yt - ops[pc].p[2];
            if (((dx * dx + dy * dy) > p0) || (fabs(dy) > ops[pc].p[3])) st = ops[pc].stop;
            break;
        }
        case OP_PIPE_S: {
            const double dx = t.xs - p1, dy = t.ys - ops[pc].p[2];
            if ((dx * dx + dy * dy) > p0) st = ops[pc].stop;
            break;
        }
        case OP_COLLIMATOR:
            if (F.coll && t.m2 > 100.0 * 100.0 && t.m2 < 200.0 * 200.0) {
                CollGeom cg;
                cg.h_entr = tab->coll.h_entr; cg.v_entr = tab->coll.v_entr;
                cg.h_exit = tab->coll.h_exit; cg.v_exit = tab->coll.v_exit;
                cg.x_off = tab->coll.x_off; cg.y_off = tab->coll.y_off;
                cg.coll_radl = tab->coll.coll_radl; cg.step_size = tab->coll.step_size;
                if (!coll_step(g, t, F, cg, F.p_spec, coll_cnt)) st = ST_COLL;
            } else {
                const double x_off = ops[pc].p[4], y_off = ops[pc].p[5];
                int s = slit(t.xs, t.ys, p0, p1, x_off, y_off);
                if (!s) {
                    project(g, t, F, ops[pc].p[6]);
                    s = slit(t.xs, t.ys, ops[pc].p[2], ops[pc].p[3], x_off, y_off);
                }
                st = s;
            }
            break;
        case OP_CHECK_T_BOX:
            if ((yt > p0) || (-yt > p1) || (-xt > ops[pc].p[2]) || (-xt < ops[pc].p[3])) st = ops[pc].stop;
            break;
        case OP_CHECK_TRAP:
            if ((fabs(t.xs) > p0) || fabs(t.ys) > (p1 + ops[pc].p[2] * (t.xs + p0))) st = ops[pc].stop;
            break;
        case OP_SHIFT:
            t.xs = t.xs + p0 * t.dxdzs;
            t.ys = t.ys + p0 * t.dydzs;
            break;
        case OP_CHECK_ABS:
            if ((fabs(t.xs) > p0) || fabs(t.ys) > p1) st = ops[pc].stop;
            break;
        case OP_CHECK_R:
            if (sqrt(t.xs * t.xs + t.ys * t.ys) > p0) st = ops[pc].stop;
            break;
        case OP_CHECK_T_SLAB:
            if ((fabs(xt - p0) > p1) || ((fabs(yt) + ops[pc].p[2] * xt) > ops[pc].p[3])) st = ops[pc].stop;
            break;
        case OP_SIEVE:
            // compiled into the RARE variants only (a run with the sieve on
            // launches them), keeping its registers out of the common ones;
            // the forced angles are also the INOUT dxdz, dydz of the event
            if constexpr (RARE) {
                if (F.sieve & ops[pc].arg) {
                    sieve_slit(g, t, (int)ops[pc].p[4], p0, p1, ops[pc].p[2], ops[pc].p[3]);
                    st64(a.o_dxdz, a.n_arr, off, t.dxdzs);
                    st64(a.o_dydz, a.n_arr, off, t.dydzs);
                }
            }
            break;
        default:
            st = ST_BADOP;  // not a magnet step (excluded by the host's check)
            break;
        }
        if (st) return st;
    }
    return ST_OK;
}

// ---------------------------------------------------------------------------
// The detector hut, compiled.  gen_hut.cpp writes the hut steps of each arm
// family as straight-line code (hut_gen.inc): one step<code, arg, stop>(k,
// p0..p7) call per step k of the program, its parameters as literals where
// they do not depend on the loaded maps.  The library checks every loaded
// program against those tables (load_arm).  Below are the steps' semantics
// (the materials of musc / musc_ext, the wire planes, boxes, scintillators,
// trigger, fit and fiducial cut of the mc_*_hut routines).
//
// Carried from step to step: the track, the random streams, resmult, the
// scintillator count and the fitted track.  The steps draw from the hut's
// stream hs (HutStreamOf); in-flight decays in the drifts draw from the arm's
// stream g.  Wire hits go to LDS (hits[plane][lane]).
// ---------------------------------------------------------------------------
template <int>
constexpr bool kNotHutStep = false;

template <int ARM, class R, bool RARE, class S>
struct Hut {
    R& g;
    S& hs;
    Track& t;
    const KArgs& a;
    Flags F;
    uint32_t off;
    double* lds_hits;
    double& resmult;
    int scincount;
    ThetaCache ms;
    double fp[4];  // fitted x_fp, x'_fp, y_fp, y'_fp

    // parameter j of program step k, for the parameters that depend on the maps
    __device__ __forceinline__ double runtime_p(int k, int j) const { return ((coptr)a.ops)[k].p[j]; }

    template <int CODE, int ARG, int STOP>
    __device__ __forceinline__ int step(int, double p0, double p1, double p2, double p3, double p4, double p5,
                                        double p6, double)
    {
        if constexpr (CODE == OP_PROJECT) {
            project(g, t, F, p0);
        } else if constexpr (CODE == OP_RESMULT_DRAW) {
            resmult = hs.next() < p0 ? 2.0 : 1.0;
        } else if constexpr (CODE == OP_RESMULT_ONE) {
            resmult = 1.0;
        } else if constexpr (CODE == OP_MUSC) {
            // musc(m2,p,radw,dydzs,dxdzs): first draw to y', second to x'
            if (F.ms && p0 != 0.0) {
                double d2, d3;
                hs.gauss_pair(d2, d3);
                const double ts = ms.theta(t, p1, p2);
                t.dydzs = t.dydzs + ts * d2;
                t.dxdzs = t.dxdzs + ts * d3;
            }
        } else if constexpr (CODE == OP_MUSC_EXT) {
            // musc_ext(...,dydzs,dxdzs,ys,xs): (x', x) then (y', y)
            if (F.ms && p0 != 0.0 && p1 > 0.0 && p0 >= 0.0) {
                double d0, d1, d2, d3;
                hs.gauss_pair(d0, d1);
                hs.gauss_pair(d2, d3);
                const double ts = ms.theta(t, p2, p3);
                const double s12 = 3.4641016151377544;  // sqrt(12.)
                t.dxdzs = t.dxdzs + ts * d0;
                t.xs = t.xs + ts * p1 * d1 / s12 + ts * p1 * d0 / 2.0;
                t.dydzs = t.dydzs + ts * d2;
                t.ys = t.ys + ts * p1 * d3 / s12 + ts * p1 * d2 / 2.0;
            }
        } else if constexpr (CODE == OP_WIRE) {
            // plane STOP (0-based); a y plane when ARG = 1
            double d2 = 0.0, d3 = 0.0;
            if (F.wcs) hs.gauss_pair(d2, d3);
            const double h = ARG ? t.ys + F.sigma * d3 * resmult : t.xs + F.sigma * d2 * resmult;
            lds_hits[STOP * kBlock + threadIdx.x] = h;
        } else if constexpr (CODE == OP_BOX) {
            if (t.xs > p0 || t.xs < p1 || t.ys > p2 || t.ys < p3) return STOP;
        } else if constexpr (CODE == OP_SCIN) {
            if (t.ys < p0 && t.ys > p1 && t.xs < p2 && t.xs > p3) ++scincount;
        } else if constexpr (CODE == OP_TRIGGER) {
            if (scincount < ARG) return STOP;
        } else if constexpr (CODE == OP_FIT) {
            const ctptr tab = (ctptr)a.tab;
            double zdc[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) zdc[k] = tab->zdc[k];
            fit_track<ARM>(F, lds_hits, zdc, fp[0], fp[1], fp[2], fp[3]);
            st64(a.x_fp, a.n_arr, off, fp[0]);
            st64(a.dx_fp, a.n_arr, off, fp[1]);
            st64(a.y_fp, a.n_arr, off, fp[2]);
            st64(a.dy_fp, a.n_arr, off, fp[3]);
        } else if constexpr (CODE == OP_BOX_ROT_H) {
            const Rot r{p0, p1, p2};
            double xr = t.xs, yr = t.ys;
            rotate_haxis(t, r, xr, yr);
            if (xr > p3 || xr < p4 || yr > p5 || yr < p6) return STOP;
        } else if constexpr (CODE == OP_FIDUCIAL) {
            // track of the fit at z = p0
            const double xcal = fp[0] + fp[1] * p0;
            const double ycal = fp[2] + fp[3] * p0;
            if (ycal > p1 || ycal < p2 || xcal > p3 || xcal < p4) return STOP;
        } else {
            static_assert(kNotHutStep<CODE>, "not a detector-hut step");
        }
        return ST_OK;
    }
};

template <int ARM, class H>
__device__ __forceinline__ int hut_steps(H& h)
{
    if constexpr (ARM == SIMC_ARM_HMS) return hut_steps_hms(h);
    else if constexpr (ARM == SIMC_ARM_SHMS) return hut_steps_shms(h);
    else if constexpr (ARM == SIMC_ARM_SOS) return hut_steps_sos(h);
    else return hut_steps_hrs(h);
}

// One event through its arm's compiled hut: ST_OK or the stop code; `drawn`
// receives the uniforms taken from the hut's own stream.
template <int ARM, class R, bool RARE>
__device__ __forceinline__ int run_hut(R& g, Track& t, const KArgs& a, const Flags& FA, uint32_t off,
                                       double* lds_hits, double& resmult, uint32_t& drawn)
{
    Flags F = FA;
    if (!RARE) {  // in-flight decay only in the RARE variant (as in the magnets)
        F.decay = 0;
        F.coll = 0;
    }
    using S = typename HutStreamOf<R>::type;
    if constexpr (HutStreamOf<R>::kOwn) {
        S hs;
        hs.init_from(g);
        Hut<ARM, R, RARE, S> h{g, hs, t, a, F, off, lds_hits, resmult, 0, ThetaCache{}, {0.0, 0.0, 0.0, 0.0}};
        const int st = hut_steps<ARM>(h);
        drawn = hs.used();
        return st;
    } else {
        Hut<ARM, R, RARE, R> h{g, g, t, a, F, off, lds_hits, resmult, 0, ThetaCache{}, {0.0, 0.0, 0.0, 0.0}};
        drawn = 0;
        return hut_steps<ARM>(h);
    }
}

// ---------------------------------------------------------------------------
// Kernels.  A launch runs in three stages, each over full wavefronts of the
// events still in flight:
//   1  target -> collimator (program steps [0, coll_pc)), from the inputs;
//   2  the magnets (steps [coll_pc, hut_pc)), from queue 1;
//   3  the detector hut, track fit and reconstruction (steps [hut_pc, end)),
//      from queue 2.
// Events stopped in stage 1 or 2 are finished there; survivors are appended
// to the next queue (track, random-stream position) in wavefront order.
// Every stage writes outputs at the event's own index, so results do not
// depend on queue order.  Each workgroup histograms stop codes in LDS and
// adds its non-empty bins to the context's device counters.
// ---------------------------------------------------------------------------

// Survivor queue (structure of arrays, capacity `cap` entries).
struct Queue {
    uint32_t* count;     // entries appended (device counter, zeroed per launch)
    uint32_t *idx, *rng, *bits;   // launch event index, stream word, dflag | setting << 1
    double *pend, *xs, *ys, *dxdzs, *dydzs, *dpps, *p, *m2, *pathlen, *decdist;
    uint32_t cap;
};

__device__ __forceinline__ SetDev setting_of(const KArgs& a, uint32_t s)
{
    return a.set ? a.set[s] : a.set0;
}

// The setting of pass index P: the last s with start <= P (binary search).
__device__ __forceinline__ uint32_t find_setting(const KArgs& a, uint64_t P)
{
    if (!a.set) return 0;
    uint32_t lo = 0, hi = a.n_set;   // invariant: start[lo] <= P < start[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.set[mid].start <= P) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ Flags flags_of(const KArgs& a, const SetDev& S)
{
    Flags F;
    F.ctau = S.ctau;
    F.p_spec = S.p_spec;
    F.sigma = S.sigma;
    F.ms = (a.flags & KF_MS) != 0;
    F.wcs = (a.flags & KF_WCS) != 0;
    F.decay = (a.flags & KF_DECAY) != 0;
    F.coll = (a.flags & KF_COLL) != 0;
    F.fit_f32 = (a.flags & KF_FIT_F32) != 0;
    F.sieve = ((a.flags & KF_SIEVE) ? 1 : 0) | ((a.flags & KF_SIEVE2) ? 2 : 0);
    return F;
}

// Identity of launch event li: its setting, the byte offset of its element
// in the event arrays and its random-stream event id.
struct EvId {
    uint32_t s, off;
    uint64_t event;
};
__device__ __forceinline__ EvId event_id(const KArgs& a, const SetDev& S, uint32_t s, uint32_t li)
{
    const uint64_t j = a.pass_base + li - S.start;
    EvId e;
    e.s = s;
    e.off = (uint32_t)(S.offset + j - a.ix_base) * 8u;
    e.event = S.event0 + j;
    return e;
}

// The reconstruction of an accepted event from the fitted focal plane in the
// output arrays (byte offset off): the target quantities, and the HRS VDC
// offset taken off y_fp afterwards.
template <int ARM>
__device__ __forceinline__ void reconstruct_event(const KArgs& a, uint32_t off)
{
    const uint32_t n = a.n_arr;
    EvOut o, r;
    o.x_fp = ld64(a.x_fp, n, off);
    o.dx_fp = ld64(a.dx_fp, n, off);
    o.y_fp = ld64(a.y_fp, n, off);
    o.dy_fp = ld64(a.dy_fp, n, off);
    recon<ARM>((cdptr)a.coef, (cuptr)a.gen, o, a.fry ? ld64(a.fry, n, off) : ld64(a.x, n, off), r);
    if (ARM == SIMC_ARM_HRSR) st64(a.y_fp, n, off, o.y_fp - ((ctptr)a.tab)->yfp_shift);
    st64(a.o_dpp, n, off, r.dpp);
    st64(a.o_y, n, off, r.y);
    st64(a.o_dxdz, n, off, r.dxdz);
    st64(a.o_dydz, n, off, r.dydz);
}

// Outputs of a finished event (byte offset off in the event arrays); an
// accepted event is reconstructed here unless RECON_LATER (the hut stage
// hands accepted events to recon_list_kernel).
template <int ARM, bool RECON_LATER = false>
__device__ __forceinline__ void finish_event(const KArgs& a, uint32_t off, int st, uint32_t used, const Track& t,
                                             double resmult)
{
    const uint32_t n = a.n_arr;
    if (st == ST_OK) {
        if (!RECON_LATER) reconstruct_event<ARM>(a, off);
    } else {
        // failure leaves the INOUT target quantities at their inputs (hms/mc_hms.f:438-440);
        // the mock sieve has already written its forced angles
        st64(a.o_dpp, n, off, ld64(a.dpp, n, off));
        st64(a.o_y, n, off, ld64(a.y, n, off));
        if (!(a.flags & (KF_SIEVE | KF_SIEVE2))) {
            st64(a.o_dxdz, n, off, ld64(a.dxdz, n, off));
            st64(a.o_dydz, n, off, ld64(a.dydz, n, off));
        }
    }
    st64(a.pathlen, n, off, t.pathlen);
    if (a.resmult) st64(a.resmult, n, off, resmult);
    if (a.m2o) st64(a.m2o, n, off, t.m2);
    if (a.used) st32(a.used, n, off >> 1, used);
    st8(a.ok, n, off >> 3, st == ST_OK ? 1 : 0);
    st8(a.stop, n, off >> 3, (uint8_t)st);
}

// Stop-code counts: the workgroup's LDS histogram serves the setting of its
// first event; events of other settings (at setting boundaries) count
// directly in device memory.
__device__ __forceinline__ void count_stop(const KArgs& a, unsigned* hist, uint32_t s_block, uint32_t s, int bin,
                                           unsigned n)
{
    if (s == s_block) atomicAdd(&hist[bin], n);
    else atomicAdd(&a.hist[(size_t)setting_of(a, s).hist * kHistSlots + bin], (unsigned long long)n);
}

// Stage kernel.  FROM_Q: events come from queue qin (else from the inputs);
// TO_Q: survivors go to queue qout (else every event is finished here); in
// the hut stage, TO_Q lists the accepted events (index and setting) for
// recon_list_kernel instead of reconstructing them here.
// The common variants are held to three wavefronts per SIMD (168 VGPRs),
// the hut stage to four (128 VGPRs: a few values spill to scratch, and the
// extra wavefront hides more of the hut's latency than the spills cost), the
// RARE variants to two (256 VGPRs, a few spills; at one wavefront a pion run
// with decay took 1.45 x as long).
template <int ARM, class R, bool RARE, int STAGE, bool FROM_Q, bool TO_Q>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(RARE ? 2 : (STAGE == STAGE_HUT ? 4 : 3)))) void stage_kernel(
    KArgs a, Queue qin, Queue qout, int32_t pc_begin, int32_t pc_end)
{
    __shared__ unsigned hist[kHistSlots];
    __shared__ uint32_t wave_n[kBlock / 64];
    __shared__ uint32_t block_base;
    __shared__ uint32_t block_set;
    __shared__ double lds_hits[STAGE == STAGE_HUT ? 12 * kBlock : 1];
    uint32_t n_in = a.count;
    if (FROM_Q) {
        n_in = *qin.count;  // complete: written by the previous launch
        if (blockIdx.x * kBlock >= n_in) return;  // whole workgroup past the queue
    }
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    for (int k = threadIdx.x; k < kHistSlots; k += blockDim.x) hist[k] = 0;
    if (threadIdx.x == 0) {
        const uint32_t i0 = blockIdx.x * kBlock;
        block_set = FROM_Q ? (ld32(qin.bits, qin.cap, i0 * 4u) >> 1) : find_setting(a, a.pass_base + i0);
    }
    __syncthreads();
    const uint32_t s_block = block_set;
    unsigned coll_cnt[3] = {0, 0, 0};
    bool survivor = false;
    uint32_t li = 0, s = s_block;
    R g;
    Track t;
    double resmult = 0.0;
    if (i < n_in) {
        if (FROM_Q) {
            const uint32_t c = qin.cap, o8 = i * 8u, o4 = i * 4u;
            li = ld32(qin.idx, c, o4);
            const uint32_t bits = ld32(qin.bits, c, o4);
            s = bits >> 1;
            const SetDev S = setting_of(a, s);
            const EvId e = event_id(a, S, s, li);
            g.init(R::kLaneKey ? S.seed : a.seed, e.event, a.tape, li * a.stride_bytes, a.stride_bytes);
            // the pending normal is stored only when there is one (bit 31)
            const uint32_t word = ld32(qin.rng, c, o4);
            g.resume(word, (word >> 31) ? ld64(qin.pend, c, o8) : 0.0);
            t.xs = ld64(qin.xs, c, o8);
            t.ys = ld64(qin.ys, c, o8);
            t.dxdzs = ld64(qin.dxdzs, c, o8);
            t.dydzs = ld64(qin.dydzs, c, o8);
            t.dpps = ld64(qin.dpps, c, o8);
            t.pathlen = ld64(qin.pathlen, c, o8);
            if constexpr (RARE) {
                // decay and collimator stepping change the momentum and mass
                t.p = ld64(qin.p, c, o8);
                t.m2 = ld64(qin.m2, c, o8);
                t.decdist = ld64(qin.decdist, c, o8);
            } else {
                // as stage 1 forms them: the same bits
                t.p = S.p_spec * (1.0 + t.dpps / 100.0);
                t.m2 = S.m2;
                t.decdist = 0.0;
            }
            t.dflag = (int)(bits & 1u);
            // resmult is drawn in the hut only: until then it is the
            // caller's INOUT value, still in place at the event's index
            resmult = a.resmult ? ld64(a.resmult, a.n_arr, e.off) : 0.0;
        } else {
            li = i;
            s = find_setting(a, a.pass_base + li);
        }
        const SetDev S = setting_of(a, s);
        const EvId e = event_id(a, S, s, li);
        const Flags F = flags_of(a, S);
        if (!FROM_Q) {
            const uint32_t n = a.n_arr, off = e.off;
            g.init(R::kLaneKey ? S.seed : a.seed, e.event, a.tape, li * a.stride_bytes, a.stride_bytes);
            t.xs = ld64(a.x, n, off);
            t.ys = ld64(a.y, n, off);
            t.dxdzs = ld64(a.dxdz, n, off);
            t.dydzs = ld64(a.dydz, n, off);
            t.dpps = ld64(a.dpp, n, off);
            t.p = S.p_spec * (1.0 + t.dpps / 100.0);
            t.m2 = S.m2;
            t.pathlen = ld64(a.pathlen, n, off);
            t.decdist = 0.0;
            t.dflag = 0;
            resmult = a.resmult ? ld64(a.resmult, n, off) : 0.0;
        }
        int st;
        uint32_t hut_drawn = 0;
        if constexpr (STAGE == STAGE_HUT) st = run_hut<ARM, R, RARE>(g, t, a, F, e.off, lds_hits, resmult, hut_drawn);
        else if (a.gen) st = run_program<ARM, R, RARE, true>(g, t, a, F, e.off, coll_cnt, pc_begin, pc_end);
        else st = run_program<ARM, R, RARE, false>(g, t, a, F, e.off, coll_cnt, pc_begin, pc_end);
        const uint32_t used = (uint32_t)g.used() + hut_drawn;
        if (g.overflow) st = ST_TAPE;
        if (STAGE == STAGE_HUT) {
            // TO_Q: accepted events go to the reconstruction list
            finish_event<ARM, TO_Q>(a, e.off, st, used, t, resmult);
            count_stop(a, hist, s_block, s, st, 1u);
            survivor = TO_Q && st == ST_OK;
        } else if (TO_Q && st == ST_OK) {
            survivor = true;
        } else {
            finish_event<ARM>(a, e.off, st, used, t, resmult);
            count_stop(a, hist, s_block, s, st, 1u);
        }
        if (coll_cnt[0] | coll_cnt[1] | coll_cnt[2]) {
            for (int k = 0; k < 3; ++k)
                if (coll_cnt[k]) count_stop(a, hist, s_block, s, kHistBins + k, coll_cnt[k]);
        }
    }

    if (TO_Q) {
        // Queue slots: rank within the wavefront, wavefront offset within
        // the workgroup, one counter update per workgroup.
        const uint64_t mask = __ballot(survivor);
        const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
        const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) wave_n[wave] = (uint32_t)__popcll(mask);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (int w = 0; w < kBlock / 64; ++w) tot += wave_n[w];
            block_base = tot ? atomicAdd(qout.count, tot) : 0u;
        }
        __syncthreads();
        if (survivor) {
            uint32_t slot = block_base + rank;
            for (int w = 0; w < wave; ++w) slot += wave_n[w];
            const uint32_t c = qout.cap, o8 = slot * 8u, o4 = slot * 4u;
            st32(qout.idx, c, o4, li);
            st32(qout.bits, c, o4, (uint32_t)t.dflag | (s << 1));
            if (STAGE != STAGE_HUT) {
                // 60 bytes per entry (84 in the RARE variants), plus 8 when
                // a Box-Muller normal is pending
                const uint32_t word = g.state_word();
                st32(qout.rng, c, o4, word);
                if (word >> 31) st64(qout.pend, c, o8, g.pending);
                st64(qout.xs, c, o8, t.xs);
                st64(qout.ys, c, o8, t.ys);
                st64(qout.dxdzs, c, o8, t.dxdzs);
                st64(qout.dydzs, c, o8, t.dydzs);
                st64(qout.dpps, c, o8, t.dpps);
                st64(qout.pathlen, c, o8, t.pathlen);
                if constexpr (RARE) {
                    st64(qout.p, c, o8, t.p);
                    st64(qout.m2, c, o8, t.m2);
                    st64(qout.decdist, c, o8, t.decdist);
                }
            }
        }
    }
    __syncthreads();
    const uint32_t hist_index = setting_of(a, s_block).hist;
    for (int k = threadIdx.x; k < kHistSlots; k += blockDim.x)
        if (hist[k]) atomicAdd(&a.hist[(size_t)hist_index * kHistSlots + k], (unsigned long long)hist[k]);
}

// Reconstruction of the accepted events the hut stage listed (one per
// lane): the COSY reconstruction polynomial alone needs few registers, so it
// runs at full occupancy here instead of holding the hut stage's.
template <int ARM>
__global__ __launch_bounds__(kBlock) void recon_list_kernel(KArgs a, Queue acc)
{
    const uint32_t n_in = *acc.count;
    if (blockIdx.x * kBlock >= n_in) return;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_in) return;
    const uint32_t li = ld32(acc.idx, acc.cap, i * 4u);
    const uint32_t s = ld32(acc.bits, acc.cap, i * 4u) >> 1;
    const SetDev S = setting_of(a, s);
    reconstruct_event<ARM>(a, event_id(a, S, s, li).off);
}

// ---------------------------------------------------------------------------
// Adjoint form: loss and gradients of an optics fit over a fixed set of
// events, with the events' random draws replayed (their Philox streams or
// tape slots) and the survival mask held fixed.
//
//   model:  delta' = 100 ((1 + delta/100) / (1 + s_p) - 1)   (momentum scale s_p)
//           dy/dz' = dy/dz - dth                                (central-angle offset)
//           transport through magnets and hut, FP64 track fit
//           X = x_fp + ox, Y = y_fp + oy, X' = dx_fp + ox', Y' = dy_fp + oy'
//           reconstruction with coefficients C (all 4 x n_terms)
//   loss:   L = 1/N sum_events [ (delta_r - delta)^2 + (x'_r - x')^2 +
//                                (y'_r - y')^2 + (y_r - y)^2 ]
//   grads:  dL/dC, dL/d(ox, oy, ox', oy', s_p, dth)
//
// One event per lane runs the step program with the value and two tangent
// directions (d/ds_p, d/dth) of the track; apertures are not tested (mask
// fixed); the hut's scattering kicks do not depend on the track; the fit is
// a fixed linear map of the hits.  The reconstruction gives its sums and
// their gradient in the focal-plane inputs; dL/dC is accumulated with each
// lane owning reconstruction terms and looping over the events of its
// workgroup's tile.  Sums are normalised by N on the device at the end.
// ---------------------------------------------------------------------------
struct AdjArgs {
    SetDev set0;
    int32_t flags;
    uint64_t seed;
    uint32_t n;                    // events in the arrays
    const ArmTables* tab;
    const double* coef;
    const Op* ops;
    const double *dpp, *x, *y, *dxdz, *dydz, *fry;
    const uint32_t* sel;           // selected events, *n_sel of them
    const uint32_t* n_sel;
    const double* tape;
    uint32_t stride_bytes;
    const double* rcoef;           // recon coefficients [n_terms][4]
    const uint8_t* rexp;           // recon exponents [n_terms][5]
    uint32_t n_terms;
    double off[6];                 // ox, oy, ox', oy' (cm, rad), s_p, dth (rad)
    double* g_coef;                // sums over events of dL_e/dC, [n_terms][4]
    double* hw;                    // per selected event: h0..h4, w0..w3 (SoA, n entries each)
    double* ft;                    // per selected event: fit (4), its tangents (4 + 4), validity
    double* g_out;                 // [0] sum L_e, [1..6] sums of dL_e/d offsets, [7] events excluded,
                                   // [8] events used (set at the end)
};
constexpr int kAdjTermsPerLane = 2;   // n_terms <= 512

__global__ void select_events(const uint8_t* mask, uint32_t n, uint32_t* sel, uint32_t* n_sel)
{
    __shared__ uint32_t wave_n[kBlock / 64];
    __shared__ uint32_t base;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool keep = i < n && mask[i] != 0;
    const uint64_t m = __ballot(keep);
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
    if (lane == 0) wave_n[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < kBlock / 64; ++w) tot += wave_n[w];
        base = tot ? atomicAdd(n_sel, tot) : 0u;
    }
    __syncthreads();
    if (keep) {
        uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        for (int w = 0; w < wave; ++w) slot += wave_n[w];
        sel[slot] = i;
    }
}

// The step program of one event with tangents, up to the track fit.
// Returns false when a hit of the fit is exactly zero (LFIT would skip it and
// the fit would not be the fixed linear map), which excludes the event.
template <int ARM, class R>
__device__ bool adjoint_track(R& g, Track& t, double (&ta)[5], double (&tb)[5], const AdjArgs& a, const Flags& F,
                              double* lds_hits, double (&fp)[4], double (&fa)[4], double (&fb)[4])
{
    const cdptr coef = (cdptr)a.coef;
    const coptr ops = (coptr)a.ops;
    const ctptr tab = (ctptr)a.tab;
    double resmult = 0.0;
    ThetaCache ms;
    // the detector hut draws from its own stream (as in run_hut)
    using HS = typename HutStreamOf<R>::type;
    HS hs_own;
    bool in_hut = false;
    // fit tangents: slope / intercept of the x and y plane sets, two directions
    double sax = 0, sbx = 0, say = 0, sby = 0, iax = 0, ibx = 0, iay = 0, iby = 0;
    for (int pc = 0;; ++pc) {
        const int code = ops[pc].code;
        if (code == OP_END) break;
        if constexpr (HutStreamOf<R>::kOwn) {
            if (!in_hut && op_stages(code) == STAGE_HUT) {
                in_hut = true;
                hs_own.init_from(g);
            }
        }
        const double p0 = ops[pc].p[0], p1 = ops[pc].p[1];
        int ng = 0;
        if (code == OP_MUSC) ng = (F.ms && p0 != 0.0) ? 2 : 0;
        else if (code == OP_MUSC_EXT) ng = (F.ms && p0 != 0.0 && p1 > 0.0 && p0 >= 0.0) ? 4 : 0;
        else if (code == OP_WIRE) ng = F.wcs ? 2 : 0;
        double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
        if (in_hut) {
            if constexpr (HutStreamOf<R>::kOwn) {
                if (ng == 4) hs_own.gauss_pair(d0, d1);
                if (ng) hs_own.gauss_pair(d2, d3);
            }
        } else {
            for (int q = 0; q < ng; ++q) {
                const double v = gauss1(g, 99.0);
                d0 = d1; d1 = d2; d2 = d3; d3 = v;
            }
        }
        switch (code) {
        case OP_PROJECT:
        case OP_COLLIMATOR: {  // the collimator's thickness is a drift for an event that passes
            const double L = code == OP_PROJECT ? p0 : ops[pc].p[6];
            project(g, t, F, L);
            ta[0] = ta[0] + ta[1] * L; ta[2] = ta[2] + ta[3] * L;
            tb[0] = tb[0] + tb[1] * L; tb[2] = tb[2] + tb[3] * L;
            break;
        }
        case OP_TRANSP: {
            const double r[5] = {t.xs, t.dxdzs * 1000.0, t.ys, t.dydzs * 1000.0, t.dpps};
            const double ra[5] = {ta[0], ta[1] * 1000.0, ta[2], ta[3] * 1000.0, ta[4]};
            const double rb[5] = {tb[0], tb[1] * 1000.0, tb[2], tb[3] * 1000.0, tb[4]};
            double o[5], oa[5], ob[5];
            poly_fwd_tan<ARM>(ops[pc].arg, coef, r, ra, rb, o, oa, ob);
            t.xs = o[0]; t.dxdzs = o[1] / 1000.0; t.ys = o[2]; t.dydzs = o[3] / 1000.0;
            t.pathlen = t.pathlen + (p0 + -o[4]);
            ta[0] = oa[0]; ta[1] = oa[1] / 1000.0; ta[2] = oa[2]; ta[3] = oa[3] / 1000.0;
            tb[0] = ob[0]; tb[1] = ob[1] / 1000.0; tb[2] = ob[2]; tb[3] = ob[3] / 1000.0;
            break;
        }
        case OP_RESMULT_DRAW: {
            double u;
            if constexpr (HutStreamOf<R>::kOwn) u = hs_own.next();
            else u = g.next();
            resmult = u < p0 ? 2.0 : 1.0;
            break;
        }
        case OP_RESMULT_ONE:
            resmult = 1.0;
            break;
        case OP_MUSC:
            if (ng) {
                const double ts = ms.theta(t, p1, ops[pc].p[2]);
                t.dydzs = t.dydzs + ts * d2;
                t.dxdzs = t.dxdzs + ts * d3;
            }
            break;
        case OP_MUSC_EXT:
            if (ng) {
                const double ts = ms.theta(t, ops[pc].p[2], ops[pc].p[3]);
                const double s12 = 3.4641016151377544;  // sqrt(12.)
                t.dxdzs = t.dxdzs + ts * d0;
                t.xs = t.xs + ts * p1 * d1 / s12 + ts * p1 * d0 / 2.0;
                t.dydzs = t.dydzs + ts * d2;
                t.ys = t.ys + ts * p1 * d3 / s12 + ts * p1 * d2 / 2.0;
            }
            break;
        case OP_WIRE: {
            const int plane = ops[pc].stop;
            const bool yplane = ops[pc].arg != 0;
            const double h = yplane ? t.ys + F.sigma * d3 * resmult : t.xs + F.sigma * d2 * resmult;
            lds_hits[plane * kBlock + threadIdx.x] = h;
            const double wa = tab->fw_a[plane], wb = tab->fw_b[plane];
            if (yplane) {
                say += wa * ta[2]; sby += wa * tb[2]; iay += wb * ta[2]; iby += wb * tb[2];
            } else {
                sax += wa * ta[0]; sbx += wa * tb[0]; iax += wb * ta[0]; ibx += wb * tb[0];
            }
            break;
        }
        case OP_FIT: {
            double zdc[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) zdc[k] = tab->zdc[k];
            Flags F64 = F;
            F64.fit_f32 = 0;
            fit_track(F64, lds_hits, zdc, fp[0], fp[1], fp[2], fp[3]);
            fa[0] = iax; fa[1] = sax; fa[2] = iay; fa[3] = say;
            fb[0] = ibx; fb[1] = sbx; fb[2] = iby; fb[3] = sby;
            bool all = true;
#pragma unroll
            for (int k = 0; k < 12; ++k) all = all && lds_hits[k * kBlock + threadIdx.x] != 0.0;
            return all;
        }
        default:  // aperture tests, rotations, scintillators: no change of the track
            break;
        }
    }
    return false;  // a program without a fit
}

// Pass 1: one selected event per lane -- transport with tangents up to the
// track fit.  Writes the fitted track, its two tangents and a validity flag
// (SoA [13][n_sel] in ft).
template <int ARM, class R>
__global__ __launch_bounds__(kBlock) void adjoint_kernel(AdjArgs a)
{
    __shared__ double lds_hits[12 * kBlock];
    const uint32_t n_sel = *a.n_sel;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_sel) return;
    Flags F;
    F.ctau = 0.0;
    F.p_spec = a.set0.p_spec;
    F.sigma = a.set0.sigma;
    F.ms = (a.flags & KF_MS) != 0;
    F.wcs = (a.flags & KF_WCS) != 0;
    F.decay = 0;
    F.coll = 0;
    F.fit_f32 = 0;
    F.sieve = 0;
    const double sp = a.off[4], dth = a.off[5];
    const uint32_t ev = a.sel[i];
    const uint32_t n = a.n, off = ev * 8u;
    R g;
    g.init(a.seed, a.set0.event0 + ev, a.tape, ev * a.stride_bytes, a.stride_bytes);
    const double dpp = ld64(a.dpp, n, off);
    Track t;
    t.xs = ld64(a.x, n, off);
    t.ys = ld64(a.y, n, off);
    t.dxdzs = ld64(a.dxdz, n, off);
    t.dydzs = ld64(a.dydz, n, off) - dth;
    t.dpps = sp == 0.0 ? dpp : 100.0 * ((1.0 + dpp / 100.0) / (1.0 + sp) - 1.0);
    t.p = a.set0.p_spec * (1.0 + dpp / 100.0);
    t.m2 = a.set0.m2;
    t.pathlen = 0.0;
    t.decdist = 0.0;
    t.dflag = 0;
    double ta[5] = {0, 0, 0, 0, -100.0 * (1.0 + dpp / 100.0) / ((1.0 + sp) * (1.0 + sp))};
    double tb[5] = {0, 0, 0, -1.0, 0};
    double fp[4], fa[4], fb[4];
    const bool ok = adjoint_track<ARM>(g, t, ta, tb, a, F, lds_hits, fp, fa, fb) && !g.overflow;
    const uint32_t c = a.n, o8 = i * 8u;
    for (int k = 0; k < 4; ++k) {
        st64(a.ft + (size_t)k * a.n, c, o8, fp[k]);
        st64(a.ft + (size_t)(4 + k) * a.n, c, o8, fa[k]);
        st64(a.ft + (size_t)(8 + k) * a.n, c, o8, fb[k]);
    }
    st64(a.ft + (size_t)12 * a.n, c, o8, ok ? 1.0 : 0.0);
}

// Pass 2: per selected event, the reconstruction with the offsets and its
// gradient, the event's loss and offset gradients (summed per workgroup),
// and the reconstruction inputs h and weights w = dL_e/d(sums) for pass 3
// (both zero for an excluded event).
template <int ARM>
__global__ __launch_bounds__(kBlock) void adjoint_recon_kernel(AdjArgs a)
{
    __shared__ double red[8 * kBlock];
    const uint32_t n_sel = *a.n_sel;
    if (blockIdx.x * kBlock >= n_sel) return;  // whole workgroup past the selection
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // loss, 6 offsets, excluded
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n_sel) {
        const uint32_t c = a.n, o8 = i * 8u;
        double wk[4] = {0, 0, 0, 0}, hv[5] = {0, 0, 0, 0, 0};
        if (ld64(a.ft + (size_t)12 * a.n, c, o8) == 0.0) {
            acc[7] = 1.0;
        } else {
            const uint32_t ev = a.sel[i];
            const uint32_t n = a.n, off = ev * 8u;
            const double dpp = ld64(a.dpp, n, off), dxdz = ld64(a.dxdz, n, off), dydz = ld64(a.dydz, n, off);
            const double yt = ld64(a.y, n, off);
            double fp[4], fa[4], fb[4];
            for (int k = 0; k < 4; ++k) {
                fp[k] = ld64(a.ft + (size_t)k * a.n, c, o8);
                fa[k] = ld64(a.ft + (size_t)(4 + k) * a.n, c, o8);
                fb[k] = ld64(a.ft + (size_t)(8 + k) * a.n, c, o8);
            }
            double h4 = (a.fry ? ld64(a.fry, n, off) : ld64(a.x, n, off)) / 100.0;
            if (fabs(h4) <= 1.e-30) h4 = 1.e-30;
            hv[0] = (fp[0] + a.off[0]) / 100.0;
            hv[1] = fp[1] + a.off[2];
            hv[2] = (fp[2] + a.off[1]) / 100.0;
            hv[3] = fp[3] + a.off[3];
            hv[4] = h4;
            double s0, s1, s2, s3;
            if (ARM == SIMC_ARM_HMS)
                cosy_hms_recon_full((cdptr)a.rcoef, hv[0], hv[1], hv[2], hv[3], hv[4], s0, s1, s2, s3);
            else
                cosy_shms_recon_full((cdptr)a.rcoef, hv[0], hv[1], hv[2], hv[3], hv[4], s0, s1, s2, s3);
            // residuals: x'_r = s0, y_r = 100 s1, y'_r = s2, delta_r = 100 s3
            const double r0 = s0 - dxdz, r1 = s1 * 100.0 - yt, r2 = s2 - dydz, r3 = s3 * 100.0 - dpp;
            acc[0] = r0 * r0 + r1 * r1 + r2 * r2 + r3 * r3;
            wk[0] = 2.0 * r0;
            wk[1] = 200.0 * r1;
            wk[2] = 2.0 * r2;
            wk[3] = 200.0 * r3;
            // dL_e/dh_i = d/dh_i sum_k w_k s_k
            double dh0, dh1, dh2, dh3;
            if (ARM == SIMC_ARM_HMS)
                cosy_hms_recon_wgrad((cdptr)a.rcoef, hv[0], hv[1], hv[2], hv[3], hv[4], wk[0], wk[1], wk[2], wk[3],
                                     dh0, dh1, dh2, dh3);
            else
                cosy_shms_recon_wgrad((cdptr)a.rcoef, hv[0], hv[1], hv[2], hv[3], hv[4], wk[0], wk[1], wk[2], wk[3],
                                      dh0, dh1, dh2, dh3);
            acc[1] = dh0 / 100.0;   // ox
            acc[2] = dh2 / 100.0;   // oy
            acc[3] = dh1;           // ox'
            acc[4] = dh3;           // oy'
            acc[5] = dh0 * fa[0] / 100.0 + dh1 * fa[1] + dh2 * fa[2] / 100.0 + dh3 * fa[3];   // s_p
            acc[6] = dh0 * fb[0] / 100.0 + dh1 * fb[1] + dh2 * fb[2] / 100.0 + dh3 * fb[3];   // dth
        }
        for (int v = 0; v < 5; ++v) st64(a.hw + (size_t)v * a.n, c, o8, hv[v]);
        for (int k = 0; k < 4; ++k) st64(a.hw + (size_t)(5 + k) * a.n, c, o8, wk[k]);
    }
    for (int q = 0; q < 8; ++q) red[q * kBlock + threadIdx.x] = acc[q];
    __syncthreads();
    for (int stride = kBlock / 2; stride > 0; stride >>= 1) {
        if ((int)threadIdx.x < stride)
            for (int q = 0; q < 8; ++q) red[q * kBlock + threadIdx.x] += red[q * kBlock + threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x < 8 && red[threadIdx.x * kBlock] != 0.0) atomicAdd(&a.g_out[threadIdx.x], red[threadIdx.x * kBlock]);
}

// Pass 3: dL/dC[t][k] = sum_e w_k(e) m_t(h_e).  Each lane owns up to
// kAdjTermsPerLane terms; a workgroup walks tiles of kAdjTile events: each
// lane tabulates h_v^p (p <= 6) of one event in LDS, then every lane forms
// its terms' monomials from the table.  Sums go to the device per workgroup.
constexpr int kAdjTile = 128;
constexpr int kAdjMaxExp = 6;
__global__ __launch_bounds__(kBlock) void adjoint_coef_kernel(AdjArgs a)
{
    __shared__ double pw[kAdjTile][5][kAdjMaxExp + 1];
    __shared__ double wt[kAdjTile][4];
    const uint32_t n_sel = *a.n_sel;
    const uint32_t n_tiles = (n_sel + kAdjTile - 1) / kAdjTile;
    uint32_t own_t[kAdjTermsPerLane];
    uint8_t own_e[kAdjTermsPerLane][5];
    double gc[kAdjTermsPerLane][4];
    for (int j = 0; j < kAdjTermsPerLane; ++j) {
        own_t[j] = threadIdx.x + j * kBlock;
        for (int v = 0; v < 5; ++v)
            own_e[j][v] = own_t[j] < a.n_terms ? a.rexp[own_t[j] * 5 + v] : (uint8_t)0;
        for (int k = 0; k < 4; ++k) gc[j][k] = 0.0;
    }
    const uint32_t c = a.n;
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint32_t e0 = tile * kAdjTile;
        const uint32_t ne = n_sel - e0 < (uint32_t)kAdjTile ? n_sel - e0 : (uint32_t)kAdjTile;
        if (threadIdx.x < kAdjTile) {
            const uint32_t e = threadIdx.x;
            const bool have = e < ne;
            const uint32_t o8 = (e0 + e) * 8u;
            for (int v = 0; v < 5; ++v) {
                const double h = have ? ld64(a.hw + (size_t)v * a.n, c, o8) : 0.0;
                double q = 1.0;
                for (int p = 0; p <= kAdjMaxExp; ++p) {
                    pw[e][v][p] = q;
                    q *= h;
                }
            }
            for (int k = 0; k < 4; ++k) wt[e][k] = have ? ld64(a.hw + (size_t)(5 + k) * a.n, c, o8) : 0.0;
        }
        __syncthreads();
        for (uint32_t e = 0; e < ne; ++e) {
            const double w0 = wt[e][0], w1 = wt[e][1], w2 = wt[e][2], w3 = wt[e][3];
#pragma unroll
            for (int j = 0; j < kAdjTermsPerLane; ++j) {
                const double m = pw[e][0][own_e[j][0]] * pw[e][1][own_e[j][1]] * pw[e][2][own_e[j][2]] *
                                 pw[e][3][own_e[j][3]] * pw[e][4][own_e[j][4]];
                gc[j][0] = __builtin_fma(w0, m, gc[j][0]);
                gc[j][1] = __builtin_fma(w1, m, gc[j][1]);
                gc[j][2] = __builtin_fma(w2, m, gc[j][2]);
                gc[j][3] = __builtin_fma(w3, m, gc[j][3]);
            }
        }
        __syncthreads();
    }
    for (int j = 0; j < kAdjTermsPerLane; ++j)
        if (own_t[j] < a.n_terms)
            for (int k = 0; k < 4; ++k)
                if (gc[j][k] != 0.0) atomicAdd(&a.g_coef[own_t[j] * 4 + k], gc[j][k]);
}

// Divide the sums by the number of events used, N = selected - excluded;
// g_out[8] receives N.  (g_out[7], the exclusions, is read by every
// workgroup and left as it is.)
__global__ void adjoint_finish(double* g_coef, uint32_t n_coef, double* g_out, const uint32_t* n_sel)
{
    const double n_used = (double)*n_sel - g_out[7];
    const double inv = n_used > 0.0 ? 1.0 / n_used : 0.0;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_coef) g_coef[i] *= inv;
    if (i == 0) {
        for (int q = 0; q < 7; ++q) g_out[q] *= inv;
        g_out[8] = n_used;
    }
}

// Queue counter reset (a kernel rather than a memset, so that a captured
// launch sequence replays correctly).
// the two survivor queues' counters and the accepted list's
__global__ void zero_queue_counts(uint32_t* p)
{
    if (threadIdx.x < 3) p[threadIdx.x] = 0u;
}

__global__ void zero_u64(unsigned long long* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

template <int ARM>
__global__ __launch_bounds__(256) void recon_kernel(int64_t n, const double* coef, const uint32_t* gen,
                                                     const double* x_fp,
                                                     const double* dx_fp, const double* y_fp,
                                                     const double* dy_fp, const double* fry, double* delta,
                                                     double* dxdz, double* dydz, double* ytar)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        EvOut o, r;
        o.x_fp = x_fp[i]; o.dx_fp = dx_fp[i]; o.y_fp = y_fp[i]; o.dy_fp = dy_fp[i];
        recon<ARM>((cdptr)coef, (cuptr)gen, o, fry ? fry[i] : 0.0, r);
        delta[i] = r.dpp; dxdz[i] = r.dxdz; dydz[i] = r.dydz; ytar[i] = r.y;
    }
}

// ---------------------------------------------------------------------------
// Target step: one event per lane (simc_target.h has the routines).
// ---------------------------------------------------------------------------
struct PrepArgs {
    PrepDev P;
    uint64_t seed, event0;
    uint32_t n, stride_bytes;
    const double* tape;
    const double *x, *y, *z, *E, *p, *Ev, *theta, *delta, *xptar, *yptar, *dyp, *dxp;
    double *dpp, *ox, *oy, *oz, *dxdz, *dydz, *fry, *eloss, *teff, *d1, *d2;
    int32_t* used;
};

template <class R, int NARM>
__global__ __launch_bounds__(256) void target_prep_kernel(PrepArgs a)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    const PrepDev& P = a.P;
    R g;
    g.init(a.seed, a.event0 + i, a.tape, i * a.stride_bytes, a.stride_bytes);
    const double tz = a.z[i];
    const double zpos = tz - P.zoffset;
    double eloss, teff, d1 = 0.0, d2 = 0.0;
    if (NARM == 1) {  // simc.f:1364-1369 with event.f:194-196
        eloss = trip_beam(g, P, zpos, teff);
        if (!P.using_eloss) eloss = 0.0;
        if (P.mc_smear) target_musc(g, a.E[i], 1.0, teff, d1, d2);
        if (g.overflow) eloss = teff = d1 = d2 = __builtin_nan("");
        if (a.eloss) a.eloss[i] = eloss;
        if (a.teff) a.teff[i] = teff;
        if (a.d1) a.d1[i] = d1;
        if (a.d2) a.d2[i] = d2;
        if (a.used) a.used[i] = g.overflow ? -1 : g.used();
        return;
    }
    const double E = a.E[i], p = a.p[i];
    eloss = trip_scattered(g, P, zpos, a.Ev ? a.Ev[i] : E, a.theta[i], teff);  // event.f:1031-1040
    if (!P.using_eloss) eloss = 0.0;
    double delta;
    if (NARM == 3) {  // simc.f:1379-1396
        if (P.using_eloss) {
            const double e = E - eloss;
            delta = (sqrt(fabs(e * e - P.mass * P.mass)) - P.p_spec) / P.p_spec * 100.0;
        } else {
            delta = a.delta[i];
        }
        if (P.mc_smear) target_musc(g, p, p / E, teff, d1, d2);
    } else {  // simc.f:1653-1662
        delta = 100.0 * (E - eloss - P.coulomb - P.p_spec) / P.p_spec;
        if (P.mc_smear) target_musc(g, p, 1.0, teff, d1, d2);
    }
    const double din1 = a.dyp ? a.dyp[i] : 0.0;
    const double din2 = a.dxp ? a.dxp[i] : 0.0;
    const double yptar = a.yptar[i] + d1 + din1;
    const double xptar = a.xptar[i] + d2 + din2 * P.cos_th;
    // TRANSPORT coordinates, offsets, drift to z = 0 (simc.f:1409-1441)
    const double tx = a.x[i], ty = a.y[i];
    double x = -ty;
    double y = -(tx * P.cos_th) - tz * P.sin_th * P.sin_phi;
    double z = tz * P.cos_th + tx * P.sin_th * P.sin_phi;
    x = x - P.off[0];
    y = y - P.off[1];
    z = z - P.off[2];
    double dx = xptar - P.off[3];
    double dy = yptar - P.off[4];
    x = x - z * dx;
    y = y - z * dy;
    if (g.overflow) delta = x = y = dx = dy = eloss = teff = d1 = d2 = __builtin_nan("");
    if (a.dpp) a.dpp[i] = delta;
    if (a.ox) a.ox[i] = x;
    if (a.oy) a.oy[i] = y;
    if (a.oz) a.oz[i] = 0.0;
    if (a.dxdz) a.dxdz[i] = dx;
    if (a.dydz) a.dydz[i] = dy;
    if (a.fry) a.fry[i] = x;
    if (a.eloss) a.eloss[i] = eloss;
    if (a.teff) a.teff[i] = teff;
    if (a.d1) a.d1[i] = d1;
    if (a.d2) a.d2[i] = d2;
    if (a.used) a.used[i] = g.overflow ? -1 : g.used();
}

// ---------------------------------------------------------------------------
// Event generator, H(e,e'p) elastic (simc_gen.h): one event per lane; every
// output column may be absent.
// ---------------------------------------------------------------------------
constexpr int kHepCols = 27;

struct GenArgs {
    HepDev H;
    PrepDev P;
    uint64_t seed, event0;
    uint32_t n, stride_bytes;
    const double* tape;
    double* col[kHepCols];
    uint8_t* ok;
    int32_t* used;
};

template <class R>
__global__ __launch_bounds__(256) void gen_hep_kernel(GenArgs a)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    R g;
    g.init(a.seed, a.event0 + i, a.tape, i * a.stride_bytes, a.stride_bytes);
    HepEvent v;
    gen_hep(g, a.H, a.P, v);
    const double c[kHepCols] = {v.x, v.y, v.z, v.rasterx, v.rastery, v.Ein, v.eloss, v.teff,
                                v.e_xptar, v.e_yptar, v.e_theta, v.e_phi, v.e_E, v.e_delta,
                                v.p_xptar, v.p_yptar, v.p_theta, v.p_phi, v.p_P, v.p_E, v.p_delta,
                                v.Q2, v.nu, v.q, v.jacobian, v.sigcc, v.weight};
    const double nan = __builtin_nan("");
#pragma unroll
    for (int k = 0; k < kHepCols; ++k)
        if (a.col[k]) a.col[k][i] = g.overflow ? nan : c[k];
    if (a.ok) a.ok[i] = (v.ok && !g.overflow) ? 1 : 0;
    if (a.used) a.used[i] = g.overflow ? -1 : g.used();
}

// The generator with radiative tails: the beam's path (Pb), the electron's
// and the proton's window losses (Pe, Pp), the radiative run constants.
constexpr int kRadCols = 15;
struct GenRadArgs {
    HepDev H;
    PrepDev Pb, Pe, Pp;
    RadDev Rd;
    uint64_t seed, event0;
    uint32_t n, stride_bytes;
    const double* tape;
    double* col[kHepCols];
    double* rcol[kRadCols];
    int32_t* ntail;
    uint8_t* ok;
    int32_t* used;
};

// Held to three wavefronts per SIMD (168 VGPRs, about 260 values spilled):
// at two (its unconstrained 276 registers) the kernel waited on its
// elementary functions' latency and took 1.8 x as long.
template <class R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void gen_hep_rad_kernel(GenRadArgs a)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    R g;
    g.init(a.seed, a.event0 + i, a.tape, i * a.stride_bytes, a.stride_bytes);
    HepEvent v;
    RadOut ro;
    RadEvent ev;
    double hardcorfac;
    gen_hep_rad(g, a.H, a.Pb, a.Pe, a.Pp, a.Rd, v, ro, ev, hardcorfac);
    const double nan = __builtin_nan("");
    const double c[kHepCols] = {v.x, v.y, v.z, v.rasterx, v.rastery, v.Ein, v.eloss, v.teff,
                                v.e_xptar, v.e_yptar, v.e_theta, v.e_phi, v.e_E, v.e_delta,
                                v.p_xptar, v.p_yptar, v.p_theta, v.p_phi, v.p_P, v.p_E, v.p_delta,
                                v.Q2, v.nu, v.q, v.jacobian, v.sigcc, v.weight};
#pragma unroll
    for (int k = 0; k < kHepCols; ++k)
        if (a.col[k]) a.col[k][i] = g.overflow ? nan : c[k];
    const double r[kRadCols] = {ro.Eg[0], ro.Eg[1], ro.Eg[2], ro.gen_weight, ro.orig_Ein, ro.orig_e_E,
                                ro.orig_e_delta, ro.orig_p_E, ro.orig_p_P, ro.orig_p_delta, ev.eloss[1], ev.teff[1],
                                ev.eloss[2], ev.teff[2], hardcorfac};
#pragma unroll
    for (int k = 0; k < kRadCols; ++k)
        if (a.rcol[k]) a.rcol[k][i] = g.overflow ? nan : r[k];
    if (a.ntail) a.ntail[i] = ro.ntail;
    if (a.ok) a.ok[i] = (v.ok && !g.overflow) ? 1 : 0;
    if (a.used) a.used[i] = g.overflow ? -1 : g.used();
}

// ---------------------------------------------------------------------------
// Target-field tracking (trg_track.f).  Each workgroup stages the field map
// (2 x 51 x 51 doubles, 41.6 KB) in LDS; one event per lane.
// ---------------------------------------------------------------------------
struct FieldFromArgs {
    const double* map;  // B_z nodes, then B_r nodes
    double stht, ctht;
    uint32_t n;
    double *x, *y, *z, *dx, *dy;
    const double* mom;
    double mass;
    uint8_t* ok;
};

__device__ __forceinline__ void stage_field_map(double* s, const double* map)
{
    for (int k = threadIdx.x; k < 2 * kTrgNodes; k += blockDim.x) s[k] = map[k];
    __syncthreads();
}

__global__ __launch_bounds__(256) void field_from_kernel(FieldFromArgs a)
{
    __shared__ double s[2 * kTrgNodes];
    stage_field_map(s, a.map);
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    const TrgMap M{s, s + kTrgNodes, a.stht, a.ctht};
    double x = a.x[i], y = a.y[i], z = a.z[i], dx = a.dx[i], dy = a.dy[i];
    const bool ok = track_from_tgt(M, x, y, z, dx, dy, a.mom[i], a.mass);
    a.x[i] = x;
    a.y[i] = y;
    a.z[i] = z;
    a.dx[i] = dx;
    a.dy[i] = dy;
    if (a.ok) a.ok[i] = ok ? 1 : 0;
}

// track_to_tgt (trg_track.f:738-876) in rounds.  The reference's loop
// alternates plane searches (RK4, a few hundred steps) with the arm's
// reconstruction, and events leave it after 1 to 10 iterations.  One lane
// per event through the whole loop would run every wave as long as its
// slowest lane (3.3 times the mean path on HMS events), so the loop is cut
// at the reconstruction: a start kernel does the first search and the
// iteration's searches up to the reconstruction; then, per round, a
// reconstruction kernel and a search kernel run over the list of events
// still iterating, the search kernel appending the events that continue to
// the next round's list.  Per event the operations and their order are the
// reference's, so the rounds change nothing in the results.
struct FieldToArgs {
    const double* map;
    double stht, ctht;
    uint32_t n;
    const double* coef;  // the arm's packed coefficients (reconstruction)
    const uint32_t* gen;  // or its run-time polynomial tables
    const double *x_fp, *dx_fp, *y_fp, *dy_fp;
    double *delta, *y, *dx, *dy;
    const double *frx, *fry, *mom;
    double mass, ctheta, stheta;
    uint8_t* ok;         // INOUT (nullptr: all true, not written)
    // per-event loop state (workspace)
    double *xx, *mom0, *delx;
    uint8_t* st_ok;
    uint8_t* iter;
    const uint32_t* list_in;   // events of this round
    const uint32_t* count_in;
    uint32_t* list_out;        // events of the next round
    uint32_t* count_out;
};

__device__ __forceinline__ void field_energy(double mom, double mass, double& vel, double& eng)
{
    const double cc = 29.9792458;
    vel = fabs(mom) / sqrt(mom * mom + mass * mass) * cc;
    eng = copysign(1., mom) * sqrt(mom * mom + mass * mass);
}

// The loop test and, while it holds, the two searches to z = 0 and the new
// vertical position (trg_track.f:803-821); otherwise the final values
// (:864-873).  vT is the track at the beam plane.
__device__ __forceinline__ void field_to_iterate(const TrgMap& M, const FieldToArgs& a, uint32_t i, double (&vT)[6],
                                                 double eng, double xx, double delx, int n, bool ok)
{
    const double fry = a.fry[i];
    if ((delx > .0001) && (n < 10) && ok) {
        delx = fabs(-fry - vT[0]);
        double vTx[6];
        vTx[0] = -fry;
#pragma unroll
        for (int k = 1; k < 6; ++k) vTx[k] = vT[k];
        trg_track_to_plane(M, vT, eng, 1., 0., 0., 1., 0., ok);
        trg_track_to_plane(M, vTx, eng, 1., 0., 0., 1., 0., ok);
        xx = xx + fmin(1., fmax(-1., (vTx[0] - vT[0])));
        a.xx[i] = xx;
        a.delx[i] = delx;
        a.iter[i] = (uint8_t)n;
        a.st_ok[i] = ok ? 1 : 0;
        // append to the next round's list: one atomic per wave (the lanes
        // active here), each lane at its rank among them
        const uint64_t act = __ballot(1);
        const int lane = __lane_id();
        const int leader = __ffsll((unsigned long long)act) - 1;
        const uint32_t rank = (uint32_t)__popcll(act & ((1ull << lane) - 1ull));
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(a.count_out, (uint32_t)__popcll(act));
        base = __shfl(base, leader);
        a.list_out[base + rank] = i;
        return;
    }
    if (delx > .2) ok = false;
    a.dy[i] = vT[4] / vT[5];
    a.dx[i] = vT[3] / vT[5];
    a.y[i] = vT[1];
    if (a.ok) a.ok[i] = ok ? 1 : 0;
}

// Round 0: the first search from z = 100 cm to the beam plane (:768-787).
__global__ __launch_bounds__(256) void field_to_start(FieldToArgs a)
{
    __shared__ double s[2 * kTrgNodes];
    stage_field_map(s, a.map);
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    const TrgMap M{s, s + kTrgNodes, a.stht, a.ctht};
    const double delta = a.delta[i], y = a.y[i], dx = a.dx[i], dy = a.dy[i];
    const double fry = a.fry[i];
    const double mom = a.mom[i];
    bool ok = a.ok ? a.ok[i] != 0 : true;
    double vel, eng;
    field_energy(mom, a.mass, vel, eng);
    a.mom0[i] = mom / (1.e0 + delta / 100.e0);
    double vT[6];
    trg_start(vT, -fry + 100. * dx, y + 100. * dy, 100., dx, dy, vel, false);
    trg_track_to_plane(M, vT, eng, 1., 0., -a.ctheta, a.stheta, a.frx[i], ok);
    field_to_iterate(M, a, i, vT, eng, -fry, 1., 0, ok);
}

// A round's reconstruction, mc_*_recon(delta, dy, dx, y, xx) on the event's
// focal-plane track (:836-846).
template <int ARM>
__global__ __launch_bounds__(256) void field_to_recon(FieldToArgs a)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= *a.count_in) return;
    const uint32_t i = a.list_in[j];
    EvOut o, r;
    o.x_fp = a.x_fp[i];
    o.dx_fp = a.dx_fp[i];
    o.y_fp = a.y_fp[i];
    o.dy_fp = a.dy_fp[i];
    recon<ARM>((cdptr)a.coef, (cuptr)a.gen, o, a.xx[i], r);
    a.delta[i] = r.dpp;
    a.dy[i] = r.dydz;
    a.dx[i] = r.dxdz;
    a.y[i] = r.y;
}

// A round's search back to the beam plane from the reconstructed track
// (:848-865), then the loop test and the next searches.
__global__ __launch_bounds__(256) void field_to_search(FieldToArgs a)
{
    __shared__ double s[2 * kTrgNodes];
    const uint32_t count = *a.count_in;
    if (blockIdx.x * 256u >= count) return;  // the whole workgroup is past the list
    stage_field_map(s, a.map);
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= count) return;
    const uint32_t i = a.list_in[j];
    const TrgMap M{s, s + kTrgNodes, a.stht, a.ctht};
    const double delta = a.delta[i], y = a.y[i], dx = a.dx[i], dy = a.dy[i], xx = a.xx[i];
    bool ok = a.st_ok[i] != 0;
    const double mom = a.mom0[i] * (1.e0 + delta / 100.e0);
    double vel, eng;
    field_energy(mom, a.mass, vel, eng);
    double vT[6];
    trg_start(vT, xx + 100. * dx, y + 100. * dy, 100., dx, dy, vel, false);
    trg_track_to_plane(M, vT, eng, 1., 0., -a.ctheta, a.stheta, a.frx[i], ok);
    field_to_iterate(M, a, i, vT, eng, xx, a.delx[i], (int)a.iter[i] + 1, ok);
}

__global__ void zero_u32(uint32_t* p) { p[0] = 0u; }

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail(SIMC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

namespace simc {
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace simc

constexpr int kNumArms = 5;  // HMS, SHMS, SOS, HRS-R, HRS-L (arm_index)

// The two survivor queues of one arm and their counters (device memory,
// grown on demand; the owner frees them).
struct QueuePair {
    void* d_queue[2] = {};
    uint32_t* d_qcount = nullptr;   // the two queue counters
    uint32_t cap = 0;
};

struct simc_ctx {
    simc::ForwardMap fwd[kNumArms];
    simc::ReconMap rec[kNumArms];
    double* d_coef[kNumArms] = {};   // null: the arm's maps are not loaded
    // run-time polynomial tables of an arm whose maps differ in structure from
    // the compiled polynomials (null: the compiled ones serve)
    uint32_t* d_gen[kNumArms] = {};
    simc::GenericPoly gen[kNumArms];
    Op* d_ops[kNumArms] = {};
    ArmTables* d_tab[kNumArms] = {};
    std::vector<Op> ops[kNumArms];
    int n_ops[kNumArms] = {};
    double zdc[kNumArms][12];
    unsigned long long* d_hist[kNumArms] = {};
    int lead_pc[kNumArms] = {};      // first step after the leading sieve steps
    int coll_pc[kNumArms] = {};      // first step after the collimator
    int hut_pc[kNumArms] = {};       // first step of the detector hut
    // survivor queues between the stages of plain runs (per arm)
    QueuePair qp[kNumArms];
    std::mutex mu;                   // guards queue growth
    // reconstruction exponents per arm (adjoint form), [n_terms][5]
    uint8_t* d_rexp[kNumArms] = {};
    // adjoint scratch: selected-event list and its counter (grown on demand)
    uint32_t* d_sel = nullptr;
    uint32_t* d_nsel = nullptr;
    double* d_hw = nullptr;          // 22 doubles per selected event (pass-1 and pass-2 outputs)
    int64_t sel_cap = 0;
    // staging buffers of the host-pointer entry point (grown on demand)
    std::mutex host_mu;
    void* d_stage = nullptr;
    size_t stage_bytes = 0;
    // pipeline of the host-array arm entry point: two slots, each a pinned
    // host buffer, a device buffer, a stream, its survivor queues and an event
    struct HostPipe {
        hipStream_t st[2] = {};
        hipEvent_t done[2] = {};
        void* pinned[2] = {};
        void* dev[2] = {};
        size_t bytes = 0;  // per slot
        QueuePair qp[2];
    } hp;
    int device = 0;
    int n_cu = 256;
};

// A prepared multi-setting run: settings grouped into passes of one arm and
// one kernel variant; each pass is one event range split into launch chunks.
struct simc_batch {
    struct Pass {
        int arm;
        bool rare;
        int32_t flags;      // KF_* bits shared by the pass's settings
        uint64_t seed;      // Philox key of the pass's first setting
        bool lane_key;      // its settings' keys differ: each lane reads its setting's key
        uint64_t total;     // events of the pass
        uint32_t first, n;  // its settings in d_set
    };
    simc_ctx* ctx = nullptr;
    // the batch's own survivor queues (per arm), sized for its launch tile:
    // a captured run keeps addressing them while the context's plain runs
    // grow or replace theirs
    QueuePair qp[kNumArms];
    int n_settings = 0;
    std::vector<int> arm_of;               // arm of each setting (input order)
    std::vector<Pass> passes;
    void* d_set = nullptr;                 // SetDev table, pass by pass
    unsigned long long* d_hist = nullptr;  // one histogram per setting (its input order)
    uint32_t n_arr = 0;                    // event-array length the settings address
    int64_t chunk = (int64_t)1 << 24;
};

namespace {

int arm_index(int arm)
{
    return arm == SIMC_ARM_HMS ? 0 : arm == SIMC_ARM_SHMS ? 1 : arm == SIMC_ARM_SOS ? 2
         : arm == SIMC_ARM_HRSR ? 3 : arm == SIMC_ARM_HRSL ? 4 : -1;
}

bool is_hrs(int arm) { return arm == SIMC_ARM_HRSR || arm == SIMC_ARM_HRSL; }

// STOP counters of an arm's common block
int ncounters_of(int arm)
{
    return arm == SIMC_ARM_HMS ? SIMC_HMS_NCOUNTERS : arm == SIMC_ARM_SOS ? SIMC_SOS_NCOUNTERS
         : is_hrs(arm)          ? SIMC_HRS_NCOUNTERS : SIMC_SHMS_NCOUNTERS;
}


ArmTables make_tables(const simc_ctx* ctx, int arm)
{
    ArmTables T;
    std::memset(&T, 0, sizeof T);
    if (arm == SIMC_ARM_HMS)
        T.coll = CollGeom{4.575, 11.646, 4.759, 12.114, 0.000, +0.028, 0.41753, 6.30 / 20};  // hms/mc_hms_coll.f:22-53
    else
        T.coll = CollGeom{8.50, 12.50, 8.65, 12.85, 0.000, +0.000, 0.42084, 6.35 / 20};      // shms/mc_shms_coll.f:23-54
    // VDC offsets: hrsr/mc_hrsr.f:524, hrsl/mc_hrsl.f:525
    T.yfp_shift = arm == SIMC_ARM_HRSR ? 0.48 : arm == SIMC_ARM_HRSL ? 0.78 : 0.0;
    std::memcpy(T.zdc, ctx->zdc[arm_index(arm)], sizeof T.zdc);
    // LFIT (cern/lfit.f:22-59) over x planes 1,3,4,6,7,9,10,12 and y planes
    // 2,5,8,11: A = sum (z - zm)(h - hm) / sum (z - zm)^2, B = hm - A zm.
    const int xpl[8] = {0, 2, 3, 5, 6, 8, 9, 11}, ypl[4] = {1, 4, 7, 10};
    auto weights = [&](const int* pl, int np) {
        double zm = 0.0, szz = 0.0;
        for (int j = 0; j < np; ++j) zm += T.zdc[pl[j]];
        zm /= np;
        for (int j = 0; j < np; ++j) szz += (T.zdc[pl[j]] - zm) * (T.zdc[pl[j]] - zm);
        for (int j = 0; j < np; ++j) {
            T.fw_a[pl[j]] = (T.zdc[pl[j]] - zm) / szz;
            T.fw_b[pl[j]] = 1.0 / np - zm * T.fw_a[pl[j]];
        }
    };
    weights(xpl, 8);
    weights(ypl, 4);
    return T;
}

// Algorithmic FP64 operations of one pass through a class's polynomial
// (cls = 0: recon) as the arm's maps are evaluated.
int64_t class_flops(const simc_ctx* ctx, int arm, int cls)
{
    const int ai = arm_index(arm);
    if (ctx->d_gen[ai]) return cls >= 0 && cls <= 32 ? ctx->gen[ai].flops[cls] : 0;
    return simc::poly_flops(arm, cls);
}

// The hut stage runs the compiled hut of the arm's kernel family
// (hut_gen.inc): the loaded program's hut steps must be that sequence, with
// the same literal parameters.
int check_compiled_hut(int arm, const simcprog::Program& P, int hut)
{
    const int family = is_hrs(arm) ? SIMC_ARM_HRSR : arm;
    const HutGenTable* g = nullptr;
    for (const HutGenTable& x : kHutGen)
        if (x.kernel_arm == family) g = &x;
    if (!g || g->first != hut || g->first + g->n != (int)P.ops.size())
        return fail(SIMC_E_INTERNAL, "the arm's program does not match its compiled detector hut");
    for (int k = 0; k < g->n; ++k) {
        const Op& o = P.ops[g->first + k];
        bool same = o.code == g->seq[k].code && o.arg == g->seq[k].arg && o.stop == g->seq[k].stop;
        for (int j = 0; j < 8; ++j)
            if (g->lit_mask[k] & (1u << j)) same = same && std::memcmp(&o.p[j], &g->lit[k][j], sizeof(double)) == 0;
        if (!same)
            return fail(SIMC_E_INTERNAL, "step " + std::to_string(g->first + k) +
                                             " of the arm's program differs from its compiled detector hut");
    }
    return SIMC_OK;
}

int load_arm(simc_ctx* ctx, int arm, const std::string& fwd_path, const std::string& rec_path)
{
    const int ai = arm_index(arm);
    std::string msg;
    int rc = simc::read_forward_map(fwd_path, ctx->fwd[ai], msg);
    if (rc) return fail(rc, msg);
    // hms/mc_hms.f:188, shms/mc_shms.f:344, sos/mc_sos.f:164, hrsl/mc_hrsl.f:151
    const int want = arm == SIMC_ARM_HMS || is_hrs(arm) ? 12 : arm == SIMC_ARM_SOS ? 10 : 32;
    if (ctx->fwd[ai].n_classes != want)
        return fail(SIMC_E_NCLASS, fwd_path + ": " + std::to_string(ctx->fwd[ai].n_classes) +
                                       " transport classes, expected " + std::to_string(want));
    rc = simc::read_recon_map(rec_path, ctx->rec[ai], msg);
    if (rc) return fail(rc, msg);
    std::vector<double> packed;
    rc = simc::pack_arm(arm, ctx->fwd[ai], ctx->rec[ai], packed, msg);
    std::vector<uint8_t> gen_bytes;
    if (rc == SIMC_E_MAPSTRUCT) {
        // another tune or fit: evaluate its polynomials from run-time tables
        rc = simc::build_generic(ctx->fwd[ai], ctx->rec[ai], ctx->gen[ai], msg);
        if (rc) return fail(rc, msg);
        ctx->gen[ai].pack(gen_bytes);
        packed.assign(1, 0.0);
    } else if (rc) {
        return fail(rc, msg);
    }
    simcprog::Program P;
    simcprog::build_program(arm, ctx->fwd[ai], P, ctx->zdc[ai]);
    ctx->n_ops[ai] = (int)P.ops.size();
    ctx->ops[ai] = P.ops;
    // Split point of the two stages; every step must belong to its stage.
    const int hut = simcprog::hut_begin(P);
    for (int k = 0; k < (int)P.ops.size(); ++k) {
        const int want = k < hut ? STAGE_MAGNETS : STAGE_HUT;
        if (!(op_stages(P.ops[k].code) & want))
            return fail(SIMC_E_INTERNAL, "step " + std::to_string(k) + " outside its stage");
    }
    ctx->hut_pc[ai] = hut;
    int lead = 0;
    while (lead < hut && P.ops[lead].code == OP_SIEVE) ++lead;
    ctx->lead_pc[ai] = lead;
    rc = check_compiled_hut(arm, P, hut);
    if (rc) return rc;
    // Stage 1 ends after the collimator; an arm without one (the HRS: slit
    // apertures only) ends it before its first COSY segment.
    int coll = 0;
    while (coll < hut && P.ops[coll].code != OP_COLLIMATOR) ++coll;
    if (coll < hut) {
        ctx->coll_pc[ai] = coll + 1;
    } else {
        int first = 0;
        while (first < hut && P.ops[first].code != OP_TRANSP) ++first;
        if (first == 0 || first == hut) return fail(SIMC_E_INTERNAL, "no stage split before the hut");
        ctx->coll_pc[ai] = first;
    }
    HIPCHK(hipMalloc(&ctx->d_coef[ai], packed.size() * sizeof(double)));
    HIPCHK(hipMemcpy(ctx->d_coef[ai], packed.data(), packed.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!gen_bytes.empty()) {
        HIPCHK(hipMalloc(&ctx->d_gen[ai], gen_bytes.size()));
        HIPCHK(hipMemcpy(ctx->d_gen[ai], gen_bytes.data(), gen_bytes.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&ctx->d_ops[ai], P.ops.size() * sizeof(Op)));
    HIPCHK(hipMemcpy(ctx->d_ops[ai], P.ops.data(), P.ops.size() * sizeof(Op), hipMemcpyHostToDevice));
    const ArmTables T = make_tables(ctx, arm);
    HIPCHK(hipMalloc(&ctx->d_tab[ai], sizeof(ArmTables)));
    HIPCHK(hipMemcpy(ctx->d_tab[ai], &T, sizeof(ArmTables), hipMemcpyHostToDevice));
    {
        const auto& terms = ctx->rec[ai].terms;
        std::vector<uint8_t> ex(terms.size() * 5);
        for (size_t t = 0; t < terms.size(); ++t)
            for (int v = 0; v < 5; ++v) ex[t * 5 + v] = terms[t].e[v];
        HIPCHK(hipMalloc(&ctx->d_rexp[ai], ex.size()));
        HIPCHK(hipMemcpy(ctx->d_rexp[ai], ex.data(), ex.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&ctx->d_hist[ai], kHistSlots * sizeof(unsigned long long)));
    HIPCHK(hipMemset(ctx->d_hist[ai], 0, kHistSlots * sizeof(unsigned long long)));
    return SIMC_OK;
}

// The two survivor queues of arm index ai, room for `cap` entries each.
constexpr size_t kQueueEntryBytes = 10 * 8 + 3 * 4;

void queue_view(void* base, uint32_t c, uint32_t* count, Queue& q)
{
    double* d = (double*)base;
    double** dp[10] = {&q.pend, &q.xs, &q.ys, &q.dxdzs, &q.dydzs, &q.dpps, &q.p, &q.m2, &q.pathlen,
                       &q.decdist};
    for (int k = 0; k < 10; ++k) *dp[k] = d + (size_t)k * c;
    uint32_t* u = (uint32_t*)(d + (size_t)10 * c);
    q.idx = u;
    q.rng = u + c;
    q.bits = u + 2 * (size_t)c;
    q.count = count;
    q.cap = c;
}

// Queue views of a pair that holds at least `cap` entries per queue (grown
// here; growing waits for the device, whose earlier launches may use them).
int ensure_pair(QueuePair& qp, uint32_t cap, Queue& q1, Queue& q2)
{
    if (!qp.d_qcount) HIPCHK(hipMalloc(&qp.d_qcount, 256));
    if (qp.cap < cap) {
        if (qp.d_queue[0] || qp.d_queue[1]) {
            HIPCHK(hipDeviceSynchronize());
            for (int j = 0; j < 2; ++j) {
                if (qp.d_queue[j]) HIPCHK(hipFree(qp.d_queue[j]));
                qp.d_queue[j] = nullptr;
            }
            qp.cap = 0;
        }
        for (int j = 0; j < 2; ++j) HIPCHK(hipMalloc(&qp.d_queue[j], (size_t)cap * kQueueEntryBytes));
        qp.cap = cap;
    }
    queue_view(qp.d_queue[0], qp.cap, qp.d_qcount, q1);
    queue_view(qp.d_queue[1], qp.cap, qp.d_qcount + 1, q2);
    return SIMC_OK;
}

void free_pair(QueuePair& qp)
{
    for (int j = 0; j < 2; ++j)
        if (qp.d_queue[j]) (void)hipFree(qp.d_queue[j]);
    if (qp.d_qcount) (void)hipFree(qp.d_qcount);
    qp = QueuePair();
}

int ensure_queues(simc_ctx* ctx, int ai, uint32_t cap, Queue& q1, Queue& q2)
{
    std::lock_guard<std::mutex> lock(ctx->mu);
    return ensure_pair(ctx->qp[ai], cap, q1, q2);
}

}  // namespace

// Scratch of the allocation-free entry point (simc_arm_run_noalloc): the two
// survivor queues of one launch of up to SIMC_NOALLOC_MAX_EVENTS events, after
// 256 bytes that hold their two counters.  Each device holds its own copy for
// the life of the process; every call on that device shares it.
constexpr size_t kScratchQueueBytes = (size_t)SIMC_NOALLOC_MAX_EVENTS * (10 * 8 + 3 * 4);
static_assert(kScratchQueueBytes % 256 == 0, "queue arrays stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[256 + 2 * kScratchQueueBytes];

namespace {

static_assert(kScratchQueueBytes == (size_t)SIMC_NOALLOC_MAX_EVENTS * kQueueEntryBytes, "queue entry layout");

// Device address of `scratch` per device, taken once with the device current
// (the array's name in host code is a stand-in, not a device address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(SIMC_E_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return SIMC_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SIMC_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
    return SIMC_OK;
}

// Queue views of the scratch array of `device` (scratch_init has run).
int scratch_queues(int device, Queue& q1, Queue& q2)
{
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (device >= 0 && device < kMaxDevices) p = g_scratch_ptr[device];
    }
    if (!p) return fail(SIMC_E_INTERNAL, "scratch address not set up for this device");
    char* base = (char*)p;
    uint32_t* counts = (uint32_t*)base;
    queue_view(base + 256, SIMC_NOALLOC_MAX_EVENTS, counts, q1);
    queue_view(base + 256 + kScratchQueueBytes, SIMC_NOALLOC_MAX_EVENTS, counts + 1, q2);
    return SIMC_OK;
}

}  // namespace

namespace {

// Stop-code histogram (kHistSlots bins) -> the STOP counters of the arm's
// common block, in its order.
void hist_to_counters(int arm, const unsigned long long* h, int64_t* c, int ncounters)
{
    for (int i = 0; i < ncounters; ++i) c[i] = 0;
    int64_t total = 0;
    for (int i = 0; i < kHistBins; ++i) total += (int64_t)h[i];
    auto H = [&](int code) { return (int64_t)h[code]; };
    if (arm == SIMC_ARM_HMS) {
        // order of common /spectrometer1/, hms/struct_hms.inc:34-42
        c[0] = total;
        c[1] = H(ST_SLIT_HOR) + (int64_t)h[kHistBins + 0];
        c[2] = H(ST_SLIT_VERT) + (int64_t)h[kHistBins + 1];
        c[3] = H(ST_SLIT_OCT) + (int64_t)h[kHistBins + 2];
        for (int k = 0; k < 9; ++k) c[4 + k] = H(ST_H_Q1_IN + k);
        c[13] = H(ST_H_D1_IN);
        c[14] = H(ST_H_D1_OUT) + H(ST_H_PIPE0) + H(ST_H_PIPE1) + H(ST_H_PIPE2) + H(ST_H_PIPE3);
        c[16] = H(ST_H_DC1);
        c[17] = H(ST_H_DC2);
        c[18] = H(ST_H_SCIN);
        c[19] = 0;  // hSTOP_cal: the calorimeter cut is commented out, hms/mc_hms_hut.f:584-588
        c[20] = H(ST_OK);
        c[15] = c[16] + c[17] + c[18] + c[20];
        c[21] = H(ST_COLL);
    } else if (arm == SIMC_ARM_SOS) {
        // order of common /spectrometer2/, sos/struct_sos.inc:9-17
        c[0] = total;
        c[1] = H(ST_SLIT_HOR);
        c[2] = H(ST_SLIT_VERT);
        c[3] = H(ST_SLIT_OCT);
        for (int k = 0; k < 7; ++k) c[4 + k] = H(ST_SO_QUAD_IN + k);
        c[11] = H(ST_SO_EXIT_FLANGE) + H(ST_SO_EXIT_APERTURE) + H(ST_SO_EXIT_WINDOW);
        c[13] = H(ST_SO_DC1);
        c[14] = H(ST_SO_DC2);
        c[15] = H(ST_SO_SCIN);
        c[16] = 0;  // sSTOP_cal: the calorimeter cut is commented out, sos/mc_sos_hut.f:519-538
        c[17] = H(ST_OK);
        c[12] = c[13] + c[14] + c[15] + c[17];
    } else if (is_hrs(arm)) {
        // order of common /spectrometer3/ and /spectrometer4/, hrsr/struct_hrsr.inc:24-40
        c[0] = total;
        c[1] = H(ST_R_PIPE_65) + H(ST_R_PIPE_80) + H(ST_SLIT_HOR);
        c[2] = H(ST_SLIT_VERT);
        c[3] = H(ST_R_Q1_PRE) + H(ST_R_Q1_IN);
        c[4] = H(ST_R_Q1_MID);
        c[5] = H(ST_R_Q1_OUT) + H(ST_R_Q1_POST);
        c[6] = H(ST_R_Q2_PRE) + H(ST_R_Q2_IN);
        c[7] = H(ST_R_Q2_MID);
        c[8] = H(ST_R_Q2_OUT) + H(ST_R_Q2_POST1) + H(ST_R_Q2_POST2);
        c[9] = H(ST_R_Q3_PRE) + H(ST_R_Q3_IN);
        c[10] = H(ST_R_Q3_MID);
        c[11] = H(ST_R_Q3_OUT) + H(ST_R_Q3_POST) + H(ST_R_WINDOW);
        c[12] = H(ST_R_D1_PRE) + H(ST_R_D1_IN);
        c[13] = H(ST_R_D1_OUT) + H(ST_R_D1_POST_R) + H(ST_R_D1_POST_BOX);
        c[15] = H(ST_R_DC1);
        c[16] = H(ST_R_DC2);
        c[17] = H(ST_R_S1);
        c[18] = H(ST_R_S2);
        c[19] = 0;  // the calorimeter cuts are commented out, hrsl/mc_hrsl_hut.f:463-490
        c[20] = H(ST_OK);
        c[14] = c[15] + c[16] + c[17] + c[18] + c[20];
    } else {
        // order of common /spectrometer_shms/, shms/struct_shms.inc:26-44
        c[0] = total;
        c[1] = H(ST_OK);
        c[5] = H(ST_SLIT_HOR) + (int64_t)h[kHistBins + 0];
        c[6] = H(ST_SLIT_VERT) + (int64_t)h[kHistBins + 1];
        c[7] = H(ST_SLIT_OCT) + (int64_t)h[kHistBins + 2];
        for (int k = 0; k < 4; ++k) c[8 + k] = H(ST_S_HB_IN + k);
        for (int k = 0; k < 15; ++k) c[12 + k] = H(ST_S_Q1_IN + k);
        for (int k = 0; k < 12; ++k) c[27 + k] = H(ST_S_D1_IN + k);
        c[42] = H(ST_S_DC1);
        c[43] = H(ST_S_DC2);
        c[46] = H(ST_S_S1X) + H(ST_S_S1Y);
        c[47] = H(ST_S_S2Y);
        c[48] = H(ST_S_S2X);
        c[49] = H(ST_S_CAL) + H(ST_S_CALFID);
        c[41] = c[42] + c[43] + c[46] + c[47] + c[48] + c[49] + c[1];
        c[51] = H(ST_COLL);
    }
}

}  // namespace

extern "C" {

const char* simc_last_error(void) { return g_err.c_str(); }
const char* simc_version(void) { return "simc_gpu 0.1 (gfx950)"; }

int simc_ctx_create_files(const char* hms_forward, const char* hms_recon, const char* shms_forward,
                          const char* shms_recon, simc_ctx** out)
{
    if (!out || !hms_forward || !hms_recon || !shms_forward || !shms_recon)
        return fail(SIMC_E_ARG, "simc_ctx_create_files: NULL argument");
    *out = nullptr;
    simc_ctx* ctx = new (std::nothrow) simc_ctx();
    if (!ctx) return fail(SIMC_E_NOMEM, "out of host memory");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        delete ctx;
        return fail(SIMC_E_HIP, "no HIP device available");
    }
    if (hipGetDevice(&ctx->device) != hipSuccess ||
        hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) {
        delete ctx;
        return fail(SIMC_E_HIP, "cannot query the current HIP device");
    }
    int rc = scratch_init(ctx->device);
    if (!rc) rc = load_arm(ctx, SIMC_ARM_HMS, hms_forward, hms_recon);
    if (!rc) rc = load_arm(ctx, SIMC_ARM_SHMS, shms_forward, shms_recon);
    if (rc) {
        std::string keep = g_err;
        simc_ctx_destroy(ctx);
        g_err = keep;
        return rc;
    }
    *out = ctx;
    return SIMC_OK;
}

int simc_ctx_create(const char* map_dir, simc_ctx** out)
{
    if (!map_dir) return fail(SIMC_E_ARG, "simc_ctx_create: NULL map_dir");
    const std::string d(map_dir);
    int rc = simc_ctx_create_files((d + "/hms/forward_cosy.dat").c_str(), (d + "/hms/recon_cosy.dat").c_str(),
                                   (d + "/shms/shms_forward.dat").c_str(), (d + "/shms/shms_recon.dat").c_str(),
                                   out);
    if (rc) return rc;
    // the SOS and HRS maps are optional: loaded when present (sos/mc_sos.f:160-165,
    // hrsl/mc_hrsl.f:148-153; the reference reads them on an arm's first call)
    const struct { int arm; const char *fwd, *rec; } optional[] = {
        {SIMC_ARM_SOS, "/sos/forward_cosy.dat", "/sos/recon_cosy.dat"},
        {SIMC_ARM_HRSR, "/hrsr/hrs_forward_cosy.dat", "/hrsr/hrs_recon_cosy.dat"},
        {SIMC_ARM_HRSL, "/hrsl/hrs_forward_cosy.dat", "/hrsl/hrs_recon_cosy.dat"},
    };
    for (const auto& o : optional) {
        const std::string sf = d + o.fwd, sr = d + o.rec;
        FILE* f = std::fopen(sf.c_str(), "r");
        if (!f) continue;
        std::fclose(f);
        rc = simc_ctx_load_arm(*out, o.arm, sf.c_str(), sr.c_str());
        if (rc) {
            std::string keep = g_err;
            simc_ctx_destroy(*out);
            *out = nullptr;
            g_err = keep;
            return rc;
        }
    }
    return SIMC_OK;
}

int simc_ctx_load_arm(simc_ctx* ctx, int arm, const char* forward, const char* recon)
{
    if (!ctx || !forward || !recon) return fail(SIMC_E_ARG, "simc_ctx_load_arm: NULL argument");
    const int ai = arm_index(arm);
    if (ai < 0) return fail(SIMC_E_ARG, "arm must be SIMC_ARM_HMS (1), _SOS (2), _HRSR (3), _HRSL (4) or _SHMS (5)");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipDeviceSynchronize());  // runs of the arm may still use its tables
    void* old[6] = {ctx->d_coef[ai], ctx->d_ops[ai], ctx->d_tab[ai], ctx->d_rexp[ai], ctx->d_hist[ai],
                    ctx->d_gen[ai]};
    ctx->d_coef[ai] = nullptr; ctx->d_ops[ai] = nullptr; ctx->d_tab[ai] = nullptr;
    ctx->d_rexp[ai] = nullptr; ctx->d_hist[ai] = nullptr; ctx->d_gen[ai] = nullptr;
    for (void* p : old)
        if (p) (void)hipFree(p);
    return load_arm(ctx, arm, forward, recon);
}

void simc_ctx_destroy(simc_ctx* ctx)
{
    if (!ctx) return;
    for (int i = 0; i < kNumArms; ++i) {
        if (ctx->d_coef[i]) (void)hipFree(ctx->d_coef[i]);
        if (ctx->d_gen[i]) (void)hipFree(ctx->d_gen[i]);
        if (ctx->d_ops[i]) (void)hipFree(ctx->d_ops[i]);
        if (ctx->d_tab[i]) (void)hipFree(ctx->d_tab[i]);
        if (ctx->d_hist[i]) (void)hipFree(ctx->d_hist[i]);
        free_pair(ctx->qp[i]);
    }
    if (ctx->d_stage) (void)hipFree(ctx->d_stage);
    for (int j = 0; j < 2; ++j) {
        if (ctx->hp.pinned[j]) (void)hipHostFree(ctx->hp.pinned[j]);
        if (ctx->hp.dev[j]) (void)hipFree(ctx->hp.dev[j]);
        if (ctx->hp.done[j]) (void)hipEventDestroy(ctx->hp.done[j]);
        if (ctx->hp.st[j]) (void)hipStreamDestroy(ctx->hp.st[j]);
        free_pair(ctx->hp.qp[j]);
    }
    for (int i = 0; i < kNumArms; ++i)
        if (ctx->d_rexp[i]) (void)hipFree(ctx->d_rexp[i]);
    if (ctx->d_sel) (void)hipFree(ctx->d_sel);
    if (ctx->d_nsel) (void)hipFree(ctx->d_nsel);
    if (ctx->d_hw) (void)hipFree(ctx->d_hw);
    delete ctx;
}

}  // extern "C"

namespace {

int check_settings(int arm, const simc_settings* s)
{
    if (arm_index(arm) < 0)
        return fail(SIMC_E_ARG, "arm must be SIMC_ARM_HMS (1), _SOS (2), _HRSR (3), _HRSL (4) or _SHMS (5)");
    if (!(s->p_spec > 0.0) || !(s->m2 >= 0.0)) return fail(SIMC_E_ARG, "p_spec must be > 0 and m2 >= 0");
    if (s->decay_flag && !(s->ctau > 0.0)) return fail(SIMC_E_ARG, "decay_flag needs ctau > 0");
    if (s->sieve < 0 || s->sieve > 3 || ((s->sieve & 2) && arm != SIMC_ARM_SHMS))
        return fail(SIMC_E_ARG, "sieve must be 0 or 1, or for the SHMS 2 (in front of the bender) or 3 (both)");
    return SIMC_OK;
}

// Collimator stepping applies to m2 in (100^2, 200^2) (hms/mc_hms.f:206,
// shms/mc_shms.f:503; the SOS and the HRS have none); m2 changes only through
// a decay.
bool coll_stepping(int arm, const simc_settings* s)
{
    return (arm == SIMC_ARM_HMS || arm == SIMC_ARM_SHMS) && s->using_coll;
}
bool needs_rare(int arm, const simc_settings* s)
{
    return s->decay_flag || (coll_stepping(arm, s) && s->m2 > 100.0 * 100.0 && s->m2 < 200.0 * 200.0) ||
           s->sieve != 0;
}

SetDev make_setdev(int arm, const simc_settings* s, uint64_t seed, uint64_t event0, uint64_t start,
                   uint64_t offset, uint32_t hist)
{
    SetDev d;
    std::memset(&d, 0, sizeof d);
    d.p_spec = s->p_spec;
    d.m2 = s->m2;
    d.ctau = s->ctau;
    // hdc_sigma (hms/mc_hms_hut.f:143-144, hrsl/mc_hrsl_hut.f:113-114), sdc_sigma
    // (sos/mc_sos_hut.f:112-113), shms/hut.inc:125-126
    d.sigma = s->wire_sigma > 0.0 ? s->wire_sigma
            : arm == SIMC_ARM_SHMS ? 0.020 : is_hrs(arm) ? 0.0225 : 0.030;
    d.flags = (s->ms_flag ? KF_MS : 0) | (s->wcs_flag ? KF_WCS : 0) | (s->decay_flag ? KF_DECAY : 0) |
              (coll_stepping(arm, s) ? KF_COLL : 0) | (s->fit_mode == SIMC_FIT_REF_F32 ? KF_FIT_F32 : 0) |
              ((s->sieve & 1) ? KF_SIEVE : 0) | ((s->sieve & 2) ? KF_SIEVE2 : 0);
    d.start = start;
    d.seed = seed;
    d.event0 = event0;
    d.offset = offset;
    d.hist = hist;
    return d;
}

bool outputs_complete(const simc_events_in* in, const simc_events_out* out)
{
    return in->dpp && in->x && in->y && in->dxdz && in->dydz && out->dpp && out->y && out->dxdz &&
           out->dydz && out->x_fp && out->dx_fp && out->y_fp && out->dy_fp && out->pathlen && out->ok &&
           out->stop;
}

// One launch chunk: counter reset, the three stages and the reconstruction on
// `st`; `ev` (five events or nullptr) marks the kernel boundaries.
int launch_stages(simc_ctx* ctx, int arm, bool rare, bool tape, const KArgs& a, const Queue& q1, const Queue& q2,
                  hipStream_t st, hipEvent_t* ev, bool lane_key = false)
{
    const int ai = arm_index(arm);
    const int32_t coll = ctx->coll_pc[ai], hut = ctx->hut_pc[ai];
    // the leading sieve steps are no-ops in the common variants
    const int32_t first = rare ? 0 : ctx->lead_pc[ai];
    const dim3 G((a.count + kBlock - 1) / kBlock), B(kBlock);
    // the accepted list of the hut stage reuses queue 1 (consumed by then)
    // with the third counter
    Queue acc = q1;
    acc.count = q1.count + 2;
    hipLaunchKernelGGL(zero_queue_counts, dim3(1), dim3(64), 0, st, q1.count);
    if (ev) HIPCHK(hipEventRecord(ev[0], st));
#define LAUNCH3(ARMV, RV, RARE)                                                                             \
    do {                                                                                                    \
        hipLaunchKernelGGL((stage_kernel<ARMV, RV, RARE, STAGE_MAGNETS, false, true>), G, B, 0, st, a, q2, q1, \
                           first, coll);                                                                    \
        if (ev) HIPCHK(hipEventRecord(ev[1], st));                                                          \
        hipLaunchKernelGGL((stage_kernel<ARMV, RV, RARE, STAGE_MAGNETS, true, true>), G, B, 0, st, a, q1, q2,  \
                           coll, hut);                                                                      \
        if (ev) HIPCHK(hipEventRecord(ev[2], st));                                                          \
        hipLaunchKernelGGL((stage_kernel<ARMV, RV, RARE, STAGE_HUT, true, true>), G, B, 0, st, a, q2, acc,     \
                           hut, 0x7fffffff);                                                                \
        if (ev) HIPCHK(hipEventRecord(ev[3], st));                                                          \
        hipLaunchKernelGGL((recon_list_kernel<ARMV>), G, B, 0, st, a, acc);                                 \
        if (ev) HIPCHK(hipEventRecord(ev[4], st));                                                          \
    } while (0)
#define LAUNCH1(ARMV, RV) do { if (rare) LAUNCH3(ARMV, RV, true); else LAUNCH3(ARMV, RV, false); } while (0)
#define LAUNCHR(ARMV)                                                                                       \
    do {                                                                                                    \
        if (tape) LAUNCH1(ARMV, TapeStream);                                                                \
        else if (lane_key) LAUNCH1(ARMV, PhiloxLaneStream);                                                 \
        else LAUNCH1(ARMV, PhiloxStream);                                                                   \
    } while (0)
    if (arm == SIMC_ARM_HMS) LAUNCHR(SIMC_ARM_HMS);
    else if (arm == SIMC_ARM_SOS) LAUNCHR(SIMC_ARM_SOS);
    else if (is_hrs(arm)) LAUNCHR(SIMC_ARM_HRSR);  // one kernel family for both HRS arms (their tables differ)
    else LAUNCHR(SIMC_ARM_SHMS);
#undef LAUNCHR
#undef LAUNCH1
#undef LAUNCH3
    HIPCHK(hipGetLastError());
    return SIMC_OK;
}

// fixed: the queue pair to use (room for n events), or nullptr for the
// context's queues of the arm (grown here as needed).
int arm_run(simc_ctx* ctx, int arm, const simc_settings* s, const simc_events_in* in, simc_events_out* out,
            int64_t n, const simc_rng* rng, void* stream, void* const* stage_events, const Queue* fixed = nullptr)
{
    if (!ctx || !s || !in || !out || !rng) return fail(SIMC_E_ARG, "simc_arm_run: NULL argument");
    const int ai = arm_index(arm);
    int rc = check_settings(arm, s);
    if (rc) return rc;
    if (!ctx->d_coef[ai]) return fail(SIMC_E_ARG, "the arm's maps are not loaded (simc_ctx_load_arm)");
    if (n < 0) return fail(SIMC_E_ARG, "negative event count");
    if (n == 0) return SIMC_OK;
    if (!outputs_complete(in, out)) return fail(SIMC_E_ARG, "missing event input or output array");
    if (rng->kind != SIMC_RNG_PHILOX && rng->kind != SIMC_RNG_TAPE) return fail(SIMC_E_ARG, "unknown rng kind");
    if (rng->kind == SIMC_RNG_TAPE && (!rng->tape || rng->stride <= 0))
        return fail(SIMC_E_ARG, "tape rng needs a tape and a positive stride");
    KArgs a;
    std::memset(&a, 0, sizeof a);
    a.set0 = make_setdev(arm, s, rng->seed, rng->event0, 0, 0, 0);
    a.set = nullptr;
    a.n_set = 1;
    a.flags = a.set0.flags;
    a.seed = rng->seed;
    a.tab = ctx->d_tab[ai];
    a.coef = ctx->d_coef[ai];
    a.gen = ctx->d_gen[ai];
    a.ops = ctx->d_ops[ai];
    a.hist = ctx->d_hist[ai];
    const bool tape = rng->kind == SIMC_RNG_TAPE;
    const bool rare = needs_rare(arm, s);
    // Launches are cut so that every byte offset of a launch fits 32 bits
    // and each survivor queue stays at most 2^24 entries (1.7 GB).
    int64_t chunk = (int64_t)1 << 24;
    if (tape) {
        const int64_t sb = rng->stride * 8;
        if (sb > 0x7fffffffLL) return fail(SIMC_E_ARG, "tape stride too large");
        const int64_t c = 0xffffffffLL / sb;
        if (c < chunk) chunk = c;
        a.stride_bytes = (uint32_t)sb;
    }
    Queue q1, q2;
    if (fixed) {
        if ((int64_t)fixed[0].cap < std::min<int64_t>(n, chunk))
            return fail(SIMC_E_INTERNAL, "arm_run: fixed queues smaller than the launch");
        q1 = fixed[0];
        q2 = fixed[1];
    } else {
        rc = ensure_queues(ctx, ai, (uint32_t)(n < chunk ? n : chunk), q1, q2);
        if (rc) return rc;
    }
    hipStream_t st = (hipStream_t)stream;
    for (int64_t b = 0; b < n; b += chunk) {
        const int64_t cnt = n - b < chunk ? n - b : chunk;
        a.count = (uint32_t)cnt;
        a.pass_base = (uint64_t)b;
        a.ix_base = (uint64_t)b;
        a.n_arr = (uint32_t)cnt;
        a.dpp = in->dpp + b; a.x = in->x + b; a.y = in->y + b; a.dxdz = in->dxdz + b; a.dydz = in->dydz + b;
        a.fry = in->fry ? in->fry + b : nullptr;
        a.o_dpp = out->dpp + b; a.o_y = out->y + b; a.o_dxdz = out->dxdz + b; a.o_dydz = out->dydz + b;
        a.x_fp = out->x_fp + b; a.dx_fp = out->dx_fp + b; a.y_fp = out->y_fp + b; a.dy_fp = out->dy_fp + b;
        a.pathlen = out->pathlen + b;
        a.resmult = out->resmult ? out->resmult + b : nullptr;
        a.m2o = out->m2 ? out->m2 + b : nullptr;
        a.ok = out->ok + b;
        a.stop = out->stop + b;
        a.used = out->used ? out->used + b : nullptr;
        a.tape = tape ? rng->tape + b * rng->stride : nullptr;
        // stage_events: boundaries of the first chunk's four kernels
        hipEvent_t* ev = (stage_events && b == 0) ? (hipEvent_t*)stage_events : nullptr;
        rc = launch_stages(ctx, arm, rare, tape, a, q1, q2, st, ev);
        if (rc) return rc;
    }
    return SIMC_OK;
}

}  // namespace

extern "C" {

int simc_arm_run(simc_ctx* ctx, int arm, const simc_settings* s, const simc_events_in* in,
                 simc_events_out* out, int64_t n, const simc_rng* rng, void* stream)
{
    return arm_run(ctx, arm, s, in, out, n, rng, stream, nullptr);
}

int simc_arm_run_noalloc(simc_ctx* ctx, int arm, const simc_settings* s, const simc_events_in* in,
                         simc_events_out* out, int64_t n, const simc_rng* rng, void* stream)
{
    if (!ctx) return fail(SIMC_E_ARG, "simc_arm_run_noalloc: NULL context");
    if (n > (int64_t)SIMC_NOALLOC_MAX_EVENTS)
        return fail(SIMC_E_ARG, "simc_arm_run_noalloc: " + std::to_string(n) + " events exceed the scratch capacity of " +
                                    std::to_string(SIMC_NOALLOC_MAX_EVENTS) + " events per call");
    Queue q[2];
    int rc = scratch_queues(ctx->device, q[0], q[1]);
    if (rc) return rc;
    return arm_run(ctx, arm, s, in, out, n, rng, stream, nullptr, q);
}

int simc_arm_run_timed(simc_ctx* ctx, int arm, const simc_settings* s, const simc_events_in* in,
                       simc_events_out* out, int64_t n, const simc_rng* rng, void* stream,
                       void* const* stage_events)
{
    if (!stage_events)
        return fail(SIMC_E_ARG, "simc_arm_run_timed: five events required");
    for (int k = 0; k < 5; ++k)
        if (!stage_events[k]) return fail(SIMC_E_ARG, "simc_arm_run_timed: five events required");
    return arm_run(ctx, arm, s, in, out, n, rng, stream, stage_events);
}

int simc_arm_run_host(simc_ctx* ctx, int arm, const simc_settings* s, const simc_events_in* in,
                      simc_events_out* out, int64_t n, const simc_rng* rng)
{
    if (!ctx || !s || !in || !out || !rng) return fail(SIMC_E_ARG, "simc_arm_run_host: NULL argument");
    if (n < 0) return fail(SIMC_E_ARG, "negative event count");
    if (n > ((int64_t)1 << 31)) return fail(SIMC_E_ARG, "simc_arm_run_host: at most 2^31 events per call");
    if (n == 0) return SIMC_OK;
    if (rng->kind == SIMC_RNG_TAPE && (!rng->tape || rng->stride <= 0))
        return fail(SIMC_E_ARG, "tape rng needs a tape and a positive stride");
    const int ai = arm_index(arm);
    if (ai < 0 || !ctx->d_coef[ai]) return arm_run(ctx, arm, s, in, out, n, rng, nullptr, nullptr);  // its error
    std::lock_guard<std::mutex> lock(ctx->host_mu);
    HIPCHK(hipSetDevice(ctx->device));
    // Columns of a chunk, in the order of the slot's layout: read only (the
    // inputs), read and written (the fitted track, kept where the fit is not
    // reached, the accumulated path length, and resmult, kept by events that
    // stop before the hut), written only.  One H2D copy
    // covers the first two groups, one D2H copy the last two.
    struct Col { const void* src; void* dst; size_t esz; };
    std::vector<Col> cols;
    const double* hin[7] = {in->dpp, in->x, in->y, in->z, in->dxdz, in->dydz, in->fry};
    for (const double* p : hin)
        if (p) cols.push_back({p, nullptr, 8});
    const size_t n_in = cols.size();
    double* hio[6] = {out->x_fp, out->dx_fp, out->y_fp, out->dy_fp, out->pathlen, out->resmult};
    for (double* p : hio)
        if (p) cols.push_back({p, p, 8});
    const size_t n_io = cols.size() - n_in;
    double* hw[5] = {out->dpp, out->y, out->dxdz, out->dydz, out->m2};
    for (double* p : hw)
        if (p) cols.push_back({nullptr, p, 8});
    cols.push_back({nullptr, out->ok, 1});    // always computed; copied back when asked for
    cols.push_back({nullptr, out->stop, 1});
    if (out->used) cols.push_back({nullptr, out->used, 4});
    const bool tape = rng->kind == SIMC_RNG_TAPE;
    const size_t tape_ev = tape ? (size_t)rng->stride * 8 : 0;
    size_t per_ev = tape_ev;
    for (const Col& c : cols) per_ev += c.esz;
    // chunk: 2^20 events, fewer for long tapes (a slot stays near 256 MB)
    int64_t chunk = std::min<int64_t>(n, (int64_t)1 << 20);
    chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, (int64_t)((256u << 20) / per_ev)));
    auto rounded = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t slot = rounded(tape_ev * (size_t)chunk);
    for (const Col& c : cols) slot += rounded(c.esz * (size_t)chunk);
    auto& hp = ctx->hp;
    for (int j = 0; j < 2; ++j) {
        if (!hp.st[j]) HIPCHK(hipStreamCreateWithFlags(&hp.st[j], hipStreamNonBlocking));
        if (!hp.done[j]) HIPCHK(hipEventCreateWithFlags(&hp.done[j], hipEventDisableTiming));
    }
    if (hp.bytes < slot) {
        HIPCHK(hipDeviceSynchronize());
        for (int j = 0; j < 2; ++j) {
            if (hp.pinned[j]) HIPCHK(hipHostFree(hp.pinned[j]));
            if (hp.dev[j]) HIPCHK(hipFree(hp.dev[j]));
            hp.pinned[j] = hp.dev[j] = nullptr;
        }
        hp.bytes = 0;
        for (int j = 0; j < 2; ++j) {
            HIPCHK(hipHostMalloc(&hp.pinned[j], slot, hipHostMallocDefault));
            HIPCHK(hipMalloc(&hp.dev[j], slot));
        }
        hp.bytes = slot;
    }
    Queue q1[2], q2[2];
    for (int j = 0; j < 2; ++j) {
        int rc = ensure_pair(hp.qp[j], (uint32_t)chunk, q1[j], q2[j]);
        if (rc) return rc;
    }
    // Copies between the caller's pageable arrays and a pinned slot, split
    // over host threads (the pageable side is the bottleneck of this entry).
    const int nthr = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    auto par_copy = [&](std::vector<std::pair<void*, const void*>>& dst_src, std::vector<size_t>& len) {
        size_t total = 0;
        for (size_t b : len) total += b;
        if (total < ((size_t)4 << 20) || nthr == 1) {
            for (size_t k = 0; k < len.size(); ++k) std::memcpy(dst_src[k].first, dst_src[k].second, len[k]);
            return;
        }
        auto work = [&](int t) {
            for (size_t k = 0; k < len.size(); ++k) {
                const size_t lo = len[k] * t / nthr, hi = len[k] * (t + 1) / nthr;
                if (hi > lo) std::memcpy((char*)dst_src[k].first + lo, (const char*)dst_src[k].second + lo, hi - lo);
            }
        };
        std::vector<std::thread> th;
        for (int t = 1; t < nthr; ++t) th.emplace_back(work, t);
        work(0);
        for (auto& x : th) x.join();
    };
    struct Pending { int64_t c0 = -1, m = 0; };
    Pending pend[2];
    // offsets of the columns in a slot for chunk size m (tape last)
    auto layout = [&](int64_t m, std::vector<size_t>& off, size_t& tape_off, size_t& in_end, size_t& out_beg,
                      size_t& out_end) {
        off.resize(cols.size());
        size_t at = 0;
        for (size_t k = 0; k < cols.size(); ++k) {
            off[k] = at;
            at += rounded(cols[k].esz * (size_t)m);
            if (k + 1 == n_in + n_io) in_end = at;
        }
        if (n_in + n_io == 0) in_end = 0;
        out_beg = n_in < cols.size() ? off[n_in] : at;
        out_end = at;
        tape_off = at;
    };
    auto drain = [&](int j) -> int {
        if (pend[j].c0 < 0) return SIMC_OK;
        HIPCHK(hipEventSynchronize(hp.done[j]));
        std::vector<size_t> off;
        size_t tape_off, in_end, out_beg, out_end;
        layout(pend[j].m, off, tape_off, in_end, out_beg, out_end);
        std::vector<std::pair<void*, const void*>> ds;
        std::vector<size_t> len;
        for (size_t k = n_in; k < cols.size(); ++k)
            if (cols[k].dst) {
                ds.push_back({(char*)cols[k].dst + cols[k].esz * (size_t)pend[j].c0, (char*)hp.pinned[j] + off[k]});
                len.push_back(cols[k].esz * (size_t)pend[j].m);
            }
        par_copy(ds, len);
        pend[j].c0 = -1;
        return SIMC_OK;
    };
    int rc = SIMC_OK;
    int j = 0;
    for (int64_t c0 = 0; c0 < n && rc == SIMC_OK; c0 += chunk, j ^= 1) {
        const int64_t m = std::min<int64_t>(chunk, n - c0);
        rc = drain(j);
        if (rc) break;
        std::vector<size_t> off;
        size_t tape_off, in_end, out_beg, out_end;
        layout(m, off, tape_off, in_end, out_beg, out_end);
        std::vector<std::pair<void*, const void*>> ds;
        std::vector<size_t> len;
        for (size_t k = 0; k < n_in + n_io; ++k) {
            ds.push_back({(char*)hp.pinned[j] + off[k], (const char*)cols[k].src + cols[k].esz * (size_t)c0});
            len.push_back(cols[k].esz * (size_t)m);
        }
        if (tape) {
            ds.push_back({(char*)hp.pinned[j] + tape_off, rng->tape + (size_t)c0 * (size_t)rng->stride});
            len.push_back(tape_ev * (size_t)m);
        }
        par_copy(ds, len);
        char* d = (char*)hp.dev[j];
        hipStream_t st = hp.st[j];
        if (in_end) HIPCHK(hipMemcpyAsync(d, hp.pinned[j], in_end, hipMemcpyHostToDevice, st));
        if (tape)
            HIPCHK(hipMemcpyAsync(d + tape_off, (char*)hp.pinned[j] + tape_off, tape_ev * (size_t)m,
                                  hipMemcpyHostToDevice, st));
        auto dcol = [&](const void* host, bool is_out) -> void* {
            if (!host) return nullptr;
            for (size_t k = is_out ? n_in : 0; k < cols.size(); ++k)
                if ((is_out ? (const void*)cols[k].dst : cols[k].src) == host) return d + off[k];
            return nullptr;
        };
        simc_events_in ein;
        ein.dpp = (const double*)dcol(in->dpp, false);
        ein.x = (const double*)dcol(in->x, false);
        ein.y = (const double*)dcol(in->y, false);
        ein.z = (const double*)dcol(in->z, false);
        ein.dxdz = (const double*)dcol(in->dxdz, false);
        ein.dydz = (const double*)dcol(in->dydz, false);
        ein.fry = (const double*)dcol(in->fry, false);
        simc_events_out eout;
        eout.dpp = (double*)dcol(out->dpp, true);
        eout.y = (double*)dcol(out->y, true);
        eout.dxdz = (double*)dcol(out->dxdz, true);
        eout.dydz = (double*)dcol(out->dydz, true);
        eout.x_fp = (double*)dcol(out->x_fp, true);
        eout.dx_fp = (double*)dcol(out->dx_fp, true);
        eout.y_fp = (double*)dcol(out->y_fp, true);
        eout.dy_fp = (double*)dcol(out->dy_fp, true);
        eout.pathlen = (double*)dcol(out->pathlen, true);
        eout.resmult = (double*)dcol(out->resmult, true);
        eout.m2 = (double*)dcol(out->m2, true);
        // ok and stop are staged even when the caller does not ask for them
        const size_t k_ok = cols.size() - (out->used ? 3 : 2);
        eout.ok = (uint8_t*)(d + off[k_ok]);
        eout.stop = (uint8_t*)(d + off[k_ok + 1]);
        eout.used = out->used ? (int32_t*)(d + off[k_ok + 2]) : nullptr;
        simc_rng drng = *rng;
        if (tape) drng.tape = (const double*)(d + tape_off);
        else drng.event0 = rng->event0 + (uint64_t)c0;
        Queue fq[2] = {q1[j], q2[j]};  // the slot's own queues: two chunks are in flight at once
        rc = arm_run(ctx, arm, s, &ein, &eout, m, &drng, (void*)st, nullptr, fq);
        if (rc) break;
        HIPCHK(hipMemcpyAsync((char*)hp.pinned[j] + out_beg, d + out_beg, out_end - out_beg, hipMemcpyDeviceToHost,
                              st));
        HIPCHK(hipEventRecord(hp.done[j], st));
        pend[j].c0 = c0;
        pend[j].m = m;
    }
    for (int k = 0; k < 2; ++k) {
        const int r2 = drain(j ^ 1);
        j ^= 1;
        if (!rc) rc = r2;
    }
    if (rc) {
        (void)hipDeviceSynchronize();
        return rc;
    }
    return SIMC_OK;
}

int simc_recon_run_host(simc_ctx* ctx, int arm, int64_t n, const double* x_fp, const double* dx_fp,
                        const double* y_fp, const double* dy_fp, const double* fry, double* delta,
                        double* dxdz_tar, double* dydz_tar, double* y_tar)
{
    if (!ctx) return fail(SIMC_E_ARG, "simc_recon_run_host: NULL context");
    if (n < 0) return fail(SIMC_E_ARG, "negative event count");
    if (n == 0) return SIMC_OK;
    if (!x_fp || !dx_fp || !y_fp || !dy_fp || !delta || !dxdz_tar || !dydz_tar || !y_tar)
        return fail(SIMC_E_ARG, "simc_recon_run_host: NULL array");
    std::lock_guard<std::mutex> lock(ctx->host_mu);
    const size_t nd = (size_t)n * 8, need = 9 * ((nd + 255) & ~(size_t)255);
    if (ctx->stage_bytes < need) {
        if (ctx->d_stage) {
            HIPCHK(hipDeviceSynchronize());
            HIPCHK(hipFree(ctx->d_stage));
            ctx->d_stage = nullptr;
            ctx->stage_bytes = 0;
        }
        HIPCHK(hipMalloc(&ctx->d_stage, need));
        ctx->stage_bytes = need;
    }
    double* d[9];
    for (int k = 0; k < 9; ++k) d[k] = (double*)((char*)ctx->d_stage + k * ((nd + 255) & ~(size_t)255));
    const double* hin[5] = {x_fp, dx_fp, y_fp, dy_fp, fry};
    for (int k = 0; k < 5; ++k)
        if (hin[k]) HIPCHK(hipMemcpy(d[k], hin[k], nd, hipMemcpyHostToDevice));
    int rc = simc_recon_run(ctx, arm, n, d[0], d[1], d[2], d[3], fry ? d[4] : nullptr, d[5], d[6], d[7], d[8],
                            nullptr);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    double* hout[4] = {delta, dxdz_tar, dydz_tar, y_tar};
    for (int k = 0; k < 4; ++k) HIPCHK(hipMemcpy(hout[k], d[5 + k], nd, hipMemcpyDeviceToHost));
    return SIMC_OK;
}

int simc_hms_run(simc_ctx* ctx, const simc_settings* s, const simc_events_in* in, simc_events_out* out,
                 int64_t n, const simc_rng* rng, void* stream)
{
    return simc_arm_run(ctx, SIMC_ARM_HMS, s, in, out, n, rng, stream);
}

int simc_shms_run(simc_ctx* ctx, const simc_settings* s, const simc_events_in* in, simc_events_out* out,
                  int64_t n, const simc_rng* rng, void* stream)
{
    return simc_arm_run(ctx, SIMC_ARM_SHMS, s, in, out, n, rng, stream);
}

int simc_sos_run(simc_ctx* ctx, const simc_settings* s, const simc_events_in* in, simc_events_out* out,
                 int64_t n, const simc_rng* rng, void* stream)
{
    return simc_arm_run(ctx, SIMC_ARM_SOS, s, in, out, n, rng, stream);
}

int simc_hrsr_run(simc_ctx* ctx, const simc_settings* s, const simc_events_in* in, simc_events_out* out,
                  int64_t n, const simc_rng* rng, void* stream)
{
    return simc_arm_run(ctx, SIMC_ARM_HRSR, s, in, out, n, rng, stream);
}

int simc_hrsl_run(simc_ctx* ctx, const simc_settings* s, const simc_events_in* in, simc_events_out* out,
                  int64_t n, const simc_rng* rng, void* stream)
{
    return simc_arm_run(ctx, SIMC_ARM_HRSL, s, in, out, n, rng, stream);
}

int simc_recon_run(simc_ctx* ctx, int arm, int64_t n, const double* x_fp, const double* dx_fp,
                   const double* y_fp, const double* dy_fp, const double* fry, double* delta,
                   double* dxdz_tar, double* dydz_tar, double* y_tar, void* stream)
{
    if (!ctx) return fail(SIMC_E_ARG, "simc_recon_run: NULL context");
    const int ai = arm_index(arm);
    if (ai < 0) return fail(SIMC_E_ARG, "arm must be 1 (HMS), 2 (SOS), 3 (HRS-R), 4 (HRS-L) or 5 (SHMS)");
    if (!ctx->d_coef[ai]) return fail(SIMC_E_ARG, "the arm's maps are not loaded (simc_ctx_load_arm)");
    if (n < 0) return fail(SIMC_E_ARG, "negative event count");
    if (n == 0) return SIMC_OK;
    if (!x_fp || !dx_fp || !y_fp || !dy_fp || !delta || !dxdz_tar || !dydz_tar || !y_tar)
        return fail(SIMC_E_ARG, "simc_recon_run: NULL array");
    const int64_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffff) return fail(SIMC_E_ARG, "too many events for one launch");
    hipStream_t st = (hipStream_t)stream;
    if (arm == SIMC_ARM_HMS)
        hipLaunchKernelGGL(recon_kernel<SIMC_ARM_HMS>, dim3(blocks), dim3(256), 0, st, n, ctx->d_coef[ai], ctx->d_gen[ai], x_fp,
                           dx_fp, y_fp, dy_fp, fry, delta, dxdz_tar, dydz_tar, y_tar);
    else if (arm == SIMC_ARM_SOS)
        hipLaunchKernelGGL(recon_kernel<SIMC_ARM_SOS>, dim3(blocks), dim3(256), 0, st, n, ctx->d_coef[ai], ctx->d_gen[ai], x_fp,
                           dx_fp, y_fp, dy_fp, fry, delta, dxdz_tar, dydz_tar, y_tar);
    else if (is_hrs(arm))
        hipLaunchKernelGGL(recon_kernel<SIMC_ARM_HRSR>, dim3(blocks), dim3(256), 0, st, n, ctx->d_coef[ai], ctx->d_gen[ai], x_fp,
                           dx_fp, y_fp, dy_fp, fry, delta, dxdz_tar, dydz_tar, y_tar);
    else
        hipLaunchKernelGGL(recon_kernel<SIMC_ARM_SHMS>, dim3(blocks), dim3(256), 0, st, n, ctx->d_coef[ai], ctx->d_gen[ai], x_fp,
                           dx_fp, y_fp, dy_fp, fry, delta, dxdz_tar, dydz_tar, y_tar);
    HIPCHK(hipGetLastError());
    return SIMC_OK;
}

// ---------------------------------------------------------------------------
// Target step (host side)
// ---------------------------------------------------------------------------
namespace {

// Setting-constant terms of enerloss_new for one material (enerloss_new.f:23-31,51-55)
MaterialDev material(double dens, double zeff, double aeff)
{
    const double me = 0.51099906;
    const double I = zeff == 1.0 ? 21.8e-06 : (16.0 * std::pow(zeff, 0.9)) * 1.0e-06;
    const double hnup = 28.816e-06 * std::sqrt(dens * zeff / aeff);
    MaterialDev m;
    m.rho = dens;
    m.k_mp = 0.1536e-03 * zeff / aeff;
    m.l0 = std::log(me / (I * I)) + 1.063;
    m.k_log = 0.1536 * zeff / aeff;
    m.k_chi = 0.307075 / 2.0 * zeff / aeff;
    m.co = std::log(hnup) - std::log(I) + 0.5;
    m.co27 = std::fabs(m.co / 27.0);
    return m;
}

// Radiation length of the target nucleus, g/cm^2 (init.f:18-42)
double target_X0(double A, double Z)
{
    const double alpha = 1.0 / 137.0359895;  // constants.inc:44
    double L1 = std::log(184.15) - std::log(Z) / 3.0;
    double L2 = std::log(1194.) - 2.0 * std::log(Z) / 3.0;
    if (Z == 1.0) {
        L1 = 5.31;
        L2 = 6.144;
    }
    const double za2 = (Z * alpha) * (Z * alpha);
    const double fc = za2 * (1.202 + za2 * (-1.0369 + za2 * 1.008 / (za2 + 1)));
    const long nA = std::lround(A);
    if (nA == 1) return 61.28;
    if (nA == 2) return 122.4;
    if (nA == 4) return 94.32;
    return 716.405 * A / Z / (Z * (L1 - fc) + L2);
}

int make_prep(const simc_target* t, const simc_prep_settings* s, const simc_vertex_in* in, PrepDev& P)
{
    if (s->narm < 1 || s->narm > 3) return fail(SIMC_E_ARG, "narm must be 1 (beam), 2 (electron) or 3 (hadron)");
    if (!(t->rho > 0.0) || !(t->length > 0.0) || !(t->A > 0.0) || !(t->Z > 0.0))
        return fail(SIMC_E_ARG, "target needs positive A, Z, rho and length");
    if (!(s->mass > 0.0)) return fail(SIMC_E_ARG, "particle mass must be positive");
    if (!in->z) return fail(SIMC_E_ARG, "the vertex z array is required");
    if (s->narm == 1) {
        if (s->mc_smear && !in->E) return fail(SIMC_E_ARG, "beam multiple scattering needs E (= Ein)");
        if (!(s->beam_energy > s->mass)) return fail(SIMC_E_ARG, "beam_energy must exceed the mass");
    } else {
        if (!in->x || !in->y || !in->E || !in->p || !in->theta || !in->xptar || !in->yptar)
            return fail(SIMC_E_ARG, "vertex arrays x, y, z, E, p, theta, xptar, yptar are required");
        if (s->narm == 3 && !s->using_eloss && !in->delta)
            return fail(SIMC_E_ARG, "a hadron arm without using_eloss reads the vertex delta");
        if (!(s->p_spec > 0.0)) return fail(SIMC_E_ARG, "p_spec must be positive");
    }
    std::memset(&P, 0, sizeof P);
    P.mat[MAT_TARGET] = material(t->rho, t->Z, t->A);
    P.mat[MAT_AL] = material(2.70, 13., 26.98);       // target.inc:6-8
    P.mat[MAT_AIR] = material(0.00121, 7.2, 14.4);    // target.inc:24-26
    P.mat[MAT_KEVLAR] = material(0.74, 2.67, 4.67);   // target.inc:18-20
    P.mat[MAT_MYLAR] = material(1.39, 4.545, 8.735);  // target.inc:12-14
    const double X0 = t->X0 > 0.0 ? t->X0 : target_X0(t->A, t->Z);
    P.x0_cm[MAT_TARGET] = X0 / t->rho;                // init.f:43
    P.x0_cm[MAT_AL] = 24.01 / 2.70;
    P.x0_cm[MAT_AIR] = 36.66 / 0.00121;
    P.x0_cm[MAT_KEVLAR] = 55.2 / 0.74;
    P.x0_cm[MAT_MYLAR] = 39.95 / 1.39;
    P.length = t->length;
    P.angle = t->angle;
    P.cos_angle = std::fabs(std::cos(t->angle));
    P.zoffset = t->zoffset;
    P.can = t->can;
    P.liquid = t->Z < 2.4 ? 1 : 0;
    P.narm = s->narm;
    P.using_eloss = s->using_eloss != 0;
    P.mc_smear = s->mc_smear != 0;
    const double inch_cm = 2.54;
    if (s->narm == 1) {  // target.f:36-45
        if (P.liquid) {
            if (t->can == 1) P.s_al = 0.0028 * inch_cm;
            else if (t->can == 2) P.s_al = 0.0050 * inch_cm;
            else if (t->can == 3) P.s_al = 0.013;
        }
    } else {  // target.f:71-101, :184-217
        P.angle_sign = -1.0;
        switch (s->spectrometer) {
        case 1:
            P.s_al = 0.016 * inch_cm; P.s_air = 15; P.s_kevlar = 0.015 * inch_cm; P.s_mylar = 0.005 * inch_cm;
            P.angle_sign = 1.0;
            break;
        case 2:
            P.s_al = 0.008 * inch_cm; P.s_air = 15; P.s_kevlar = 0.005 * inch_cm; P.s_mylar = 0.003 * inch_cm;
            break;
        case 3:
        case 4:
            P.s_al = 0.013 * inch_cm; P.s_air = 15; P.s_kevlar = 0. * inch_cm; P.s_mylar = 0.010 * inch_cm;
            break;
        case 5:
        case 6:
            P.s_al = (0.02 + 0.01) * inch_cm; P.s_air = 57.27; P.s_kevlar = 0.0; P.s_mylar = 0.0;
            break;
        default:
            return fail(SIMC_E_ARG, "spectrometer must be 1 (HMS), 2 (SOS), 3/4 (HRS) or 5/6 (SHMS)");
        }
    }
    P.p_spec = s->p_spec;
    P.cos_th = std::cos(s->th_spec);
    P.sin_th = std::sin(s->th_spec);
    P.sin_phi = std::sin(s->phi);
    P.mass = s->mass;
    P.coulomb = s->coulomb;
    P.beam_energy = s->beam_energy;
    for (int k = 0; k < 5; ++k) P.off[k] = s->offset[k];
    P.ln10 = std::log(10.);
    return SIMC_OK;
}

void launch_prep(bool tape, int narm, const PrepArgs& a, hipStream_t st)
{
    using Ph = PhiloxStreamT<kTargetTag>;
    const dim3 G((a.n + 255) / 256), B(256);
    if (tape) {
        if (narm == 1) hipLaunchKernelGGL((target_prep_kernel<TapeStream, 1>), G, B, 0, st, a);
        else if (narm == 2) hipLaunchKernelGGL((target_prep_kernel<TapeStream, 2>), G, B, 0, st, a);
        else hipLaunchKernelGGL((target_prep_kernel<TapeStream, 3>), G, B, 0, st, a);
    } else {
        if (narm == 1) hipLaunchKernelGGL((target_prep_kernel<Ph, 1>), G, B, 0, st, a);
        else if (narm == 2) hipLaunchKernelGGL((target_prep_kernel<Ph, 2>), G, B, 0, st, a);
        else hipLaunchKernelGGL((target_prep_kernel<Ph, 3>), G, B, 0, st, a);
    }
}

// simc_hep_settings -> the device form (setting constants with the
// reference's libm calls)
void make_hep(const simc_hep_settings* h, HepDev& H)
{
    std::memset(&H, 0, sizeof H);
    H.Ebeam = h->Ebeam;
    H.dEbeam = h->dEbeam;
    H.xwid = h->xwid;
    H.ywid = h->ywid;
    H.xoffset = h->xoffset;
    H.yoffset = h->yoffset;
    H.fr1 = h->fr1;
    H.fr2 = h->fr2;
    H.coulomb = h->using_coulomb ? h->coulomb : 0.0;
    H.e_yp0 = h->e_yptar[0];
    H.e_ypw = h->e_yptar[1] - h->e_yptar[0];
    H.e_xp0 = h->e_xptar[0];
    H.e_xpw = h->e_xptar[1] - h->e_xptar[0];
    H.e_cos = std::cos(h->e_theta);
    H.e_sin = std::sin(h->e_theta);
    H.e_sinph = std::sin(h->e_phi);
    H.e_phi = h->e_phi;
    H.e_P = h->e_P;
    H.p_x0 = std::sin(h->p_theta) * std::cos(h->p_phi);
    H.p_y0 = std::sin(h->p_theta) * std::sin(h->p_phi);
    H.p_z0 = std::cos(h->p_theta);
    H.p_P = h->p_P;
    H.Mh = h->Mh;
    H.Mh2 = h->Mh * h->Mh;
    H.fr_pattern = h->fr_pattern;
    H.using_eloss = h->using_eloss != 0;
}

void hep_columns(const simc_hep_out* out, double* cols[kHepCols])
{
    double* const c[kHepCols] = {out->x, out->y, out->z, out->rasterx, out->rastery, out->Ein, out->eloss,
                                 out->teff, out->e_xptar, out->e_yptar, out->e_theta, out->e_phi, out->e_E,
                                 out->e_delta, out->p_xptar, out->p_yptar, out->p_theta, out->p_phi,
                                 out->p_P, out->p_E, out->p_delta, out->Q2, out->nu, out->q,
                                 out->jacobian, out->sigcc, out->weight};
    for (int k = 0; k < kHepCols; ++k) cols[k] = c[k];
}

// The window losses of a scattered particle for complete_ev's
// trip_thru_target(narm, ...) (target.f:51-246): the spectrometer's windows
// and the particle's mass.
int scattered_prep(const simc_target* t, int narm, int spectrometer, double mass, double p_spec, int using_eloss,
                   PrepDev& P)
{
    simc_prep_settings s;
    std::memset(&s, 0, sizeof s);
    s.narm = narm;
    s.spectrometer = spectrometer;
    s.using_eloss = using_eloss;
    s.mass = mass;
    s.p_spec = p_spec;
    static const double present = 0.0;  // make_prep checks that the vertex columns exist
    simc_vertex_in vin;
    std::memset(&vin, 0, sizeof vin);
    vin.x = vin.y = vin.z = vin.E = vin.p = vin.theta = vin.xptar = vin.yptar = vin.delta = &present;
    return make_prep(t, &s, &vin, P);
}

}  // namespace

int simc_target_prep(simc_ctx* ctx, const simc_target* t, const simc_prep_settings* s, const simc_vertex_in* in,
                     simc_prep_out* out, int64_t n, const simc_rng* rng, void* stream)
{
    if (!ctx || !t || !s || !in || !out || !rng) return fail(SIMC_E_ARG, "simc_target_prep: NULL argument");
    if (n < 0) return fail(SIMC_E_ARG, "negative event count");
    if (n == 0) return SIMC_OK;
    if (rng->kind != SIMC_RNG_PHILOX && rng->kind != SIMC_RNG_TAPE)
        return fail(SIMC_E_ARG, "unknown rng kind");
    if (rng->kind == SIMC_RNG_TAPE && (!rng->tape || rng->stride <= 0 || rng->stride > (1 << 24)))
        return fail(SIMC_E_ARG, "tape rng needs a tape and a stride in [1, 2^24]");
    PrepArgs a;
    std::memset(&a, 0, sizeof a);
    if (int rc = make_prep(t, s, in, a.P)) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    const uint64_t stride_bytes = rng->kind == SIMC_RNG_TAPE ? (uint64_t)rng->stride * 8u : 0u;
    // one launch covers at most 2^30 events, and a tape slot offset must fit 32 bits
    int64_t per = (int64_t)1 << 30;
    if (stride_bytes) per = std::min<int64_t>(per, (int64_t)(0xffffffffull / stride_bytes));
    a.seed = rng->seed;
    a.stride_bytes = (uint32_t)stride_bytes;
    for (int64_t at = 0; at < n; at += per) {
        const int64_t m = std::min(per, n - at);
        auto sh = [&](auto* ptr) { return ptr ? ptr + at : ptr; };
        a.n = (uint32_t)m;
        a.event0 = rng->event0 + (uint64_t)at;
        a.tape = rng->kind == SIMC_RNG_TAPE ? rng->tape + at * rng->stride : nullptr;
        a.x = sh(in->x); a.y = sh(in->y); a.z = sh(in->z); a.E = sh(in->E); a.p = sh(in->p);
        a.Ev = sh(in->E_vertex); a.theta = sh(in->theta); a.delta = sh(in->delta);
        a.xptar = sh(in->xptar); a.yptar = sh(in->yptar); a.dyp = sh(in->dang_yp); a.dxp = sh(in->dang_xp);
        a.dpp = sh(out->dpp); a.ox = sh(out->x); a.oy = sh(out->y); a.oz = sh(out->z);
        a.dxdz = sh(out->dxdz); a.dydz = sh(out->dydz); a.fry = sh(out->fry); a.eloss = sh(out->eloss);
        a.teff = sh(out->teff); a.d1 = sh(out->dang1); a.d2 = sh(out->dang2); a.used = sh(out->used);
        launch_prep(rng->kind == SIMC_RNG_TAPE, s->narm, a, st);
        HIPCHK(hipGetLastError());
    }
    return SIMC_OK;
}

int simc_target_prep_host(simc_ctx* ctx, const simc_target* t, const simc_prep_settings* s,
                          const simc_vertex_in* in, simc_prep_out* out, int64_t n, const simc_rng* rng)
{
    if (!ctx || !t || !s || !in || !out || !rng) return fail(SIMC_E_ARG, "simc_target_prep_host: NULL argument");
    if (n < 0) return fail(SIMC_E_ARG, "negative event count");
    if (n > ((int64_t)1 << 31)) return fail(SIMC_E_ARG, "simc_target_prep_host: at most 2^31 events per call");
    if (n == 0) return SIMC_OK;
    if (rng->kind == SIMC_RNG_TAPE && (!rng->tape || rng->stride <= 0 || rng->stride > (1 << 24)))
        return fail(SIMC_E_ARG, "tape rng needs a tape and a stride in [1, 2^24]");
    std::lock_guard<std::mutex> lock(ctx->host_mu);
    const double* hin[12] = {in->x, in->y, in->z, in->E, in->p, in->E_vertex, in->theta, in->delta,
                             in->xptar, in->yptar, in->dang_yp, in->dang_xp};
    double* hout[11] = {out->dpp, out->x, out->y, out->z, out->dxdz, out->dydz, out->fry,
                        out->eloss, out->teff, out->dang1, out->dang2};
    const size_t nd = (size_t)n * 8;
    const size_t tape_bytes = rng->kind == SIMC_RNG_TAPE ? (size_t)n * (size_t)rng->stride * 8 : 0;
    auto rounded = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
    const size_t need = 23 * rounded(nd) + rounded(4 * (size_t)n) + rounded(tape_bytes);
    if (ctx->stage_bytes < need) {
        if (ctx->d_stage) {
            HIPCHK(hipDeviceSynchronize());
            HIPCHK(hipFree(ctx->d_stage));
            ctx->d_stage = nullptr;
            ctx->stage_bytes = 0;
        }
        HIPCHK(hipMalloc(&ctx->d_stage, need));
        ctx->stage_bytes = need;
    }
    char* base = (char*)ctx->d_stage;
    size_t at = 0;
    auto carve = [&](size_t bytes) {
        char* p = base + at;
        at += rounded(bytes);
        return p;
    };
    const double* din[12];
    for (int k = 0; k < 12; ++k) {
        din[k] = nullptr;
        if (hin[k]) {
            double* d = (double*)carve(nd);
            HIPCHK(hipMemcpy(d, hin[k], nd, hipMemcpyHostToDevice));
            din[k] = d;
        }
    }
    double* dout[11];
    for (int k = 0; k < 11; ++k) dout[k] = hout[k] ? (double*)carve(nd) : nullptr;
    int32_t* d_used = (int32_t*)carve(4 * (size_t)n);  // always staged: tape overflow check
    simc_rng drng = *rng;
    if (tape_bytes) {
        double* d = (double*)carve(tape_bytes);
        HIPCHK(hipMemcpy(d, rng->tape, tape_bytes, hipMemcpyHostToDevice));
        drng.tape = d;
    }
    if (at > need) return fail(SIMC_E_INTERNAL, "simc_target_prep_host: staging layout exceeds its buffer");
    simc_vertex_in vin = {din[0], din[1], din[2], din[3], din[4], din[5],
                          din[6], din[7], din[8], din[9], din[10], din[11]};
    simc_prep_out pout = {dout[0], dout[1], dout[2], dout[3], dout[4], dout[5], dout[6],
                          dout[7], dout[8], dout[9], dout[10], d_used};
    int rc = simc_target_prep(ctx, t, s, &vin, &pout, n, &drng, nullptr);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    for (int k = 0; k < 11; ++k)
        if (hout[k]) HIPCHK(hipMemcpy(hout[k], dout[k], nd, hipMemcpyDeviceToHost));
    std::vector<int32_t> used((size_t)n);
    HIPCHK(hipMemcpy(used.data(), d_used, 4 * (size_t)n, hipMemcpyDeviceToHost));
    if (out->used) std::memcpy(out->used, used.data(), 4 * (size_t)n);
    for (int64_t i = 0; i < n; ++i)
        if (used[(size_t)i] < 0) return fail(SIMC_E_TAPE, "event " + std::to_string(i) + " ran out of its tape slot");
    return SIMC_OK;
}

int simc_gen_hep(simc_ctx* ctx, const simc_target* t, const simc_hep_settings* h, simc_hep_out* out, int64_t n,
                 const simc_rng* rng, void* stream)
{
    if (!ctx || !t || !h || !out || !rng) return fail(SIMC_E_ARG, "simc_gen_hep: NULL argument");
    if (n < 0) return fail(SIMC_E_ARG, "negative event count");
    if (n == 0) return SIMC_OK;
    if (rng->kind != SIMC_RNG_PHILOX && rng->kind != SIMC_RNG_TAPE)
        return fail(SIMC_E_ARG, "unknown rng kind");
    if (rng->kind == SIMC_RNG_TAPE && (!rng->tape || rng->stride <= 0 || rng->stride > (1 << 24)))
        return fail(SIMC_E_ARG, "tape rng needs a tape and a stride in [1, 2^24]");
    if (h->fr_pattern < 0 || h->fr_pattern > 3)
        return fail(SIMC_E_ARG, "fr_pattern must be 0 (none), 1 (bedpost), 2 (circle) or 3 (flat square)");
    if (!(h->Mh > 0.0) || !(h->e_P > 0.0) || !(h->p_P > 0.0))
        return fail(SIMC_E_ARG, "Mh and the central momenta must be positive");
    GenArgs a;
    std::memset(&a, 0, sizeof a);
    // the beam's trip_thru_target (narm 1, energy Ebeam, electron mass, event.f:194-196)
    simc_prep_settings s;
    std::memset(&s, 0, sizeof s);
    s.narm = SIMC_TARGET_BEAM;
    s.spectrometer = 1;
    s.using_eloss = h->using_eloss;
    s.mass = kMe;
    s.beam_energy = h->Ebeam;
    static const double z_present = 0.0;  // make_prep checks that a z column exists
    simc_vertex_in vin;
    std::memset(&vin, 0, sizeof vin);
    vin.z = &z_present;
    if (int rc = make_prep(t, &s, &vin, a.P)) return rc;
    make_hep(h, a.H);
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    const uint64_t stride_bytes = rng->kind == SIMC_RNG_TAPE ? (uint64_t)rng->stride * 8u : 0u;
    int64_t per = (int64_t)1 << 30;
    if (stride_bytes) per = std::min<int64_t>(per, (int64_t)(0xffffffffull / stride_bytes));
    a.seed = rng->seed;
    a.stride_bytes = (uint32_t)stride_bytes;
    double* cols[kHepCols];
    hep_columns(out, cols);
    for (int64_t at = 0; at < n; at += per) {
        const int64_t m = std::min(per, n - at);
        a.n = (uint32_t)m;
        a.event0 = rng->event0 + (uint64_t)at;
        a.tape = rng->kind == SIMC_RNG_TAPE ? rng->tape + at * rng->stride : nullptr;
        for (int k = 0; k < kHepCols; ++k) a.col[k] = cols[k] ? cols[k] + at : nullptr;
        a.ok = out->ok ? out->ok + at : nullptr;
        a.used = out->used ? out->used + at : nullptr;
        const dim3 G((unsigned)((m + 255) / 256)), B(256);
        if (rng->kind == SIMC_RNG_TAPE) hipLaunchKernelGGL((gen_hep_kernel<TapeStream>), G, B, 0, st, a);
        else hipLaunchKernelGGL((gen_hep_kernel<PhiloxStreamT<kGenTag>>), G, B, 0, st, a);
        HIPCHK(hipGetLastError());
    }
    return SIMC_OK;
}

int simc_gen_hep_rad(simc_ctx* ctx, const simc_target* t, const simc_hep_settings* h, const simc_rad_settings* r,
                     simc_hep_out* out, simc_rad_out* rout, int64_t n, const simc_rng* rng, void* stream)
{
    if (!ctx || !t || !h || !r || !out || !rng) return fail(SIMC_E_ARG, "simc_gen_hep_rad: NULL argument");
    if (n < 0) return fail(SIMC_E_ARG, "negative event count");
    if (n == 0) return SIMC_OK;
    if (rng->kind != SIMC_RNG_PHILOX && rng->kind != SIMC_RNG_TAPE)
        return fail(SIMC_E_ARG, "unknown rng kind");
    if (rng->kind == SIMC_RNG_TAPE && (!rng->tape || rng->stride <= 0 || rng->stride > (1 << 24)))
        return fail(SIMC_E_ARG, "tape rng needs a tape and a stride in [1, 2^24]");
    if (h->fr_pattern < 0 || h->fr_pattern > 3)
        return fail(SIMC_E_ARG, "fr_pattern must be 0 (none), 1 (bedpost), 2 (circle) or 3 (flat square)");
    if (!(h->Mh > 0.0) || !(h->e_P > 0.0) || !(h->p_P > 0.0))
        return fail(SIMC_E_ARG, "Mh and the central momenta must be positive");
    if (r->rad_flag < 0 || r->rad_flag > 3) return fail(SIMC_E_ARG, "rad_flag must be 0..3 (radc.f:141-159)");
    if (r->extrad_flag < 0 || r->extrad_flag > 3) return fail(SIMC_E_ARG, "extrad_flag must be 0..3");
    if (r->intcor_mode != 0 && r->intcor_mode != 1) return fail(SIMC_E_ARG, "intcor_mode must be 0 or 1");
    // radc_init's default (init.f:621-629)
    const int extrad = r->extrad_flag != 0 ? r->extrad_flag : (r->rad_flag == 0 ? 3 : 1);
    // peaked_rad_weight with rad_flag 1 and the Friedrich form reads saved
    // locals that no call of that run sets (radc.f:560-637)
    if (r->rad_flag == 1 && extrad == 3)
        return fail(SIMC_E_ARG, "rad_flag 1 with extrad_flag 3 reads undefined values in the reference");
    GenRadArgs a;
    std::memset(&a, 0, sizeof a);
    {
        simc_prep_settings s;
        std::memset(&s, 0, sizeof s);
        s.narm = SIMC_TARGET_BEAM;
        s.spectrometer = 1;
        s.using_eloss = h->using_eloss;
        s.mass = kMe;
        s.beam_energy = h->Ebeam;
        static const double z_present = 0.0;
        simc_vertex_in vin;
        std::memset(&vin, 0, sizeof vin);
        vin.z = &z_present;
        if (int rc = make_prep(t, &s, &vin, a.Pb)) return rc;
    }
    if (int rc = scattered_prep(t, SIMC_TARGET_ELECTRON, r->e_arm, kMe, h->e_P, h->using_eloss, a.Pe)) return rc;
    if (int rc = scattered_prep(t, SIMC_TARGET_HADRON, r->p_arm, h->Mh, h->p_P, h->using_eloss, a.Pp)) return rc;
    make_hep(h, a.H);
    RadDev& D = a.Rd;
    D.rad_flag = r->rad_flag;
    D.extrad_flag = extrad;
    D.intcor_mode = r->intcor_mode;
    D.use_expon = r->use_expon;
    D.use_offshell_rad = r->use_offshell_rad != 0;
    for (int k = 0; k < 3; ++k) D.doing_tail[k] = r->doing_tail[k] != 0;
    D.hardwired = r->Egamma_gen_max > 0.01;
    D.dE_edge_test = r->dE_edge_test;
    D.Egamma_gen_max = r->Egamma_gen_max;
    for (int k = 0; k < 2; ++k) {
        D.edge_e_E[k] = r->edge_e_E[k];
        D.edge_p_E[k] = r->edge_p_E[k];
        D.edge_Em[k] = r->edge_Em[k];
        D.vertex_Trec[k] = r->vertex_Trec[k];
    }
    D.Egamma_tot_max = r->Egamma_tot_max;
    D.Egamma1_max = D.doing_tail[0] ? r->Egamma_tot_max : 0.0;  // init.f:422-424
    if (r->etatzai > 0.0) {
        D.etatzai = r->etatzai;
    } else {  // radc_init with target_init's L1, L2 (init.f:19-24, :634)
        double L1 = std::log(184.15) - std::log(t->Z) / 3.0, L2 = std::log(1194.) - 2. * std::log(t->Z) / 3.0;
        if (t->Z == 1) {
            L1 = 5.31;
            L2 = 6.144;
        }
        D.etatzai = (12.0 + (t->Z + 1.) / (t->Z * L1 + L2)) / 9.0;
    }
    D.targ_A = t->A;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    const uint64_t stride_bytes = rng->kind == SIMC_RNG_TAPE ? (uint64_t)rng->stride * 8u : 0u;
    int64_t per = (int64_t)1 << 30;
    if (stride_bytes) per = std::min<int64_t>(per, (int64_t)(0xffffffffull / stride_bytes));
    a.seed = rng->seed;
    a.stride_bytes = (uint32_t)stride_bytes;
    double* cols[kHepCols];
    hep_columns(out, cols);
    simc_rad_out none;
    std::memset(&none, 0, sizeof none);
    const simc_rad_out* ro = rout ? rout : &none;
    double* const rcols[kRadCols] = {ro->Egamma1, ro->Egamma2, ro->Egamma3, ro->gen_weight, ro->orig_Ein,
                                     ro->orig_e_E, ro->orig_e_delta, ro->orig_p_E, ro->orig_p_P, ro->orig_p_delta,
                                     ro->eloss2, ro->teff2, ro->eloss3, ro->teff3, ro->hardcorfac};
    for (int64_t at = 0; at < n; at += per) {
        const int64_t m = std::min(per, n - at);
        a.n = (uint32_t)m;
        a.event0 = rng->event0 + (uint64_t)at;
        a.tape = rng->kind == SIMC_RNG_TAPE ? rng->tape + at * rng->stride : nullptr;
        for (int k = 0; k < kHepCols; ++k) a.col[k] = cols[k] ? cols[k] + at : nullptr;
        for (int k = 0; k < kRadCols; ++k) a.rcol[k] = rcols[k] ? rcols[k] + at : nullptr;
        a.ntail = ro->ntail ? ro->ntail + at : nullptr;
        a.ok = out->ok ? out->ok + at : nullptr;
        a.used = out->used ? out->used + at : nullptr;
        const dim3 G((unsigned)((m + 255) / 256)), B(256);
        if (rng->kind == SIMC_RNG_TAPE) hipLaunchKernelGGL((gen_hep_rad_kernel<TapeStream>), G, B, 0, st, a);
        else hipLaunchKernelGGL((gen_hep_rad_kernel<PhiloxStreamT<kGenTag>>), G, B, 0, st, a);
        HIPCHK(hipGetLastError());
    }
    return SIMC_OK;
}

int simc_gen_hep_host(simc_ctx* ctx, const simc_target* t, const simc_hep_settings* h, simc_hep_out* out,
                      int64_t n, const simc_rng* rng)
{
    if (!ctx || !t || !h || !out || !rng) return fail(SIMC_E_ARG, "simc_gen_hep_host: NULL argument");
    if (n < 0) return fail(SIMC_E_ARG, "negative event count");
    if (n > ((int64_t)1 << 31)) return fail(SIMC_E_ARG, "simc_gen_hep_host: at most 2^31 events per call");
    if (n == 0) return SIMC_OK;
    if (rng->kind == SIMC_RNG_TAPE && (!rng->tape || rng->stride <= 0 || rng->stride > (1 << 24)))
        return fail(SIMC_E_ARG, "tape rng needs a tape and a stride in [1, 2^24]");
    std::lock_guard<std::mutex> lock(ctx->host_mu);
    double* hcol[kHepCols] = {out->x, out->y, out->z, out->rasterx, out->rastery, out->Ein, out->eloss,
                              out->teff, out->e_xptar, out->e_yptar, out->e_theta, out->e_phi, out->e_E,
                              out->e_delta, out->p_xptar, out->p_yptar, out->p_theta, out->p_phi,
                              out->p_P, out->p_E, out->p_delta, out->Q2, out->nu, out->q,
                              out->jacobian, out->sigcc, out->weight};
    const size_t nd = (size_t)n * 8;
    const size_t tape_bytes = rng->kind == SIMC_RNG_TAPE ? (size_t)n * (size_t)rng->stride * 8 : 0;
    auto rounded = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
    const size_t need = kHepCols * rounded(nd) + rounded((size_t)n) + rounded(4 * (size_t)n) + rounded(tape_bytes);
    if (ctx->stage_bytes < need) {
        if (ctx->d_stage) {
            HIPCHK(hipDeviceSynchronize());
            HIPCHK(hipFree(ctx->d_stage));
            ctx->d_stage = nullptr;
            ctx->stage_bytes = 0;
        }
        HIPCHK(hipMalloc(&ctx->d_stage, need));
        ctx->stage_bytes = need;
    }
    char* base = (char*)ctx->d_stage;
    size_t at = 0;
    auto carve = [&](size_t bytes) {
        char* p = base + at;
        at += rounded(bytes);
        return p;
    };
    double* dcol[kHepCols];
    for (int k = 0; k < kHepCols; ++k) dcol[k] = hcol[k] ? (double*)carve(nd) : nullptr;
    uint8_t* d_ok = out->ok ? (uint8_t*)carve((size_t)n) : nullptr;
    int32_t* d_used = (int32_t*)carve(4 * (size_t)n);  // always staged: tape overflow check
    simc_rng drng = *rng;
    if (tape_bytes) {
        double* d = (double*)carve(tape_bytes);
        HIPCHK(hipMemcpy(d, rng->tape, tape_bytes, hipMemcpyHostToDevice));
        drng.tape = d;
    }
    if (at > need) return fail(SIMC_E_INTERNAL, "simc_gen_hep_host: staging layout exceeds its buffer");
    simc_hep_out dout = {dcol[0], dcol[1], dcol[2], dcol[3], dcol[4], dcol[5], dcol[6], dcol[7], dcol[8],
                         dcol[9], dcol[10], dcol[11], dcol[12], dcol[13], dcol[14], dcol[15], dcol[16], dcol[17],
                         dcol[18], dcol[19], dcol[20], dcol[21], dcol[22], dcol[23], dcol[24], dcol[25],
                         dcol[26], d_ok, d_used};
    int rc = simc_gen_hep(ctx, t, h, &dout, n, &drng, nullptr);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    for (int k = 0; k < kHepCols; ++k)
        if (hcol[k]) HIPCHK(hipMemcpy(hcol[k], dcol[k], nd, hipMemcpyDeviceToHost));
    if (out->ok) HIPCHK(hipMemcpy(out->ok, d_ok, (size_t)n, hipMemcpyDeviceToHost));
    std::vector<int32_t> used((size_t)n);
    HIPCHK(hipMemcpy(used.data(), d_used, 4 * (size_t)n, hipMemcpyDeviceToHost));
    if (out->used) std::memcpy(out->used, used.data(), 4 * (size_t)n);
    for (int64_t i = 0; i < n; ++i)
        if (used[(size_t)i] < 0) return fail(SIMC_E_TAPE, "event " + std::to_string(i) + " ran out of its tape slot");
    return SIMC_OK;
}

}  // extern "C"

// A loaded target field (trgInit): the map on the device and the field
// angles of the two arms.
struct simc_field {
    double* d_map = nullptr;   // B_z nodes, then B_r nodes (2 x 2601)
    double stht[2] = {}, ctht[2] = {};   // [0]: spect -1 (electron arm), [1]: spect +1
    int device = 0;
    std::mutex mu;             // guards the host-pointer staging
    void* d_stage = nullptr;
    size_t stage_bytes = 0;
    // track_to_tgt loop state and event lists (grown on demand; calls on one
    // field must be ordered)
    void* d_work = nullptr;
    int64_t work_cap = 0;
};

namespace {

int field_angles(const simc_field* f, int spect, double& st, double& ct)
{
    if (spect != -1 && spect != 1) return fail(SIMC_E_ARG, "spect must be -1 (electron arm) or +1 (hadron arm)");
    st = f->stht[spect == -1 ? 0 : 1];
    ct = f->ctht[spect == -1 ? 0 : 1];
    return SIMC_OK;
}

// Staging of host arrays for the host-pointer entry points.
struct Stager {
    simc_field* f;
    size_t at = 0;
    char* base = nullptr;
    static size_t rounded(size_t b) { return (b + 255) & ~(size_t)255; }
    int reserve(size_t need)
    {
        if (f->stage_bytes < need) {
            if (f->d_stage) {
                HIPCHK(hipDeviceSynchronize());
                HIPCHK(hipFree(f->d_stage));
                f->d_stage = nullptr;
                f->stage_bytes = 0;
            }
            HIPCHK(hipMalloc(&f->d_stage, need));
            f->stage_bytes = need;
        }
        base = (char*)f->d_stage;
        return SIMC_OK;
    }
    template <class T>
    T* carve(size_t bytes)
    {
        T* p = (T*)(base + at);
        at += rounded(bytes);
        return p;
    }
};

}  // namespace

extern "C" {

int simc_field_create(const char* map, double theta_e, double phi_e, double theta_p, double phi_p,
                      simc_field** out)
{
    if (!out) return fail(SIMC_E_ARG, "simc_field_create: NULL out");
    *out = nullptr;
    (void)phi_e;
    (void)phi_p;  // B_phi is treated as 0 (trg_track.f:255-256)
    std::vector<double> nodes(2 * kTrgNodes, 0.0);
    bool blank = true;  // trgInit compares the blank-padded name with ' ' (:300)
    for (const char* c = map; c && *c; ++c) blank = blank && *c == ' ';
    if (blank) {  // uniform 5 T over 26 cm in z and 16 cm in r (trg_track.f:311-323)
        for (int ir = 0; ir < kTrgN; ++ir)
            for (int iz = 0; iz < kTrgN; ++iz)
                nodes[ir * kTrgN + iz] = (2. * ir <= 16. && 2. * iz <= 26.) ? 5.0 : 0.0;
    } else if (std::strcmp(map, "0") != 0) {  // the map file, z fastest (trg_track.f:300-310)
        FILE* fp = std::fopen(map, "r");
        if (!fp) return fail(SIMC_E_ARG, std::string("cannot open the field map ") + map);
        char line[512];
        for (int k = 0; k < kTrgNodes; ++k) {
            double v[7];
            if (!std::fgets(line, sizeof line, fp) ||
                std::sscanf(line, "%lf %lf %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5],
                            &v[6]) != 7) {
                std::fclose(fp);
                return fail(SIMC_E_MAPFMT, std::string("field map line ") + std::to_string(k + 1) + " of " + map +
                                               ": expected 7 numbers");
            }
            nodes[k] = v[2];
            nodes[kTrgNodes + k] = v[3];
        }
        std::fclose(fp);
    }
    simc_field* f = new (std::nothrow) simc_field();
    if (!f) return fail(SIMC_E_NOMEM, "out of host memory");
    const double pi180 = 3.141592653 / 180.;  // trg_track.f:266, :271-281
    f->stht[0] = std::sin(theta_e * pi180);
    f->ctht[0] = std::cos(theta_e * pi180);
    f->stht[1] = std::sin(theta_p * pi180);
    f->ctht[1] = std::cos(theta_p * pi180);
    if (hipGetDevice(&f->device) != hipSuccess ||
        hipMalloc(&f->d_map, nodes.size() * sizeof(double)) != hipSuccess ||
        hipMemcpy(f->d_map, nodes.data(), nodes.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        if (f->d_map) (void)hipFree(f->d_map);
        delete f;
        return fail(SIMC_E_HIP, "simc_field_create: device allocation or copy failed");
    }
    *out = f;
    return SIMC_OK;
}

void simc_field_destroy(simc_field* f)
{
    if (!f) return;
    if (f->d_map) (void)hipFree(f->d_map);
    if (f->d_stage) (void)hipFree(f->d_stage);
    delete f;
}

int simc_track_from_tgt(const simc_field* f, int spect, int64_t n, double* x, double* y, double* z, double* dx,
                        double* dy, const double* mom, double mass, uint8_t* ok, void* stream)
{
    if (!f) return fail(SIMC_E_ARG, "simc_track_from_tgt: NULL field");
    if (n < 0 || n > ((int64_t)1 << 31)) return fail(SIMC_E_ARG, "simc_track_from_tgt: n must be in [0, 2^31]");
    if (n == 0) return SIMC_OK;
    if (!x || !y || !z || !dx || !dy || !mom) return fail(SIMC_E_ARG, "simc_track_from_tgt: NULL argument");
    FieldFromArgs a;
    std::memset(&a, 0, sizeof a);
    if (int rc = field_angles(f, spect, a.stht, a.ctht)) return rc;
    HIPCHK(hipSetDevice(f->device));
    a.map = f->d_map;
    a.n = (uint32_t)n;
    a.x = x; a.y = y; a.z = z; a.dx = dx; a.dy = dy;
    a.mom = mom;
    a.mass = mass;
    a.ok = ok;
    hipLaunchKernelGGL(field_from_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a);
    HIPCHK(hipGetLastError());
    return SIMC_OK;
}

int simc_track_to_tgt(simc_ctx* ctx, const simc_field* f, int arm, int spect, int64_t n, const double* x_fp,
                      const double* dx_fp, const double* y_fp, const double* dy_fp, double* delta, double* y,
                      double* dx, double* dy, const double* frx, const double* fry, const double* mom, double mass,
                      double ctheta, double stheta, uint8_t* ok, void* stream)
{
    if (!ctx || !f) return fail(SIMC_E_ARG, "simc_track_to_tgt: NULL context or field");
    const int ai = arm_index(arm);
    if (ai < 0) return fail(SIMC_E_ARG, "arm must be 1 (HMS), 2 (SOS), 3 (HRS-R), 4 (HRS-L) or 5 (SHMS)");
    if (!ctx->d_coef[ai]) return fail(SIMC_E_ARG, "the arm's maps are not loaded (simc_ctx_load_arm)");
    if (n < 0 || n > ((int64_t)1 << 31)) return fail(SIMC_E_ARG, "simc_track_to_tgt: n must be in [0, 2^31]");
    if (n == 0) return SIMC_OK;
    if (!x_fp || !dx_fp || !y_fp || !dy_fp || !delta || !y || !dx || !dy || !frx || !fry || !mom)
        return fail(SIMC_E_ARG, "simc_track_to_tgt: NULL argument");
    FieldToArgs a;
    std::memset(&a, 0, sizeof a);
    if (int rc = field_angles(f, spect, a.stht, a.ctht)) return rc;
    HIPCHK(hipSetDevice(f->device));
    simc_field* fw = const_cast<simc_field*>(f);
    // workspace: xx, mom0, delx (doubles), two event lists, two counters, ok, iteration
    auto rounded = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t nn = (size_t)n;
    const size_t need = 3 * rounded(8 * nn) + 2 * rounded(4 * nn) + 256 + 2 * rounded(nn);
    if (fw->work_cap < n) {
        if (fw->d_work) {
            HIPCHK(hipDeviceSynchronize());  // earlier calls may still use it
            HIPCHK(hipFree(fw->d_work));
            fw->d_work = nullptr;
            fw->work_cap = 0;
        }
        HIPCHK(hipMalloc(&fw->d_work, need));
        fw->work_cap = n;
    }
    char* w = (char*)fw->d_work;
    size_t at = 0;
    auto carve = [&](size_t b) {
        char* p = w + at;
        at += rounded(b);
        return p;
    };
    a.xx = (double*)carve(8 * nn);
    a.mom0 = (double*)carve(8 * nn);
    a.delx = (double*)carve(8 * nn);
    uint32_t* list[2] = {(uint32_t*)carve(4 * nn), (uint32_t*)carve(4 * nn)};
    uint32_t* count = (uint32_t*)carve(256);
    a.st_ok = (uint8_t*)carve(nn);
    a.iter = (uint8_t*)carve(nn);
    if (at > need) return fail(SIMC_E_INTERNAL, "simc_track_to_tgt: workspace layout exceeds its buffer");
    a.map = f->d_map;
    a.n = (uint32_t)n;
    a.coef = ctx->d_coef[ai];
    a.gen = ctx->d_gen[ai];
    a.x_fp = x_fp; a.dx_fp = dx_fp; a.y_fp = y_fp; a.dy_fp = dy_fp;
    a.delta = delta; a.y = y; a.dx = dx; a.dy = dy;
    a.frx = frx; a.fry = fry; a.mom = mom;
    a.mass = mass; a.ctheta = ctheta; a.stheta = stheta;
    a.ok = ok;
    const dim3 G((uint32_t)((n + 255) / 256)), B(256);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, st, count);
    a.list_out = list[0];
    a.count_out = count;
    hipLaunchKernelGGL(field_to_start, G, B, 0, st, a);
    // at most 10 iterations (trg_track.f:800): after round r every event
    // still listed has made r + 1 of them
    for (int r = 0; r < 10; ++r) {
        const int in = r & 1;
        a.list_in = list[in];
        a.count_in = count + in;
        a.list_out = list[in ^ 1];
        a.count_out = count + (in ^ 1);
        hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, st, a.count_out);
        if (arm == SIMC_ARM_HMS) hipLaunchKernelGGL(field_to_recon<SIMC_ARM_HMS>, G, B, 0, st, a);
        else if (arm == SIMC_ARM_SOS) hipLaunchKernelGGL(field_to_recon<SIMC_ARM_SOS>, G, B, 0, st, a);
        else if (is_hrs(arm)) hipLaunchKernelGGL(field_to_recon<SIMC_ARM_HRSR>, G, B, 0, st, a);
        else hipLaunchKernelGGL(field_to_recon<SIMC_ARM_SHMS>, G, B, 0, st, a);
        hipLaunchKernelGGL(field_to_search, G, B, 0, st, a);
    }
    HIPCHK(hipGetLastError());
    return SIMC_OK;
}

int simc_track_from_tgt_host(simc_field* f, int spect, int64_t n, double* x, double* y, double* z, double* dx,
                             double* dy, const double* mom, double mass, uint8_t* ok)
{
    if (!f || !x || !y || !z || !dx || !dy || !mom) return fail(SIMC_E_ARG, "simc_track_from_tgt_host: NULL argument");
    if (n < 0 || n > ((int64_t)1 << 31)) return fail(SIMC_E_ARG, "simc_track_from_tgt_host: n must be in [0, 2^31]");
    if (n == 0) return SIMC_OK;
    std::lock_guard<std::mutex> lock(f->mu);
    HIPCHK(hipSetDevice(f->device));
    const size_t nd = (size_t)n * 8;
    Stager S{f};
    if (int rc = S.reserve(6 * Stager::rounded(nd) + Stager::rounded((size_t)n))) return rc;
    double* h[5] = {x, y, z, dx, dy};
    double* d[5];
    for (int k = 0; k < 5; ++k) {
        d[k] = S.carve<double>(nd);
        HIPCHK(hipMemcpy(d[k], h[k], nd, hipMemcpyHostToDevice));
    }
    double* dm = S.carve<double>(nd);
    HIPCHK(hipMemcpy(dm, mom, nd, hipMemcpyHostToDevice));
    uint8_t* dok = S.carve<uint8_t>((size_t)n);
    int rc = simc_track_from_tgt(f, spect, n, d[0], d[1], d[2], d[3], d[4], dm, mass, dok, nullptr);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    for (int k = 0; k < 5; ++k) HIPCHK(hipMemcpy(h[k], d[k], nd, hipMemcpyDeviceToHost));
    if (ok) HIPCHK(hipMemcpy(ok, dok, (size_t)n, hipMemcpyDeviceToHost));
    return SIMC_OK;
}

int simc_track_to_tgt_host(simc_ctx* ctx, simc_field* f, int arm, int spect, int64_t n, const double* x_fp,
                           const double* dx_fp, const double* y_fp, const double* dy_fp, double* delta, double* y,
                           double* dx, double* dy, const double* frx, const double* fry, const double* mom,
                           double mass, double ctheta, double stheta, uint8_t* ok)
{
    if (!ctx || !f || !x_fp || !dx_fp || !y_fp || !dy_fp || !delta || !y || !dx || !dy || !frx || !fry || !mom)
        return fail(SIMC_E_ARG, "simc_track_to_tgt_host: NULL argument");
    if (n < 0 || n > ((int64_t)1 << 31)) return fail(SIMC_E_ARG, "simc_track_to_tgt_host: n must be in [0, 2^31]");
    if (n == 0) return SIMC_OK;
    std::lock_guard<std::mutex> lock(f->mu);
    HIPCHK(hipSetDevice(f->device));
    const size_t nd = (size_t)n * 8;
    Stager S{f};
    if (int rc = S.reserve(11 * Stager::rounded(nd) + Stager::rounded((size_t)n))) return rc;
    const double* hin[7] = {x_fp, dx_fp, y_fp, dy_fp, frx, fry, mom};
    const double* din[7];
    for (int k = 0; k < 7; ++k) {
        double* dd = S.carve<double>(nd);
        HIPCHK(hipMemcpy(dd, hin[k], nd, hipMemcpyHostToDevice));
        din[k] = dd;
    }
    double* hio[4] = {delta, y, dx, dy};
    double* dio[4];
    for (int k = 0; k < 4; ++k) {
        dio[k] = S.carve<double>(nd);
        HIPCHK(hipMemcpy(dio[k], hio[k], nd, hipMemcpyHostToDevice));
    }
    uint8_t* dok = S.carve<uint8_t>((size_t)n);
    if (ok) HIPCHK(hipMemcpy(dok, ok, (size_t)n, hipMemcpyHostToDevice));
    else HIPCHK(hipMemset(dok, 1, (size_t)n));
    int rc = simc_track_to_tgt(ctx, f, arm, spect, n, din[0], din[1], din[2], din[3], dio[0], dio[1], dio[2], dio[3],
                               din[4], din[5], din[6], mass, ctheta, stheta, dok, nullptr);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    for (int k = 0; k < 4; ++k) HIPCHK(hipMemcpy(hio[k], dio[k], nd, hipMemcpyDeviceToHost));
    if (ok) HIPCHK(hipMemcpy(ok, dok, (size_t)n, hipMemcpyDeviceToHost));
    return SIMC_OK;
}

int simc_counters_get(simc_ctx* ctx, int arm, int64_t* c, int ncounters)
{
    if (!ctx || !c) return fail(SIMC_E_ARG, "simc_counters_get: NULL argument");
    const int ai = arm_index(arm);
    if (ai < 0) return fail(SIMC_E_ARG, "arm must be 1 (HMS), 2 (SOS), 3 (HRS-R), 4 (HRS-L) or 5 (SHMS)");
    if (!ctx->d_coef[ai]) return fail(SIMC_E_ARG, "the arm's maps are not loaded (simc_ctx_load_arm)");
    const int need = ncounters_of(arm);
    if (ncounters < need) return fail(SIMC_E_ARG, "counter array too short");
    unsigned long long h[kHistSlots];
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(h, ctx->d_hist[ai], sizeof h, hipMemcpyDeviceToHost));
    hist_to_counters(arm, h, c, ncounters);
    return SIMC_OK;
}

int simc_counters_reset(simc_ctx* ctx, int arm)
{
    if (!ctx) return fail(SIMC_E_ARG, "simc_counters_reset: NULL context");
    const int ai = arm_index(arm);
    if (ai < 0) return fail(SIMC_E_ARG, "arm must be 1 (HMS), 2 (SOS), 3 (HRS-R), 4 (HRS-L) or 5 (SHMS)");
    if (!ctx->d_coef[ai]) return fail(SIMC_E_ARG, "the arm's maps are not loaded (simc_ctx_load_arm)");
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemset(ctx->d_hist[ai], 0, kHistSlots * sizeof(unsigned long long)));
    HIPCHK(hipDeviceSynchronize());
    return SIMC_OK;
}

int simc_map_class_info(simc_ctx* ctx, int arm, int cls, double* driftdist, int32_t* adrift, int32_t* n_terms)
{
    if (!ctx) return fail(SIMC_E_ARG, "NULL context");
    const int ai = arm_index(arm);
    if (ai < 0) return fail(SIMC_E_ARG, "arm must be 1 (HMS), 2 (SOS), 3 (HRS-R), 4 (HRS-L) or 5 (SHMS)");
    if (!ctx->d_coef[ai]) return fail(SIMC_E_ARG, "the arm's maps are not loaded (simc_ctx_load_arm)");
    const simc::ForwardMap& f = ctx->fwd[ai];
    if (cls < 1 || cls > f.n_classes) return fail(SIMC_E_ARG, "class out of range");
    if (driftdist) *driftdist = f.driftdist[cls - 1];
    if (adrift) *adrift = f.adrift[cls - 1];
    if (n_terms) *n_terms = (int32_t)f.classes[cls - 1].size();
    return SIMC_OK;
}

int simc_stop_flops_stage(simc_ctx* ctx, int arm, int stage, int64_t* flops, int nflops)
{
    if (!ctx || !flops) return fail(SIMC_E_ARG, "simc_stop_flops: NULL argument");
    const int ai = arm_index(arm);
    if (ai < 0) return fail(SIMC_E_ARG, "arm must be 1 (HMS), 2 (SOS), 3 (HRS-R), 4 (HRS-L) or 5 (SHMS)");
    if (!ctx->d_coef[ai]) return fail(SIMC_E_ARG, "the arm's maps are not loaded (simc_ctx_load_arm)");
    if (stage < 0 || stage > 3) return fail(SIMC_E_ARG, "stage must be 0 (all), 1, 2 or 3 (reconstruction)");
    if (nflops < kHistBins) return fail(SIMC_E_ARG, "flops array needs 256 entries");
    for (int i = 0; i < nflops; ++i) flops[i] = 0;
    int64_t acc = 0;
    const std::vector<Op>& ops = ctx->ops[ai];
    for (int pc = 0; pc < (int)ops.size(); ++pc) {
        const Op& o = ops[pc];
        const int in_stage = pc < ctx->coll_pc[ai] ? 1 : pc < ctx->hut_pc[ai] ? 2 : 3;
        if (o.code == OP_TRANSP && (stage == 0 || stage == in_stage)) acc += class_flops(ctx, arm, o.arg);
        if (o.code == OP_COLLIMATOR) {
            flops[ST_SLIT_HOR] = flops[ST_SLIT_VERT] = flops[ST_SLIT_OCT] = flops[ST_COLL] = acc;
        } else if (o.stop > 0 && o.code != OP_WIRE && o.stop < kHistBins) {
            flops[o.stop] = acc;
        }
        if (o.code == OP_END) flops[ST_OK] = acc + (stage == 0 || stage == 3 ? class_flops(ctx, arm, 0) : 0);
    }
    return SIMC_OK;
}

int simc_stop_flops_hut(simc_ctx* ctx, int arm, int ms_flag, int wcs_flag, int64_t* flops, int nflops)
{
    if (!ctx || !flops) return fail(SIMC_E_ARG, "simc_stop_flops_hut: NULL argument");
    const int ai = arm_index(arm);
    if (ai < 0) return fail(SIMC_E_ARG, "arm must be 1 (HMS), 2 (SOS), 3 (HRS-R), 4 (HRS-L) or 5 (SHMS)");
    if (!ctx->d_coef[ai]) return fail(SIMC_E_ARG, "the arm's maps are not loaded (simc_ctx_load_arm)");
    if (nflops < kHistBins) return fail(SIMC_E_ARG, "flops array needs 256 entries");
    for (int i = 0; i < nflops; ++i) flops[i] = 0;
    const std::vector<Op>& ops = ctx->ops[ai];
    const int64_t normal = 4;  // per deviate: half a Box-Muller pair
    int planes = 0;
    for (int pc = ctx->hut_pc[ai]; pc < (int)ops.size(); ++pc) planes += ops[pc].code == OP_WIRE;
    int64_t acc = 0;
    for (int pc = ctx->hut_pc[ai]; pc < (int)ops.size(); ++pc) {
        const Op& o = ops[pc];
        switch (o.code) {
        case OP_PROJECT: acc += 11; break;
        case OP_MUSC:
            if (ms_flag && o.p[0] != 0.0) acc += 18 + 2 * normal;
            break;
        case OP_MUSC_EXT:
            if (ms_flag && o.p[0] != 0.0 && o.p[1] > 0.0) acc += 36 + 4 * normal;
            break;
        case OP_WIRE: acc += 3 + (wcs_flag ? 2 * normal : 0); break;
        case OP_FIT: acc += 11 * planes + 18; break;
        case OP_BOX_ROT_H: acc += 13; break;
        case OP_FIDUCIAL: acc += 4; break;
        default: break;  // comparisons and counters
        }
        if (o.code == OP_END) flops[ST_OK] = acc;
        else if (o.stop > 0 && o.code != OP_WIRE && o.stop < kHistBins) flops[o.stop] = acc;
    }
    return SIMC_OK;
}

int simc_stop_flops(simc_ctx* ctx, int arm, int64_t* flops, int nflops)
{
    return simc_stop_flops_stage(ctx, arm, 0, flops, nflops);
}

int simc_recon_coefficients(simc_ctx* ctx, int arm, double* coef, int32_t* n_terms)
{
    if (!ctx) return fail(SIMC_E_ARG, "simc_recon_coefficients: NULL context");
    const int ai = arm_index(arm);
    if (ai < 0) return fail(SIMC_E_ARG, "arm must be 1 (HMS), 2 (SOS), 3 (HRS-R), 4 (HRS-L) or 5 (SHMS)");
    if (!ctx->d_coef[ai]) return fail(SIMC_E_ARG, "the arm's maps are not loaded (simc_ctx_load_arm)");
    const auto& terms = ctx->rec[ai].terms;
    if (n_terms) *n_terms = (int32_t)terms.size();
    if (coef)
        for (size_t t = 0; t < terms.size(); ++t)
            for (int k = 0; k < 4; ++k) coef[t * 4 + k] = terms[t].c[k];
    return SIMC_OK;
}

int simc_adjoint_run(simc_ctx* ctx, int arm, const simc_settings* s, const simc_events_in* in,
                     const uint8_t* mask, int64_t n, const simc_rng* rng, const double* recon_coef,
                     const double* offsets, double* grad_coef, double* out9, void* stream)
{
    if (!ctx || !s || !in || !mask || !rng || !recon_coef || !offsets || !grad_coef || !out9)
        return fail(SIMC_E_ARG, "simc_adjoint_run: NULL argument");
    int rc = check_settings(arm, s);
    if (rc) return rc;
    const int ai = arm_index(arm);
    if (arm != SIMC_ARM_HMS && arm != SIMC_ARM_SHMS)
        return fail(SIMC_E_ARG, "the adjoint form covers the HMS and the SHMS");
    if (s->sieve) return fail(SIMC_E_ARG, "the adjoint form has no mock sieve slit");
    if (needs_rare(arm, s)) return fail(SIMC_E_ARG, "the adjoint form has no decay or collimator stepping");
    if (ctx->d_coef[ai] && ctx->d_gen[ai])
        return fail(SIMC_E_MAPSTRUCT, "the adjoint form needs maps of the compiled structure");
    if (!in->dpp || !in->x || !in->y || !in->dxdz || !in->dydz) return fail(SIMC_E_ARG, "missing event input array");
    if (n <= 0 || n > ((int64_t)1 << 29)) return fail(SIMC_E_ARG, "simc_adjoint_run: n must be in [1, 2^29]");
    if (rng->kind != SIMC_RNG_PHILOX && rng->kind != SIMC_RNG_TAPE) return fail(SIMC_E_ARG, "unknown rng kind");
    const bool tape = rng->kind == SIMC_RNG_TAPE;
    AdjArgs a;
    std::memset(&a, 0, sizeof a);
    if (tape) {
        if (!rng->tape || rng->stride <= 0) return fail(SIMC_E_ARG, "tape rng needs a tape and a positive stride");
        if ((double)n * (double)rng->stride * 8.0 >= 4294967296.0)
            return fail(SIMC_E_ARG, "simc_adjoint_run: tape larger than 4 GB");
        a.stride_bytes = (uint32_t)(rng->stride * 8);
        a.tape = rng->tape;
    }
    const int n_terms = (int)ctx->rec[ai].terms.size();
    if (n_terms > kAdjTermsPerLane * kBlock) return fail(SIMC_E_MAPSTRUCT, "too many reconstruction terms");
    for (const auto& term : ctx->rec[ai].terms)
        for (int v = 0; v < 5; ++v)
            if (term.e[v] > kAdjMaxExp) return fail(SIMC_E_MAPSTRUCT, "reconstruction exponent above 6");
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        if (ctx->sel_cap < n) {
            if (ctx->d_sel || ctx->d_hw) {
                HIPCHK(hipDeviceSynchronize());
                if (ctx->d_sel) HIPCHK(hipFree(ctx->d_sel));
                if (ctx->d_hw) HIPCHK(hipFree(ctx->d_hw));
                ctx->d_sel = nullptr;
                ctx->d_hw = nullptr;
                ctx->sel_cap = 0;
            }
            HIPCHK(hipMalloc(&ctx->d_sel, (size_t)n * 4));
            HIPCHK(hipMalloc(&ctx->d_hw, (size_t)n * 22 * 8));
            ctx->sel_cap = n;
        }
        if (!ctx->d_nsel) HIPCHK(hipMalloc(&ctx->d_nsel, 256));
    }
    a.set0 = make_setdev(arm, s, rng->seed, rng->event0, 0, 0, 0);
    a.flags = a.set0.flags;
    a.seed = rng->seed;
    a.n = (uint32_t)n;
    a.tab = ctx->d_tab[ai];
    a.coef = ctx->d_coef[ai];
    a.ops = ctx->d_ops[ai];
    a.dpp = in->dpp; a.x = in->x; a.y = in->y; a.dxdz = in->dxdz; a.dydz = in->dydz; a.fry = in->fry;
    a.sel = ctx->d_sel;
    a.n_sel = ctx->d_nsel;
    a.rcoef = recon_coef;
    a.rexp = ctx->d_rexp[ai];
    a.n_terms = (uint32_t)n_terms;
    for (int q = 0; q < 6; ++q) a.off[q] = offsets[q];
    a.g_coef = grad_coef;
    a.g_out = out9;
    a.hw = ctx->d_hw;
    a.ft = ctx->d_hw + (size_t)9 * n;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t nc = (uint32_t)n_terms * 4;
    hipLaunchKernelGGL(zero_u64, dim3((nc + 255) / 256), dim3(256), 0, st, (unsigned long long*)grad_coef, nc);
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(256), 0, st, (unsigned long long*)out9, 9u);
    hipLaunchKernelGGL(zero_queue_counts, dim3(1), dim3(64), 0, st, ctx->d_nsel);
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(select_events, dim3(blocks), dim3(kBlock), 0, st, mask, (uint32_t)n, ctx->d_sel, ctx->d_nsel);
#define ADJ(ARMV, RV) hipLaunchKernelGGL((adjoint_kernel<ARMV, RV>), dim3(blocks), dim3(kBlock), 0, st, a)
    if (arm == SIMC_ARM_HMS) {
        if (tape) ADJ(SIMC_ARM_HMS, TapeStream); else ADJ(SIMC_ARM_HMS, PhiloxStream);
    } else {
        if (tape) ADJ(SIMC_ARM_SHMS, TapeStream); else ADJ(SIMC_ARM_SHMS, PhiloxStream);
    }
#undef ADJ
    if (arm == SIMC_ARM_HMS)
        hipLaunchKernelGGL((adjoint_recon_kernel<SIMC_ARM_HMS>), dim3(blocks), dim3(kBlock), 0, st, a);
    else
        hipLaunchKernelGGL((adjoint_recon_kernel<SIMC_ARM_SHMS>), dim3(blocks), dim3(kBlock), 0, st, a);
    const unsigned cgrid = std::min<unsigned>((unsigned)((n + kAdjTile - 1) / kAdjTile), (unsigned)ctx->n_cu * 4u);
    hipLaunchKernelGGL(adjoint_coef_kernel, dim3(cgrid), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL(adjoint_finish, dim3((nc + 255) / 256), dim3(256), 0, st, grad_coef, nc, out9, ctx->d_nsel);
    HIPCHK(hipGetLastError());
    return SIMC_OK;
}

int simc_batch_create(simc_ctx* ctx, const simc_setting* settings, int n_settings, simc_batch** out)
{
    if (!ctx || !settings || !out || n_settings <= 0)
        return fail(SIMC_E_ARG, "simc_batch_create: NULL argument or no settings");
    *out = nullptr;
    std::vector<std::pair<int64_t, int64_t>> ranges;
    for (int i = 0; i < n_settings; ++i) {
        const simc_setting& q = settings[i];
        int rc = check_settings(q.arm, &q.s);
        if (rc) return fail(rc, "setting " + std::to_string(i) + ": " + g_err);
        if (!ctx->d_coef[arm_index(q.arm)])
            return fail(SIMC_E_ARG, "setting " + std::to_string(i) + ": the arm's maps are not loaded");
        if (q.offset < 0 || q.count < 0 || q.offset + q.count > ((int64_t)1 << 29))
            return fail(SIMC_E_ARG, "setting " + std::to_string(i) + ": event range outside [0, 2^29)");
        if (q.count) ranges.push_back({q.offset, q.offset + q.count});
    }
    std::sort(ranges.begin(), ranges.end());
    for (size_t k = 1; k < ranges.size(); ++k)
        if (ranges[k].first < ranges[k - 1].second)
            return fail(SIMC_E_ARG, "simc_batch_create: settings' event ranges overlap");
    simc_batch* b = new (std::nothrow) simc_batch();
    if (!b) return fail(SIMC_E_NOMEM, "out of host memory");
    b->ctx = ctx;
    b->n_settings = n_settings;
    for (int i = 0; i < n_settings; ++i) b->arm_of.push_back(settings[i].arm);
    std::vector<SetDev> table;
    uint64_t max_total = 0;
    // passes: one per (arm, kernel variant, flag set), settings in input
    // order; settings with different Philox keys share a pass and its lanes
    // read their own setting's key
    std::vector<bool> placed(n_settings, false);
    for (int i0 = 0; i0 < n_settings; ++i0) {
        const simc_setting& q0 = settings[i0];
        if (placed[i0] || q0.count == 0) continue;
        const int32_t f0 = make_setdev(q0.arm, &q0.s, 0, 0, 0, 0, 0).flags;
        const bool r0 = needs_rare(q0.arm, &q0.s);
        simc_batch::Pass pass{q0.arm, r0, f0, q0.seed, false, 0, (uint32_t)table.size(), 0};
        for (int i = i0; i < n_settings; ++i) {
            const simc_setting& q = settings[i];
            if (placed[i] || q.count == 0 || q.arm != q0.arm || needs_rare(q.arm, &q.s) != r0 ||
                make_setdev(q.arm, &q.s, 0, 0, 0, 0, 0).flags != f0)
                continue;
            placed[i] = true;
            if (q.seed != q0.seed) pass.lane_key = true;
            table.push_back(make_setdev(q.arm, &q.s, q.seed, q.event0, pass.total, (uint64_t)q.offset, (uint32_t)i));
            pass.total += (uint64_t)q.count;
            ++pass.n;
            b->n_arr = std::max(b->n_arr, (uint32_t)(q.offset + q.count));
        }
        b->passes.push_back(pass);
        max_total = std::max(max_total, pass.total);
    }
    auto cleanup = [&](int rc) {
        std::string keep = g_err;
        simc_batch_destroy(b);
        g_err = keep;
        return rc;
    };
    if (!table.empty()) {
        if (hipMalloc(&b->d_set, table.size() * sizeof(SetDev)) != hipSuccess ||
            hipMemcpy(b->d_set, table.data(), table.size() * sizeof(SetDev), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(SIMC_E_HIP, "simc_batch_create: settings table"));
    }
    if (hipMalloc(&b->d_hist, (size_t)n_settings * kHistSlots * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(b->d_hist, 0, (size_t)n_settings * kHistSlots * sizeof(unsigned long long)) != hipSuccess)
        return cleanup(fail(SIMC_E_HIP, "simc_batch_create: histograms"));
    // queues for the default chunk, so that runs allocate nothing (graph capture)
    b->chunk = std::min<int64_t>(b->chunk, std::max<int64_t>((int64_t)max_total, 1));
    int rc = simc_batch_set_chunk(b, b->chunk);
    if (rc) return cleanup(rc);
    *out = b;
    return SIMC_OK;
}

int simc_batch_set_chunk(simc_batch* b, int64_t chunk)
{
    if (!b) return fail(SIMC_E_ARG, "simc_batch_set_chunk: NULL batch");
    if (chunk < kBlock || chunk > ((int64_t)1 << 24))
        return fail(SIMC_E_ARG, "chunk must be in [256, 2^24] events");
    Queue q1, q2;
    for (const auto& pass : b->passes) {
        const uint32_t cap = (uint32_t)std::min<uint64_t>((uint64_t)chunk, pass.total);
        int rc = ensure_pair(b->qp[arm_index(pass.arm)], cap, q1, q2);
        if (rc) return rc;
    }
    b->chunk = chunk;
    return SIMC_OK;
}

int simc_batch_run(simc_batch* b, const simc_events_in* in, simc_events_out* out, void* stream)
{
    if (!b || !in || !out) return fail(SIMC_E_ARG, "simc_batch_run: NULL argument");
    if (!outputs_complete(in, out)) return fail(SIMC_E_ARG, "missing event input or output array");
    simc_ctx* ctx = b->ctx;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t nh = (uint32_t)b->n_settings * kHistSlots;
    hipLaunchKernelGGL(zero_u64, dim3((nh + 255) / 256), dim3(256), 0, st, b->d_hist, nh);
    for (const auto& pass : b->passes) {
        const int ai = arm_index(pass.arm);
        const QueuePair& qp = b->qp[ai];
        if (qp.cap < std::min<uint64_t>((uint64_t)b->chunk, pass.total))
            return fail(SIMC_E_INTERNAL, "simc_batch_run: queues smaller than the chunk");
        Queue q1, q2;
        queue_view(qp.d_queue[0], qp.cap, qp.d_qcount, q1);
        queue_view(qp.d_queue[1], qp.cap, qp.d_qcount + 1, q2);
        KArgs a;
        std::memset(&a, 0, sizeof a);
        a.set = (const SetDev*)b->d_set + pass.first;
        a.n_set = pass.n;
        a.flags = pass.flags;
        a.seed = pass.seed;
        a.ix_base = 0;
        a.n_arr = b->n_arr;
        a.tab = ctx->d_tab[ai];
        a.coef = ctx->d_coef[ai];
        a.gen = ctx->d_gen[ai];
        a.ops = ctx->d_ops[ai];
        a.hist = b->d_hist;
        a.dpp = in->dpp; a.x = in->x; a.y = in->y; a.dxdz = in->dxdz; a.dydz = in->dydz; a.fry = in->fry;
        a.o_dpp = out->dpp; a.o_y = out->y; a.o_dxdz = out->dxdz; a.o_dydz = out->dydz;
        a.x_fp = out->x_fp; a.dx_fp = out->dx_fp; a.y_fp = out->y_fp; a.dy_fp = out->dy_fp;
        a.pathlen = out->pathlen; a.resmult = out->resmult; a.m2o = out->m2;
        a.ok = out->ok; a.stop = out->stop; a.used = out->used;
        for (uint64_t c0 = 0; c0 < pass.total; c0 += (uint64_t)b->chunk) {
            a.pass_base = c0;
            a.count = (uint32_t)std::min<uint64_t>((uint64_t)b->chunk, pass.total - c0);
            int rc = launch_stages(ctx, pass.arm, pass.rare, false, a, q1, q2, st, nullptr, pass.lane_key);
            if (rc) return rc;
        }
    }
    return SIMC_OK;
}

int simc_batch_passes(simc_batch* b, int32_t* n_passes, int32_t* n_lane_key)
{
    if (!b) return fail(SIMC_E_ARG, "simc_batch_passes: NULL batch");
    int32_t lk = 0;
    for (const auto& pass : b->passes) lk += pass.lane_key ? 1 : 0;
    if (n_passes) *n_passes = (int32_t)b->passes.size();
    if (n_lane_key) *n_lane_key = lk;
    return SIMC_OK;
}

int simc_batch_counters(simc_batch* b, int setting, int64_t* c, int ncounters)
{
    if (!b || !c) return fail(SIMC_E_ARG, "simc_batch_counters: NULL argument");
    if (setting < 0 || setting >= b->n_settings) return fail(SIMC_E_ARG, "setting index out of range");
    const int arm = b->arm_of[setting];
    const int need = ncounters_of(arm);
    if (ncounters < need) return fail(SIMC_E_ARG, "counter array too short");
    unsigned long long h[kHistSlots];
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(h, b->d_hist + (size_t)setting * kHistSlots, sizeof h, hipMemcpyDeviceToHost));
    hist_to_counters(arm, h, c, ncounters);
    return SIMC_OK;
}

int simc_batch_tune(simc_batch* b, const simc_events_in* in, simc_events_out* out, void* stream,
                    int64_t* best_chunk, double* best_ms)
{
    if (!b || !in || !out) return fail(SIMC_E_ARG, "simc_batch_tune: NULL argument");
    uint64_t max_total = 0;
    for (const auto& pass : b->passes) max_total = std::max(max_total, pass.total);
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int64_t best = b->chunk;
    float best_t = 1e30f;
    int rc = SIMC_OK;
    for (int lg = 18; lg <= 24 && rc == SIMC_OK; ++lg) {
        const int64_t chunk = (int64_t)1 << lg;
        if (lg > 18 && (uint64_t)(chunk >> 1) >= max_total) break;  // larger chunks launch the same
        rc = simc_batch_set_chunk(b, chunk);
        if (rc) break;
        rc = simc_batch_run(b, in, out, stream);  // warm-up
        float tmin = 1e30f;
        for (int r = 0; r < 3 && rc == SIMC_OK; ++r) {
            if (hipEventRecord(e0, st) != hipSuccess) rc = fail(SIMC_E_HIP, "hipEventRecord");
            if (!rc) rc = simc_batch_run(b, in, out, stream);
            if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
                rc = fail(SIMC_E_HIP, "hipEventSynchronize");
            float ms = 0;
            if (!rc && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) tmin = std::min(tmin, ms);
        }
        if (!rc && tmin < best_t) {
            best_t = tmin;
            best = chunk;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    rc = simc_batch_set_chunk(b, best);
    if (rc) return rc;
    if (best_chunk) *best_chunk = best;
    if (best_ms) *best_ms = best_t;
    return SIMC_OK;
}

void simc_batch_destroy(simc_batch* b)
{
    if (!b) return;
    if (b->d_set) (void)hipFree(b->d_set);
    if (b->d_hist) (void)hipFree(b->d_hist);
    for (int i = 0; i < kNumArms; ++i) free_pair(b->qp[i]);
    delete b;
}

int64_t simc_poly_flops_full(simc_ctx* ctx, int arm)
{
    if (!ctx || arm_index(arm) < 0) return 0;
    static const int hms_cls[] = {2, 3, 5, 6, 8, 9, 11, 0};
    static const int shms_cls[] = {3, 7, 8, 12, 13, 17, 18, 23, 24, 25, 26, 27, 28, 29, 30, 0};
    static const int sos_cls[] = {2, 3, 4, 5, 6, 7, 8, 9, 0};
    static const int hrs_cls[] = {2, 3, 5, 6, 8, 10, 11, 0};
    int64_t f = 0;
    if (!ctx->d_coef[arm_index(arm)]) return 0;
    if (arm == SIMC_ARM_HMS)
        for (int c : hms_cls) f += class_flops(ctx, arm, c);
    else if (arm == SIMC_ARM_SOS)
        for (int c : sos_cls) f += class_flops(ctx, arm, c);
    else if (is_hrs(arm))
        for (int c : hrs_cls) f += class_flops(ctx, arm, c);
    else
        for (int c : shms_cls) f += class_flops(ctx, arm, c);
    return f;
}

}  // extern "C"
