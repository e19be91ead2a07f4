// Device-side building blocks of the spectrometer path (gfx950, FP64).
//
// One event per lane.  Every routine restates one reference routine (cited)
// with the same operation order; the file is compiled with
// -ffp-contract=off so that only the explicit __builtin_fma calls of the
// generated COSY polynomials fuse.  Uniform draws come from a per-event
// counter-based Philox4x32-10 stream or from a per-event tape.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "simc_bm_tables.h"

namespace simcdev {

// ---------------------------------------------------------------------------
// Event arrays are addressed through buffer resources (base and size in
// SGPRs, one 32-bit byte offset per lane); accesses past the array size read
// zero and write nothing.
// ---------------------------------------------------------------------------
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int kBufWord3 = 0x00020000;  // raw 32-bit data format, gfx9 family

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, kBufWord3);
}
__device__ __forceinline__ double ld64(const double* p, uint32_t n, uint32_t off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc(p, n * 8u), off, 0, 0));
}
__device__ __forceinline__ uint32_t ld32(const uint32_t* p, uint32_t n, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b32(rsrc(p, n * 4u), off, 0, 0);
}
__device__ __forceinline__ void st64(double* p, uint32_t n, uint32_t off, double v)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rsrc(p, n * 8u), off, 0, 0);
}
__device__ __forceinline__ void st32(void* p, uint32_t n, uint32_t off, uint32_t v)
{
    __builtin_amdgcn_raw_buffer_store_b32(v, rsrc(p, n * 4u), off, 0, 0);
}
__device__ __forceinline__ void st8(uint8_t* p, uint32_t n, uint32_t off, uint8_t v)
{
    __builtin_amdgcn_raw_buffer_store_b8(v, rsrc(p, n), off, 0, 0);
}

// constants.inc:23-42
constexpr double kPi = 3.141592653589793;
constexpr double kMpi = 139.57018;
constexpr double kMmu = 105.6583755;
constexpr double kMk = 493.677;

// ---------------------------------------------------------------------------
// Uniform sources
// ---------------------------------------------------------------------------

// Philox4x32-10 (Salmon et al., SC'11; Random123 constants).  Uniform k of an
// event is word k%4 of the block with counter (k/4, event lo, event hi, tag),
// mapped to (w + 1/2) 2^-32.
constexpr uint32_t kPhiloxTag = 0x53494d43u;  // "SIMC"
constexpr uint32_t kTargetTag = 0x54475453u;  // "TGTS"
constexpr uint32_t kHutTag = 0x48555453u;     // "HUTS": the detector huts' draws

template <bool LANE_KEY = false>
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4])
{
    // A launch-uniform key: without this empty statement the compiler forms
    // the 20 round keys once per kernel and keeps them in SGPRs across the
    // whole step loop, where they are spilled to VGPR lanes and read back
    // (v_readlane) in every block.  Made opaque here, they are recomputed by
    // scalar adds in each call instead.  A per-lane key (events of several
    // settings in one wavefront) is kept opaque in VGPRs for the same reason.
    // A launch-uniform key read back from memory (a stream passed to an
    // outlined function) arrives in VGPRs: readfirstlane moves it to SGPRs
    // (a copy where it is already there).
    if (LANE_KEY) {
        asm("" : "+v"(k0), "+v"(k1));
    } else {
        k0 = __builtin_amdgcn_readfirstlane(k0);
        k1 = __builtin_amdgcn_readfirstlane(k1);
        asm("" : "+s"(k0), "+s"(k1));
    }
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// ---------------------------------------------------------------------------
// Elementary functions of the Box-Muller normal generator.  Both are fixed
// sequences of IEEE operations (explicit fma, exact reductions) over the
// tables of simc_bm_tables.h, so that the host restatement produces
// bit-identical values.
// ---------------------------------------------------------------------------
struct __attribute__((aligned(16))) BmLogEntry {
    double nhi;
    float nlo, invc;
};
static __constant__ BmLogEntry kBmLogTab[SIMC_BM_LOG_N] = SIMC_BM_LOG_TABLE;
static __constant__ double kBmSinCosTab[SIMC_BM_SINCOS_N][2] = SIMC_BM_SINCOS_TABLE;

// Natural logarithm of u in (0, 1) (normal numbers): u = m 2^e with m in
// [sqrt(1/2), sqrt(2)); c = 1 + k/64 the nearest table point, invc = 1/c to
// 7 bits, so t = m invc - 1 is exact (|t| < 0.0155); log m = -log(invc) +
// log1p(t), log1p by its series to t^9 (truncation below 2^-56 relative);
// then e ln2 in two parts (ln2_hi has 21 significant bits, so e ln2_hi is
// exact) and the table's two-part -log(invc).
__device__ __forceinline__ double bm_log(double u)
{
    double m = __builtin_amdgcn_frexp_mant(u);  // [1/2, 1)
    int e = __builtin_amdgcn_frexp_exp(u);
    if (m < 0x1.6a09e667f3bcdp-1) {
        m = m * 2.0;
        e = e - 1;
    }
    const double kd = rint((m - 1.0) * 64.0);
    const BmLogEntry L = kBmLogTab[(int)kd + SIMC_BM_LOG_K0];
    const double t = __builtin_fma(m, (double)L.invc, -1.0);
    double q = 0x1.c71c71c71c71cp-4;                 // 1/9
    q = __builtin_fma(t, q, -0x1p-3);                // -1/8
    q = __builtin_fma(t, q, 0x1.2492492492492p-3);   // 1/7
    q = __builtin_fma(t, q, -0x1.5555555555555p-3);  // -1/6
    q = __builtin_fma(t, q, 0x1.999999999999ap-3);   // 1/5
    q = __builtin_fma(t, q, -0x1p-2);                // -1/4
    q = __builtin_fma(t, q, 0x1.5555555555555p-2);   // 1/3
    q = __builtin_fma(t, q, -0x1p-1);                // -1/2
    const double p = __builtin_fma(t * t, q, t);
    const double de = (double)e;
    return __builtin_fma(de, 0x1.62e42p-1, L.nhi + (p + __builtin_fma(de, 0x1.fdf473de6af28p-22, (double)L.nlo)));
}

// sin and cos of 2 pi u for u in [0, 1): x = 64u (exact), k = rint(x),
// r = x - k in [-1/2, 1/2] (exact), d = r pi/32; sin d and cos d by their
// Taylor series to d^9 and d^8 (|d| <= pi/64, truncation below 2^-62), then
// the rotation by the table's sin and cos of 2 pi k/64.
__device__ __forceinline__ void bm_sincos2pi(double u, double& sn, double& cs)
{
    const double x = u * 64.0;
    const double kd = rint(x);
    const double r = x - kd;
    const double d = r * 0x1.921fb54442d18p-4;
    const double z = d * d;
    double ps = 0x1.71de3a556c734p-19;                 // 1/9!
    ps = __builtin_fma(z, ps, -0x1.a01a01a01a01ap-13);  // -1/7!
    ps = __builtin_fma(z, ps, 0x1.1111111111111p-7);    // 1/5!
    ps = __builtin_fma(z, ps, -0x1.5555555555555p-3);   // -1/3!
    const double sd = __builtin_fma(d * z, ps, d);
    double pc = 0x1.a01a01a01a01ap-16;                 // 1/8!
    pc = __builtin_fma(z, pc, -0x1.6c16c16c16c17p-10);  // -1/6!
    pc = __builtin_fma(z, pc, 0x1.5555555555555p-5);    // 1/4!
    pc = __builtin_fma(z, pc, -0x1p-1);                 // -1/2
    const double cd = __builtin_fma(z, pc, 1.0);
    const int k = (int)kd & 63;
    const double sa = kBmSinCosTab[k][0], ca = kBmSinCosTab[k][1];
    sn = __builtin_fma(sa, cd, ca * sd);
    cs = __builtin_fma(ca, cd, -(sa * sd));
}

// TAG is the fourth counter word: streams of different tags are independent
// (the arms use kPhiloxTag, the target step kTargetTag).
template <uint32_t TAG, bool LANE_KEY = false>
struct PhiloxStreamT {
    // LANE_KEY: the key (seed) may differ between the lanes of a wavefront
    static constexpr bool kLaneKey = LANE_KEY;
    uint32_t k0, k1, e0, e1;
    uint32_t pos;
    uint32_t w[4];
    int overflow;
    int have;        // a second Box-Muller normal is pending
    double pending;
    // inlined: an outlined init would leave the key in memory, no longer
    // provably uniform for the SGPR constraint of philox4x32_10
    __device__ __forceinline__ void init(uint64_t seed, uint64_t event, const double*, uint32_t, uint32_t)
    {
        k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
        e0 = (uint32_t)event; e1 = (uint32_t)(event >> 32);
        pos = 0; overflow = 0; have = 0; pending = 0.0;
    }
    // Stream position for a later resume: uniforms drawn, and the pending
    // Box-Muller value.
    __device__ __forceinline__ uint32_t state_word() const { return pos | ((uint32_t)have << 31); }
    // The block's words are consumed from w[0], shifting the rest down: an
    // index by the position would put w in scratch memory (one round trip
    // per uniform), the shift is three register moves.
    __device__ __forceinline__ void shift()
    {
        w[0] = w[1];
        w[1] = w[2];
        w[2] = w[3];
    }
    // Resume at a saved position (re-forms the current Philox block when the
    // position is inside one).
    __device__ __forceinline__ void resume(uint32_t word, double pend)
    {
        pos = word & 0x7fffffffu;
        have = (int)(word >> 31);
        pending = pend;
        if (pos & 3u) {
            philox4x32_10<LANE_KEY>(pos >> 2, e0, e1, TAG, k0, k1, w);
            if ((pos & 3u) >= 1u) shift();
            if ((pos & 3u) >= 2u) shift();
            if ((pos & 3u) >= 3u) shift();
        }
    }
    __device__ __forceinline__ double next()
    {
        if ((pos & 3u) == 0) philox4x32_10<LANE_KEY>(pos >> 2, e0, e1, TAG, k0, k1, w);
        ++pos;
        const uint32_t v = w[0];
        shift();
        return __builtin_fma((double)v, 0x1p-32, 0x1p-33);  // (v + 1/2) 2^-32, exact
    }
    // Normal deviates for the counter-based stream: Box-Muller pairs
    // (u1, u2) -> sqrt(-2 ln u1) (cos 2 pi u2, sin 2 pi u2), both used, the
    // cosine value first.  There is no rejection step, so all lanes of a
    // wavefront consume uniforms in step and every Philox block is formed once
    // for the whole wavefront.  |g| <= sqrt(-2 ln 2^-33) = 6.8 always, below
    // every truncation used; a value beyond nsigmax would be skipped.
    __device__ __forceinline__ double gauss(double nsigmax)
    {
        for (;;) {
            if (have) {
                have = 0;
                if (fabs(pending) <= nsigmax) return pending;
                continue;
            }
            const double u1 = next();
            const double u2 = next();
            const double r = sqrt(-2.0 * bm_log(u1));
            double sn, cs;
            bm_sincos2pi(u2, sn, cs);
            pending = r * sn;
            have = 1;
            const double g1 = r * cs;
            if (fabs(g1) <= nsigmax) return g1;
        }
    }
    // Two consecutive gauss(nsigmax) draws for nsigmax >= 7, where no
    // Box-Muller value is ever rejected (|g| <= 6.8): with no value pending
    // that is exactly one pair, cosine value first.
    __device__ __forceinline__ void gauss_pair(double& a, double& b)
    {
        if (have) {
            a = gauss(99.0);
            b = gauss(99.0);
            return;
        }
        const double u1 = next();
        const double u2 = next();
        const double r = sqrt(-2.0 * bm_log(u1));
        double sn, cs;
        bm_sincos2pi(u2, sn, cs);
        a = r * cs;
        b = r * sn;
        pending = b;
    }
    __device__ __forceinline__ int32_t used() const { return (int32_t)pos; }
};
using PhiloxStream = PhiloxStreamT<kPhiloxTag>;
using PhiloxLaneStream = PhiloxStreamT<kPhiloxTag, true>;

// The detector hut's own counter-based stream: uniform k of an event's hut
// is word k%4 of the Philox block (k/4, event lo, event hi, kHutTag) under
// the event's key.  Every hut step draws one uniform (resmult) or Box-Muller
// pairs (both values used, cosine first), so nothing is pending between
// steps.  The draws depend on launch-uniform flags only: the position is the
// same in every lane of a wavefront, so block refills and word selection are
// scalar branches.
template <bool LANE_KEY>
struct HutPhilox {
    uint32_t k0, k1, e0, e1;
    uint32_t pos;
    uint32_t w[4];
    template <uint32_t TAG>
    __device__ __forceinline__ void init_from(const PhiloxStreamT<TAG, LANE_KEY>& g)
    {
        k0 = g.k0; k1 = g.k1; e0 = g.e0; e1 = g.e1;
        pos = 0;
    }
    __device__ __forceinline__ double next()
    {
        const uint32_t j = pos & 3u;
        if (j == 0) philox4x32_10<LANE_KEY>(pos >> 2, e0, e1, kHutTag, k0, k1, w);
        ++pos;
        const uint32_t v = j == 0 ? w[0] : j == 1 ? w[1] : j == 2 ? w[2] : w[3];
        return __builtin_fma((double)v, 0x1p-32, 0x1p-33);  // (v + 1/2) 2^-32, exact
    }
    __device__ __forceinline__ void gauss_pair(double& a, double& b)
    {
        const double u1 = next();
        const double u2 = next();
        const double r = sqrt(-2.0 * bm_log(u1));
        double sn, cs;
        bm_sincos2pi(u2, sn, cs);
        a = r * cs;
        b = r * sn;
    }
    __device__ __forceinline__ uint32_t used() const { return pos; }
};

// The stream the detector hut draws from: its own for the counter-based
// source; a tape (the reference's sequence) continues through the hut.
template <class R>
struct HutStreamOf {
    using type = R;
    static constexpr bool kOwn = false;
};
template <uint32_t TAG, bool LANE_KEY>
struct HutStreamOf<PhiloxStreamT<TAG, LANE_KEY>> {
    using type = HutPhilox<LANE_KEY>;
    static constexpr bool kOwn = true;
};

template <class R>
__device__ __forceinline__ double polar_gauss(R& g, double nsigmax);

// Tape of the launch: event slot at byte offset `base`, `stride` bytes long.
// Normal deviates use the reference's Marsaglia polar method (gauss1.f).
struct TapeStream {
    static constexpr bool kLaneKey = false;
    const double* tape;
    uint32_t base, stride, pos;
    int overflow;
    int have;          // unused (kept for a common state layout)
    double pending;
    __device__ __forceinline__ void init(uint64_t, uint64_t, const double* t, uint32_t slot_off, uint32_t s)
    {
        tape = t; base = slot_off; stride = s; pos = 0; overflow = 0; have = 0; pending = 0.0;
    }
    __device__ __forceinline__ uint32_t state_word() const { return pos; }
    __device__ __forceinline__ void resume(uint32_t word, double) { pos = word; }
    __device__ __forceinline__ double next()
    {
        if (pos >= stride) { overflow = 1; pos += 8u; return 0.5; }
        const double v = __builtin_bit_cast(
            double, __builtin_amdgcn_raw_buffer_load_b64(rsrc(tape, 0xffffffffu), base + pos, 0, 0));
        pos += 8u;
        return v;
    }
    __device__ __forceinline__ double gauss(double nsigmax) { return polar_gauss(*this, nsigmax); }
    __device__ __forceinline__ void gauss_pair(double& a, double& b)
    {
        a = polar_gauss(*this, 99.0);
        b = polar_gauss(*this, 99.0);
    }
    __device__ __forceinline__ int32_t used() const { return (int32_t)(pos / 8u); }
};

// ---------------------------------------------------------------------------
// Per-event track state: common /track/ plus the decay bookkeeping.
// ---------------------------------------------------------------------------
struct Track {
    double xs, ys, dxdzs, dydzs, dpps;
    double p, m2, pathlen, decdist;
    int dflag;
};

// Per-event settings of the run (one kinematic setting).
struct Flags {
    double ctau;     // decay length c tau (cm)
    double p_spec;   // central momentum (MeV/c)
    double sigma;    // wire-chamber resolution (cm)
    int ms, wcs, decay, coll, fit_f32;
    int sieve;       // the run's mock sieve options: bit 0 use_sieve, bit 1 the SHMS's use_2sieve
};

// Rotation constants tan/sin/cos of rotang * 0.017453292 (the truncated
// degree constant of shared/rotate_haxis.f:41), evaluated on the host with
// the same libm calls the reference makes.
struct Rot {
    double tan_th, sin_th, cos_th;
};

// ---------------------------------------------------------------------------
// gauss1, gauss1.f:17-27: Marsaglia polar method, second normal discarded.
// ---------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ double polar_gauss(R& g, double nsigmax)
{
    for (;;) {
        const double u1 = g.next();
        const double u2 = g.next();
        if (g.overflow) return 0.0;
        const double v1 = 2.0 * u1 - 1.0;
        const double v2 = 2.0 * u2 - 1.0;
        const double s = v1 * v1 + v2 * v2;
        if (s > 1.0 || s == 0.0) continue;
        const double r = v1 * sqrt(-(2.0 * log(s) / s));
        if (fabs(r) > nsigmax) continue;
        return r;
    }
}

// A normal deviate from the event's stream (polar method on a tape,
// Box-Muller pairs on the counter-based stream).
template <class R>
__device__ __forceinline__ double gauss1(R& g, double nsigmax)
{
    return g.gauss(nsigmax);
}

// theta_sigma of musc / musc_ext (shared/musc.f:48-51, shared/musc_ext.f:36-39)
__device__ __forceinline__ double theta_sigma(double m2, double p, double rad_len)
{
    const double Es = 13.6, epsilon = 0.088;
    const double beta = p / sqrt(m2 + p * p);
    return Es / p / beta * sqrt(rad_len) * (1.0 + epsilon * log10(rad_len / (beta * beta)));
}

// theta_sigma for the detector-hut steps, whose radiation lengths are
// constants: sqrt(rad_len) and log10(rad_len) come with the step, the track
// factors Es/p/beta and log10(beta^2) are formed once per momentum, and
// log10(rad_len/beta^2) is taken as log10(rad_len) - log10(beta^2) (the same
// value to a few units in the last place).
struct ThetaCache {
    double p = -1.0, m2 = -1.0, a = 0.0, lb = 0.0;
    __device__ __forceinline__ double theta(const Track& t, double sqrt_rl, double log10_rl)
    {
        const double Es = 13.6, epsilon = 0.088;
        if (t.p != p || t.m2 != m2) {
            p = t.p;
            m2 = t.m2;
            const double beta = p / sqrt(m2 + p * p);
            a = Es / p / beta;
            lb = log10(beta * beta);
        }
        return a * sqrt_rl * (1.0 + epsilon * (log10_rl - lb));
    }
};

// The mock sieve slit (hms/mc_hms.f:155-168, shms/mc_shms.f:276-310,
// sos/mc_sos.f:122-137, hrsl/mc_hrsl.f:122-132): the track from the target
// point is forced through the nearest hole of a sieve at z, at an offset drawn
// inside the hole (two uniforms), by changing its angles.  kind 0: holes on a
// (px, py) grid; 1: the SOS's rows (middle three at 1.016 cm, outer ones on a
// 2.54 cm pitch from 1.016); 2: the HRS's (the offset is drawn, not applied).
// NINT is round half away from zero.
template <class R>
__device__ __forceinline__ void sieve_slit(R& g, Track& t, int kind, double z, double px, double py, double r)
{
    const double x = t.xs, y = t.ys;
    double xt = x + z * t.dxdzs;
    double yt = y + z * t.dydzs;
    xt = px * round(xt / px);
    if (kind == 1) {
        if (fabs(yt) < 1.524) yt = 1.016 * round(yt / 1.016);
        else yt = yt / fabs(yt) * (2.54 * round((fabs(yt) - 1.016) / 2.54) + 1.016);
    } else {
        yt = py * round(yt / py);
    }
    const double rt = r * sqrt(g.next());
    const double tht = 2.0 * kPi * g.next();
    if (kind != 2) {
        xt = xt + rt * cos(tht);
        yt = yt + rt * sin(tht);
    }
    t.dxdzs = (xt - x) / z;
    t.dydzs = (yt - y) / z;
}

// musc(m2, p, rad_len, dth, dph), shared/musc.f:43-57
template <class R>
__device__ __forceinline__ void musc(R& g, const Track& t, double rad_len, double& dth, double& dph)
{
    if (rad_len == 0.0) return;
    const double ts = theta_sigma(t.m2, t.p, rad_len);
    dth = dth + ts * gauss1(g, 99.0);
    dph = dph + ts * gauss1(g, 99.0);
}

// musc_ext(m2, p, rad_len, x_len, dph, dth, y, x), shared/musc_ext.f:25-51.
// Every call site passes (dydzs, dxdzs, ys, xs).
template <class R>
__device__ __forceinline__ void musc_ext(R& g, Track& t, double rad_len, double x_len)
{
    if (rad_len == 0.0) return;
    if (x_len <= 0.0 || rad_len < 0.0) return;
    const double ts = theta_sigma(t.m2, t.p, rad_len);
    const double s12 = 3.4641016151377544;  // sqrt(12.)
    double g1 = gauss1(g, 99.0);
    double g2 = gauss1(g, 99.0);
    t.dxdzs = t.dxdzs + ts * g1;
    t.xs = t.xs + ts * x_len * g2 / s12 + ts * x_len * g1 / 2.0;
    g1 = gauss1(g, 99.0);
    g2 = gauss1(g, 99.0);
    t.dydzs = t.dydzs + ts * g1;
    t.ys = t.ys + ts * x_len * g2 / s12 + ts * x_len * g1 / 2.0;
}

// loren, loren.f:18-23
__device__ __forceinline__ void loren(double gam, double bx, double by, double bz, double e,
                                      double x, double y, double z, double& pxf, double& pyf,
                                      double& pzf, double& pf1)
{
    const double gam1 = gam * gam / (1.0 + gam);
    pxf = (1.0 + gam1 * (bx * bx)) * x + gam1 * bx * (by * y + bz * z) - gam * bx * e;
    pyf = (1.0 + gam1 * (by * by)) * y + gam1 * by * (bx * x + bz * z) - gam * by * e;
    pzf = (1.0 + gam1 * (bz * bz)) * z + gam1 * bz * (by * y + bx * x) - gam * bz * e;
    pf1 = sqrt(pxf * pxf + pyf * pyf + pzf * pzf);
}

// Decay products and boost: shared/project.f:72-109, shared/transp.f:146-184.
// m_kk is the final mass on the kaon two-body branch (Mk in the first half of
// transp, transp.f:158; Mpi elsewhere).
template <class R>
__device__ void decay_products(R& g, Track& t, double beta, double gamma, double p_spec, double m_kk)
{
    const double rph = g.next() * 2.0 * kPi;
    const double rth1 = g.next() * 2.0 - 1.0;
    const double rth = acos(rth1);
    double pr = 0.0, m_final = kMmu;
    if (fabs(sqrt(t.m2) - kMpi) < 2.0) pr = 29.783;
    if (fabs(sqrt(t.m2) - kMk) < 2.0) {
        if (g.next() < 0.7) {
            pr = 235.5;
        } else {
            pr = sqrt(kMk * kMk / 4.0 - kMpi * kMpi);
            m_final = m_kk;
        }
    }
    const double er = sqrt(m_final * m_final + pr * pr);
    const double pxr = pr * sin(rth) * cos(rph);
    const double pyr = pr * sin(rth) * sin(rph);
    const double pzr = pr * cos(rth);
    const double s = sqrt(1.0 + t.dxdzs * t.dxdzs + t.dydzs * t.dydzs);
    const double bx = -(beta * t.dxdzs / s);
    const double by = -(beta * t.dydzs / s);
    const double bz = -(beta * 1.0 / s);
    double pxf, pyf, pzf, pf;
    loren(gamma, bx, by, bz, er, pxr, pyr, pzr, pxf, pyf, pzf, pf);
    t.dxdzs = pxf / pzf;
    t.dydzs = pyf / pzf;
    t.dpps = 100.0 * (pf / p_spec - 1.0);
    t.p = pf;
    t.m2 = m_final * m_final;
}

// project, shared/project.f:38-119 applied to (xs, ys)
template <class R>
__device__ __forceinline__ void project(R& g, Track& t, const Flags& F, double z_drift)
{
    if (!F.decay || t.dflag) {
        t.pathlen = t.pathlen + z_drift * sqrt(1.0 + t.dxdzs * t.dxdzs + t.dydzs * t.dydzs);
        t.xs = t.xs + t.dxdzs * z_drift;
        t.ys = t.ys + t.dydzs * z_drift;
        return;
    }
    const double p_spec = t.p / (1.0 + t.dpps / 100.0);
    const double beta = t.p / sqrt(t.p * t.p + t.m2);
    const double gamma = 1.0 / sqrt(1.0 - beta * beta);
    const double dlen = F.ctau * beta * gamma;
    const double z_decay = -1.0 * dlen * log(1.0 - g.next());
    const double sq = sqrt(1.0 + t.dxdzs * t.dxdzs + t.dydzs * t.dydzs);
    if (z_decay > z_drift * sq) {
        t.decdist = t.decdist + z_drift * sq;
        t.pathlen = t.pathlen + z_drift * sq;
        t.xs = t.xs + t.dxdzs * z_drift;
        t.ys = t.ys + t.dydzs * z_drift;
    } else {
        t.dflag = 1;
        t.decdist = t.decdist + z_decay;
        t.pathlen = t.pathlen + z_decay;
        t.xs = t.xs + t.dxdzs * z_decay / sq;
        t.ys = t.ys + t.dydzs * z_decay / sq;
        decay_products(g, t, beta, gamma, p_spec, kMpi);
        const double sq2 = sqrt(1.0 + t.dxdzs * t.dxdzs + t.dydzs * t.dydzs);
        const double tmpdrift = z_drift - z_decay / sq2;
        t.pathlen = t.pathlen + tmpdrift * sq2;
        t.xs = t.xs + t.dxdzs * tmpdrift;
        t.ys = t.ys + t.dydzs * tmpdrift;
    }
}

// transp, shared/transp.f:134-279.  POLY(r0..r4, o0..o4) evaluates the class.
template <class R, class POLY>
__device__ __forceinline__ void transp(R& g, Track& t, const Flags& F, double zd, POLY poly)
{
    double p_spec = 0.0, beta = 0.0, gamma = 0.0, z_decay = 0.0;
    const bool check = F.decay && !t.dflag;
    if (check) {
        p_spec = t.p / (1.0 + t.dpps / 100.0);
        beta = t.p / sqrt(t.p * t.p + t.m2);
        gamma = 1.0 / sqrt(1.0 - beta * beta);
        const double dlen = F.ctau * beta * gamma;
        z_decay = -1.0 * dlen * log(1.0 - g.next());
        if (z_decay <= zd / 2.0) {
            t.dflag = 1;
            t.decdist = t.decdist + z_decay;
            decay_products(g, t, beta, gamma, p_spec, kMk);
        }
    }
    double o0, o1, o2, o3, o4;
    poly(t.xs, t.dxdzs * 1000.0, t.ys, t.dydzs * 1000.0, t.dpps, o0, o1, o2, o3, o4);
    t.xs = o0;
    t.dxdzs = o1 / 1000.0;
    t.ys = o2;
    t.dydzs = o3 / 1000.0;
    const double delta_z = -o4;
    if (check && !t.dflag) {
        if (z_decay > zd + delta_z) {
            t.decdist = t.decdist + (zd + delta_z);
        } else {
            t.dflag = 1;
            t.decdist = t.decdist + z_decay;
            decay_products(g, t, beta, gamma, p_spec, kMpi);
        }
    }
    t.pathlen = t.pathlen + (zd + delta_z);
}

// rotate_haxis, shared/rotate_haxis.f:43-61: new (xp0, yp0) from (xs, ys)
__device__ __forceinline__ void rotate_haxis(const Track& t, const Rot& r, double& xp0, double& yp0)
{
    const double alpha = t.dxdzs, beta = t.dydzs;
    const double alpha_p = (alpha + r.tan_th) / (1.0 - alpha * r.tan_th);
    const double beta_p = beta / (r.cos_th - alpha * r.sin_th);
    const double xi = xp0;
    xp0 = xi * (r.cos_th + alpha_p * r.sin_th);
    yp0 = yp0 + xi * beta_p * r.sin_th;
}

// rotate_vaxis, shared/rotate_vaxis.f:42-57
__device__ __forceinline__ void rotate_vaxis(const Track& t, const Rot& r, double& xp0, double& yp0)
{
    const double alpha = t.dydzs, beta = t.dxdzs;
    const double alpha_p = (alpha + r.tan_th) / (1.0 - alpha * r.tan_th);
    const double beta_p = beta / (r.cos_th - alpha * r.sin_th);
    const double yi = yp0;
    yp0 = yi * (r.cos_th + alpha_p * r.sin_th);
    xp0 = xp0 + yi * beta_p * r.sin_th;
}

// enerloss_new, enerloss_new.f:19-82
template <class R>
__device__ double enerloss_new(R& g, double len, double dens, double zeff, double aeff,
                               double epart, double mpart, int typeflag)
{
    const double me = 0.51099906;
    const double thick = len * dens;
    const double gamma = epart / mpart;
    const double beta = sqrt(1.0 - 1.0 / (gamma * gamma));
    const double I = zeff == 1.0 ? 21.8e-06 : (16.0 * pow(zeff, 0.9)) * 1.0e-06;
    const double hnup = 28.816e-06 * sqrt(dens * zeff / aeff);
    const double log10bg = log(beta * gamma) / log(10.0);
    const double CO = log(hnup) - log(I) + 0.5;
    double denscorr;
    if (log10bg < 0.0) denscorr = 0.0;
    else if (log10bg < 3.0) {
        const double t3 = 3.0 - log10bg;
        denscorr = CO + log(10.0) * log10bg + fabs(CO / 27.0) * (t3 * t3 * t3);
    } else if (log10bg < 4.7) denscorr = CO + log(10.0) * log10bg;
    else denscorr = CO + log(10.0) * 4.7;
    double E;
    if (thick <= 0.0) {
        E = 0.0;
    } else {
        const double b2 = beta * beta;
        const double mp_new = 0.1536e-03 * zeff / aeff * thick / b2 *
                              (log(me / (I * I)) + 1.063 + 2.0 * log(gamma * beta) +
                               log(0.1536 * zeff / aeff * thick / b2) - b2 - denscorr);
        const double mp = mp_new * 1000.0;
        const double chsi = 0.307075 / 2.0 * zeff / aeff * thick / b2;
        double x = 0.0;
        if (typeflag == 1) x = fabs(gauss1(g, 10.0));
        else if (typeflag == 2) x = 3.0;
        else if (typeflag == 3) x = 0.0067;
        else if (typeflag == 4) x = 1.0;
        const double lambda = x > 0.0 ? -2.0 * log(x) : 100000.0;
        E = lambda * chsi + mp;
    }
    if (E > (epart - mpart)) E = (epart - mpart) - 0.0000001;
    return E;
}

// sigA * 1.d-27 of pion_coll_absorb.f:223,229: the DOUBLE PRECISION literal is
// a 16-byte real under -fdefault-real-8, so the product is the correctly
// rounded value of sigA times the quad-precision 1e-27, formed here from a
// double-double split of that constant with an exact FMA product.
__device__ __forceinline__ double times_1em27_quad(double a)
{
    const double hi = 1.0e-27;
    const double lo = -3.849486974919184e-44;  // 1e-27 - hi, to double precision
    const double p = a * hi;
    const double e = __builtin_fma(a, hi, -p);
    return p + (e + a * lo);
}

// pion_coll_absorb, hms/pion_coll_absorb.f:54-90: transmission probability
__device__ double pion_coll_absorb(double ppi, double thick)
{
    const double T[14] = {85.0, 125.0, 165.0, 205.0, 245.0, 315.0, 584.02, 711.95, 870.12,
                          1227.57, 1446.58, 1865.29, 2858.0, 4159.0};
    const double sigreac[14] = {26.03, 84.47, 117.3, 117.4, 101.9, 69.58, 42.5, 44.7, 47.9,
                                46.5, 45.2, 39.6, 35.34, 33.15};
    const double qreac[14] = {0.948, 0.659, 0.56, 0.5342, 0.5452, 0.60796, 0.699, 0.689,
                              0.679, 0.683, 0.688, 0.704, 0.7483, 0.7705};
    const double Navagadro = 6.0221367e+23, mpi = 139.56995;
    const double mate_dens = 17.0, mate_A = 171.57;
    const double Epi = sqrt(ppi * ppi + mpi * mpi);
    const double Tpi = Epi - mpi;
    double sigA = 0.0;
    for (int i = 0; i < 13; ++i) {
        if (Tpi > T[i] && Tpi <= T[i + 1]) {
            const double Thi = T[i + 1], Tlo = T[i];
            const double sigAhi = sigreac[i + 1] * pow(mate_A, qreac[i + 1]);
            const double sigAlo = sigreac[i] * pow(mate_A, qreac[i]);
            sigA = (sigAlo * (Thi - Tpi) + sigAhi * (Tpi - Tlo)) / (Thi - Tlo);
            sigA = times_1em27_quad(sigA);
        }
    }
    if (Tpi > T[13]) sigA = times_1em27_quad(sigreac[13] * pow(mate_A, qreac[13]));
    const double lambdai = mate_dens * Navagadro * sigA / mate_A;
    return exp(-(0.0 + thick * lambdai));
}

// mc_hms_coll / mc_shms_coll, hms/mc_hms_coll.f:50-136, shms/mc_shms_coll.f.
// Returns success; per-step slit hits go to cnt_step[0..2] (hor, vert, oct).
struct CollGeom {
    double h_entr, v_entr, h_exit, v_exit, x_off, y_off, coll_radl, step_size;
};

template <class R>
__device__ bool coll_step(R& g, Track& t, const Flags& F, const CollGeom& G, double p_spec,
                          unsigned* cnt_step)
{
    const double coll_dens = 17.0, zcoll = 69.45, acoll = 171.56797;
    const int nstep = 20;
    double h_step = G.h_entr, v_step = G.v_entr;
    double epart = sqrt(t.p * t.p + t.m2);
    for (int n = 1; n <= nstep; ++n) {
        bool step_flag = false;
        if (fabs(t.ys - G.y_off) > h_step) { step_flag = true; ++cnt_step[0]; }
        if (fabs(t.xs - G.x_off) > v_step) { step_flag = true; ++cnt_step[1]; }
        if (fabs(t.xs - G.x_off) > (-(v_step / h_step * fabs(t.ys - G.y_off)) + 3.0 * v_step / 2.0)) {
            step_flag = true;
            ++cnt_step[2];
        }
        const double thick = G.step_size;
        if (step_flag) {
            if (t.m2 > 12000.0 && t.m2 < 20000.0) {
                const double trans = pion_coll_absorb(t.p, thick);
                if (g.next() > trans) return false;
            }
            musc(g, t, thick / G.coll_radl, t.dydzs, t.dxdzs);
            epart = sqrt(t.p * t.p + t.m2);
            const double eloss = enerloss_new(g, thick, coll_dens, zcoll, acoll, epart, sqrt(t.m2), 1);
            epart = epart - eloss;
            if (epart < sqrt(t.m2)) return false;
            t.p = sqrt(epart * epart - t.m2);
            t.dpps = 100.0 * (t.p / p_spec - 1.0);
        }
        project(g, t, F, G.step_size);
        h_step = h_step + (G.h_exit - G.h_entr) / nstep;
        v_step = v_step + (G.v_exit - G.v_entr) / nstep;
    }
    return true;
}

// Fixed-collimator octagon test (hms/mc_hms.f:225-253, shms/mc_shms.f:524-565).
// Returns 0 inside, else the stop code 1 (hor), 2 (vert), 3 (oct).
__device__ __forceinline__ int slit(double x, double y, double h, double v, double x_off, double y_off)
{
    if (fabs(y - y_off) > h) return 1;
    if (fabs(x - x_off) > v) return 2;
    if (fabs(x - x_off) > (-(v / h * fabs(y - y_off)) + 3.0 * v / 2.0)) return 3;
    return 0;
}

// LFIT, cern/lfit.f:22-59 with KEY = 0.  T is float for the reference's
// REAL*4 hits (sums in FP64) or double for the FP64 fit.
template <class T, int L>
__device__ __forceinline__ void lfit(const T (&X)[L], const T (&Y)[L], T& A, T& B)
{
    double count = 0.0, sumx = 0.0, sumy = 0.0, sumxy = 0.0, sumxx = 0.0;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        if (Y[j] == (T)0) continue;
        sumx = sumx + (double)X[j];
        sumy = sumy + (double)Y[j];
        count = count + 1.0;
    }
    if (count <= 1.0) { A = (T)0; B = (T)0; return; }
    const double ymed = sumy / count;
    const double xmed = sumx / count;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        if (Y[j] == (T)0) continue;
        const double sx = (double)X[j] - xmed;
        const double sy = (double)Y[j] - ymed;
        sumxy = sumxy + sx * sy;
        sumxx = sumxx + sx * sx;
    }
    if (sumxx == 0.0) { A = (T)0; B = (T)0; return; }
    A = (T)(sumxy / sumxx);
    B = (T)(ymed - (double)A * xmed);
}

}  // namespace simcdev
