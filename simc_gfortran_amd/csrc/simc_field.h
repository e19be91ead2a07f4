// Target-field tracking on the device: trg_track.f's trgField, trgDeriv,
// trgRK4 and trgTrackToPlane for one event per lane.
//
// The field map (51 x 51 nodes of B_z and B_r on a 2 cm grid in |z'| and r,
// z' along the field axis) is staged in LDS by the kernel; the arithmetic is
// the reference's, operation for operation, so that with FP contraction off
// every lane reproduces the reference's doubles.
#pragma once

#include <hip/hip_runtime.h>

namespace simcdev {

constexpr int kTrgN = 51;                 // nz = nr = 51, trg_track.f:257-258
constexpr int kTrgNodes = kTrgN * kTrgN;  // B(iz, ir) at [ir * 51 + iz]

// The map of one launch: B_z, B_r (LDS) and the field angle of the arm
// (sin, cos of theta * pi/180 with pi = 3.141592653, trg_track.f:271-281).
struct TrgMap {
    const double* bz;
    const double* br;
    double stht, ctht;
};

// trgField, trg_track.f:328-418.  The grid starts at 0 with spacing 2, so
// zz(i) = 2 (i - 1) and (z - zz(1)) / (zz(2) - zz(1)) = z / 2 exactly.
// Points at |z'| or r >= 200 cm are outside the grid (field 0) before the
// integer conversion, which keeps it in range.
__device__ __forceinline__ void trg_field(const TrgMap& M, double x0, double x1, double x2, double& B0,
                                          double& B1, double& B2)
{
    const double y1 = M.stht * x2 + M.ctht * x1;
    const double y2 = M.ctht * x2 - M.stht * x1;
    const double z = fabs(y2);
    const double r = sqrt(x0 * x0 + y1 * y1);
    B0 = 0.0;
    B1 = 0.0;
    B2 = 0.0;
    if (!(z < 200.0) || !(r < 200.0)) return;
    const int i = (int)(z / 2.0) + 1;
    const int j = (int)(r / 2.0) + 1;
    if ((i + 1 > kTrgN) || (i < 1) || (j + 1 > kTrgN) || (j < 1)) return;
    const int i0 = i - 1, j0 = j - 1;
    const double az = ((z - 2.0 * (double)i0) / 2.0);
    const double ar = ((r - 2.0 * (double)j0) / 2.0);
    const double* bz = M.bz + j0 * kTrgN + i0;
    double a0 = az * (bz[1] - bz[0]) + bz[0];
    double a1 = az * (bz[kTrgN + 1] - bz[kTrgN]) + bz[kTrgN];
    const double Bz = (ar * (a1 - a0) + a0);
    if (r > 0.0) {
        const double* br = M.br + j0 * kTrgN + i0;
        a0 = az * (br[1] - br[0]) + br[0];
        a1 = az * (br[kTrgN + 1] - br[kTrgN]) + br[kTrgN];
        double b = (ar * (a1 - a0) + a0) / r;
        if (y2 < 0.0) b = -b;
        const double Bx = b * x0, By = b * y1;
        B0 = Bx;
        B1 = -M.stht * Bz + M.ctht * By;
        B2 = M.ctht * Bz + M.stht * By;
    } else {
        B1 = -M.stht * Bz;
        B2 = M.ctht * Bz;
    }
}

// trgDeriv, trg_track.f:423-459: (v, (v x B) * factor); the integral terms
// u(7..9) are not carried (trgRK4 does not use them).
__device__ __forceinline__ void trg_deriv(const TrgMap& M, const double (&u)[6], double (&d)[6], double factor)
{
    double B0, B1, B2;
    trg_field(M, u[0], u[1], u[2], B0, B1, B2);
    d[0] = u[3];
    d[1] = u[4];
    d[2] = u[5];
    const double c0 = u[4] * B2 - u[5] * B1;
    const double c1 = u[5] * B0 - u[3] * B2;
    const double c2 = u[3] * B1 - u[4] * B0;
    d[3] = c0 * factor;
    d[4] = c1 * factor;
    d[5] = c2 * factor;
}

// trgRK4, trg_track.f:461-498
__device__ __forceinline__ void trg_rk4(const TrgMap& M, const double (&u0)[6], double (&u1)[6], double h,
                                        double factor)
{
    double ut[6], dudt[6], dut[6], dum[6];
    const double hh = h * 0.5;
    const double h6 = h / 6.;
    trg_deriv(M, u0, dudt, factor);
#pragma unroll
    for (int i = 0; i < 6; i++) ut[i] = u0[i] + hh * dudt[i];
    trg_deriv(M, ut, dut, factor);
#pragma unroll
    for (int i = 0; i < 6; i++) ut[i] = u0[i] + hh * dut[i];
    trg_deriv(M, ut, dum, factor);
#pragma unroll
    for (int i = 0; i < 6; i++) {
        ut[i] = u0[i] + h * dum[i];
        dum[i] = dut[i] + dum[i];
    }
    trg_deriv(M, ut, dut, factor);
#pragma unroll
    for (int i = 0; i < 6; i++) u1[i] = u0[i] + h6 * (dudt[i] + dut[i] + 2. * dum[i]);
}

__device__ __forceinline__ double sgn1(double v) { return copysign(1.0, v); }

// Step cap of one plane search.  The reference's bound is
// int(max(dist0, 10 dl) / dl) * 10 (trg_track.f:209); a start beyond 1e5
// step lengths from the plane (1 km at dl = 1 cm) is capped there, so the
// integer conversion stays in range and every lane ends.
constexpr double kTrgMaxDist = 1.0e5;

// trgTrackToPlane, trg_track.f:154-238: step (time step -dl/|v|, reversed
// when the first step moves away) until a x + b y + c z + d = 0 is crossed,
// then interpolate u linearly between the two last states.
__device__ __forceinline__ void trg_track_to_plane(const TrgMap& M, double (&u)[6], double E, double dl, double a,
                                                   double b, double c, double d, bool& ok)
{
    if (!ok) return;
    const double n = 1 / sqrt(a * a + b * b + c * c);
    const double an = a * n, bn = b * n, cn = c * n, dn = d * n;
    const double factor = 90. / E;
    double ts = -dl / sqrt(u[3] * u[3] + u[4] * u[4] + u[5] * u[5]);
    double u0[6], u1[6];
    double dist0 = u[0] * an + u[1] * bn + u[2] * cn + dn;
    const double maxdist = fmax(fabs(dist0) * 4., 1.0);
    trg_rk4(M, u, u1, ts, factor);
    double dist1 = u1[0] * an + u1[1] * bn + u1[2] * cn + dn;
    if ((sgn1(dist0) == sgn1(dist1)) && (fabs(dist0) < fabs(dist1))) ts = -ts;
    const int max_steps = (int)fmin(fmax(dist0, 10. * dl) / dl, kTrgMaxDist) * 10;
    int steps = 0;
    if (sgn1(dist0) == sgn1(dist1)) {
        dist1 = dist0;
        while ((sgn1(dist0) == sgn1(dist1)) && ok) {
            trg_rk4(M, u1, u0, ts, factor);
            dist0 = u0[0] * an + u0[1] * bn + u0[2] * cn + dn;
            if (sgn1(dist0) == sgn1(dist1)) {
                trg_rk4(M, u0, u1, ts, factor);
                dist1 = u1[0] * an + u1[1] * bn + u1[2] * cn + dn;
            }
            ok = (fabs(dist1) < maxdist) && steps < max_steps;
            steps = steps + 1;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 6; i++) u0[i] = u[i];
    }
    if (ok) {
#pragma unroll
        for (int i = 0; i < 6; i++) u[i] = u0[i] + (u1[i] - u0[i]) * dist0 / (dist0 - dist1);
    }
}

// Start vector of a track from position and slopes (trg_track.f:628-636,
// :776-781; the two routines sum the slope squares in opposite orders).
__device__ __forceinline__ void trg_start(double (&vT)[6], double x, double y, double z, double dx, double dy,
                                          double vel, bool dx_first)
{
    vT[0] = x;
    vT[1] = y;
    vT[2] = z;
    vT[5] = dx_first ? vel / sqrt(1 + dx * dx + dy * dy) : vel / sqrt(1 + dy * dy + dx * dx);
    vT[3] = dx * vT[5];
    vT[4] = dy * vT[5];
}

// track_from_tgt, trg_track.f:591-658: to the plane z = 0, then to z = 100 cm;
// the status of the first search is not kept.
__device__ __forceinline__ bool track_from_tgt(const TrgMap& M, double& x, double& y, double& z, double& dx,
                                               double& dy, double mom, double mass)
{
    const double cc = 29.9792458;
    const double vel = fabs(mom) / sqrt(mom * mom + mass * mass) * cc;
    const double eng = copysign(1., mom) * sqrt(mom * mom + mass * mass);
    double vT[6];
    trg_start(vT, x, y, z, dx, dy, vel, true);
    bool ok = true;
    trg_track_to_plane(M, vT, eng, 1., 0., 0., 1., 0., ok);
    ok = true;
    trg_track_to_plane(M, vT, eng, 1., 0., 0., 1., -100., ok);
    x = vT[0];
    y = vT[1];
    z = vT[2];
    dx = vT[3] / vT[5];
    dy = vT[4] / vT[5];
    return ok;
}

}  // namespace simcdev
