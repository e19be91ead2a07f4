// Step programs of the spectrometer arms.
//
// An arm (mc_hms, mc_shms, mc_sos, mc_hrsr, mc_hrsl and their detector huts) is a fixed sequence of
// drifts, COSY map segments, aperture tests and detector-hut steps.  The host
// writes that sequence once per context as an array of Op records; one
// kernel per arm walks it for every event (all lanes of a wavefront execute
// the same op, so dispatch is uniform and the code of each step exists once).
// Every numeric parameter is evaluated on the host with the reference's own
// expressions (cited where the table is built, simc_kernels.hip).
#pragma once

#include <stdint.h>

namespace simcdev {

enum OpCode : int32_t {
    OP_END = 0,       // accepted: reconstruct and stop
    OP_PROJECT,       // drift p0 (shared/project.f)
    OP_TRANSP,        // COSY class in slot arg, central length p0 (shared/transp.f)
    OP_CHECK_R2,      // stop if xs^2 + ys^2 > p0
    OP_ROT_H,         // (xt,yt) = rotate_haxis(xs,ys; tan p0, sin p1, cos p2) + (p3, p4)
    OP_ROT_V,         // (xt,yt) = rotate_vaxis(xs,ys; tan p0, sin p1, cos p2) + (p3, p4)
    OP_CHECK_T_R2,    // stop if xt^2 + yt^2 > p0
    OP_CHECK_T_HB,    // stop if xt^2 > p0 or yt > p1 or yt < p2
    OP_HIT_DIPOLE_T,  // stop if hit_dipole(xt, yt) (HMS dipole shape)
    OP_PIPE_T,        // stop if (xt-p1)^2 + (yt-p2)^2 > p0 or |yt-p2| > p3
    OP_PIPE_S,        // stop if (xs-p1)^2 + (ys-p2)^2 > p0
    OP_COLLIMATOR,    // pions/muons: step through (arm geometry); else octagons
                      // (h p0, v p1) at entry, drift p6, (h p2, v p3) at exit; offsets p4, p5
    OP_RESMULT_DRAW,  // resmult = 2 if uniform < p0 else 1   (hms/mc_hms_hut.f:292-298)
    OP_RESMULT_ONE,   // resmult = 1                          (shms/mc_shms_hut.f:80)
    OP_MUSC,          // musc with rad_len p0 (p1 = sqrt p0, p2 = log10 p0)
    OP_MUSC_EXT,      // musc_ext with rad_len p0, length p1 (p2 = sqrt p0, p3 = log10 p0)
    OP_WIRE,          // wire-plane hit of plane `stop` (0-11); arg = 1: y plane
    OP_BOX,           // stop if xs > p0 or xs < p1 or ys > p2 or ys < p3
    OP_SCIN,          // count++ if ys < p0 and ys > p1 and xs < p2 and xs > p3
    OP_TRIGGER,       // stop if count < arg
    OP_FIT,           // straight-line fit of the 12 planes -> x_fp, x'_fp, y_fp, y'_fp
    OP_FIDUCIAL,      // track at z = p0: stop if yc > p1 or yc < p2 or xc > p3 or xc < p4
    // magnet-stage steps of the SOS (sos/mc_sos.f:236-358)
    OP_CHECK_T_BOX,   // stop if yt > p0 or -yt > p1 or -xt > p2 or -xt < p3
    OP_CHECK_TRAP,    // stop if |xs| > p0 or |ys| > p1 + p2 (xs + p0)
    OP_SHIFT,         // xs += p0 x', ys += p0 y' (no path length, no decay)
    OP_CHECK_ABS,     // stop if |xs| > p0 or |ys| > p1
    // magnet-stage steps of the HRS (hrsl/mc_hrsl.f:157-499)
    OP_CHECK_R,       // stop if sqrt(xs^2 + ys^2) > p0
    OP_CHECK_T_SLAB,  // stop if |xt - p0| > p1 or |yt| + p2 xt > p3
    // hut steps of the HRS (hrsl/mc_hrsl_hut.f:289-304)
    OP_BOX_ROT_H,     // (xt,yt) = rotate_haxis(xs,ys; tan p0, sin p1, cos p2);
                      // stop if xt > p3 or xt < p4 or yt > p5 or yt < p6
    OP_NCODES
};

// The two stages a program is run in: the magnets (target to the end of the
// dipole: COSY segments and apertures) and the detector hut (materials, wire
// planes, fit, trigger).  Each stage's interpreter is compiled with its own
// steps only; the host checks every program against this table.
enum : int { STAGE_MAGNETS = 1, STAGE_HUT = 2 };
constexpr int op_stages(int code)
{
    return code == OP_END || code == OP_PROJECT                                  ? (STAGE_MAGNETS | STAGE_HUT)
         : code <= OP_COLLIMATOR || (code >= OP_CHECK_T_BOX && code <= OP_CHECK_T_SLAB) ? STAGE_MAGNETS
                                                                                 : STAGE_HUT;
}

struct Op {
    int32_t code;
    int32_t arg;
    int32_t stop;
    int32_t pad;
    double p[8];
};

}  // namespace simcdev
