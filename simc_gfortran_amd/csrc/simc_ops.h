// Step programs of the spectrometer arms.
//
// An arm (mc_hms, mc_shms, mc_sos, mc_hrsr, mc_hrsl and their detector huts) is a fixed sequence of
// drifts, COSY map segments, aperture tests and detector-hut steps.  The host
// writes that sequence once per context as an array of Op records
// (simc_programs.h, every numeric parameter formed with the reference's own
// expressions, cited there).  The magnet stages walk the records (all lanes
// of a wavefront execute the same op, so dispatch is uniform and the code of
// each step exists once); the detector hut runs as straight-line code
// generated from the same builders (gen_hut.cpp), checked against the loaded
// records.
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
    // the mock sieve slit (the program's first steps)
    OP_SIEVE,         // when the run's sieve option is arg: force the track through the nearest
                      // hole of a sieve at z = p0 (pitch p1, p2; hole radius p3; layout p4)
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
         : code <= OP_COLLIMATOR || (code >= OP_CHECK_T_BOX && code <= OP_CHECK_T_SLAB) ||
                   code == OP_SIEVE                                              ? STAGE_MAGNETS
                                                                                 : STAGE_HUT;
}

struct Op {
    int32_t code;
    int32_t arg;
    int32_t stop;
    int32_t pad;
    double p[8];
};

// ---------------------------------------------------------------------------
// Stop codes (one per exit point; 0 = accepted)
// ---------------------------------------------------------------------------
enum : int {
    ST_OK = 0, ST_SLIT_HOR = 1, ST_SLIT_VERT = 2, ST_SLIT_OCT = 3, ST_COLL = 4,
    // HMS
    ST_H_Q1_IN = 10, ST_H_Q1_MID, ST_H_Q1_OUT, ST_H_Q2_IN, ST_H_Q2_MID, ST_H_Q2_OUT,
    ST_H_Q3_IN, ST_H_Q3_MID, ST_H_Q3_OUT, ST_H_D1_IN, ST_H_D1_OUT, ST_H_PIPE0, ST_H_PIPE1,
    ST_H_PIPE2, ST_H_PIPE3, ST_H_DC1, ST_H_DC2, ST_H_SCIN,
    // SHMS
    ST_S_HB_IN = 40, ST_S_HB_MEN, ST_S_HB_MEX, ST_S_HB_OUT, ST_S_Q1_IN, ST_S_Q1_MEN,
    ST_S_Q1_MID, ST_S_Q1_MEX, ST_S_Q1_OUT, ST_S_Q2_IN, ST_S_Q2_MEN, ST_S_Q2_MID, ST_S_Q2_MEX,
    ST_S_Q2_OUT, ST_S_Q3_IN, ST_S_Q3_MEN, ST_S_Q3_MID, ST_S_Q3_MEX, ST_S_Q3_OUT, ST_S_D1_IN,
    ST_S_D1_FLR, ST_S_D1_MEN, ST_S_D1_MID1, ST_S_D1_MID2, ST_S_D1_MID3, ST_S_D1_MID4,
    ST_S_D1_MID5, ST_S_D1_MID6, ST_S_D1_MID7, ST_S_D1_MEX, ST_S_D1_OUT, ST_S_DC1, ST_S_DC2,
    ST_S_S1X, ST_S_S1Y, ST_S_S2X, ST_S_S2Y, ST_S_CAL, ST_S_CALFID,
    // SOS
    ST_SO_QUAD_IN = 100, ST_SO_QUAD_MID, ST_SO_QUAD_OUT, ST_SO_BM01_IN, ST_SO_BM01_OUT,
    ST_SO_BM02_IN, ST_SO_BM02_OUT, ST_SO_EXIT_FLANGE, ST_SO_EXIT_APERTURE, ST_SO_EXIT_WINDOW,
    ST_SO_DC1, ST_SO_DC2, ST_SO_SCIN,
    // HRS (both arms)
    ST_R_PIPE_65 = 120, ST_R_PIPE_80, ST_R_Q1_PRE, ST_R_Q1_IN, ST_R_Q1_MID, ST_R_Q1_OUT,
    ST_R_Q1_POST, ST_R_Q2_PRE, ST_R_Q2_IN, ST_R_Q2_MID, ST_R_Q2_OUT, ST_R_Q2_POST1, ST_R_Q2_POST2,
    ST_R_D1_PRE, ST_R_D1_IN, ST_R_D1_OUT, ST_R_D1_POST_R, ST_R_D1_POST_BOX, ST_R_Q3_PRE, ST_R_Q3_IN,
    ST_R_Q3_MID, ST_R_Q3_OUT, ST_R_Q3_POST, ST_R_WINDOW, ST_R_DC1, ST_R_DC2, ST_R_S1, ST_R_S2,
    ST_TAPE = 250, ST_BADOP = 251
};

// One step of a compiled detector hut: its op code, arg and stop code.
struct HutStep {
    int32_t code, arg, stop;
};

}  // namespace simcdev
