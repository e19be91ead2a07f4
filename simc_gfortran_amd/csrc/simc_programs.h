// Step programs of the spectrometer arms, host side: the builders that write
// each arm's sequence of Op records (see simc_ops.h) from the reference's own
// constants and expressions.  Used by the library when an arm's maps are
// loaded and by gen_hut.cpp, which compiles the detector-hut part of every
// program into straight-line device code.
#pragma once

#include <cmath>
#include <cstring>
#include <initializer_list>
#include <vector>

#include "../../include/simc_gpu.h"
#include "simc_maps.h"
#include "simc_ops.h"

namespace simcprog {

using namespace simcdev;

// Step-program builder.  All parameters are formed with the reference's own
// expressions in double precision, so they equal the reference's values.
struct Program {
    std::vector<Op> ops;
    Op& add(int code, std::initializer_list<double> p = {}, int stop = 0, int arg = 0)
    {
        Op o;
        std::memset(&o, 0, sizeof o);
        o.code = code;
        o.stop = stop;
        o.arg = arg;
        int k = 0;
        for (double v : p) o.p[k++] = v;
        // multiple scattering steps also carry sqrt(rad_len) and log10(rad_len)
        if (code == OP_MUSC || code == OP_MUSC_EXT) {
            const double rl = o.p[0];
            const int at = code == OP_MUSC ? 1 : 2;
            o.p[at] = rl > 0.0 ? std::sqrt(rl) : 0.0;
            o.p[at + 1] = rl > 0.0 ? std::log10(rl) : 0.0;
        }
        ops.push_back(o);
        return ops.back();
    }
    // rotate_haxis / rotate_vaxis of the track into a tilted aperture plane
    // (shared/rotate_haxis.f:41-51: truncated degree constant, libm trig of
    // the product), then the aperture-centre offsets.
    void rot(int code, double rotang, double xoff, double yoff)
    {
        const double raddeg = 0.017453292;
        const double r = rotang * raddeg;
        add(code, {std::tan(r), std::sin(r), std::cos(r), xoff, yoff});
    }
};

constexpr double kInf = HUGE_VAL;

// Wire-chamber planes of one chamber: hms/mc_hms_hut.f:341-370 (the SHMS loop,
// shms/mc_shms_hut.f:172-205, and the SOS loop, sos/mc_sos_hut.f:298-333, are
// the same with their own constants).  Bit iplane-1 of y_planes marks the
// planes that measure y: 2 and 5 in the HMS and SHMS, 1, 3 and 5 in the SOS.
inline void chamber(Program& P, int jchamber, double cath_radw, double hdc_thick, double hdc_radlen,
             double cath_thick, double wire_radw, double wire_thick, unsigned y_planes = 0x12u)
{
    for (int iplane = 1; iplane <= 6; ++iplane) {
        P.add(OP_MUSC, {cath_radw});
        double drift = 0.5 * hdc_thick;
        P.add(OP_MUSC, {drift / hdc_radlen});
        drift = drift + cath_thick;
        P.add(OP_PROJECT, {drift});
        P.add(OP_MUSC, {wire_radw});
        P.add(OP_WIRE, {}, (jchamber - 1) * 6 + iplane - 1, (y_planes >> (iplane - 1)) & 1u);
        drift = 0.5 * hdc_thick;
        P.add(OP_MUSC, {drift / hdc_radlen});
        drift = drift + wire_thick;
        P.add(OP_PROJECT, {drift});
    }
}

// mc_hms (hms/mc_hms.f:150-441) followed by mc_hms_hut (hms/mc_hms_hut.f:292-601)
inline void build_hms(const simc::ForwardMap& f, Program& P, double zdc[12])
{
    const double z_off = +40.17, z_entr = 126.2e0 + z_off, z_exit = z_entr + 6.3e0;  // :84-89
    const double z_dip1 = 64.77e0, z_dip2 = z_dip1 + 297.18e0, z_dip3 = z_dip2 + 115.57e0;  // :92-94
    const double rq1 = 20.50 * 20.50, rq2 = 30.22 * 30.22, rq3 = 30.22 * 30.22;  // apertures_hms.inc:5-7
    const double* dd = f.driftdist.data();  // driftdist(spectr, k) = dd[k-1]
    P.add(OP_SIEVE, {z_entr, 2.54, 1.524, 0.254, 0.0}, 0, 1);                          // :155-168
    P.add(OP_PROJECT, {z_entr});                                                        // :201-202
    P.add(OP_COLLIMATOR, {4.575, 11.646, 4.759, 12.114, 0.000, +0.028, z_exit - z_entr});  // :206-254
    P.add(OP_PROJECT, {dd[0] - z_exit});                                                // :259-260
    P.add(OP_CHECK_R2, {rq1}, ST_H_Q1_IN);
    P.add(OP_TRANSP, {125.233e0}, 0, 2);  P.add(OP_CHECK_R2, {rq1}, ST_H_Q1_MID);    // :268-272
    P.add(OP_TRANSP, {62.617e0}, 0, 3);   P.add(OP_CHECK_R2, {rq1}, ST_H_Q1_OUT);    // :276-280
    P.add(OP_PROJECT, {dd[3]});           P.add(OP_CHECK_R2, {rq2}, ST_H_Q2_IN);     // :285-290
    P.add(OP_TRANSP, {143.90e0}, 0, 5);   P.add(OP_CHECK_R2, {rq2}, ST_H_Q2_MID);
    P.add(OP_TRANSP, {71.95e0}, 0, 6);    P.add(OP_CHECK_R2, {rq2}, ST_H_Q2_OUT);
    P.add(OP_PROJECT, {dd[6]});           P.add(OP_CHECK_R2, {rq3}, ST_H_Q3_IN);     // :311-316
    P.add(OP_TRANSP, {143.8e0}, 0, 8);    P.add(OP_CHECK_R2, {rq3}, ST_H_Q3_MID);
    P.add(OP_TRANSP, {71.9e0}, 0, 9);     P.add(OP_CHECK_R2, {rq3}, ST_H_Q3_OUT);
    P.add(OP_PROJECT, {dd[9]});                                                         // :338-339
    P.rot(OP_ROT_H, -6.0e0, 0.0, 0.0);    P.add(OP_HIT_DIPOLE_T, {}, ST_H_D1_IN);    // :340-346
    P.add(OP_TRANSP, {526.053e0}, 0, 11);                                               // :351
    P.rot(OP_ROT_H, 6.0e0, 0.0, 0.0);     P.add(OP_HIT_DIPOLE_T, {}, ST_H_D1_OUT);   // :352-358
    P.add(OP_PIPE_T, {30.48 * 30.48, 2.8, 0.0, 20.5232}, ST_H_PIPE0);                  // :363-367
    P.add(OP_PROJECT, {z_dip1});           P.add(OP_PIPE_S, {1145.518, 2.8, 0.0}, ST_H_PIPE1);
    P.add(OP_PROJECT, {z_dip2 - z_dip1});  P.add(OP_PIPE_S, {1512.2299, 2.8, 0.0}, ST_H_PIPE2);
    P.add(OP_PROJECT, {z_dip3 - z_dip2});  P.add(OP_PIPE_S, {2162.9383, 2.8, 0.0}, ST_H_PIPE3);

    // detector hut, hms/mc_hms_hut.f (constants :57-232)
    const double zinit = -(dd[11] - z_dip3);  // mc_hms.f:413-415
    const double hfoil_exit_radlen = 8.90, hfoil_exit_thick = 0.011 * 2.54, hair_radlen = 30420.;
    const double hdc_entr_radlen = 28.7, hdc_entr_thick = 0.001 * 2.54;
    const double hdc_radlen = 16700.0, hdc_thick = 1.8;
    const double hdc_wire_radlen = 0.35, hdc_wire_thick = 0.0000049;
    const double hdc_cath_radlen = 7.2, hdc_cath_thick = 0.000177;
    const double hdc_exit_radlen = 28.7, hdc_exit_thick = 0.001 * 2.54;
    const double haer_entr_radlen = 8.90, haer_entr_thick = 0.15, haer_radlen = 150.0, haer_thick = 9.0;
    const double haer_air_radlen = 30420.0, haer_air_thick = 16.0, haer_exit_radlen = 8.90, haer_exit_thick = 0.1;
    const double hscin_radlen = 42.4, hscin_thick = 1.067;
    const double hcer_entr_radlen = 8.90, hcer_entr_thick = 0.040 * 2.54, hcer_radlen = 9620.0;
    const double hcer_mir_radlen = 400.0, hcer_mir_thick = 2.0;
    const double hcer_exit_radlen = 8.90, hcer_exit_thick = 0.040 * 2.54;
    const int hdc_nr_plan = 6;
    const double hdc_1_zpos = -52.1084, hdc_2_zpos = 29.2608;
    const double hdc_del_plane = hdc_thick + hdc_wire_thick + hdc_cath_thick;
    const double haer_zentrance = 35.699, haer_zexit = 60.949;
    const double hscin_1x_zpos = 77.830, hscin_1y_zpos = 97.520, hscin_2x_zpos = 298.820, hscin_2y_zpos = 318.510;
    const double hcer_zentrance = 110.000, hcer_zmirror = 230.000, hcer_zexit = 265.000;
    const double hcal_4ta_zpos = 371.69;
    // scintillator box of :487-490 (reused for the second hodoscope, :537-551)
    const double scin[4] = {37.75 + -1.3, -37.75 + -1.3, 60.25 + -1.3, -60.25 + -1.3};
    double drift;

    P.add(OP_RESMULT_DRAW, {0.15});                                               // :292-298
    P.add(OP_PROJECT, {(hdc_1_zpos - 25.000) - zinit});                           // :322-323
    P.add(OP_MUSC, {hfoil_exit_thick / hfoil_exit_radlen});
    drift = (hdc_1_zpos - 0.5 * hdc_nr_plan * hdc_del_plane) - (hdc_1_zpos - 25.000);
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_MUSC, {hdc_entr_thick / hdc_entr_radlen});                           // chamber 1, :337-381
    chamber(P, 1, hdc_cath_thick / hdc_cath_radlen, hdc_thick, hdc_radlen, hdc_cath_thick,
            hdc_wire_thick / hdc_wire_radlen, hdc_wire_thick);
    P.add(OP_MUSC, {hdc_exit_thick / hdc_exit_radlen});
    P.add(OP_BOX, {56.5 - 1.670, -56.5 - 1.670, 26.0 - 1.443, -26.0 - 1.443}, ST_H_DC1);
    P.add(OP_MUSC, {hdc_cath_thick / hdc_cath_radlen});
    drift = (hdc_2_zpos - 0.5 * hdc_nr_plan * hdc_del_plane) - (hdc_1_zpos + 0.5 * hdc_nr_plan * hdc_del_plane);
    P.add(OP_PROJECT, {drift});                                                   // :385-389
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_MUSC, {hdc_entr_thick / hdc_entr_radlen});                           // chamber 2, :391-436
    chamber(P, 2, hdc_cath_thick / hdc_cath_radlen, hdc_thick, hdc_radlen, hdc_cath_thick,
            hdc_wire_thick / hdc_wire_radlen, hdc_wire_thick);
    P.add(OP_MUSC, {hdc_exit_thick / hdc_exit_radlen});
    P.add(OP_BOX, {56.5 - 2.758, -56.5 - 2.758, 26.0 - 2.753, -26.0 - 2.753}, ST_H_DC2);
    P.add(OP_MUSC, {hdc_cath_thick / hdc_cath_radlen});
    for (int i = 0; i < 6; ++i) {                                                 // :439-447
        zdc[i] = hdc_1_zpos + ((double)(i + 1) - 0.5 - 0.5 * hdc_nr_plan) * hdc_del_plane;
        zdc[6 + i] = hdc_2_zpos + ((double)(i + 1) - 0.5 - 0.5 * hdc_nr_plan) * hdc_del_plane;
    }
    P.add(OP_FIT);                                                                // :449-455
    drift = haer_zentrance - hdc_2_zpos - 0.5 * hdc_nr_plan * hdc_del_plane;      // :459-462
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_MUSC, {haer_entr_thick / haer_entr_radlen});
    P.add(OP_PROJECT, {haer_thick});
    P.add(OP_MUSC_EXT, {haer_thick / haer_radlen, haer_thick});
    P.add(OP_PROJECT, {haer_air_thick});
    P.add(OP_MUSC_EXT, {haer_air_thick / haer_air_radlen, haer_air_thick});
    P.add(OP_MUSC, {haer_exit_thick / haer_exit_radlen});
    drift = hscin_1x_zpos - haer_zexit;                                           // :482-492
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_SCIN, {scin[0], scin[1], scin[2], scin[3]});
    P.add(OP_MUSC, {hscin_thick / hscin_radlen});
    drift = hscin_1y_zpos - hscin_1x_zpos;                                        // :494-503
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_SCIN, {scin[0], scin[1], scin[2], scin[3]});
    P.add(OP_MUSC, {hscin_thick / hscin_radlen});
    drift = hcer_zentrance - hscin_1y_zpos;                                       // :507-513
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_MUSC, {hcer_entr_thick / hcer_entr_radlen});
    drift = hcer_zmirror - hcer_zentrance;                                        // :515-521
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hcer_radlen, drift});
    P.add(OP_MUSC, {hcer_mir_thick / hcer_mir_radlen});
    drift = hcer_zexit - hcer_zmirror;                                            // :523-529
    P.add(OP_MUSC, {hcer_exit_thick / hcer_exit_radlen});
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC, {hcer_exit_thick / hcer_exit_radlen});
    drift = hscin_2x_zpos - hcer_zexit;                                           // :533-542
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_SCIN, {scin[0], scin[1], scin[2], scin[3]});
    P.add(OP_MUSC, {hscin_thick / hscin_radlen});
    drift = hscin_2y_zpos - hscin_2x_zpos;                                        // :544-553
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_SCIN, {scin[0], scin[1], scin[2], scin[3]});
    P.add(OP_MUSC, {hscin_thick / hscin_radlen});
    P.add(OP_TRIGGER, {}, ST_H_SCIN, 3);                                          // :556-559
    drift = hcal_4ta_zpos - hscin_2y_zpos;                                        // :563-566
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_END);
}

// mc_shms (shms/mc_shms.f:255-1106) followed by mc_shms_hut (shms/mc_shms_hut.f:74-448)
// mc_sos (sos/mc_sos.f:160-372, apertures sos/apertures_sos.inc:9-26) followed
// by mc_sos_hut (sos/mc_sos_hut.f:253-503, constants :42-197)
inline void build_sos(const simc::ForwardMap& f, Program& P, double zdc[12])
{
    const double z_entr = 126.3e0, z_exit = z_entr + 6.3e0;  // :69-70
    const double r2_quad = 163.84, w_bm01 = 8.0, w_bm02 = 8.0;
    const double* dd = f.driftdist.data();
    P.add(OP_SIEVE, {z_entr, 1.524, 0.0, 0.254, 1.0}, 0, 1);                              // :122-137
    P.add(OP_PROJECT, {z_entr});                                                          // :171-172
    P.add(OP_COLLIMATOR, {7.201, 4.696, 7.567, 4.935, 0.0, 0.0, z_exit - z_entr});       // :173-204
    P.add(OP_PROJECT, {dd[0] - z_exit});  P.add(OP_CHECK_R2, {r2_quad}, ST_SO_QUAD_IN);  // :209-215
    P.add(OP_TRANSP, {35.0e0}, 0, 2);     P.add(OP_CHECK_R2, {r2_quad}, ST_SO_QUAD_MID); // :219-223
    P.add(OP_TRANSP, {35.0e0}, 0, 3);     P.add(OP_CHECK_R2, {r2_quad}, ST_SO_QUAD_OUT); // :227-231
    P.add(OP_TRANSP, {80.0e0}, 0, 4);                                                     // :236-245
    P.rot(OP_ROT_H, -45.0e0, 0.0, 0.0);
    P.add(OP_CHECK_T_BOX, {w_bm01, w_bm01 - 0.05 * 2.54, 41.32, -30.52}, ST_SO_BM01_IN);
    P.add(OP_TRANSP, {169.52e0}, 0, 5);                                                   // :250-257
    P.rot(OP_ROT_H, 45.0e0, 0.0, 0.0);
    P.add(OP_CHECK_T_BOX, {w_bm01, w_bm01, 53.05, -65.07}, ST_SO_BM01_OUT);
    P.add(OP_TRANSP, {80.80e0}, 0, 6);                                                    // :262-269
    P.rot(OP_ROT_H, 49.0e0, 0.0, 0.0);
    P.add(OP_CHECK_T_BOX, {w_bm02, w_bm02, 51.73, -66.38}, ST_SO_BM02_IN);
    P.add(OP_TRANSP, {77.06e0}, 0, 7);                                                    // :274-281
    P.rot(OP_ROT_H, 57.0e0, 0.0, 0.0);
    P.add(OP_CHECK_T_BOX, {w_bm02, w_bm02, 51.52, -55.85}, ST_SO_BM02_OUT);
    P.add(OP_TRANSP, {43.82e0}, 0, 8);                                                    // :311-319
    P.rot(OP_ROT_H, 45.0e0, 0.0, 0.0);
    P.add(OP_CHECK_T_BOX, {8.57, 8.57, 51.52, -55.85}, ST_SO_EXIT_FLANGE);
    P.add(OP_TRANSP, {44.34e0}, 0, 9);                                                    // :329-336
    P.add(OP_CHECK_TRAP, {37.694, 10.998, 0.10209}, ST_SO_EXIT_APERTURE);
    P.add(OP_SHIFT, {7.62});                                                              // :342-343
    P.add(OP_CHECK_ABS, {38.1, 12.7}, ST_SO_EXIT_WINDOW);                                 // :346-350

    // detector hut from z = -3.206267 (:370-371)
    const double zinit = -3.206267e0;
    const double sfoil_exit_radlen = 53.3, sfoil_exit_thick = 0.020 * 2.54, sfoil_exit_zpos = -3.22;
    const double sair_radlen = 30420.;
    const double sdc_radlen = 16700.0, sdc_thick = 0.61775;
    const double sdc_wire_radlen = 0.35, sdc_wire_thick = 0.0000354;
    const double sdc_cath_radlen = 28.7, sdc_cath_thick = 0.0005 * 2.54;
    const double sscin_radlen = 42.4;
    const double scer_entr_radlen = 8.90, scer_entr_thick = 0.050, scer_radlen = 4810.0;
    const double scer_mir_radlen = 400.0, scer_mir_thick = 2.0;
    const double scer_exit_radlen = 8.90, scer_exit_thick = 0.050;
    const int sdc_nr_plan = 6;
    const double sdc_1_zpos = 6.25, sdc_2_zpos = 55.77;
    const double sdc_del_plane = sdc_thick + sdc_wire_thick + sdc_cath_thick;
    const double sscin_1y_zpos = 73.61, sscin_1x_zpos = 97.11, sscin_2y_zpos = 249.51, sscin_2x_zpos = 290.81;
    const double sscin_1x_thick = 1.040, sscin_1y_thick = 1.098, sscin_2x_thick = 1.040, sscin_2y_thick = 1.098;
    const double scer_zentrance = 130.000, scer_zmirror = 155.000, scer_zexit = 160.000;
    const double scal_4ta_zpos = 346.01;
    const double scin1[4] = {18.25 + 2.25, -18.25 + 2.25, 31.75 + 2.8, -31.75 + 2.8};  // :422-425
    const double scin2[4] = {18.25 + 2.9, -18.25 + 2.9, 56.25 + 4.9, -56.25 + 4.9};    // :472-475
    double drift;

    P.add(OP_RESMULT_DRAW, {0.15});                                                 // :253-259
    // Exit foil at sfoil_exit_zpos, tilt angle 0 (:49): the drift
    // (zpos + xt tan 0) - zinit = -0.013733 is raised to 0.001 (:281-285), and
    // the air gap to the chambers starts at zpos (:292-293).
    drift = sfoil_exit_zpos - zinit;
    if (drift <= 0.001) drift = 0.001;
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC, {sfoil_exit_thick / sfoil_exit_radlen / std::cos(0.0)});
    drift = (sdc_1_zpos - 0.5 * sdc_nr_plan * sdc_del_plane) - sfoil_exit_zpos;    // :292-296
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / sair_radlen, drift});
    chamber(P, 1, sdc_cath_thick / sdc_cath_radlen, sdc_thick, sdc_radlen, sdc_cath_thick,
            sdc_wire_thick / sdc_wire_radlen, sdc_wire_thick, 0x15u);                // :298-333
    P.add(OP_BOX, {32.0 - 8.649, -32.0 - 8.649, 24.0 - -1.822, -24.0 - -1.822}, ST_SO_DC1);  // :335-341
    P.add(OP_MUSC, {sdc_cath_thick / sdc_cath_radlen});
    drift = sdc_2_zpos - sdc_1_zpos - sdc_nr_plan * sdc_del_plane;                 // :346-350
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / sair_radlen, drift});
    chamber(P, 2, sdc_cath_thick / sdc_cath_radlen, sdc_thick, sdc_radlen, sdc_cath_thick,
            sdc_wire_thick / sdc_wire_radlen, sdc_wire_thick, 0x15u);                // :352-383
    P.add(OP_BOX, {32.0 - -1.532, -32.0 - -1.532, 24.0 - -1.976, -24.0 - -1.976}, ST_SO_DC2);  // :385-391
    P.add(OP_MUSC, {sdc_cath_thick / sdc_cath_radlen});
    for (int i = 0; i < 6; ++i) {                                                   // :399-407
        zdc[i] = sdc_1_zpos + ((double)(i + 1) - 0.5 - 0.5 * sdc_nr_plan) * sdc_del_plane;
        zdc[6 + i] = sdc_2_zpos + ((double)(i + 1) - 0.5 - 0.5 * sdc_nr_plan) * sdc_del_plane;
    }
    P.add(OP_FIT);                                                                  // :409-414
    drift = sscin_1y_zpos - sdc_2_zpos - 0.5 * sdc_nr_plan * sdc_del_plane;         // :418-428
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / sair_radlen, drift});
    P.add(OP_SCIN, {scin1[0], scin1[1], scin1[2], scin1[3]});
    P.add(OP_MUSC, {sscin_1y_thick / sscin_radlen});
    drift = sscin_1x_zpos - sscin_1y_zpos;                                          // :430-439
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / sair_radlen, drift});
    P.add(OP_SCIN, {scin1[0], scin1[1], scin1[2], scin1[3]});
    P.add(OP_MUSC, {sscin_1x_thick / sscin_radlen});
    drift = scer_zentrance - sscin_1x_zpos;                                         // :443-448
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / sair_radlen, drift});
    P.add(OP_MUSC, {scer_entr_thick / scer_entr_radlen});
    drift = scer_zmirror - scer_zentrance;                                          // :451-456
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / scer_radlen, drift});
    P.add(OP_MUSC, {scer_mir_thick / scer_mir_radlen});
    drift = scer_zexit - scer_zmirror;                                              // :459-464
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / scer_radlen, drift});
    P.add(OP_MUSC, {scer_exit_thick / scer_exit_radlen});
    drift = sscin_2y_zpos - scer_zexit;                                             // :469-478
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / sair_radlen, drift});
    P.add(OP_SCIN, {scin2[0], scin2[1], scin2[2], scin2[3]});
    P.add(OP_MUSC, {sscin_2y_thick / sscin_radlen});
    drift = sscin_2x_zpos - sscin_2y_zpos;                                          // :480-489
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / sair_radlen, drift});
    P.add(OP_SCIN, {scin2[0], scin2[1], scin2[2], scin2[3]});
    P.add(OP_MUSC, {sscin_2x_thick / sscin_radlen});
    P.add(OP_TRIGGER, {}, ST_SO_SCIN, 3);                                           // :491-494
    drift = scal_4ta_zpos - sscin_2x_zpos;                                          // :500-503
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / sair_radlen, drift});
    P.add(OP_END);
}

// mc_hrsr / mc_hrsl (hrsr/mc_hrsr.f:155-526, hrsl/mc_hrsl.f:155-527, apertures
// hrsl/apertures_hrsl.inc:6-8) followed by their hut (hrsl/mc_hrsl_hut.f:229-456,
// constants :37-185).  The arms differ in the slit position and exit half width
// (hrsr/mc_hrsr.f:58,67; hrsl/mc_hrsl.f:58,67); their huts are the same.
inline void build_hrs(int arm, const simc::ForwardMap& f, Program& P, double zdc[12])
{
    const bool left = arm == SIMC_ARM_HRSL;
    const double h_entr = 3.145, v_entr = 6.090, v_exit = 6.485, h_exit = left ? 3.335 : 3.340;
    const double z_off = 0.0, z_entr = (left ? 110.9e0 : 110.0e0) + z_off, z_exit = z_entr + 8.0e0;
    const double rq1 = 15.0 * 15.0, rq2 = 30.22 * 30.22, rq3 = 30.22 * 30.22;
    const double* dd = f.driftdist.data();
    double ztmp, zdrift;
    P.add(OP_SIEVE, {z_exit, 2.50, 1.25, 0.1, 2.0}, 0, 1);  // hrsl/mc_hrsl.f:122-132, hrsr/mc_hrsr.f:122-132
    auto slit = [&](double h, double v) {  // |ys - y_off| > h, then |xs - x_off| > v (offsets 0)
        P.add(OP_CHECK_ABS, {kInf, h}, ST_SLIT_HOR);
        P.add(OP_CHECK_ABS, {v, kInf}, ST_SLIT_VERT);
    };
    auto d1_plane = [&](double ang, int stop) {  // rotated D1 aperture plane, :367-383, :389-405
        P.rot(OP_ROT_H, ang, 0.0, 0.0);
        P.add(OP_CHECK_T_SLAB, {2.500, 52.5, 0.01861, 12.5}, stop);
    };
    zdrift = 65.686; ztmp = zdrift;                                                   // :159-168
    P.add(OP_PROJECT, {zdrift});  P.add(OP_CHECK_R, {7.3787}, ST_R_PIPE_65);
    zdrift = 80.436 - ztmp; ztmp = 80.436;                                            // :170-179
    P.add(OP_PROJECT, {zdrift});  P.add(OP_CHECK_R, {7.4092}, ST_R_PIPE_80);
    P.add(OP_PROJECT, {z_entr - ztmp});  slit(h_entr, v_entr);                       // :184-199
    P.add(OP_PROJECT, {z_exit - z_entr});  slit(h_exit, v_exit);                     // :203-218
    ztmp = 135.064;                                                                   // :222-230
    P.add(OP_PROJECT, {ztmp - z_exit});  P.add(OP_CHECK_R, {12.5222}, ST_R_Q1_PRE);
    P.add(OP_PROJECT, {dd[0] - ztmp});  P.add(OP_CHECK_R2, {rq1}, ST_R_Q1_IN);       // :234-243
    P.add(OP_TRANSP, {62.75333333e0}, 0, 2);  P.add(OP_CHECK_R2, {rq1}, ST_R_Q1_MID); // :247-254
    P.add(OP_TRANSP, {31.37666667e0}, 0, 3);  P.add(OP_CHECK_R2, {rq1}, ST_R_Q1_OUT); // :258-265
    zdrift = 300.464 - 253.16; ztmp = zdrift;                                         // :270-278
    P.add(OP_PROJECT, {zdrift});  P.add(OP_CHECK_R, {14.9225}, ST_R_Q1_POST);
    zdrift = 314.464 - 300.464; ztmp = ztmp + zdrift;                                 // :281-289
    P.add(OP_PROJECT, {zdrift});  P.add(OP_CHECK_R, {20.9550}, ST_R_Q2_PRE);
    P.add(OP_PROJECT, {dd[3] - ztmp});  P.add(OP_CHECK_R2, {rq2}, ST_R_Q2_IN);       // :293-302
    P.add(OP_TRANSP, {121.77333333e0}, 0, 5);  P.add(OP_CHECK_R2, {rq2}, ST_R_Q2_MID);
    P.add(OP_TRANSP, {60.88666667e0}, 0, 6);   P.add(OP_CHECK_R2, {rq2}, ST_R_Q2_OUT);
    zdrift = 609.664 - 553.020; ztmp = zdrift;                                        // :329-337
    P.add(OP_PROJECT, {zdrift});  P.add(OP_CHECK_R, {30.0073}, ST_R_Q2_POST1);
    zdrift = 641.800 - 609.664; ztmp = ztmp + zdrift;                                 // :340-348
    P.add(OP_PROJECT, {zdrift});  P.add(OP_CHECK_R, {30.0073}, ST_R_Q2_POST2);
    zdrift = 819.489 - 641.800; ztmp = ztmp + zdrift;                                 // :351-359
    P.add(OP_PROJECT, {zdrift});  P.add(OP_CHECK_ABS, {50.0, 15.0}, ST_R_D1_PRE);
    P.add(OP_PROJECT, {dd[6] - ztmp});  d1_plane(-30.0, ST_R_D1_IN);                 // :364-383
    P.add(OP_TRANSP, {659.73445725e0}, 0, 8);  d1_plane(30.0, ST_R_D1_OUT);          // :388-405
    zdrift = 1745.33546 - 1655.83446; ztmp = zdrift;                                  // :411-426
    P.add(OP_PROJECT, {zdrift});  P.add(OP_CHECK_R, {30.3276}, ST_R_D1_POST_R);
    P.add(OP_CHECK_ABS, {50.0, 15.0}, ST_R_D1_POST_BOX);
    zdrift = 1759.00946 - 1745.33546; ztmp = ztmp + zdrift;                           // :428-437
    P.add(OP_PROJECT, {zdrift});  P.add(OP_CHECK_R, {30.3276}, ST_R_Q3_PRE);
    P.add(OP_PROJECT, {dd[8] - ztmp});  P.add(OP_CHECK_R2, {rq3}, ST_R_Q3_IN);       // :442-451
    P.add(OP_TRANSP, {121.7866667e0}, 0, 10);  P.add(OP_CHECK_R2, {rq3}, ST_R_Q3_MID);
    P.add(OP_TRANSP, {60.89333333e0}, 0, 11);  P.add(OP_CHECK_R2, {rq3}, ST_R_Q3_OUT);
    zdrift = 2080.38746 - 1997.76446; ztmp = zdrift;                                  // :477-486
    P.add(OP_PROJECT, {zdrift});  P.add(OP_CHECK_ABS, {35.56, 17.145}, ST_R_Q3_POST);
    zdrift = 2327.47246 - 2080.38746; ztmp = ztmp + zdrift;                           // :490-499
    P.add(OP_PROJECT, {zdrift});  P.add(OP_CHECK_ABS, {99.76635, 17.145}, ST_R_WINDOW);

    // the hut from zinit = -(driftdist(12) - ztmp) (:505-510)
    const double zinit = -(dd[11] - ztmp);
    const double hfoil_exit_radlen = 3.56, hfoil_exit_thick = 0.01;
    const double hair_radlen = 30420.;
    const double hdc_entr_radlen = 34.4, hdc_entr_thick = 0.00018 * 2.54;
    const double hdc_radlen = 16700.0, hdc_thick = 1.5;
    const double hdc_wire_radlen = 0.35, hdc_wire_thick = 0.0000049;
    const double hdc_cath_radlen = 7.2, hdc_cath_thick = 0.000177;
    const double hdc_exit_radlen = 34.4, hdc_exit_thick = 0.00018 * 2.54;
    const double hscin_radlen = 42.4;
    const double hcer_entr_radlen = 8.90, hcer_entr_thick = 0.040 * 2.54, hcer_radlen = 36620.0;
    const double hcer_mir_radlen = 400.0, hcer_mir_thick = 2.0;
    const double hcer_exit_radlen = 8.90, hcer_exit_thick = 0.040 * 2.54;
    const int hdc_nr_plan = 6;
    const double hdc_1_zpos = -25.0 + 25.0, hdc_2_zpos = 25.0 + 25.0;
    const double hdc_del_plane = hdc_thick + hdc_wire_thick + hdc_cath_thick;
    const double hdc_left = 14.4, hdc_right = -14.4, hdc_top = -105.6, hdc_bot = 105.6;
    const double hdc_x_offset = 0.000, hdc_y_offset = 0.000;
    const double hscin_1x_zpos = 95.0 + 25.0, hscin_2x_zpos = 288.3 + 25.0;
    const double hscin_1x_thick = 0.5 * 1.067, hscin_2x_thick = 0.5 * 1.067;
    const double hcer_zentrance = 137.0 + 25.0, hcer_zmirror = 197.0 + 25.0, hcer_zexit = 237.0 + 25.0;
    const double hcal_4ta_zpos = 407.3 + 25.0;
    const double raddeg = 0.017453292, r45 = 45.0e0 * raddeg;  // shared/rotate_haxis.f:41-51
    double drift;
    auto air = [&](double d, double radlen) {
        P.add(OP_PROJECT, {d});
        P.add(OP_MUSC_EXT, {d / radlen, d});
    };

    P.add(OP_RESMULT_ONE);                                                          // :231
    P.add(OP_MUSC, {hfoil_exit_thick / hfoil_exit_radlen});                         // :240-241
    for (int j = 1; j <= 2; ++j) {
        drift = j == 1 ? (hdc_1_zpos - 0.5 * hdc_nr_plan * hdc_del_plane) - zinit   // :247-250
                       : hdc_2_zpos - hdc_1_zpos - hdc_nr_plan * hdc_del_plane;     // :310-313
        air(drift, hair_radlen);
        P.add(OP_MUSC, {hdc_entr_thick / hdc_entr_radlen});
        chamber(P, j, hdc_cath_thick / hdc_cath_radlen, hdc_thick, hdc_radlen, hdc_cath_thick,
                hdc_wire_thick / hdc_wire_radlen, hdc_wire_thick, 0x15u);            // :256-285
        P.add(OP_MUSC, {hdc_exit_thick / hdc_exit_radlen});
        P.add(OP_BOX_ROT_H, {std::tan(r45), std::sin(r45), std::cos(r45), hdc_bot - hdc_x_offset,
                             hdc_top - hdc_x_offset, hdc_left - hdc_y_offset, hdc_right - hdc_y_offset},
              j == 1 ? ST_R_DC1 : ST_R_DC2);                                        // :289-304, :352-367
        P.add(OP_MUSC, {hdc_cath_thick / hdc_cath_radlen});
    }
    for (int i = 0; i < 6; ++i) {                                                   // :373-381
        zdc[i] = hdc_1_zpos + ((double)(i + 1) - 0.5 - 0.5 * hdc_nr_plan) * hdc_del_plane;
        zdc[6 + i] = hdc_2_zpos + ((double)(i + 1) - 0.5 - 0.5 * hdc_nr_plan) * hdc_del_plane;
    }
    P.add(OP_FIT);                                                                  // :383-389
    air(hscin_1x_zpos - hdc_2_zpos - 0.5 * hdc_nr_plan * hdc_del_plane, hair_radlen);  // :393-405
    P.add(OP_BOX, {kInf, -kInf, 18.0, -18.0}, ST_R_S1);
    P.add(OP_MUSC, {hscin_1x_thick / hscin_radlen});
    air(hcer_zentrance - hscin_1x_zpos, hair_radlen);                               // :410-416
    P.add(OP_MUSC, {hcer_entr_thick / hcer_entr_radlen});
    air(hcer_zmirror - hcer_zentrance, hcer_radlen);                                // :418-424
    P.add(OP_MUSC, {hcer_mir_thick / hcer_mir_radlen});
    air(hcer_zexit - hcer_zmirror, hcer_radlen);                                    // :426-432
    P.add(OP_MUSC, {hcer_exit_thick / hcer_exit_radlen});
    air(hscin_2x_zpos - hcer_zexit, hair_radlen);                                   // :436-448
    P.add(OP_BOX, {kInf, -kInf, 30.0, -30.0}, ST_R_S2);
    P.add(OP_MUSC, {hscin_2x_thick / hscin_radlen});
    air(hcal_4ta_zpos - hscin_2x_zpos, hair_radlen);                                // :453-456
    P.add(OP_END);
}

inline void build_shms(Program& P, double zdc[12])
{
    const double zd_hbin = 118.39, zd_hbmen = 17.61, zd_hbmex = 80.0, zd_hbout = 17.61;  // :176-179
    const double z_entr = 25.189, z_thick = 6.35;                                        // :181-182
    const double zd_q1in = 58.39, zd_q1men = 28.35, zd_q1mid = 93.65, zd_q1mex = 93.65, zd_q1out = 28.35;
    const double zd_q2in = 25.55, zd_q2men = 39.1, zd_q2mid = 79.35, zd_q2mex = 79.35, zd_q2out = 39.1;
    const double zd_q3in = 28.10, zd_q3men = 39.1, zd_q3mid = 79.35, zd_q3mex = 79.35, zd_q3out = 39.1;
    const double zd_q3d1trans = 18.00, zd_d1flare = 30.10, zd_d1men = 39.47, zd_d1mid = 36.406263;
    const double zd_d1out = 60.68, zd_fp = 307.95;
    // shms/apertures_shms.inc:9-28
    const double rhb = 11.2 * 11.2, rq1 = 20.00 * 20.00, rq2 = 30.00 * 30.00, rq3 = 30.00 * 30.00,
                 rd1 = 30.00 * 30.00;
    P.add(OP_SIEVE, {zd_hbin, 2.5, 1.64, 0.3, 0.0}, 0, 2);                        // :276-290
    P.add(OP_SIEVE, {zd_hbin + zd_hbmen + zd_hbmex + zd_hbout + z_entr, 2.5, 1.64, 0.3, 0.0}, 0, 1);  // :293-310
    P.add(OP_PROJECT, {zd_hbin});                                                 // :422-423
    P.rot(OP_ROT_V, 1.5, 0.0, 1.51);
    P.add(OP_CHECK_T_HB, {rhb, 11.75, -4.13}, ST_S_HB_IN);
    P.add(OP_PROJECT, {zd_hbmen});                                                // :441-455
    P.rot(OP_ROT_V, 1.5, 0.0, 0.98);
    P.add(OP_CHECK_T_HB, {rhb, 11.74, -5.45}, ST_S_HB_MEN);
    P.add(OP_TRANSP, {zd_hbmex}, 0, 3);                                           // :462-476
    P.rot(OP_ROT_V, -1.5, 0.0, 0.98);
    P.add(OP_CHECK_T_HB, {rhb, 11.71, -10.25}, ST_S_HB_MEX);
    P.add(OP_PROJECT, {zd_hbout});                                                // :479-495
    P.rot(OP_ROT_V, -1.5, 0.0, 1.51);
    P.add(OP_CHECK_T_HB, {rhb, 11.70, -11.71}, ST_S_HB_OUT);
    P.add(OP_PROJECT, {z_entr});                                                  // :505-506, 527-528
    P.add(OP_COLLIMATOR, {8.5, 12.5, 8.65, 12.85, 0.00, 0.00, z_thick});          // :500-568
    P.add(OP_PROJECT, {zd_q1in - z_entr - z_thick});                              // :573-585
    P.add(OP_CHECK_R2, {rq1}, ST_S_Q1_IN);
    P.add(OP_PROJECT, {zd_q1men});         P.add(OP_CHECK_R2, {rq1}, ST_S_Q1_MEN);
    P.add(OP_TRANSP, {zd_q1mid}, 0, 7);    P.add(OP_CHECK_R2, {rq1}, ST_S_Q1_MID);
    P.add(OP_TRANSP, {zd_q1mex}, 0, 8);    P.add(OP_CHECK_R2, {rq1}, ST_S_Q1_MEX);
    P.add(OP_PROJECT, {zd_q1out});         P.add(OP_CHECK_R2, {rq1}, ST_S_Q1_OUT);
    P.add(OP_PROJECT, {zd_q2in});          P.add(OP_CHECK_R2, {rq2}, ST_S_Q2_IN);   // :644-712
    P.add(OP_PROJECT, {zd_q2men});         P.add(OP_CHECK_R2, {rq2}, ST_S_Q2_MEN);
    P.add(OP_TRANSP, {zd_q2mid}, 0, 12);   P.add(OP_CHECK_R2, {rq2}, ST_S_Q2_MID);
    P.add(OP_TRANSP, {zd_q2mex}, 0, 13);   P.add(OP_CHECK_R2, {rq2}, ST_S_Q2_MEX);
    P.add(OP_PROJECT, {zd_q2out});         P.add(OP_CHECK_R2, {rq2}, ST_S_Q2_OUT);
    P.add(OP_PROJECT, {zd_q3in});          P.add(OP_CHECK_R2, {rq3}, ST_S_Q3_IN);   // :714-781
    P.add(OP_PROJECT, {zd_q3men});         P.add(OP_CHECK_R2, {rq3}, ST_S_Q3_MEN);
    P.add(OP_TRANSP, {zd_q3mid}, 0, 17);   P.add(OP_CHECK_R2, {rq3}, ST_S_Q3_MID);
    P.add(OP_TRANSP, {zd_q3mex}, 0, 18);   P.add(OP_CHECK_R2, {rq3}, ST_S_Q3_MEX);
    P.add(OP_PROJECT, {zd_q3out});         P.add(OP_CHECK_R2, {rq3}, ST_S_Q3_OUT);
    P.add(OP_PROJECT, {zd_q3d1trans});     P.add(OP_CHECK_R2, {rd1}, ST_S_D1_IN);   // :815-827
    P.add(OP_PROJECT, {zd_d1flare});                                              // :832-846
    P.rot(OP_ROT_H, 9.200, -3.5, 0.0);     P.add(OP_CHECK_T_R2, {rd1}, ST_S_D1_FLR);
    P.add(OP_PROJECT, {zd_d1men});                                                // :851-865
    P.rot(OP_ROT_H, 9.200, 2.82, 0.0);     P.add(OP_CHECK_T_R2, {rd1}, ST_S_D1_MEN);
    const double ang[8] = {6.9, 4.6, 2.3, 0.0, -2.3, -4.6, -6.9, -9.2};          // :871-1011
    const double off[8] = {8.05, 11.75, 13.96, 14.70, 13.96, 11.75, 8.05, 2.82};
    for (int k = 0; k < 8; ++k) {
        P.add(OP_TRANSP, {zd_d1mid}, 0, 23 + k);
        P.rot(OP_ROT_H, ang[k], off[k], 0.0);
        P.add(OP_CHECK_T_R2, {rd1}, ST_S_D1_MID1 + k);
    }
    P.add(OP_PROJECT, {zd_d1out});                                                // :1015-1031
    P.rot(OP_ROT_H, -9.20, -6.88, 0.0);    P.add(OP_CHECK_T_R2, {rd1}, ST_S_D1_OUT);
    const double hcer_1_zentrance = -291.700, hcer_1_zmirror = -84.900, hcer_1_zexit = -61.700;
    P.add(OP_PROJECT, {zd_fp + hcer_1_zentrance});                                // :1036-1047

    // detector hut, shms/mc_shms_hut.f with shms/hut.inc
    const double hfoil_exit_radlen = 8.89, hfoil_exit_thick = 0.020 * 2.54, hair_radlen = 30420.;
    const double hdc_entr_radlen = 28.7, hdc_entr_thick = 0.001 * 2.54;
    const double hdc_radlen = 16700.0, hdc_thick = 0.125 * 2.54;
    const double hdc_wire_radlen = 0.35, hdc_wire_thick = 0.0000354;
    const double hdc_cath_radlen = 28.6, hdc_cath_thick = 0.001 * 2.54;
    const double hdc_exit_radlen = 28.7, hdc_exit_thick = 0.001 * 2.54;
    const double hscin_radlen = 42.4, hscin_thick = 1.000 * 1.067;
    const double hcer_entr_radlen = 19.63, hcer_entr_thick = 0.002 * 2.54, hcer_1_radlen = 11700.0;
    const double hcer_mirglass_radlen = 12.29, hcer_mirglass_thick = 0.3;
    const double hcer_exit_radlen = 19.63, hcer_exit_thick = 0.002 * 2.54;
    const double hcer_2_entr_radlen = 8.90, hcer_2_entr_thick = 0.040 * 2.54, hcer_2_radlen = 1202.5;
    const double hcer_mir_radlen = 400., hcer_mir_thick = 2.00;
    const double hcer_2_exit_radlen = 8.90, hcer_2_exit_thick = 0.040 * 2.54;
    const int hdc_nr_plan = 6;
    const double hdc_1_zpos = -40.656, hdc_2_zpos = 39.332;
    const double hscin_1x_zpos = 52.1, hscin_1y_zpos = 61.7, hscin_2x_zpos = 271.4, hscin_2y_zpos = 282.4;
    const double hcer_2_zentrance = 72.600, hcer_2_zmirror = 179.400, hcer_2_zexit = 202.600;
    const double hcal_4ta_zpos = 341.0;
    const double hcal_left = 63.00, hcal_right = -63.00, hcal_top = -70.00, hcal_bottom = 70.00;
    const double hdc_del_plane = hdc_thick + hdc_wire_thick + hdc_cath_thick;     // :74
    double drift;

    P.add(OP_RESMULT_ONE);                                                        // :80
    P.add(OP_MUSC, {hfoil_exit_thick / hfoil_exit_radlen});                       // :99-120
    P.add(OP_MUSC, {hcer_entr_thick / hcer_entr_radlen});
    drift = hcer_1_zmirror - hcer_1_zentrance - hcer_mirglass_thick / 2;
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hcer_1_radlen, drift});
    P.add(OP_MUSC, {hcer_mirglass_thick / hcer_mirglass_radlen});
    drift = hcer_1_zexit - hcer_1_zmirror - hcer_mirglass_thick / 2;
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hcer_1_radlen, drift});
    P.add(OP_MUSC, {hcer_exit_thick / hcer_exit_radlen});
    drift = (hdc_1_zpos - 0.5 * hdc_nr_plan * hdc_del_plane) - hcer_1_zexit;      // :160-165
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_MUSC, {hdc_entr_thick / hdc_entr_radlen});                           // chamber 1, :168-222
    chamber(P, 1, hdc_cath_thick / hdc_cath_radlen, hdc_thick, hdc_radlen, hdc_cath_thick,
            hdc_wire_thick / hdc_wire_radlen, hdc_wire_thick);
    P.add(OP_MUSC, {hdc_exit_thick / hdc_exit_radlen});
    P.add(OP_BOX, {40. - 0.0, -40. - 0.0, 40.0 - 0.0, -40.0 - 0.0}, ST_S_DC1);
    P.add(OP_MUSC, {hdc_cath_thick / hdc_cath_radlen});
    drift = hdc_2_zpos - hdc_1_zpos - hdc_nr_plan * hdc_del_plane;                // :227-234
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_MUSC, {hdc_entr_thick / hdc_entr_radlen});                           // chamber 2, :236-287
    chamber(P, 2, hdc_cath_thick / hdc_cath_radlen, hdc_thick, hdc_radlen, hdc_cath_thick,
            hdc_wire_thick / hdc_wire_radlen, hdc_wire_thick);
    P.add(OP_MUSC, {hdc_exit_thick / hdc_exit_radlen});
    P.add(OP_BOX, {40. - 0., -40. - 0., 40.0 - 0.0, -40.0 - 0.0}, ST_S_DC2);
    P.add(OP_MUSC, {hdc_cath_thick / hdc_cath_radlen});
    drift = hscin_1x_zpos - hdc_2_zpos - 0.5 * hdc_nr_plan * hdc_del_plane;       // :291-306
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_BOX, {kInf, -kInf, 50. + 0.0, -50. + 0.0}, ST_S_S1X);
    P.add(OP_MUSC, {hscin_thick / hscin_radlen});
    drift = hscin_1y_zpos - hscin_1x_zpos;                                        // :307-321
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_BOX, {45. + 0.0, -45. + 0.0, kInf, -kInf}, ST_S_S1Y);
    P.add(OP_MUSC, {hscin_thick / hscin_radlen});
    drift = hcer_2_zentrance - hscin_1y_zpos;                                     // :326-332
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_MUSC, {hcer_2_entr_thick / hcer_2_entr_radlen});
    drift = hcer_2_zmirror - hcer_2_zentrance;                                    // :334-341
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hcer_2_radlen, drift});
    P.add(OP_MUSC, {hcer_mir_thick / hcer_mir_radlen});
    drift = hcer_2_zexit - hcer_2_zmirror;                                        // :343-349
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hcer_2_radlen, drift});
    P.add(OP_MUSC, {hcer_2_exit_thick / hcer_2_exit_radlen});
    drift = hscin_2x_zpos - hcer_2_zexit;                                         // :353-367
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_BOX, {kInf, -kInf, 55. + 0., -55. + 0.}, ST_S_S2X);
    P.add(OP_MUSC, {hscin_thick / hscin_radlen});
    drift = hscin_2y_zpos - hscin_2x_zpos;                                        // :368-382
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_BOX, {62.5 + 0., -62.5 + 0., kInf, -kInf}, ST_S_S2Y);
    P.add(OP_MUSC, {hscin_thick / hscin_radlen});
    drift = hcal_4ta_zpos - hscin_2y_zpos;                                        // :390-401
    P.add(OP_PROJECT, {drift});
    P.add(OP_MUSC_EXT, {drift / hair_radlen, drift});
    P.add(OP_BOX, {hcal_bottom, hcal_top, hcal_left, hcal_right}, ST_S_CAL);
    for (int i = 0; i < 6; ++i) {                                                 // :410-418
        zdc[i] = hdc_1_zpos + ((double)(i + 1) - 0.5 - 0.5 * hdc_nr_plan) * hdc_del_plane;
        zdc[6 + i] = hdc_2_zpos + ((double)(i + 1) - 0.5 - 0.5 * hdc_nr_plan) * hdc_del_plane;
    }
    P.add(OP_FIT);                                                                // :419-426
    P.add(OP_FIDUCIAL, {hcal_4ta_zpos, hcal_left - 5.0, hcal_right + 5.0, hcal_bottom - 5.0,
                        hcal_top + 5.0}, ST_S_CALFID);                            // :435-444
    P.add(OP_END);
}

// The program of an arm (the HRS arms differ in their constants only).
inline void build_program(int arm, const simc::ForwardMap& f, Program& P, double zdc[12])
{
    if (arm == SIMC_ARM_HMS) build_hms(f, P, zdc);
    else if (arm == SIMC_ARM_SOS) build_sos(f, P, zdc);
    else if (arm == SIMC_ARM_HRSR || arm == SIMC_ARM_HRSL) build_hrs(arm, f, P, zdc);
    else build_shms(P, zdc);
}

// Index of the first detector-hut step of a program.
inline int hut_begin(const Program& P)
{
    int hut = 0;
    while (hut < (int)P.ops.size() && op_stages(P.ops[hut].code) != STAGE_HUT) ++hut;
    return hut;
}

}  // namespace simcprog
