#define _GNU_SOURCE
/*
 * simc_oracle.c -- FP64 CPU restatement of SIMC's spectrometer transport and
 * reconstruction (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker for the HIP kernels and
 * the timed CPU baseline; it is never linked into the product library.
 *
 * Each function follows one reference routine, cited as path:line into the
 * reference tree (JeffersonLab/simc_gfortran).  Operation order, literal
 * constants (all REAL literals are 8-byte under -fdefault-real-8, Makefile:63),
 * early exits and the order in which uniforms are drawn are kept exactly, so
 * the results agree with the reference compiled from its own sources on the
 * same uniform tape.  Build with -ffp-contract=off (no FMA contraction), as
 * gfortran -O on x86-64 does not contract either.
 */
#include "simc_oracle.h"
#include "../simc_gfortran_amd/csrc/simc_bm_tables.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* constants.inc:17-42 */
#define PI_C 3.141592653589793
static const double Mpi_C = 139.57018;
static const double Mmu_C = 105.6583755;
static const double Mk_C = 493.677;

/* ========================================================================= */
/* Map files                                                                  */
/* ========================================================================= */

/* Read one record into a 132-character buffer padded with blanks, like
 * READ(chan,'(a)') into CHARACTER*132 (shared/transp.f:75,327).  Returns 0 at
 * end of file. */
static int f_read_record(FILE *f, char *buf)
{
    char tmp[4096];
    if (!fgets(tmp, sizeof tmp, f)) return 0;
    size_t n = strcspn(tmp, "\n");
    if (n > 132) n = 132;
    memset(buf, ' ', 132);
    memcpy(buf, tmp, n);
    buf[132] = 0;
    return 1;
}

/* Fortran G/F/E input editing of a w-character field with d fraction digits
 * and BLANK='NULL' (blanks ignored; an all-blank field is zero).  Without a
 * decimal point the last d mantissa digits are the fraction. */
static int f_real_field(const char *p, int w, int d, double *out)
{
    char s[64];
    int n = 0;
    for (int i = 0; i < w; i++)
        if (p[i] != ' ' && p[i] != '\t' && n < 62) s[n++] = p[i];
    s[n] = 0;
    if (n == 0) { *out = 0.0; return 0; }
    int i = 0, neg = 0;
    if (s[i] == '+' || s[i] == '-') { neg = (s[i] == '-'); i++; }
    char mant[64];
    int nm = 0, have_point = 0, frac_digits = 0, ndig = 0;
    for (; s[i]; i++) {
        if (isdigit((unsigned char)s[i])) { mant[nm++] = s[i]; ndig++; if (have_point) frac_digits++; }
        else if (s[i] == '.' && !have_point) { have_point = 1; }
        else break;
    }
    if (ndig == 0) return -1;
    long ex = 0;
    if (s[i]) {
        int j = i;
        if (s[j] == 'E' || s[j] == 'e' || s[j] == 'D' || s[j] == 'd' || s[j] == 'Q' || s[j] == 'q') j++;
        int eneg = 0;
        if (s[j] == '+' || s[j] == '-') { eneg = (s[j] == '-'); j++; }
        if (!isdigit((unsigned char)s[j])) return -1;
        for (; s[j]; j++) {
            if (!isdigit((unsigned char)s[j])) return -1;
            ex = ex * 10 + (s[j] - '0');
        }
        if (eneg) ex = -ex;
    }
    mant[nm] = 0;
    long scale = have_point ? -frac_digits : -d;
    char buf[128];
    snprintf(buf, sizeof buf, "%s%se%ld", neg ? "-" : "", mant, ex + scale);
    *out = strtod(buf, NULL);
    return 0;
}

/* I1 input editing: blank is zero. */
static int f_int1(char c, int *v)
{
    if (c == ' ') { *v = 0; return 0; }
    if (c >= '0' && c <= '9') { *v = c - '0'; return 0; }
    return -1;
}

/* List-directed read of the first real on a line (shared/transp.f:330). */
static double f_list_real(const char *p)
{
    while (*p == ' ' || *p == ',') p++;
    char tok[64];
    int n = 0;
    while (*p && *p != ' ' && *p != ',' && *p != '/' && n < 63) tok[n++] = *p++;
    tok[n] = 0;
    for (int i = 0; i < n; i++)
        if (tok[i] == 'd' || tok[i] == 'D') tok[i] = 'e';
    return strtod(tok, NULL);
}

/* transp_init / transp_init_file, shared/transp.f:287-478. */
int or_read_forward(const char *path, or_fwd_map *m)
{
    const double coeff_min = 1.0e-14; /* shared/transp.f:66 */
    memset(m, 0, sizeof *m);
    FILE *f = fopen(path, "r");
    if (!f) return OR_E_IO;
    char line[133];
    /* strip header (:325-334) */
    line[0] = '!';
    while (line[0] == '!') {
        if (!f_read_record(f, line)) { fclose(f); return OR_E_MAPFMT; }
        if (strncmp(line, "!LENGTH:", 8) == 0) m->length[0] = 100. * f_list_real(line + 9);
    }
    int n_classes = 0;
    for (int i = 0; i < OR_MAX_CLASS; i++) {
        m->n_terms[i] = 0;
        m->adrift[i] = 1;
        m->driftdist[i] = 0.;
    }
    for (;;) {
        int kk = n_classes; /* 0-based */
        if (kk >= OR_MAX_CLASS) { fclose(f); return OR_E_TOOMANY; }
        while (strncmp(line, " ---", 4) != 0) {
            int it = m->n_terms[kk]++;
            if (m->n_terms[kk] > OR_MAX_TERMS) { fclose(f); return OR_E_TOOMANY; }
            double c[5];
            int e[6];
            for (int j = 0; j < 5; j++)
                if (f_real_field(line + 1 + 14 * j, 14, 7, &c[j])) { fclose(f); return OR_E_MAPFMT; }
            for (int j = 0; j < 6; j++)
                if (f_int1(line[72 + j], &e[j])) { fclose(f); return OR_E_MAPFMT; }
            int idummy = e[4];
            for (int j = 0; j < 5; j++) m->coeff[kk][it][j] = c[j];
            m->expon[kk][it][0] = (short)e[0];
            m->expon[kk][it][1] = (short)e[1];
            m->expon[kk][it][2] = (short)e[2];
            m->expon[kk][it][3] = (short)e[3];
            m->expon[kk][it][4] = (short)e[5];
            if (idummy != 0) { /* :384-391 */
                if (c[0] != 0 || c[1] != 0 || c[2] != 0 || c[3] != 0) { fclose(f); return OR_E_TOF; }
                m->n_terms[kk]--;
            }
            if (idummy == 0 && m->adrift[kk]) { /* :399-438 */
                int e1 = e[0], e2 = e[1], e3 = e[2], e4 = e[3];
                double c1 = c[0], c2 = c[1], c3 = c[2], c4 = c[3];
                double csum = fabs(c1) + fabs(c2) + fabs(c3) + fabs(c4);
                int order = e1 + e2 + e3 + e4;
                if (order == 1) {
                    if (e1 == 1) {
                        if (fabs(c1 - 1.e0) > coeff_min) m->adrift[kk] = 0;
                        if (fabs(c2 - 0.e0) > coeff_min) m->adrift[kk] = 0;
                        if (fabs(c3 - 0.e0) > coeff_min) m->adrift[kk] = 0;
                        if (fabs(c4 - 0.e0) > coeff_min) m->adrift[kk] = 0;
                    } else if (e2 == 1) {
                        m->driftdist[kk] = 1000.e0 * c1;
                        if (fabs(c2 - 1.e0) > coeff_min) m->adrift[kk] = 0;
                        if (fabs(c3 - 0.e0) > coeff_min) m->adrift[kk] = 0;
                        if (fabs(c4 - 0.e0) > coeff_min) m->adrift[kk] = 0;
                    } else if (e3 == 1) {
                        if (fabs(c1 - 0.e0) > coeff_min) m->adrift[kk] = 0;
                        if (fabs(c2 - 0.e0) > coeff_min) m->adrift[kk] = 0;
                        if (fabs(c3 - 1.e0) > coeff_min) m->adrift[kk] = 0;
                        if (fabs(c4 - 0.e0) > coeff_min) m->adrift[kk] = 0;
                    } else if (e4 == 1) {
                        if (fabs(c1 - 0.e0) > coeff_min) m->adrift[kk] = 0;
                        if (fabs(c2 - 0.e0) > coeff_min) m->adrift[kk] = 0;
                        if (fabs(m->driftdist[kk] - 1000.e0 * c3) > coeff_min) m->adrift[kk] = 0;
                        if (fabs(c4 - 1.e0) > coeff_min) m->adrift[kk] = 0;
                    }
                } else {
                    if (fabs(csum) > coeff_min) m->adrift[kk] = 0;
                }
            }
            if (!f_read_record(f, line)) { fclose(f); return OR_E_MAPFMT; }
        }
        n_classes = kk + 1;
        /* :461-468, read until a data line; EOF ends the file */
        for (;;) {
            if (!f_read_record(f, line)) goto done;
            if (strncmp(line, "!LENGTH:", 8) == 0 && kk + 1 < OR_MAX_CLASS)
                m->length[kk + 1] = 100. * f_list_real(line + 9);
            int blank = 1;
            for (int i = 0; i < 132; i++)
                if (line[i] != ' ') { blank = 0; break; }
            if (line[0] == '!' || strncmp(line, " ---", 4) == 0 || blank) continue;
            break;
        }
    }
done:
    fclose(f);
    m->n_classes = n_classes;
    return OR_OK;
}

/* Reconstruction file reader, hms/mc_hms_recon.f:70-102 (format :144). */
int or_read_recon(const char *path, or_recon_map *m)
{
    memset(m, 0, sizeof *m);
    FILE *f = fopen(path, "r");
    if (!f) return OR_E_IO;
    char line[133];
    line[0] = '!';
    while (line[0] == '!')
        if (!f_read_record(f, line)) { fclose(f); return OR_E_MAPFMT; }
    m->n_terms = 0;
    m->max_order = 0;
    while (strncmp(line, " ---", 4) != 0) {
        int it = m->n_terms++;
        if (m->n_terms > OR_MAX_TERMS) { fclose(f); return OR_E_TOOMANY; }
        for (int j = 0; j < 4; j++)
            if (f_real_field(line + 1 + 16 * j, 16, 9, &m->coeff[it][j])) { fclose(f); return OR_E_MAPFMT; }
        int ord = 0;
        for (int j = 0; j < 5; j++) {
            int e;
            if (f_int1(line[66 + j], &e)) { fclose(f); return OR_E_MAPFMT; }
            m->expon[it][j] = (short)e;
            ord += e;
        }
        if (!f_read_record(f, line)) { fclose(f); return OR_E_MAPFMT; }
        if (ord > m->max_order) m->max_order = ord;
    }
    fclose(f);
    return OR_OK;
}

or_maps *or_maps_load(const char *dir)
{
    or_maps *M = (or_maps *)calloc(1, sizeof *M);
    M->hms_fwd = (or_fwd_map *)calloc(1, sizeof(or_fwd_map));
    M->shms_fwd = (or_fwd_map *)calloc(1, sizeof(or_fwd_map));
    M->hms_rec = (or_recon_map *)calloc(1, sizeof(or_recon_map));
    M->shms_rec = (or_recon_map *)calloc(1, sizeof(or_recon_map));
    char p[4096];
    int rc = 0;
    snprintf(p, sizeof p, "%s/hms/forward_cosy.dat", dir);
    rc |= or_read_forward(p, M->hms_fwd);
    snprintf(p, sizeof p, "%s/hms/recon_cosy.dat", dir);
    rc |= or_read_recon(p, M->hms_rec);
    snprintf(p, sizeof p, "%s/shms/shms_forward.dat", dir);
    rc |= or_read_forward(p, M->shms_fwd);
    snprintf(p, sizeof p, "%s/shms/shms_recon.dat", dir);
    rc |= or_read_recon(p, M->shms_rec);
    if (rc || M->hms_fwd->n_classes != 12 || M->shms_fwd->n_classes != 32) {
        or_maps_free(M);
        return NULL;
    }
    /* SOS maps are optional (sos/mc_sos.f:160-165 reads them on first use) */
    snprintf(p, sizeof p, "%s/sos/forward_cosy.dat", dir);
    FILE *probe = fopen(p, "r");
    if (probe) {
        fclose(probe);
        M->sos_fwd = (or_fwd_map *)calloc(1, sizeof(or_fwd_map));
        M->sos_rec = (or_recon_map *)calloc(1, sizeof(or_recon_map));
        rc |= or_read_forward(p, M->sos_fwd);
        snprintf(p, sizeof p, "%s/sos/recon_cosy.dat", dir);
        rc |= or_read_recon(p, M->sos_rec);
        if (rc || M->sos_fwd->n_classes != 10) {
            or_maps_free(M);
            return NULL;
        }
    }
    /* HRS maps are optional too (hrs?/mc_hrs?.f:148-153): hrsr/ for arm 3, hrsl/ for arm 4 */
    for (int k = 0; k < 2; k++) {
        const char *sub = k ? "hrsl" : "hrsr";
        snprintf(p, sizeof p, "%s/%s/hrs_forward_cosy.dat", dir, sub);
        probe = fopen(p, "r");
        if (!probe) continue;
        fclose(probe);
        or_fwd_map *f = (or_fwd_map *)calloc(1, sizeof(or_fwd_map));
        or_recon_map *r = (or_recon_map *)calloc(1, sizeof(or_recon_map));
        if (k) { M->hrsl_fwd = f; M->hrsl_rec = r; }
        else { M->hrsr_fwd = f; M->hrsr_rec = r; }
        rc |= or_read_forward(p, f);
        snprintf(p, sizeof p, "%s/%s/hrs_recon_cosy.dat", dir, sub);
        rc |= or_read_recon(p, r);
        if (rc || f->n_classes != 12) {
            or_maps_free(M);
            return NULL;
        }
    }
    return M;
}

void or_maps_free(or_maps *M)
{
    if (!M) return;
    free(M->hms_fwd);
    free(M->shms_fwd);
    free(M->hms_rec);
    free(M->shms_rec);
    free(M->sos_fwd);
    free(M->sos_rec);
    free(M->hrsr_fwd);
    free(M->hrsr_rec);
    free(M->hrsl_fwd);
    free(M->hrsl_rec);
    free(M);
}

/* ========================================================================= */
/* Random numbers                                                             */
/* ========================================================================= */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:
 * as easy as 1, 2, 3"; Random123 multipliers and Weyl key increments).  This
 * is the counter-based generator of the GPU path; it has no counterpart in the
 * reference, which uses RANLUX. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform number k of event `event`: counter (k/4, event lo, event hi, 0x53494d43),
 * word k%4, mapped to (w + 1/2) 2^-32 in (0,1). */
#define OR_PHILOX_TAG 0x53494d43u
double or_philox_uniform(uint32_t key0, uint32_t key1, uint64_t event, uint64_t k)
{
    uint32_t ctr[4] = {(uint32_t)(k >> 2), (uint32_t)event, (uint32_t)(event >> 32), OR_PHILOX_TAG};
    uint32_t key[2] = {key0, key1};
    uint32_t o[4];
    or_philox4x32_10(ctr, key, o);
    return ((double)o[k & 3] + 0.5) * 2.3283064365386963e-10;
}

/* RANLUX, cern/ranlux.f:3-311 (F. James 1993), REAL state is 8-byte under
 * -fdefault-real-8. */
static const int ndskip[5] = {0, 24, 73, 199, 365};

static void ranlux_go(or_ranlux *r, int lux, int ins)
{
    /* RLUXGO (cern/ranlux.f:217-272), K1 = K2 = 0 */
    if (lux < 0) r->luxlev = 3;
    else if (lux <= 4) r->luxlev = lux;
    else r->luxlev = 4;
    r->nskip = ndskip[r->luxlev];
    r->in24 = 0;
    int jseed = ins > 0 ? ins : 314159265;
    r->twom24 = 1.;
    int iseeds[25];
    for (int i = 1; i <= 24; i++) {
        r->twom24 = r->twom24 * 0.5;
        int k = jseed / 53668;
        jseed = 40014 * (jseed - k * 53668) - k * 12211;
        if (jseed < 0) jseed = jseed + 2147483563;
        iseeds[i] = jseed % (1 << 24);
    }
    r->twom12 = r->twom24 * 4096.;
    for (int i = 1; i <= 24; i++) {
        r->seeds[i] = (double)iseeds[i] * r->twom24;
        r->next[i] = i - 1;
    }
    r->next[1] = 24;
    r->i24 = 24;
    r->j24 = 10;
    r->carry = 0.;
    if (r->seeds[24] == 0.) r->carry = r->twom24;
    r->kount = 0;
}

or_ranlux *or_ranlux_new(int lux, int seed)
{
    or_ranlux *r = (or_ranlux *)calloc(1, sizeof *r);
    ranlux_go(r, lux, seed);
    r->latest = 0;
    return r;
}

/* RANLUX(RVEC, LENV), cern/ranlux.f:112-148 */
void or_ranlux_fill(or_ranlux *r, double *rvec, int len)
{
    for (int iv = 0; iv < len; iv++) {
        double uni = r->seeds[r->j24] - r->seeds[r->i24] - r->carry;
        if (uni < 0.) {
            uni = uni + 1.0;
            r->carry = r->twom24;
        } else {
            r->carry = 0.;
        }
        r->seeds[r->i24] = uni;
        r->i24 = r->next[r->i24];
        r->j24 = r->next[r->j24];
        rvec[iv] = uni;
        if (uni < r->twom12) {
            rvec[iv] = rvec[iv] + r->twom24 * r->seeds[r->j24];
            if (rvec[iv] == 0.) rvec[iv] = r->twom24 * r->twom24;
        }
        r->in24 = r->in24 + 1;
        if (r->in24 == 24) {
            r->in24 = 0;
            r->kount += r->nskip;
            for (int isk = 1; isk <= r->nskip; isk++) {
                uni = r->seeds[r->j24] - r->seeds[r->i24] - r->carry;
                if (uni < 0.) {
                    uni = uni + 1.0;
                    r->carry = r->twom24;
                } else {
                    r->carry = 0.;
                }
                r->seeds[r->i24] = uni;
                r->i24 = r->next[r->i24];
                r->j24 = r->next[r->j24];
            }
        }
    }
    r->kount += len;
}

/* RLUXUT, cern/ranlux.f:196-203 */
void or_ranlux_save(const or_ranlux *r, int32_t isdext[25])
{
    for (int i = 1; i <= 24; i++) isdext[i - 1] = (int32_t)(r->seeds[i] * 4096. * 4096.);
    int v = r->i24 + 100 * r->j24 + 10000 * r->in24 + 1000000 * r->luxlev;
    if (r->carry > 0.) v = -v;
    isdext[24] = v;
}

/* RLUXIN, cern/ranlux.f:156-194 */
void or_ranlux_restore(or_ranlux *r, const int32_t isdext[25])
{
    r->twom24 = 1.;
    for (int i = 1; i <= 24; i++) {
        r->next[i] = i - 1;
        r->twom24 = r->twom24 * 0.5;
    }
    r->next[1] = 24;
    r->twom12 = r->twom24 * 4096.;
    for (int i = 1; i <= 24; i++) r->seeds[i] = (double)isdext[i - 1] * r->twom24;
    r->carry = 0.;
    if (isdext[24] < 0) r->carry = r->twom24;
    int isd = abs(isdext[24]);
    r->i24 = isd % 100;
    isd /= 100;
    r->j24 = isd % 100;
    isd /= 100;
    r->in24 = isd % 100;
    isd /= 100;
    r->luxlev = isd;
    if (r->luxlev <= 4) r->nskip = ndskip[r->luxlev];
    else if (r->luxlev >= 24) r->nskip = r->luxlev - 24;
    else { r->nskip = ndskip[4]; r->luxlev = 4; }
}

/* grnd(), call_ranlux.f:58-73 (uses rvec(1..999), refills when latest=1000) */
static double ranlux_grnd(or_ranlux *r)
{
    if (r->latest <= 0 || r->latest >= 1000) {
        or_ranlux_fill(r, r->rvec + 1, 1000);
        r->latest = 1;
    }
    double v = r->rvec[r->latest];
    r->latest = r->latest + 1;
    return v;
}

double or_grnd(or_rng *g)
{
    g->used++;
    switch (g->kind) {
    case 0: {
        if (g->pos >= g->stride) { g->overflow = 1; return 0.5; }
        return g->tape[g->pos++];
    }
    case 1: {
        uint64_t k = (uint64_t)g->pos++;
        if ((k & 3) == 0) {
            uint32_t ctr[4] = {(uint32_t)(k >> 2), (uint32_t)g->event, (uint32_t)(g->event >> 32),
                               g->tag ? g->tag : OR_PHILOX_TAG};
            or_philox4x32_10(ctr, g->key, g->buf);
        }
        return ((double)g->buf[k & 3] + 0.5) * 2.3283064365386963e-10;
    }
    default:
        return ranlux_grnd(g->rl);
    }
}

/* The detector hut's own stream of the counter-based source: the event's
 * key and event id with the fourth counter word OR_HUT_TAG, from uniform 0.
 * Every hut step draws one uniform (resmult) or Box-Muller pairs, so no
 * normal is pending between steps.  A tape (the reference's own sequence)
 * and the RANLUX source continue through the hut: *store is unused and g is
 * returned. */
static or_rng *or_hut_stream(or_rng *g, or_rng *store)
{
    if (g->kind != 1) return g;
    *store = *g;
    store->tag = OR_HUT_TAG;
    store->pos = 0;
    store->used = 0;
    store->have = 0;
    store->pending = 0.0;
    return store;
}

/* ========================================================================= */
/* Elementary routines                                                        */
/* ========================================================================= */

/* Integer power by binary exponentiation, the algorithm of libgcc's
 * __powidf2 / compiler-rt's __powidf2, which gfortran and flang call for
 * REAL**INTEGER with a variable exponent (shared/transp.f:210,
 * hms/mc_hms_recon.f:123-125). */
double or_powi(double x, int m)
{
    unsigned int n = m < 0 ? -(unsigned int)m : (unsigned int)m;
    double y = (n % 2) ? x : 1;
    while (n >>= 1) {
        x = x * x;
        if (n % 2) y = y * x;
    }
    return m < 0 ? 1 / y : y;
}

/* Elementary functions of the Box-Muller generator of the philox stream,
 * the same sequence of IEEE operations over the same tables
 * (simc_bm_tables.h, written by gen_bm_tables.py) as the GPU library, so
 * both produce identical bits.
 *
 * or_bm_log: log u for u in (0, 1): u = m 2^e, m in [sqrt(1/2), sqrt(2));
 * c = 1 + k/64 the nearest table point and invc = 1/c to 7 bits, so
 * t = m invc - 1 is exact; log m = -log(invc) + log1p(t) (series to t^9);
 * then e ln2 in two parts (ln2_hi has 21 significant bits). */
typedef struct {
    double nhi;
    float nlo, invc;
} bm_log_entry;
static const bm_log_entry bm_log_tab[SIMC_BM_LOG_N] = SIMC_BM_LOG_TABLE;
static const double bm_sincos_tab[SIMC_BM_SINCOS_N][2] = SIMC_BM_SINCOS_TABLE;

double or_bm_log(double u)
{
    int e;
    double m = frexp(u, &e); /* [1/2, 1) */
    if (m < 0x1.6a09e667f3bcdp-1) {
        m = m * 2.0;
        e = e - 1;
    }
    double kd = nearbyint((m - 1.0) * 64.0);
    const bm_log_entry *L = &bm_log_tab[(int)kd + SIMC_BM_LOG_K0];
    double t = fma(m, (double)L->invc, -1.0);
    double q = 0x1.c71c71c71c71cp-4;
    q = fma(t, q, -0x1p-3);
    q = fma(t, q, 0x1.2492492492492p-3);
    q = fma(t, q, -0x1.5555555555555p-3);
    q = fma(t, q, 0x1.999999999999ap-3);
    q = fma(t, q, -0x1p-2);
    q = fma(t, q, 0x1.5555555555555p-2);
    q = fma(t, q, -0x1p-1);
    double p = fma(t * t, q, t);
    double de = (double)e;
    return fma(de, 0x1.62e42p-1, L->nhi + (p + fma(de, 0x1.fdf473de6af28p-22, (double)L->nlo)));
}

/* sin and cos of 2 pi u, u in [0, 1): x = 64u, k = rint(x), r = x - k in
 * [-1/2, 1/2]; d = r pi/32, Taylor series of sin d (to d^9) and cos d (to
 * d^8), then the rotation by the table's sin and cos of 2 pi k/64. */
void or_bm_sincos2pi(double u, double *sn, double *cs)
{
    double x = u * 64.0;
    double kd = nearbyint(x);
    double r = x - kd;
    double d = r * 0x1.921fb54442d18p-4;
    double z = d * d;
    double ps = 0x1.71de3a556c734p-19;
    ps = fma(z, ps, -0x1.a01a01a01a01ap-13);
    ps = fma(z, ps, 0x1.1111111111111p-7);
    ps = fma(z, ps, -0x1.5555555555555p-3);
    double sd = fma(d * z, ps, d);
    double pc = 0x1.a01a01a01a01ap-16;
    pc = fma(z, pc, -0x1.6c16c16c16c17p-10);
    pc = fma(z, pc, 0x1.5555555555555p-5);
    pc = fma(z, pc, -0x1p-1);
    double cd = fma(z, pc, 1.0);
    int k = (int)kd & 63;
    double sa = bm_sincos_tab[k][0], ca = bm_sincos_tab[k][1];
    *sn = fma(sa, cd, ca * sd);
    *cs = fma(ca, cd, -(sa * sd));
}

/* Normal deviates of the counter-based (philox) stream: Box-Muller pairs,
 * both values used.  (u1, u2) -> sqrt(-2 ln u1) (cos 2 pi u2, sin 2 pi u2);
 * the cosine value is returned first, the sine value on the next call.
 * Values beyond nsigmax are skipped like gauss1 rejects them. */
static double or_gauss_bm(or_rng *g, double nsigmax)
{
    for (;;) {
        if (g->have) {
            g->have = 0;
            if (fabs(g->pending) <= nsigmax) return g->pending;
            continue;
        }
        double u1 = or_grnd(g);
        double u2 = or_grnd(g);
        double r = sqrt(-2.0 * or_bm_log(u1));
        double sn, cs;
        or_bm_sincos2pi(u2, &sn, &cs);
        g->pending = r * sn;
        g->have = 1;
        double g1 = r * cs;
        if (fabs(g1) <= nsigmax) return g1;
    }
}

/* gauss1, gauss1.f:1-30 (Marsaglia polar; second normal discarded).  The
 * tape and RANLUX sources use it as the reference does; the philox source
 * uses Box-Muller pairs (no rejection, so SIMD lanes stay in step). */
double or_gauss1(or_rng *g, double nsigmax)
{
    if (g->kind == 1) return or_gauss_bm(g, nsigmax);
    for (;;) {
        double u1 = or_grnd(g);
        double u2 = or_grnd(g);
        if (g->overflow) return 0.0;
        double v1 = 2.0 * u1 - 1.0;
        double v2 = 2.0 * u2 - 1.0;
        double s = v1 * v1 + v2 * v2;
        if (s > 1. || s == 0) continue;
        double r = v1 * sqrt(-(2. * log(s) / s));
        if (fabs(r) > nsigmax) continue;
        return r;
    }
}

/* LFIT, cern/lfit.f:11-60.  X, Y, A, B, E are REAL*4; the sums are the
 * implicit REAL type, 8-byte under -fdefault-real-8.  Pinned bit for bit
 * against the reference's own lfit.f (built into oracle/_ref) on seeded REAL*4
 * point sets (tests/golden/lfit/ref_lfit.npz). */
void or_lfit_f32(const float *X, const float *Y, int L, int KEY, float *A, float *B, float *E)
{
    if (L - 2 < 0) goto l25;
    {
        double count = 0.0, sumx = 0.0, sumy = 0.0, sumxy = 0.0, sumxx = 0.0, sumyy = 0.0;
        for (int j = 0; j < L; j++) {
            if (Y[j] == 0. && KEY == 0) continue;
            sumx = sumx + (double)X[j];
            sumy = sumy + (double)Y[j];
            count = count + 1.0;
        }
        if (count <= 1.) goto l25;
        double ymed = sumy / count;
        double xmed = sumx / count;
        for (int j = 0; j < L; j++) {
            if (Y[j] == 0. && KEY == 0) continue;
            double scartx = (double)X[j] - xmed;
            double scarty = (double)Y[j] - ymed;
            sumxy = sumxy + scartx * scarty;
            sumxx = sumxx + scartx * scartx;
            sumyy = sumyy + scarty * scarty;
        }
        if (sumxx == 0.) goto l25;
        *A = (float)(sumxy / sumxx);
        *B = (float)(ymed - (double)*A * xmed);
        if (count < 3.) goto l101;
        *E = (float)((sumyy - sumxy * (double)*A) / (count - 2.0));
        return;
    }
l25:
    *A = 0.0f;
    *B = 0.0f;
l101:
    *E = 0.0f;
}

/* The same fit carried out entirely in FP64 (the `fp64` track-fit mode). */
void or_lfit_f64(const double *X, const double *Y, int L, int KEY, double *A, double *B, double *E)
{
    if (L - 2 < 0) goto l25;
    {
        double count = 0.0, sumx = 0.0, sumy = 0.0, sumxy = 0.0, sumxx = 0.0, sumyy = 0.0;
        for (int j = 0; j < L; j++) {
            if (Y[j] == 0. && KEY == 0) continue;
            sumx = sumx + X[j];
            sumy = sumy + Y[j];
            count = count + 1.0;
        }
        if (count <= 1.) goto l25;
        double ymed = sumy / count;
        double xmed = sumx / count;
        for (int j = 0; j < L; j++) {
            if (Y[j] == 0. && KEY == 0) continue;
            double scartx = X[j] - xmed;
            double scarty = Y[j] - ymed;
            sumxy = sumxy + scartx * scarty;
            sumxx = sumxx + scartx * scartx;
            sumyy = sumyy + scarty * scarty;
        }
        if (sumxx == 0.) goto l25;
        *A = sumxy / sumxx;
        *B = ymed - *A * xmed;
        if (count < 3.) goto l101;
        *E = (sumyy - sumxy * *A) / (count - 2.0);
        return;
    }
l25:
    *A = 0.0;
    *B = 0.0;
l101:
    *E = 0.0;
}

/* loren, loren.f:1-26 */
void or_loren(double gam, double bx, double by, double bz, double e, double x, double y,
              double z, double *ef1, double *pxf, double *pyf, double *pzf, double *pf1)
{
    double gam1 = gam * gam / (1. + gam);
    *ef1 = gam * (e - bx * x - by * y - bz * z);
    *pxf = (1 + gam1 * (bx * bx)) * x + gam1 * bx * (by * y + bz * z) - gam * bx * e;
    *pyf = (1 + gam1 * (by * by)) * y + gam1 * by * (bx * x + bz * z) - gam * by * e;
    *pzf = (1 + gam1 * (bz * bz)) * z + gam1 * bz * (by * y + bx * x) - gam * bz * e;
    *pf1 = sqrt((*pxf) * (*pxf) + (*pyf) * (*pyf) + (*pzf) * (*pzf));
}

/* enerloss_new, enerloss_new.f:1-85 */
void or_enerloss_new(or_rng *g, double len, double dens, double zeff, double aeff,
                     double epart, double mpart, int typeflag, double *eloss)
{
    const double me = 0.51099906;
    double thick = len * dens;
    double gamma = epart / mpart;
    double beta = sqrt(1. - 1. / (gamma * gamma));
    double I;
    if (zeff == 1) I = 21.8e-06;
    else I = (16. * pow(zeff, 0.9)) * 1.0e-06;
    double hnup = 28.816e-06 * sqrt(dens * zeff / aeff);
    double log10bg = log(beta * gamma) / log(10.);
    double CO = log(hnup) - log(I) + 0.5;
    double denscorr;
    if (log10bg < 0.) denscorr = 0.;
    else if (log10bg < 3.) {
        double t3 = 3. - log10bg;
        denscorr = CO + log(10.) * log10bg + fabs(CO / 27.) * (t3 * t3 * t3);
    } else if (log10bg < 4.7) denscorr = CO + log(10.) * log10bg;
    else denscorr = CO + log(10.) * 4.7;
    double E;
    if (thick <= 0.) {
        E = 0.;
    } else {
        double b2 = beta * beta;
        double Eloss_mp_new = 0.1536e-03 * zeff / aeff * thick / b2 *
                              (log(me / (I * I)) + 1.063 + 2. * log(gamma * beta) +
                               log(0.1536 * zeff / aeff * thick / b2) - b2 - denscorr);
        double Eloss_mp = Eloss_mp_new * 1000.;
        double chsi = 0.307075 / 2. * zeff / aeff * thick / b2;
        double x = 0.;
        if (typeflag == 1) x = fabs(or_gauss1(g, 10.0e0));
        else if (typeflag == 2) x = 3;
        else if (typeflag == 3) x = 0.0067;
        else if (typeflag == 4) x = 1;
        double lambda;
        if (x > 0.0) lambda = -2.0 * log(x);
        else lambda = 100000.;
        E = lambda * chsi + Eloss_mp;
    }
    if (E > (epart - mpart)) E = (epart - mpart) - 0.0000001;
    *eloss = E;
}

/* pion_coll_absorb, hms/pion_coll_absorb.f:1-92.  The d-exponent literal
 * 1.d-27 is DOUBLE PRECISION, which -fdefault-real-8 promotes to a 16-byte
 * real, so that product is formed in quad precision (:76,82). */
double or_pion_coll_absorb(double ppi, double thick)
{
    static const double T[14] = {85.0, 125.0, 165.0, 205.0, 245.0, 315.0, 584.02, 711.95, 870.12,
                                 1227.57, 1446.58, 1865.29, 2858.0, 4159.0};
    static const double sigreac[14] = {26.03, 84.47, 117.3, 117.4, 101.9, 69.58, 42.5, 44.7, 47.9,
                                       46.5, 45.2, 39.6, 35.34, 33.15};
    static const double qreac[14] = {0.948, 0.659, 0.56, 0.5342, 0.5452, 0.60796, 0.699, 0.689,
                                     0.679, 0.683, 0.688, 0.704, 0.7483, 0.7705};
    const double Navagadro = 6.0221367e+23;
    const double mpi = 139.56995;
    double mate_dens = 17.0, mate_A = 171.57;
    double Epi = sqrt(ppi * ppi + mpi * mpi);
    double Tpi = Epi - mpi;
    double sigA = 0.0;
    for (int i = 0; i < 13; i++) {
        if (Tpi > T[i] && Tpi <= T[i + 1]) {
            double Thi = T[i + 1], Tlo = T[i];
            double sigAhi = sigreac[i + 1] * pow(mate_A, qreac[i + 1]);
            double sigAlo = sigreac[i] * pow(mate_A, qreac[i]);
            sigA = (sigAlo * (Thi - Tpi) + sigAhi * (Tpi - Tlo)) / (Thi - Tlo);
            sigA = (double)((__float128)sigA * 1.0e-27Q);
        }
    }
    if (Tpi > T[13]) {
        sigA = sigreac[13] * pow(mate_A, qreac[13]);
        sigA = (double)((__float128)sigA * 1.0e-27Q);
    }
    double sum = 0.0;
    double lambdai = mate_dens * Navagadro * sigA / mate_A;
    sum = sum + thick * lambdai;
    return exp(-sum);
}

/* rotate_haxis, shared/rotate_haxis.f:1-64 (truncated degree constant :41) */
void or_rotate_haxis(const or_track *t, double rotang, double *xp0, double *yp0)
{
    const double raddeg = 0.017453292;
    double rotang_rad = rotang * raddeg;
    double tan_th = tan(rotang_rad), sin_th = sin(rotang_rad), cos_th = cos(rotang_rad);
    double alpha = t->dxdzs, beta = t->dydzs;
    double alpha_p = (alpha + tan_th) / (1. - alpha * tan_th);
    double beta_p = beta / (cos_th - alpha * sin_th);
    double xi = *xp0;
    *xp0 = xi * (cos_th + alpha_p * sin_th);
    *yp0 = *yp0 + xi * beta_p * sin_th;
}

/* rotate_vaxis, shared/rotate_vaxis.f:1-60 */
void or_rotate_vaxis(const or_track *t, double rotang, double *xp0, double *yp0)
{
    const double raddeg = 0.017453292;
    double rotang_rad = rotang * raddeg;
    double tan_th = tan(rotang_rad), sin_th = sin(rotang_rad), cos_th = cos(rotang_rad);
    double alpha = t->dydzs, beta = t->dxdzs;
    double alpha_p = (alpha + tan_th) / (1. - alpha * tan_th);
    double beta_p = beta / (cos_th - alpha * sin_th);
    double yi = *yp0;
    *yp0 = yi * (cos_th + alpha_p * sin_th);
    *xp0 = *xp0 + yi * beta_p * sin_th;
}

/* musc, shared/musc.f:1-58 */
static void musc(or_rng *g, double m2, double p, double rad_len, double *dth, double *dph)
{
    const double Es = 13.6, epsilon = 0.088, nsig_max = 99.0e0;
    if (rad_len == 0) return;
    double beta = p / sqrt(m2 + p * p);
    double theta_sigma = Es / p / beta * sqrt(rad_len) * (1 + epsilon * log10(rad_len / (beta * beta)));
    *dth = *dth + theta_sigma * or_gauss1(g, nsig_max);
    *dph = *dph + theta_sigma * or_gauss1(g, nsig_max);
}

/* musc_ext, shared/musc_ext.f:1-54 (argument order dph, dth, y, x) */
static void musc_ext(or_rng *g, double m2, double p, double rad_len, double x_len, double *dph,
                     double *dth, double *y, double *x)
{
    const double Es = 13.6, epsilon = 0.088, nsig_max = 99.0e0;
    if (rad_len == 0) return;
    if (x_len <= 0 || rad_len < 0) return; /* the reference stops here (:146-151) */
    double beta = p / sqrt(m2 + p * p);
    double theta_sigma = Es / p / beta * sqrt(rad_len) * (1 + epsilon * log10(rad_len / (beta * beta)));
    double g1 = or_gauss1(g, nsig_max);
    double g2 = or_gauss1(g, nsig_max);
    *dth = *dth + theta_sigma * g1;
    *x = *x + theta_sigma * x_len * g2 / sqrt(12.) + theta_sigma * x_len * g1 / 2.;
    g1 = or_gauss1(g, nsig_max);
    g2 = or_gauss1(g, nsig_max);
    *dph = *dph + theta_sigma * g1;
    *y = *y + theta_sigma * x_len * g2 / sqrt(12.) + theta_sigma * x_len * g1 / 2.;
}

/* Shared decay kinematics of project.f:72-109 and transp.f:146-184.
 * m_kk is the final-state mass assigned on the kaon two-body branch (Mk in the
 * first half of transp, Mpi elsewhere, transp.f:158 vs :251, project.f:84). */
static void decay_products(or_rng *g, or_track *t, double beta, double gamma, double p_spec,
                           double *m2, double *ph, double m_kk)
{
    double rph = or_grnd(g) * 2. * PI_C;
    double rth1 = or_grnd(g) * 2. - 1.;
    double rth = acos(rth1);
    double pr = 0.;
    double m_final = Mmu_C;
    if (fabs(sqrt(*m2) - Mpi_C) < 2) pr = 29.783;
    if (fabs(sqrt(*m2) - Mk_C) < 2) {
        if (or_grnd(g) < 0.7) {
            pr = 235.5;
        } else {
            pr = sqrt(Mk_C * Mk_C / 4. - Mpi_C * Mpi_C);
            m_final = m_kk;
        }
    }
    /* pr == 0 (a mass that cannot decay) stops the reference program; the
     * decay then proceeds at rest here. */
    double er = sqrt(m_final * m_final + pr * pr);
    double pxr = pr * sin(rth) * cos(rph);
    double pyr = pr * sin(rth) * sin(rph);
    double pzr = pr * cos(rth);
    double s = sqrt(1. + t->dxdzs * t->dxdzs + t->dydzs * t->dydzs);
    double bx = -(beta * t->dxdzs / s);
    double by = -(beta * t->dydzs / s);
    double bz = -(beta * 1. / s);
    double ef, pf, pxf, pyf, pzf;
    or_loren(gamma, bx, by, bz, er, pxr, pyr, pzr, &ef, &pxf, &pyf, &pzf, &pf);
    t->dxdzs = pxf / pzf;
    t->dydzs = pyf / pzf;
    t->dpps = 100. * (pf / p_spec - 1.);
    *ph = pf;
    *m2 = m_final * m_final;
    t->Mh2_final = *m2;
}

/* project, shared/project.f:1-122.  x_new/y_new alias the track or temporaries. */
static void project(or_rng *g, or_track *t, const or_settings *S, double *x_new, double *y_new,
                    double z_drift, int decay_flag, int *dflag, double *m2, double *ph,
                    double *pathlen)
{
    if (!decay_flag || *dflag) {
        *pathlen = *pathlen + z_drift * sqrt(1 + t->dxdzs * t->dxdzs + t->dydzs * t->dydzs);
        *x_new = *x_new + t->dxdzs * z_drift;
        *y_new = *y_new + t->dydzs * z_drift;
        return;
    }
    double p_spec = *ph / (1. + t->dpps / 100.);
    double beta = *ph / sqrt((*ph) * (*ph) + *m2);
    double gamma = 1. / sqrt(1. - beta * beta);
    double dlen = S->ctau * beta * gamma;
    double z_decay = -1. * dlen * log(1 - or_grnd(g));
    double sq = sqrt(1 + t->dxdzs * t->dxdzs + t->dydzs * t->dydzs);
    if (z_decay > z_drift * sq) {
        t->decdist = t->decdist + z_drift * sq;
        *pathlen = *pathlen + z_drift * sq;
        *x_new = *x_new + t->dxdzs * z_drift;
        *y_new = *y_new + t->dydzs * z_drift;
    } else {
        *dflag = 1;
        t->decdist = t->decdist + z_decay;
        *pathlen = *pathlen + z_decay;
        *x_new = *x_new + t->dxdzs * z_decay / sq;
        *y_new = *y_new + t->dydzs * z_decay / sq;
        decay_products(g, t, beta, gamma, p_spec, m2, ph, Mpi_C);
        double sq2 = sqrt(1 + t->dxdzs * t->dxdzs + t->dydzs * t->dydzs);
        double tmpdrift = z_drift - z_decay / sq2;
        *pathlen = *pathlen + tmpdrift * sq2;
        *x_new = *x_new + t->dxdzs * tmpdrift;
        *y_new = *y_new + t->dydzs * tmpdrift;
    }
}

/* COSY sum of one forward class, shared/transp.f:199-218 */
void or_transp_eval(const or_fwd_map *m, int cls, const double ray[5], double sum[5])
{
    int k = cls - 1;
    for (int i = 0; i < 5; i++) sum[i] = 0.;
    for (int i = 0; i < m->n_terms[k]; i++) {
        double term = 1.0e0;
        for (int j = 0; j < 5; j++) {
            double temp = 1.0e0;
            if (m->expon[k][i][j] != 0.) temp = or_powi(ray[j], m->expon[k][i][j]);
            term = term * temp;
        }
        sum[0] = sum[0] + term * m->coeff[k][i][0];
        sum[1] = sum[1] + term * m->coeff[k][i][1];
        sum[2] = sum[2] + term * m->coeff[k][i][2];
        sum[3] = sum[3] + term * m->coeff[k][i][3];
        sum[4] = sum[4] + term * m->coeff[k][i][4];
    }
}

/* transp, shared/transp.f:1-281 */
static void transp(const or_fwd_map *m, or_rng *g, or_track *t, const or_settings *S, int cls,
                   int decay_flag, int *dflag, double *m2, double *ph, double zd, double *pathlen)
{
    double p_spec = 0, beta = 0, gamma = 0, z_decay = 0;
    if (decay_flag && !*dflag) { /* :134-187 */
        p_spec = *ph / (1. + t->dpps / 100.);
        beta = *ph / sqrt((*ph) * (*ph) + *m2);
        gamma = 1. / sqrt(1. - beta * beta);
        double dlen = S->ctau * beta * gamma;
        z_decay = -1. * dlen * log(1 - or_grnd(g));
        if (z_decay <= zd / 2) {
            *dflag = 1;
            t->decdist = t->decdist + z_decay;
            decay_products(g, t, beta, gamma, p_spec, m2, ph, Mk_C);
        }
    }
    double ray[5], sum[5];
    ray[0] = t->xs;
    ray[1] = t->dxdzs * 1000.e0;
    ray[2] = t->ys;
    ray[3] = t->dydzs * 1000.e0;
    ray[4] = t->dpps;
    or_transp_eval(m, cls, ray, sum);
    t->xs = sum[0];
    t->dxdzs = sum[1] / 1000.e0;
    t->ys = sum[2];
    t->dydzs = sum[3] / 1000.e0;
    double delta_z = -sum[4];
    if (decay_flag && !*dflag) { /* :231-276 */
        if (z_decay > zd + delta_z) {
            t->decdist = t->decdist + (zd + delta_z);
        } else {
            *dflag = 1;
            t->decdist = t->decdist + z_decay;
            decay_products(g, t, beta, gamma, p_spec, m2, ph, Mpi_C);
        }
    }
    *pathlen = *pathlen + (zd + delta_z);
}

/* COSY reconstruction sum, hms/mc_hms_recon.f:104-137 (all five powers taken) */
void or_recon_eval(const or_recon_map *m, const double hut[5], double sum[4])
{
    for (int i = 0; i < 4; i++) sum[i] = 0.;
    for (int i = 0; i < m->n_terms; i++) {
        double term = or_powi(hut[0], m->expon[i][0]) * or_powi(hut[1], m->expon[i][1]) *
                      or_powi(hut[2], m->expon[i][2]) * or_powi(hut[3], m->expon[i][3]) *
                      or_powi(hut[4], m->expon[i][4]);
        sum[0] = sum[0] + term * m->coeff[i][0];
        sum[1] = sum[1] + term * m->coeff[i][1];
        sum[2] = sum[2] + term * m->coeff[i][2];
        sum[3] = sum[3] + term * m->coeff[i][3];
    }
}

/* mc_hms_recon / mc_shms_recon (hms/mc_hms_recon.f:110-137,
 * shms/mc_shms_recon.f:124-160): identical arithmetic. */
static void recon(const or_recon_map *m, const or_track *t, double fry, double *delta_p,
                  double *delta_t, double *delta_phi, double *y_tgt)
{
    double hut[5], sum[4];
    hut[0] = t->xs / 100.;
    hut[1] = t->dxdzs;
    hut[2] = t->ys / 100.;
    hut[3] = t->dydzs;
    hut[4] = fry / 100.;
    if (fabs(hut[4]) <= 1.e-30) hut[4] = 1.e-30;
    or_recon_eval(m, hut, sum);
    *delta_phi = sum[0];
    *y_tgt = sum[1] * 100.;
    *delta_t = sum[2];
    *delta_p = sum[3] * 100.;
}

/* Straight-line fit of the 12 chamber planes, hms/mc_hms_hut.f:439-455 and
 * shms/mc_shms_hut.f:410-426.  In fit_f32 mode the hits, plane positions and
 * fitted parameters are REAL*4 as in the reference. */
typedef struct {
    double xdc[12], ydc[12], zdc[12];
} hut_hits;

static void fit_planes(const or_settings *S, const hut_hits *h, double *x_fp, double *dx_fp,
                       double *y_fp, double *dy_fp)
{
    if (S->fit_f32) {
        float xdc[12], ydc[12], zdc[12], a, b, e;
        for (int i = 0; i < 12; i++) {
            xdc[i] = (float)h->xdc[i];
            ydc[i] = (float)h->ydc[i];
            zdc[i] = (float)h->zdc[i];
        }
        or_lfit_f32(zdc, xdc, 12, 0, &a, &b, &e);
        *dx_fp = (double)a;
        *x_fp = (double)b;
        or_lfit_f32(zdc, ydc, 12, 0, &a, &b, &e);
        *dy_fp = (double)a;
        *y_fp = (double)b;
    } else {
        double a, b, e;
        or_lfit_f64(h->zdc, h->xdc, 12, 0, &a, &b, &e);
        *dx_fp = a;
        *x_fp = b;
        or_lfit_f64(h->zdc, h->ydc, 12, 0, &a, &b, &e);
        *dy_fp = a;
        *y_fp = b;
    }
}

/* ========================================================================= */
/* HMS                                                                        */
/* ========================================================================= */

/* mc_hms_coll, hms/mc_hms_coll.f:1-147 (h/v/x/y offsets, 20 steps of 6.30/20) */
static int coll_step(or_rng *g, or_track *t, const or_settings *S, int64_t *cnt, int arm,
                     double *m2, double *p, double p_spec, int decay_flag, int *dflag,
                     double *pathlen)
{
    double h_entr, v_entr, h_exit, v_exit, x_off, y_off, coll_radl, step_size;
    const double coll_dens = 17.0, zcoll = 69.45, acoll = 171.56797;
    const int nstep = 20;
    int ih, iv, io;
    if (arm == 1) {
        h_entr = 4.575; v_entr = 11.646; h_exit = 4.759; v_exit = 12.114;
        x_off = 0.000; y_off = +0.028; coll_radl = 0.41753; step_size = 6.30 / nstep;
        ih = H_SLIT_HOR; iv = H_SLIT_VERT; io = H_SLIT_OCT;
    } else {
        h_entr = 8.50; v_entr = 12.50; h_exit = 8.65; v_exit = 12.85;
        x_off = 0.000; y_off = +0.000; coll_radl = 0.42084; step_size = 6.35 / nstep;
        ih = S_SLIT_HOR; iv = S_SLIT_VERT; io = S_SLIT_OCT;
    }
    double h_step = h_entr, v_step = v_entr;
    double epart = sqrt((*p) * (*p) + *m2);
    for (int n = 1; n <= nstep; n++) {
        int step_flag = 0;
        if (fabs(t->ys - y_off) > h_step) { step_flag = 1; cnt[ih]++; }
        if (fabs(t->xs - x_off) > v_step) { step_flag = 1; cnt[iv]++; }
        if (fabs(t->xs - x_off) > (-(v_step / h_step * fabs(t->ys - y_off)) + 3 * v_step / 2)) {
            step_flag = 1;
            cnt[io]++;
        }
        double thick = step_size;
        if (step_flag) {
            if (*m2 > 12000. && *m2 < 20000.) {
                double trans = or_pion_coll_absorb(*p, thick);
                double rantemp = or_grnd(g);
                if (rantemp > trans) return 0;
            }
            double coll_radw = thick / coll_radl;
            musc(g, *m2, *p, coll_radw, &t->dydzs, &t->dxdzs);
            epart = sqrt((*p) * (*p) + *m2);
            double eloss;
            or_enerloss_new(g, thick, coll_dens, zcoll, acoll, epart, sqrt(*m2), 1, &eloss);
            epart = epart - eloss;
            if (epart < sqrt(*m2)) return 0;
            *p = sqrt(epart * epart - *m2);
            t->dpps = 100. * (*p / p_spec - 1.);
        }
        project(g, t, S, &t->xs, &t->ys, step_size, decay_flag, dflag, m2, p, pathlen);
        h_step = h_step + (h_exit - h_entr) / nstep;
        v_step = v_step + (v_exit - v_entr) / nstep;
    }
    return 1;
}

/* hit_dipole, hms/mc_hms.f:445-492 with hms/apertures_hms.inc:20-32 */
static int hms_hit_dipole(double x, double y)
{
    const double x_d1 = 34.29, y_d1 = 12.07, x_d2 = 27.94, y_d2 = 18.42, x_d3 = 13.97,
                 y_d3 = 18.95, x_d4 = 1.956, y_d4 = 20.32, x_d5 = 27.94, y_d5 = 12.065,
                 r_d5 = 6.35, a_d6 = -0.114, b_d6 = 20.54;
    double xl = fabs(x), yl = fabs(y);
    int c1 = (xl <= x_d1) && (yl <= y_d1);
    int c2 = (xl <= x_d2) && (yl <= y_d2);
    int c3 = (xl <= x_d3) && (yl <= y_d3);
    int c4 = (xl <= x_d4) && (yl <= y_d4);
    int c5 = ((xl - x_d5) * (xl - x_d5) + (yl - y_d5) * (yl - y_d5)) <= r_d5 * r_d5;
    int c6 = (xl >= x_d4) && (xl <= x_d3) && ((yl - a_d6 * xl - b_d6) <= 0.0);
    return !(c1 || c2 || c3 || c4 || c5 || c6);
}

/* mc_hms_hut, hms/mc_hms_hut.f:1-608.  Returns ok_hut; *fit_done is set once
 * x_fp..dy_fp have been written (:452-455). */
static int hms_hut(or_rng *g, or_rng *hg, or_track *t, const or_settings *S, int64_t *cnt, double *m2,
                   double *p, double *x_fp, double *dx_fp, double *y_fp, double *dy_fp,
                   int ms_flag, int wcs_flag, int decay_flag, int *dflag, double *resmult,
                   double zinit, double *pathlen, int *stop_code, int *fit_done)
{
    const double hfoil_exit_radlen = 8.90, hfoil_exit_thick = 0.011 * 2.54;
    const double hair_radlen = 30420.;
    const double hdc_entr_radlen = 28.7, hdc_entr_thick = 0.001 * 2.54;
    const double hdc_radlen = 16700.0, hdc_thick = 1.8;
    const double hdc_wire_radlen = 0.35, hdc_wire_thick = 0.0000049;
    const double hdc_cath_radlen = 7.2, hdc_cath_thick = 0.000177;
    const double hdc_exit_radlen = 28.7, hdc_exit_thick = 0.001 * 2.54;
    const double haer_entr_radlen = 8.90, haer_entr_thick = 0.15;
    const double haer_radlen = 150.0, haer_thick = 9.0;
    const double haer_air_radlen = 30420.0, haer_air_thick = 16.0;
    const double haer_exit_radlen = 8.90, haer_exit_thick = 0.1;
    const double hscin_radlen = 42.4;
    const double hcer_entr_radlen = 8.90, hcer_entr_thick = 0.040 * 2.54;
    const double hcer_radlen = 9620.0;
    const double hcer_mir_radlen = 400.0, hcer_mir_thick = 2.0;
    const double hcer_exit_radlen = 8.90, hcer_exit_thick = 0.040 * 2.54;
    const int hdc_nr_plan = 6;
    const double hdc_1_zpos = -52.1084, hdc_2_zpos = 29.2608;
    const double hdc_del_plane = hdc_thick + hdc_wire_thick + hdc_cath_thick;
    const double hdc_1_left = 26.0, hdc_1_right = -26.0, hdc_1y_offset = 1.443, hdc_1_top = -56.5,
                 hdc_1_bot = 56.5, hdc_1x_offset = 1.670;
    const double hdc_2_left = 26.0, hdc_2_right = -26.0, hdc_2y_offset = 2.753, hdc_2_top = -56.5,
                 hdc_2_bot = 56.5, hdc_2x_offset = 2.758;
    const double haer_zentrance = 35.699, haer_zexit = 60.949;
    const double hscin_1x_zpos = 77.830, hscin_1y_zpos = 97.520, hscin_2x_zpos = 298.820,
                 hscin_2y_zpos = 318.510;
    const double hscin_1x_thick = 1.067, hscin_1y_thick = 1.067, hscin_2x_thick = 1.067,
                 hscin_2y_thick = 1.067;
    const double hscin_1x_left = 37.75, hscin_1x_right = -37.75, hscin_1x_offset = -1.3;
    const double hscin_1y_top = -60.25, hscin_1y_bot = 60.25, hscin_1y_offset = -1.3;
    const double hcer_zentrance = 110.000, hcer_zmirror = 230.000, hcer_zexit = 265.000;
    const double hcal_4ta_zpos = 371.69;
    const double nsig_max = 99.0e0;
    const int scintrig = 3;
    double sigma = S->wire_sigma > 0 ? S->wire_sigma : 0.030; /* hdc_sigma, :143-144 */

    double radw, drift, tmpran1, tmpran2;
    hut_hits h;
    int scincount;

    double tmpran = or_grnd(hg); /* :292-298 */
    double tmplim = 0.15;
    if (tmpran < tmplim) *resmult = 2.0;
    else *resmult = 1.0;
    for (int i = 0; i < 12; i++) { h.xdc[i] = 0.; h.ydc[i] = 0.; }
    scincount = 0;

    drift = (hdc_1_zpos - 25.000) - zinit;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    radw = hfoil_exit_thick / hfoil_exit_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);

    drift = (hdc_1_zpos - 0.5 * hdc_nr_plan * hdc_del_plane) - (hdc_1_zpos - 25.000);
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);

    for (int jchamber = 1; jchamber <= 2; jchamber++) {
        if (jchamber == 2) {
            drift = (hdc_2_zpos - 0.5 * hdc_nr_plan * hdc_del_plane) -
                    (hdc_1_zpos + 0.5 * hdc_nr_plan * hdc_del_plane);
            radw = drift / hair_radlen;
            project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
            if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
        }
        radw = hdc_entr_thick / hdc_entr_radlen;
        if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
        int npl_off = (jchamber - 1) * hdc_nr_plan;
        for (int iplane = 1; iplane <= hdc_nr_plan; iplane++) {
            radw = hdc_cath_thick / hdc_cath_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            drift = 0.5 * hdc_thick;
            radw = drift / hdc_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            drift = drift + hdc_cath_thick;
            project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
            radw = hdc_wire_thick / hdc_wire_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            if (wcs_flag) {
                tmpran1 = or_gauss1(hg, nsig_max);
                tmpran2 = or_gauss1(hg, nsig_max);
            } else {
                tmpran1 = 0.;
                tmpran2 = 0.;
            }
            h.xdc[npl_off + iplane - 1] = t->xs + sigma * tmpran1 * *resmult;
            h.ydc[npl_off + iplane - 1] = t->ys + sigma * tmpran2 * *resmult;
            if (iplane == 2 || iplane == 5) h.xdc[npl_off + iplane - 1] = 0.;
            else h.ydc[npl_off + iplane - 1] = 0.;
            drift = 0.5 * hdc_thick;
            radw = drift / hdc_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            drift = drift + hdc_wire_thick;
            project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
        }
        radw = hdc_exit_thick / hdc_exit_radlen;
        if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
        if (jchamber == 1) {
            if (t->xs > (hdc_1_bot - hdc_1x_offset) || t->xs < (hdc_1_top - hdc_1x_offset) ||
                t->ys > (hdc_1_left - hdc_1y_offset) || t->ys < (hdc_1_right - hdc_1y_offset)) {
                cnt[H_DC1]++;
                *stop_code = STOP_H_DC1;
                return 0;
            }
        } else {
            if (t->xs > (hdc_2_bot - hdc_2x_offset) || t->xs < (hdc_2_top - hdc_2x_offset) ||
                t->ys > (hdc_2_left - hdc_2y_offset) || t->ys < (hdc_2_right - hdc_2y_offset)) {
                cnt[H_DC2]++;
                *stop_code = STOP_H_DC2;
                return 0;
            }
        }
        radw = hdc_cath_thick / hdc_cath_radlen;
        if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    }

    for (int jchamber = 1; jchamber <= 2; jchamber++) {
        int npl_off = (jchamber - 1) * hdc_nr_plan;
        for (int iplane = 1; iplane <= hdc_nr_plan; iplane++) {
            double z0 = jchamber == 1 ? hdc_1_zpos : hdc_2_zpos;
            h.zdc[npl_off + iplane - 1] = z0 + ((double)iplane - 0.5 - 0.5 * hdc_nr_plan) * hdc_del_plane;
        }
    }
    fit_planes(S, &h, x_fp, dx_fp, y_fp, dy_fp);
    *fit_done = 1;

    drift = haer_zentrance - hdc_2_zpos - 0.5 * hdc_nr_plan * hdc_del_plane;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    radw = haer_entr_thick / haer_entr_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    drift = haer_thick;
    radw = drift / haer_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    drift = haer_air_thick;
    radw = drift / haer_air_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    radw = haer_exit_thick / haer_exit_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);

#define HMS_SCIN_IN()                                                                          \
    (t->ys < (hscin_1x_left + hscin_1y_offset) && t->ys > (hscin_1x_right + hscin_1y_offset) && \
     t->xs < (hscin_1y_bot + hscin_1x_offset) && t->xs > (hscin_1y_top + hscin_1x_offset))

    drift = hscin_1x_zpos - haer_zexit;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    if (HMS_SCIN_IN()) scincount = scincount + 1;
    radw = hscin_1x_thick / hscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);

    drift = hscin_1y_zpos - hscin_1x_zpos;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    if (HMS_SCIN_IN()) scincount = scincount + 1;
    radw = hscin_1y_thick / hscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);

    drift = hcer_zentrance - hscin_1y_zpos;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    radw = hcer_entr_thick / hcer_entr_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    drift = hcer_zmirror - hcer_zentrance;
    radw = drift / hcer_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    radw = hcer_mir_thick / hcer_mir_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    drift = hcer_zexit - hcer_zmirror;
    radw = hcer_exit_thick / hcer_exit_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    radw = hcer_exit_thick / hcer_exit_radlen; /* exit window scattering a second time, :528-529 */
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);

    drift = hscin_2x_zpos - hcer_zexit;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    if (HMS_SCIN_IN()) scincount = scincount + 1; /* first-plane limits reused, :537-540 */
    radw = hscin_2x_thick / hscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);

    drift = hscin_2y_zpos - hscin_2x_zpos;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    if (HMS_SCIN_IN()) scincount = scincount + 1;
    radw = hscin_2y_thick / hscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
#undef HMS_SCIN_IN

    if (scincount < scintrig) {
        cnt[H_SCIN]++;
        *stop_code = STOP_H_SCIN;
        return 0;
    }
    drift = hcal_4ta_zpos - hscin_2y_zpos;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    return 1;
}

/* The mock sieve slit (hms/mc_hms.f:155-168, shms/mc_shms.f:276-310,
 * sos/mc_sos.f:122-137, hrsl/mc_hrsl.f:122-132): the track from the target
 * point is forced through the nearest hole of a sieve at z_slit, at an
 * offset drawn inside the hole (two uniforms), by changing the INOUT angles.
 * kind 0: holes on a (px, py) grid; 1: the SOS's rows (middle three at
 * 1.016 cm, outer ones on a 2.54 cm pitch from 1.016); 2: the HRS's (the
 * offset is drawn but not applied).  NINT rounds half away from zero. */
static void sieve_slit(or_rng *g, int kind, double z_slit, double px, double py, double r, double x, double y,
                       double *dxdz, double *dydz)
{
    const double pi = 3.141592653589793; /* constants.inc:42 */
    double xt = x + z_slit * *dxdz;
    double yt = y + z_slit * *dydz;
    xt = px * (double)lround(xt / px);
    if (kind == 1) {
        if (fabs(yt) < 1.524) yt = 1.016 * (double)lround(yt / 1.016);
        else yt = yt / fabs(yt) * (2.54 * (double)lround((fabs(yt) - 1.016) / 2.54) + 1.016);
    } else {
        yt = py * (double)lround(yt / py);
    }
    double rt = r * sqrt(or_grnd(g));
    double tht = 2 * pi * or_grnd(g);
    if (kind != 2) {
        xt = xt + rt * cos(tht);
        yt = yt + rt * sin(tht);
    }
    *dxdz = (xt - x) / z_slit;
    *dydz = (yt - y) / z_slit;
}

/* mc_hms, hms/mc_hms.f:1-441 */
void or_mc_hms(const or_maps *M, const or_settings *S, or_rng *g, or_track *t, int64_t *cnt,
               double p_spec, double *dpp, double *x, double *y, double *z, double *dxdz,
               double *dydz, double *x_fp, double *dx_fp, double *y_fp, double *dy_fp,
               double *m2, double *resmult, double fry, int *ok_spec, double *pathlen,
               int *stop_code, int *fit_done)
{
    const or_fwd_map *F = M->hms_fwd;
    const double h_entr = 4.575, v_entr = 11.646, h_exit = 4.759, v_exit = 12.114;
    const double x_off = +0.000, y_off = +0.028, z_off = +40.17;
    const double z_entr = 126.2e0 + z_off, z_exit = z_entr + 6.3e0;
    const double z_dip1 = 64.77e0, z_dip2 = z_dip1 + 297.18e0, z_dip3 = z_dip2 + 115.57e0;
    const double r_Q1 = 20.50, r_Q2 = 30.22, r_Q3 = 30.22;
    const double x_offset_pipes = 2.8, y_offset_pipes = 0.0;
    const int decay_flag = S->decay_flag, ms_flag = S->ms_flag, wcs_flag = S->wcs_flag;

    double p, xt, yt, zdrift;
    double dpp_recon, dth_recon, dph_recon, y_recon;
    int dflag = 0, ok;
    *ok_spec = 0;
    *stop_code = STOP_OK;
    *fit_done = 0;
    cnt[H_TRIALS]++;
    if (S->sieve == 1) sieve_slit(g, 0, z_entr, 2.54, 1.524, 0.254, *x, *y, dxdz, dydz); /* :155-168 */

    t->xs = *x;
    t->ys = *y;
    t->zs = *z;
    t->dxdzs = *dxdz;
    t->dydzs = *dydz;
    t->dpps = *dpp;
    p = p_spec * (1. + t->dpps / 100.);
    if (S->mom_offset != 0.0) t->dpps = 100. * ((1. + t->dpps / 100.) / (1. + S->mom_offset) - 1.);

    zdrift = z_entr;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (S->using_coll && (*m2 > 100.0 * 100.0) && (*m2 < 200.0 * 200.0)) {
        if (!coll_step(g, t, S, cnt, 1, m2, &p, p_spec, decay_flag, &dflag, pathlen)) {
            cnt[H_COLL]++;
            *stop_code = STOP_COLL;
            return;
        }
    } else {
        if (fabs(t->ys - y_off) > h_entr) { cnt[H_SLIT_HOR]++; *stop_code = STOP_SLIT_HOR; return; }
        if (fabs(t->xs - x_off) > v_entr) { cnt[H_SLIT_VERT]++; *stop_code = STOP_SLIT_VERT; return; }
        if (fabs(t->xs - x_off) > (-(v_entr / h_entr * fabs(t->ys - y_off)) + 3 * v_entr / 2)) {
            cnt[H_SLIT_OCT]++; *stop_code = STOP_SLIT_OCT; return;
        }
        zdrift = z_exit - z_entr;
        project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
        if (fabs(t->ys - y_off) > h_exit) { cnt[H_SLIT_HOR]++; *stop_code = STOP_SLIT_HOR; return; }
        if (fabs(t->xs - x_off) > v_exit) { cnt[H_SLIT_VERT]++; *stop_code = STOP_SLIT_VERT; return; }
        if (fabs(t->xs - x_off) > (-(v_exit / h_exit * fabs(t->ys - y_off)) + 3 * v_exit / 2)) {
            cnt[H_SLIT_OCT]++; *stop_code = STOP_SLIT_OCT; return;
        }
    }

#define RQ(r) ((t->xs * t->xs + t->ys * t->ys) > (r) * (r))
    zdrift = F->driftdist[0] - z_exit;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_Q1)) { cnt[H_Q1_IN]++; *stop_code = STOP_H_Q1_IN; return; }
    transp(F, g, t, S, 2, decay_flag, &dflag, m2, &p, 125.233e0, pathlen);
    if (RQ(r_Q1)) { cnt[H_Q1_MID]++; *stop_code = STOP_H_Q1_MID; return; }
    transp(F, g, t, S, 3, decay_flag, &dflag, m2, &p, 62.617e0, pathlen);
    if (RQ(r_Q1)) { cnt[H_Q1_OUT]++; *stop_code = STOP_H_Q1_OUT; return; }

    zdrift = F->driftdist[3];
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_Q2)) { cnt[H_Q2_IN]++; *stop_code = STOP_H_Q2_IN; return; }
    transp(F, g, t, S, 5, decay_flag, &dflag, m2, &p, 143.90e0, pathlen);
    if (RQ(r_Q2)) { cnt[H_Q2_MID]++; *stop_code = STOP_H_Q2_MID; return; }
    transp(F, g, t, S, 6, decay_flag, &dflag, m2, &p, 71.95e0, pathlen);
    if (RQ(r_Q2)) { cnt[H_Q2_OUT]++; *stop_code = STOP_H_Q2_OUT; return; }

    zdrift = F->driftdist[6];
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_Q3)) { cnt[H_Q3_IN]++; *stop_code = STOP_H_Q3_IN; return; }
    transp(F, g, t, S, 8, decay_flag, &dflag, m2, &p, 143.8e0, pathlen);
    if (RQ(r_Q3)) { cnt[H_Q3_MID]++; *stop_code = STOP_H_Q3_MID; return; }
    transp(F, g, t, S, 9, decay_flag, &dflag, m2, &p, 71.9e0, pathlen);
    if (RQ(r_Q3)) { cnt[H_Q3_OUT]++; *stop_code = STOP_H_Q3_OUT; return; }
#undef RQ

    zdrift = F->driftdist[9];
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    xt = t->xs;
    yt = t->ys;
    or_rotate_haxis(t, -6.0e0, &xt, &yt);
    if (hms_hit_dipole(xt, yt)) { cnt[H_D1_IN]++; *stop_code = STOP_H_D1_IN; return; }

    transp(F, g, t, S, 11, decay_flag, &dflag, m2, &p, 526.053e0, pathlen);
    xt = t->xs;
    yt = t->ys;
    or_rotate_haxis(t, 6.0e0, &xt, &yt);
    if (hms_hit_dipole(xt, yt)) { cnt[H_D1_OUT]++; *stop_code = STOP_H_D1_OUT; return; }

    if ((((xt - x_offset_pipes) * (xt - x_offset_pipes) + (yt - y_offset_pipes) * (yt - y_offset_pipes)) >
         30.48 * 30.48) ||
        (fabs(yt - y_offset_pipes) > 20.5232)) {
        cnt[H_D1_OUT]++; *stop_code = STOP_H_PIPE0; return;
    }
#define PIPE(lim) ((((t->xs - x_offset_pipes) * (t->xs - x_offset_pipes) + \
                     (t->ys - y_offset_pipes) * (t->ys - y_offset_pipes))) > (lim))
    zdrift = z_dip1;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (PIPE(1145.518)) { cnt[H_D1_OUT]++; *stop_code = STOP_H_PIPE1; return; }
    zdrift = z_dip2 - z_dip1;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (PIPE(1512.2299)) { cnt[H_D1_OUT]++; *stop_code = STOP_H_PIPE2; return; }
    zdrift = z_dip3 - z_dip2;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (PIPE(2162.9383)) { cnt[H_D1_OUT]++; *stop_code = STOP_H_PIPE3; return; }
#undef PIPE

    cnt[H_HUT]++;
    zdrift = F->driftdist[11] - z_dip3;
    or_rng hut_store;
    or_rng *hg = or_hut_stream(g, &hut_store);
    ok = hms_hut(g, hg, t, S, cnt, m2, &p, x_fp, dx_fp, y_fp, dy_fp, ms_flag, wcs_flag, decay_flag,
                 &dflag, resmult, -zdrift, pathlen, stop_code, fit_done);
    if (hg != g) g->used += hg->used;
    if (!ok) return;

    t->xs = *x_fp;
    t->ys = *y_fp;
    t->dxdzs = *dx_fp;
    t->dydzs = *dy_fp;
    recon(M->hms_rec, t, fry, &dpp_recon, &dth_recon, &dph_recon, &y_recon);
    *dpp = dpp_recon;
    *dxdz = dph_recon;
    *dydz = dth_recon;
    *y = y_recon;
    *ok_spec = 1;
    cnt[H_SUCCESSES]++;
}

/* ========================================================================= */
/* SOS                                                                        */
/* ========================================================================= */

/* mc_sos_recon, sos/mc_sos_recon.f:113-137: every hut coordinate, not only
 * fry, is kept away from 0 (|h| <= 1e-30 -> 1e-30). */
static void recon_sos(const or_recon_map *m, const or_track *t, double fry, double *delta_p,
                      double *delta_t, double *delta_phi, double *y_tgt)
{
    double hut[5], sum[4];
    hut[0] = t->xs / 100.;
    hut[1] = t->dxdzs;
    hut[2] = t->ys / 100.;
    hut[3] = t->dydzs;
    hut[4] = fry / 100.;
    for (int i = 0; i < 5; i++)
        if (fabs(hut[i]) <= 1.e-30) hut[i] = 1.e-30;
    or_recon_eval(m, hut, sum);
    *delta_phi = sum[0];
    *y_tgt = sum[1] * 100.;
    *delta_t = sum[2];
    *delta_p = sum[3] * 100.;
}

/* mc_sos_hut, sos/mc_sos_hut.f:1-548 (constants :42-197) */
static int sos_hut(or_rng *g, or_rng *hg, or_track *t, const or_settings *S, int64_t *cnt, double *m2,
                   double *p, double *x_fp, double *dx_fp, double *y_fp, double *dy_fp,
                   int ms_flag, int wcs_flag, int decay_flag, int *dflag, double *resmult,
                   double zinit, double *pathlen, int *stop_code, int *fit_done)
{
    const double pi = 3.141592654, d_r = pi / 180.;
    const double sfoil_exit_radlen = 53.3, sfoil_exit_thick = 0.020 * 2.54;
    const double sfoil_exit_zpos = -3.22, sfoil_exit_ang = 0. * d_r;
    const double sair_radlen = 30420.;
    const double sdc_radlen = 16700.0, sdc_thick = 0.61775;
    const double sdc_wire_radlen = 0.35, sdc_wire_thick = 0.0000354;
    const double sdc_cath_radlen = 28.7, sdc_cath_thick = 0.0005 * 2.54;
    const double sscin_radlen = 42.4;
    const double scer_entr_radlen = 8.90, scer_entr_thick = 0.050;
    const double scer_radlen = 4810.0;
    const double scer_mir_radlen = 400.0, scer_mir_thick = 2.0;
    const double scer_exit_radlen = 8.90, scer_exit_thick = 0.050;
    const int sdc_nr_cham = 2, sdc_nr_plan = 6;
    const double sdc_1_zpos = 6.25, sdc_2_zpos = 55.77;
    const double sdc_del_plane = sdc_thick + sdc_wire_thick + sdc_cath_thick;
    const double sdc_1_left = 24.0, sdc_1_right = -24.0, sdc_1y_offset = -1.822, sdc_1_top = -32.0,
                 sdc_1_bot = 32.0, sdc_1x_offset = 8.649;
    const double sdc_2_left = 24.0, sdc_2_right = -24.0, sdc_2y_offset = -1.976, sdc_2_top = -32.0,
                 sdc_2_bot = 32.0, sdc_2x_offset = -1.532;
    const double sscin_1y_zpos = 73.61, sscin_1x_zpos = 97.11, sscin_2y_zpos = 249.51,
                 sscin_2x_zpos = 290.81;
    const double sscin_1x_thick = 1.040, sscin_1y_thick = 1.098, sscin_2x_thick = 1.040,
                 sscin_2y_thick = 1.098;
    const double sscin_1_left = 18.25, sscin_1_right = -18.25, sscin_1x_offset = 2.8;
    const double sscin_1_top = -31.75, sscin_1_bot = 31.75, sscin_1y_offset = 2.25;
    const double sscin_2_left = 18.25, sscin_2_right = -18.25, sscin_2x_offset = 4.9;
    const double sscin_2_top = -56.25, sscin_2_bot = 56.25, sscin_2y_offset = 2.9;
    const double scer_zentrance = 130.000, scer_zmirror = 155.000, scer_zexit = 160.000;
    const double scal_4ta_zpos = 346.01;
    const double nsig_max = 99.0e0;
    const int scintrig = 3;
    const double sigma = S->wire_sigma > 0 ? S->wire_sigma : 0.030; /* sdc_sigma, :112-113 */

    double radw, drift, tmpran1, tmpran2, xt;
    hut_hits h;
    int scincount;

    double tmpran = or_grnd(hg); /* :253-259 */
    double tmplim = 0.15;
    if (tmpran < tmplim) *resmult = 2.0;
    else *resmult = 1.0;
    for (int i = 0; i < 12; i++) { h.xdc[i] = 0.; h.ydc[i] = 0.; }
    scincount = 0;

    /* exit foil, :281-287 */
    xt = t->xs + t->dxdzs * sfoil_exit_zpos;
    drift = (sfoil_exit_zpos + xt * tan(sfoil_exit_ang)) - zinit;
    if (drift <= 0.001) drift = 0.001;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    radw = sfoil_exit_thick / sfoil_exit_radlen / cos(sfoil_exit_ang);
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);

    for (int jchamber = 1; jchamber <= sdc_nr_cham; jchamber++) {
        if (jchamber == 1) { /* :292-296 */
            drift = (sdc_1_zpos - 0.5 * sdc_nr_plan * sdc_del_plane) -
                    (sfoil_exit_zpos + xt * tan(sfoil_exit_ang));
        } else { /* :346-350 */
            drift = sdc_2_zpos - sdc_1_zpos - sdc_nr_plan * sdc_del_plane;
        }
        radw = drift / sair_radlen;
        project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
        if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
        int npl_off = (jchamber - 1) * sdc_nr_plan;
        for (int iplane = 1; iplane <= sdc_nr_plan; iplane++) {
            radw = sdc_cath_thick / sdc_cath_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            drift = 0.5 * sdc_thick;
            radw = drift / sdc_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            drift = drift + sdc_cath_thick;
            project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
            radw = sdc_wire_thick / sdc_wire_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            if (wcs_flag) {
                tmpran1 = or_gauss1(hg, nsig_max);
                tmpran2 = or_gauss1(hg, nsig_max);
            } else {
                tmpran1 = 0.;
                tmpran2 = 0.;
            }
            h.xdc[npl_off + iplane - 1] = t->xs + sigma * tmpran1 * *resmult;
            h.ydc[npl_off + iplane - 1] = t->ys + sigma * tmpran2 * *resmult;
            if (iplane == 1 || iplane == 3 || iplane == 5) h.xdc[npl_off + iplane - 1] = 0.;
            else h.ydc[npl_off + iplane - 1] = 0.;
            drift = 0.5 * sdc_thick;
            radw = drift / sdc_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            drift = drift + sdc_wire_thick;
            project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
        }
        if (jchamber == 1) { /* :335-341 */
            if (t->xs > (sdc_1_bot - sdc_1x_offset) || t->xs < (sdc_1_top - sdc_1x_offset) ||
                t->ys > (sdc_1_left - sdc_1y_offset) || t->ys < (sdc_1_right - sdc_1y_offset)) {
                cnt[SO_DC1]++;
                *stop_code = STOP_SO_DC1;
                return 0;
            }
        } else { /* :385-391 */
            if (t->xs > (sdc_2_bot - sdc_2x_offset) || t->xs < (sdc_2_top - sdc_2x_offset) ||
                t->ys > (sdc_2_left - sdc_2y_offset) || t->ys < (sdc_2_right - sdc_2y_offset)) {
                cnt[SO_DC2]++;
                *stop_code = STOP_SO_DC2;
                return 0;
            }
        }
        radw = sdc_cath_thick / sdc_cath_radlen;
        if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    }

    for (int jchamber = 1; jchamber <= sdc_nr_cham; jchamber++) { /* :399-407 */
        int npl_off = (jchamber - 1) * sdc_nr_plan;
        for (int iplane = 1; iplane <= sdc_nr_plan; iplane++) {
            double z0 = jchamber == 1 ? sdc_1_zpos : sdc_2_zpos;
            h.zdc[npl_off + iplane - 1] = z0 + ((double)iplane - 0.5 - 0.5 * sdc_nr_plan) * sdc_del_plane;
        }
    }
    fit_planes(S, &h, x_fp, dx_fp, y_fp, dy_fp); /* :409-414 */
    *fit_done = 1;

#define SOS_SCIN_IN(L, R, YO, B, T, XO)                                   \
    (t->ys < ((L) + (YO)) && t->ys > ((R) + (YO)) && t->xs < ((B) + (XO)) && \
     t->xs > ((T) + (XO)))
#define SOS_STEP(DRIFT, RADLEN)                                                          \
    do {                                                                                 \
        drift = (DRIFT);                                                                 \
        radw = drift / (RADLEN);                                                         \
        project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);      \
        if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs); \
    } while (0)

    SOS_STEP(sscin_1y_zpos - sdc_2_zpos - 0.5 * sdc_nr_plan * sdc_del_plane, sair_radlen); /* :418-428 */
    if (SOS_SCIN_IN(sscin_1_left, sscin_1_right, sscin_1y_offset, sscin_1_bot, sscin_1_top, sscin_1x_offset))
        scincount = scincount + 1;
    radw = sscin_1y_thick / sscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    SOS_STEP(sscin_1x_zpos - sscin_1y_zpos, sair_radlen); /* :430-439 */
    if (SOS_SCIN_IN(sscin_1_left, sscin_1_right, sscin_1y_offset, sscin_1_bot, sscin_1_top, sscin_1x_offset))
        scincount = scincount + 1;
    radw = sscin_1x_thick / sscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    SOS_STEP(scer_zentrance - sscin_1x_zpos, sair_radlen); /* :443-448 */
    radw = scer_entr_thick / scer_entr_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    SOS_STEP(scer_zmirror - scer_zentrance, scer_radlen); /* :451-456 */
    radw = scer_mir_thick / scer_mir_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    SOS_STEP(scer_zexit - scer_zmirror, scer_radlen); /* :459-464 */
    radw = scer_exit_thick / scer_exit_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    SOS_STEP(sscin_2y_zpos - scer_zexit, sair_radlen); /* :469-478 */
    if (SOS_SCIN_IN(sscin_2_left, sscin_2_right, sscin_2y_offset, sscin_2_bot, sscin_2_top, sscin_2x_offset))
        scincount = scincount + 1;
    radw = sscin_2y_thick / sscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    SOS_STEP(sscin_2x_zpos - sscin_2y_zpos, sair_radlen); /* :480-489 */
    if (SOS_SCIN_IN(sscin_2_left, sscin_2_right, sscin_2y_offset, sscin_2_bot, sscin_2_top, sscin_2x_offset))
        scincount = scincount + 1;
    radw = sscin_2x_thick / sscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    if (scincount < scintrig) { /* :491-494 */
        cnt[SO_SCIN]++;
        *stop_code = STOP_SO_SCIN;
        return 0;
    }
    SOS_STEP(scal_4ta_zpos - sscin_2x_zpos, sair_radlen); /* :500-503 */
#undef SOS_STEP
#undef SOS_SCIN_IN
    return 1;
}

/* mc_sos, sos/mc_sos.f:1-394 (apertures sos/apertures_sos.inc:9-26) */
void or_mc_sos(const or_maps *M, const or_settings *S, or_rng *g, or_track *t, int64_t *cnt,
               double p_spec, double *dpp, double *x, double *y, double *z, double *dxdz,
               double *dydz, double *x_fp, double *dx_fp, double *y_fp, double *dy_fp,
               double *m2, double *resmult, double fry, int *ok_spec, double *pathlen,
               int *stop_code, int *fit_done)
{
    const or_fwd_map *F = M->sos_fwd;
    const double h_entr = 7.201, v_entr = 4.696, h_exit = 7.567, v_exit = 4.935;
    const double z_entr = 126.3e0, z_exit = z_entr + 6.3e0;
    const double r2_quad = 163.84, w_bm01 = 8.0, w_bm02 = 8.0;
    const double t_bm01_in = 41.32, b_bm01_in = -30.52, t_bm01_out = 53.05, b_bm01_out = -65.07;
    const double t_bm02_in = 51.73, b_bm02_in = -66.38, t_bm02_out = 51.52, b_bm02_out = -55.85;
    const double w_exit = 8.57, t_exit = 51.52, b_exit = -55.85;
    const int decay_flag = S->decay_flag, ms_flag = S->ms_flag, wcs_flag = S->wcs_flag;

    double p, xt, yt, zdrift, tmpwidth;
    double dpp_recon, dth_recon, dph_recon, y_recon;
    int dflag = 0, ok;
    *ok_spec = 0;
    *stop_code = STOP_OK;
    *fit_done = 0;
    cnt[SO_TRIALS]++;
    if (S->sieve == 1) sieve_slit(g, 1, z_entr, 1.524, 0.0, 0.254, *x, *y, dxdz, dydz); /* sos/mc_sos.f:122-137 */

    t->xs = *x;
    t->ys = *y;
    t->zs = *z;
    t->dxdzs = *dxdz;
    t->dydzs = *dydz;
    t->dpps = *dpp;
    p = p_spec * (1. + t->dpps / 100.);

#define SO_STOP(C, CODE) do { cnt[C]++; *stop_code = (CODE); return; } while (0)
    zdrift = z_entr; /* :171-187 */
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (fabs(t->ys) > h_entr) SO_STOP(SO_SLIT_HOR, STOP_SLIT_HOR);
    if (fabs(t->xs) > v_entr) SO_STOP(SO_SLIT_VERT, STOP_SLIT_VERT);
    if (fabs(t->xs) > (-v_entr / h_entr * fabs(t->ys) + 3 * v_entr / 2))
        SO_STOP(SO_SLIT_OCT, STOP_SLIT_OCT);
    zdrift = z_exit - z_entr; /* :191-204 */
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (fabs(t->ys) > h_exit) SO_STOP(SO_SLIT_HOR, STOP_SLIT_HOR);
    if (fabs(t->xs) > v_exit) SO_STOP(SO_SLIT_VERT, STOP_SLIT_VERT);
    if (fabs(t->xs) > (-v_exit / h_exit * fabs(t->ys) + 3 * v_exit / 2))
        SO_STOP(SO_SLIT_OCT, STOP_SLIT_OCT);

    zdrift = F->driftdist[0] - z_exit; /* :209-215 */
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if ((t->xs * t->xs + t->ys * t->ys) > r2_quad) SO_STOP(SO_QUAD_IN, STOP_SO_QUAD_IN);
    transp(F, g, t, S, 2, decay_flag, &dflag, m2, &p, 35.0e0, pathlen); /* :219-223 */
    if ((t->xs * t->xs + t->ys * t->ys) > r2_quad) SO_STOP(SO_QUAD_MID, STOP_SO_QUAD_MID);
    transp(F, g, t, S, 3, decay_flag, &dflag, m2, &p, 35.0e0, pathlen); /* :227-231 */
    if ((t->xs * t->xs + t->ys * t->ys) > r2_quad) SO_STOP(SO_QUAD_OUT, STOP_SO_QUAD_OUT);

    transp(F, g, t, S, 4, decay_flag, &dflag, m2, &p, 80.0e0, pathlen); /* :236-245 */
    xt = t->xs;
    yt = t->ys;
    or_rotate_haxis(t, -45.0e0, &xt, &yt);
    if ((yt > w_bm01) || (-yt > (w_bm01 - 0.05 * 2.54)) || (-xt > t_bm01_in) || (-xt < b_bm01_in))
        SO_STOP(SO_BM01_IN, STOP_SO_BM01_IN);
    transp(F, g, t, S, 5, decay_flag, &dflag, m2, &p, 169.52e0, pathlen); /* :250-257 */
    xt = t->xs;
    yt = t->ys;
    or_rotate_haxis(t, 45.0e0, &xt, &yt);
    if ((fabs(yt) > w_bm01) || (-xt > t_bm01_out) || (-xt < b_bm01_out))
        SO_STOP(SO_BM01_OUT, STOP_SO_BM01_OUT);
    transp(F, g, t, S, 6, decay_flag, &dflag, m2, &p, 80.80e0, pathlen); /* :262-269 */
    xt = t->xs;
    yt = t->ys;
    or_rotate_haxis(t, 49.0e0, &xt, &yt);
    if ((fabs(yt) > w_bm02) || (-xt > t_bm02_in) || (-xt < b_bm02_in))
        SO_STOP(SO_BM02_IN, STOP_SO_BM02_IN);
    transp(F, g, t, S, 7, decay_flag, &dflag, m2, &p, 77.06e0, pathlen); /* :274-281 */
    xt = t->xs;
    yt = t->ys;
    or_rotate_haxis(t, 57.0e0, &xt, &yt);
    if ((fabs(yt) > w_bm02) || (-xt > t_bm02_out) || (-xt < b_bm02_out))
        SO_STOP(SO_BM02_OUT, STOP_SO_BM02_OUT);

    transp(F, g, t, S, 8, decay_flag, &dflag, m2, &p, 43.82e0, pathlen); /* :311-319 */
    xt = t->xs;
    yt = t->ys;
    or_rotate_haxis(t, 45.0e0, &xt, &yt);
    if ((fabs(yt) > w_exit) || (-xt > t_exit) || (-xt < b_exit)) SO_STOP(SO_EXIT, STOP_SO_EXIT_FLANGE);
    transp(F, g, t, S, 9, decay_flag, &dflag, m2, &p, 44.34e0, pathlen); /* :329-336 */
    tmpwidth = 10.998 + 0.10209 * (t->xs + 37.694);
    if ((fabs(t->xs) > 37.694) || fabs(t->ys) > tmpwidth) SO_STOP(SO_EXIT, STOP_SO_EXIT_APERTURE);
    t->xs = t->xs + 7.62 * t->dxdzs; /* :342-350 */
    t->ys = t->ys + 7.62 * t->dydzs;
    if ((fabs(t->xs) > 38.1) || fabs(t->ys) > 12.7) SO_STOP(SO_EXIT, STOP_SO_EXIT_WINDOW);
#undef SO_STOP

    cnt[SO_HUT]++; /* :366-372 */
    or_rng hut_store;
    or_rng *hg = or_hut_stream(g, &hut_store);
    ok = sos_hut(g, hg, t, S, cnt, m2, &p, x_fp, dx_fp, y_fp, dy_fp, ms_flag, wcs_flag, decay_flag,
                 &dflag, resmult, -3.206267e0, pathlen, stop_code, fit_done);
    if (hg != g) g->used += hg->used;
    if (!ok) return;

    t->xs = *x_fp;
    t->ys = *y_fp;
    t->dxdzs = *dx_fp;
    t->dydzs = *dy_fp;
    recon_sos(M->sos_rec, t, fry, &dpp_recon, &dth_recon, &dph_recon, &y_recon);
    *dpp = dpp_recon;
    *dxdz = dph_recon;
    *dydz = dth_recon;
    *y = y_recon;
    *ok_spec = 1;
    cnt[SO_SUCCESSES]++;
}

/* ========================================================================= */
/* HRS (HRS-R = spectrometer 3, HRS-L = spectrometer 4)                      */
/* ========================================================================= */

/* mc_hrsl_hut / mc_hrsr_hut, hrsl/mc_hrsl_hut.f:1-498 (constants :37-185); the two
 * huts are the same code with their own counters.  resmult is 1 (:231). */
static int hrs_hut(or_rng *g, or_rng *hg, or_track *t, const or_settings *S, int64_t *cnt, double *m2,
                   double *p, double *x_fp, double *dx_fp, double *y_fp, double *dy_fp,
                   int ms_flag, int wcs_flag, int decay_flag, int *dflag, double *resmult,
                   double zinit, double *pathlen, int *stop_code, int *fit_done)
{
    const double hfoil_exit_radlen = 3.56, hfoil_exit_thick = 0.01;
    const double hair_radlen = 30420.;
    const double hdc_entr_radlen = 34.4, hdc_entr_thick = 0.00018 * 2.54;
    const double hdc_radlen = 16700.0, hdc_thick = 1.5;
    const double hdc_wire_radlen = 0.35, hdc_wire_thick = 0.0000049;
    const double hdc_cath_radlen = 7.2, hdc_cath_thick = 0.000177;
    const double hdc_exit_radlen = 34.4, hdc_exit_thick = 0.00018 * 2.54;
    const double hscin_radlen = 42.4;
    const double hcer_entr_radlen = 8.90, hcer_entr_thick = 0.040 * 2.54;
    const double hcer_radlen = 36620.0;
    const double hcer_mir_radlen = 400.0, hcer_mir_thick = 2.0;
    const double hcer_exit_radlen = 8.90, hcer_exit_thick = 0.040 * 2.54;
    const int hdc_nr_cham = 2, hdc_nr_plan = 6;
    const double hdc_1_zpos = -25.0 + 25.0, hdc_2_zpos = 25.0 + 25.0;
    const double hdc_del_plane = hdc_thick + hdc_wire_thick + hdc_cath_thick;
    const double hdc_left = 14.4, hdc_right = -14.4, hdc_top = -105.6, hdc_bot = 105.6;
    const double hdc_x_offset = 0.000, hdc_y_offset = 0.000;
    const double hscin_1x_zpos = 95.0 + 25.0, hscin_2x_zpos = 288.3 + 25.0;
    const double hscin_1x_thick = 0.5 * 1.067, hscin_2x_thick = 0.5 * 1.067;
    const double hscin_1x_left = 18.0, hscin_1x_right = -18.0;
    const double hscin_2x_left = 30.0, hscin_2x_right = -30.0;
    const double hcer_zentrance = 137.0 + 25.0, hcer_zmirror = 197.0 + 25.0, hcer_zexit = 237.0 + 25.0;
    const double hcal_4ta_zpos = 407.3 + 25.0;
    const double nsig_max = 99.0e0;
    const double sigma = S->wire_sigma > 0 ? S->wire_sigma : 0.0225; /* hdc_sigma, :113-114 */

    double radw, drift, tmpran1, tmpran2, xt, yt;
    hut_hits h;

    for (int i = 0; i < 12; i++) { h.xdc[i] = 0.; h.ydc[i] = 0.; }
    *resmult = 1.0;

    radw = hfoil_exit_thick / hfoil_exit_radlen; /* exit foil at zinit, :240-241 */
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);

    for (int jchamber = 1; jchamber <= hdc_nr_cham; jchamber++) {
        if (jchamber == 1) /* :247-250 */
            drift = (hdc_1_zpos - 0.5 * hdc_nr_plan * hdc_del_plane) - zinit;
        else /* :310-313 */
            drift = hdc_2_zpos - hdc_1_zpos - hdc_nr_plan * hdc_del_plane;
        radw = drift / hair_radlen;
        project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
        if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
        radw = hdc_entr_thick / hdc_entr_radlen;
        if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
        int npl_off = (jchamber - 1) * hdc_nr_plan;
        for (int iplane = 1; iplane <= hdc_nr_plan; iplane++) { /* :256-285, :319-348 */
            radw = hdc_cath_thick / hdc_cath_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            drift = 0.5 * hdc_thick;
            radw = drift / hdc_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            drift = drift + hdc_cath_thick;
            project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
            radw = hdc_wire_thick / hdc_wire_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            if (wcs_flag) {
                tmpran1 = or_gauss1(hg, nsig_max);
                tmpran2 = or_gauss1(hg, nsig_max);
            } else {
                tmpran1 = 0.;
                tmpran2 = 0.;
            }
            h.xdc[npl_off + iplane - 1] = t->xs + sigma * tmpran1 * *resmult;
            h.ydc[npl_off + iplane - 1] = t->ys + sigma * tmpran2 * *resmult;
            if (iplane == 1 || iplane == 3 || iplane == 5) h.xdc[npl_off + iplane - 1] = 0.;
            else h.ydc[npl_off + iplane - 1] = 0.;
            drift = 0.5 * hdc_thick;
            radw = drift / hdc_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            drift = drift + hdc_wire_thick;
            project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
        }
        radw = hdc_exit_thick / hdc_exit_radlen;
        if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
        /* the VDC frame, rotated by 45 degrees (:291-304, :354-367) */
        xt = t->xs;
        yt = t->ys;
        or_rotate_haxis(t, 45.0e0, &xt, &yt);
        if (xt > (hdc_bot - hdc_x_offset) || xt < (hdc_top - hdc_x_offset) ||
            yt > (hdc_left - hdc_y_offset) || yt < (hdc_right - hdc_y_offset)) {
            cnt[jchamber == 1 ? R_DC1 : R_DC2]++;
            *stop_code = jchamber == 1 ? STOP_R_DC1 : STOP_R_DC2;
            return 0;
        }
        radw = hdc_cath_thick / hdc_cath_radlen;
        if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    }

    for (int jchamber = 1; jchamber <= hdc_nr_cham; jchamber++) { /* :373-381 */
        int npl_off = (jchamber - 1) * hdc_nr_plan;
        for (int iplane = 1; iplane <= hdc_nr_plan; iplane++) {
            double z0 = jchamber == 1 ? hdc_1_zpos : hdc_2_zpos;
            h.zdc[npl_off + iplane - 1] = z0 + ((double)iplane - 0.5 - 0.5 * hdc_nr_plan) * hdc_del_plane;
        }
    }
    fit_planes(S, &h, x_fp, dx_fp, y_fp, dy_fp); /* :383-389 */
    *fit_done = 1;

#define HRS_STEP(DRIFT, RADLEN)                                                          \
    do {                                                                                 \
        drift = (DRIFT);                                                                 \
        radw = drift / (RADLEN);                                                         \
        project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);      \
        if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs); \
    } while (0)

    HRS_STEP(hscin_1x_zpos - hdc_2_zpos - 0.5 * hdc_nr_plan * hdc_del_plane, hair_radlen); /* :393-405 */
    if (t->ys > hscin_1x_left || t->ys < hscin_1x_right) {
        cnt[R_S1]++;
        *stop_code = STOP_R_S1;
        return 0;
    }
    radw = hscin_1x_thick / hscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    HRS_STEP(hcer_zentrance - hscin_1x_zpos, hair_radlen); /* :410-416 */
    radw = hcer_entr_thick / hcer_entr_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    HRS_STEP(hcer_zmirror - hcer_zentrance, hcer_radlen); /* :418-424 */
    radw = hcer_mir_thick / hcer_mir_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    HRS_STEP(hcer_zexit - hcer_zmirror, hcer_radlen); /* :426-432 */
    radw = hcer_exit_thick / hcer_exit_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    HRS_STEP(hscin_2x_zpos - hcer_zexit, hair_radlen); /* :436-448 */
    if (t->ys > hscin_2x_left || t->ys < hscin_2x_right) {
        cnt[R_S2]++;
        *stop_code = STOP_R_S2;
        return 0;
    }
    radw = hscin_2x_thick / hscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    HRS_STEP(hcal_4ta_zpos - hscin_2x_zpos, hair_radlen); /* :453-456 */
#undef HRS_STEP
    return 1;
}

/* mc_hrsr / mc_hrsl, hrsr/mc_hrsr.f:1-550, hrsl/mc_hrsl.f:1-551 (apertures
 * hrsl/apertures_hrsl.inc:6-8).  The two arms differ in the slit position,
 * the exit slit's half width and the VDC offset subtracted from y_fp. */
void or_mc_hrs(const or_maps *M, int arm, const or_settings *S, or_rng *g, or_track *t,
               int64_t *cnt, double p_spec, double *dpp, double *x, double *y, double *z,
               double *dxdz, double *dydz, double *x_fp, double *dx_fp, double *y_fp,
               double *dy_fp, double *m2, double *resmult, double fry, int *ok_spec,
               double *pathlen, int *stop_code, int *fit_done)
{
    const int left = arm == 4;
    const or_fwd_map *F = left ? M->hrsl_fwd : M->hrsr_fwd;
    const or_recon_map *R = left ? M->hrsl_rec : M->hrsr_rec;
    const double h_entr = 3.145, v_entr = 6.090, v_exit = 6.485;
    const double h_exit = left ? 3.335 : 3.340;
    const double y_off = 0.0, x_off = 0.0, z_off = 0.0;
    const double z_entr = (left ? 110.9e0 : 110.0e0) + z_off, z_exit = z_entr + 8.0e0;
    const double r_Q1 = 15.0, r_Q2 = 30.22, r_Q3 = 30.22;
    const double vdc_offset = left ? 0.78 : 0.48;
    const int decay_flag = S->decay_flag, ms_flag = S->ms_flag, wcs_flag = S->wcs_flag;

    double p, xt, yt, zdrift, ztmp;
    double dpp_recon, dth_recon, dph_recon, y_recon;
    int dflag = 0, ok;
    *ok_spec = 0;
    *stop_code = STOP_OK;
    *fit_done = 0;
    cnt[R_TRIALS]++;
    if (S->sieve == 1) sieve_slit(g, 2, z_exit, 2.50, 1.25, 0.1, *x, *y, dxdz, dydz); /* hrsl/mc_hrsl.f:122-132 */

    t->xs = *x;
    t->ys = *y;
    t->zs = *z;
    t->dxdzs = *dxdz;
    t->dydzs = *dydz;
    t->dpps = *dpp;
    p = p_spec * (1. + t->dpps / 100.);

#define R_STOP(C, CODE) do { cnt[C]++; *stop_code = (CODE); return; } while (0)
#define R_DRIFT(D) project(g, t, S, &t->xs, &t->ys, (D), decay_flag, &dflag, m2, &p, pathlen)
#define R_RAD (sqrt(t->xs * t->xs + t->ys * t->ys))
    zdrift = 65.686; /* circular apertures before the slit box, :159-179 */
    ztmp = zdrift;
    R_DRIFT(zdrift);
    if (R_RAD > 7.3787) R_STOP(R_SLIT_HOR, STOP_R_PIPE_65);
    zdrift = 80.436 - ztmp;
    ztmp = 80.436;
    R_DRIFT(zdrift);
    if (R_RAD > 7.4092) R_STOP(R_SLIT_HOR, STOP_R_PIPE_80);

    zdrift = z_entr - ztmp; /* fixed slit, :184-218 */
    R_DRIFT(zdrift);
    if (fabs(t->ys - y_off) > h_entr) R_STOP(R_SLIT_HOR, STOP_SLIT_HOR);
    if (fabs(t->xs - x_off) > v_entr) R_STOP(R_SLIT_VERT, STOP_SLIT_VERT);
    zdrift = z_exit - z_entr;
    R_DRIFT(zdrift);
    if (fabs(t->ys - y_off) > h_exit) R_STOP(R_SLIT_HOR, STOP_SLIT_HOR);
    if (fabs(t->xs - x_off) > v_exit) R_STOP(R_SLIT_VERT, STOP_SLIT_VERT);

    ztmp = 135.064; /* :222-230 */
    zdrift = ztmp - z_exit;
    R_DRIFT(zdrift);
    if (R_RAD > 12.5222) R_STOP(R_Q1_IN, STOP_R_Q1_PRE);
    zdrift = F->driftdist[0] - ztmp; /* Q1, :234-265 */
    R_DRIFT(zdrift);
    if ((t->xs * t->xs + t->ys * t->ys) > r_Q1 * r_Q1) R_STOP(R_Q1_IN, STOP_R_Q1_IN);
    transp(F, g, t, S, 2, decay_flag, &dflag, m2, &p, 62.75333333e0, pathlen);
    if ((t->xs * t->xs + t->ys * t->ys) > r_Q1 * r_Q1) R_STOP(R_Q1_MID, STOP_R_Q1_MID);
    transp(F, g, t, S, 3, decay_flag, &dflag, m2, &p, 31.37666667e0, pathlen);
    if ((t->xs * t->xs + t->ys * t->ys) > r_Q1 * r_Q1) R_STOP(R_Q1_OUT, STOP_R_Q1_OUT);

    zdrift = 300.464 - 253.16; /* :268-290 */
    ztmp = zdrift;
    R_DRIFT(zdrift);
    if (R_RAD > 14.9225) R_STOP(R_Q1_OUT, STOP_R_Q1_POST);
    zdrift = 314.464 - 300.464;
    ztmp = ztmp + zdrift;
    R_DRIFT(zdrift);
    if (R_RAD > 20.9550) R_STOP(R_Q2_IN, STOP_R_Q2_PRE);
    zdrift = F->driftdist[3] - ztmp; /* Q2, :293-324 */
    R_DRIFT(zdrift);
    if ((t->xs * t->xs + t->ys * t->ys) > r_Q2 * r_Q2) R_STOP(R_Q2_IN, STOP_R_Q2_IN);
    transp(F, g, t, S, 5, decay_flag, &dflag, m2, &p, 121.77333333e0, pathlen);
    if ((t->xs * t->xs + t->ys * t->ys) > r_Q2 * r_Q2) R_STOP(R_Q2_MID, STOP_R_Q2_MID);
    transp(F, g, t, S, 6, decay_flag, &dflag, m2, &p, 60.88666667e0, pathlen);
    if ((t->xs * t->xs + t->ys * t->ys) > r_Q2 * r_Q2) R_STOP(R_Q2_OUT, STOP_R_Q2_OUT);

    zdrift = 609.664 - 553.020; /* :327-359 */
    ztmp = zdrift;
    R_DRIFT(zdrift);
    if (R_RAD > 30.0073) R_STOP(R_Q2_OUT, STOP_R_Q2_POST1);
    zdrift = 641.800 - 609.664;
    ztmp = ztmp + zdrift;
    R_DRIFT(zdrift);
    if (R_RAD > 30.0073) R_STOP(R_Q2_OUT, STOP_R_Q2_POST2);
    zdrift = 819.489 - 641.800;
    ztmp = ztmp + zdrift;
    R_DRIFT(zdrift);
    if (fabs(t->xs) > 50.0 || fabs(t->ys) > 15.0) R_STOP(R_D1_IN, STOP_R_D1_PRE);

    zdrift = F->driftdist[6] - ztmp; /* D1 entrance, :361-383 */
    R_DRIFT(zdrift);
    xt = t->xs;
    yt = t->ys;
    or_rotate_haxis(t, -30.0, &xt, &yt);
    if (fabs(xt - 2.500) > 52.5) R_STOP(R_D1_IN, STOP_R_D1_IN);
    if ((fabs(yt) + 0.01861 * xt) > 12.5) R_STOP(R_D1_IN, STOP_R_D1_IN);
    transp(F, g, t, S, 8, decay_flag, &dflag, m2, &p, 659.73445725e0, pathlen); /* :385-405 */
    xt = t->xs;
    yt = t->ys;
    or_rotate_haxis(t, 30.0, &xt, &yt);
    if (fabs(xt - 2.500) > 52.5) R_STOP(R_D1_OUT, STOP_R_D1_OUT);
    if ((fabs(yt) + 0.01861 * xt) > 12.5) R_STOP(R_D1_OUT, STOP_R_D1_OUT);

    zdrift = 1745.33546 - 1655.83446; /* :409-437 */
    ztmp = zdrift;
    R_DRIFT(zdrift);
    if (R_RAD > 30.3276) R_STOP(R_D1_OUT, STOP_R_D1_POST_R);
    if (fabs(t->xs) > 50.0 || fabs(t->ys) > 15.0) R_STOP(R_D1_OUT, STOP_R_D1_POST_BOX);
    zdrift = 1759.00946 - 1745.33546;
    ztmp = ztmp + zdrift;
    R_DRIFT(zdrift);
    if (R_RAD > 30.3276) R_STOP(R_Q3_IN, STOP_R_Q3_PRE);
    zdrift = F->driftdist[8] - ztmp; /* Q3, :441-472 */
    R_DRIFT(zdrift);
    if ((t->xs * t->xs + t->ys * t->ys) > r_Q3 * r_Q3) R_STOP(R_Q3_IN, STOP_R_Q3_IN);
    transp(F, g, t, S, 10, decay_flag, &dflag, m2, &p, 121.7866667e0, pathlen);
    if ((t->xs * t->xs + t->ys * t->ys) > r_Q3 * r_Q3) R_STOP(R_Q3_MID, STOP_R_Q3_MID);
    transp(F, g, t, S, 11, decay_flag, &dflag, m2, &p, 60.89333333e0, pathlen);
    if ((t->xs * t->xs + t->ys * t->ys) > r_Q3 * r_Q3) R_STOP(R_Q3_OUT, STOP_R_Q3_OUT);

    zdrift = 2080.38746 - 1997.76446; /* :475-499 */
    ztmp = zdrift;
    R_DRIFT(zdrift);
    if (fabs(t->xs) > 35.56 || fabs(t->ys) > 17.145) R_STOP(R_Q3_OUT, STOP_R_Q3_POST);
    zdrift = 2327.47246 - 2080.38746;
    ztmp = ztmp + zdrift;
    R_DRIFT(zdrift);
    if (fabs(t->xs) > 99.76635 || fabs(t->ys) > 17.145) R_STOP(R_Q3_OUT, STOP_R_WINDOW);
#undef R_RAD
#undef R_DRIFT
#undef R_STOP

    cnt[R_HUT]++; /* :503-510 */
    zdrift = F->driftdist[11] - ztmp;
    or_rng hut_store;
    or_rng *hg = or_hut_stream(g, &hut_store);
    ok = hrs_hut(g, hg, t, S, cnt, m2, &p, x_fp, dx_fp, y_fp, dy_fp, ms_flag, wcs_flag, decay_flag,
                 &dflag, resmult, -zdrift, pathlen, stop_code, fit_done);
    if (hg != g) g->used += hg->used;
    if (!ok) return;

    t->xs = *x_fp; /* :512-535: recon from the unshifted track, y_fp shifted by the VDC offset */
    t->ys = *y_fp;
    t->dxdzs = *dx_fp;
    t->dydzs = *dy_fp;
    *y_fp = *y_fp - vdc_offset;
    recon(R, t, fry, &dpp_recon, &dth_recon, &dph_recon, &y_recon);
    *dpp = dpp_recon;
    *dxdz = dph_recon;
    *dydz = dth_recon;
    *y = y_recon;
    *ok_spec = 1;
    cnt[R_SUCCESSES]++;
}

/* ========================================================================= */
/* SHMS                                                                       */
/* ========================================================================= */

/* mc_shms_hut, shms/mc_shms_hut.f:1-458 with shms/hut.inc (cer_flag forced
 * true, vac_flag false by mc_shms, shms/mc_shms.f:257-258; use_det_cut true) */
static int shms_hut(or_rng *g, or_rng *hg, or_track *t, const or_settings *S, int64_t *cnt, double *m2,
                    double *p, double *x_fp, double *dx_fp, double *y_fp, double *dy_fp,
                    int ms_flag, int wcs_flag, int decay_flag, int *dflag, double *resmult,
                    double *pathlen, int *stop_code, int *fit_done)
{
    const double hfoil_exit_radlen = 8.89, hfoil_exit_thick = 0.020 * 2.54;
    const double hair_radlen = 30420.;
    const double hdc_entr_radlen = 28.7, hdc_entr_thick = 0.001 * 2.54;
    const double hdc_radlen = 16700.0, hdc_thick = 0.125 * 2.54;
    const double hdc_wire_radlen = 0.35, hdc_wire_thick = 0.0000354;
    const double hdc_cath_radlen = 28.6, hdc_cath_thick = 0.001 * 2.54;
    const double hdc_exit_radlen = 28.7, hdc_exit_thick = 0.001 * 2.54;
    const double hscin_radlen = 42.4;
    const double hcer_entr_radlen = 19.63, hcer_entr_thick = 0.002 * 2.54;
    const double hcer_1_radlen = 11700.0;
    const double hcer_mirglass_radlen = 12.29, hcer_mirglass_thick = 0.3;
    const double hcer_exit_radlen = 19.63, hcer_exit_thick = 0.002 * 2.54;
    const double hcer_2_entr_radlen = 8.90, hcer_2_entr_thick = 0.040 * 2.54;
    const double hcer_2_radlen = 1202.5;
    const double hcer_mir_radlen = 400., hcer_mir_thick = 2.00;
    const double hcer_2_exit_radlen = 8.90, hcer_2_exit_thick = 0.040 * 2.54;
    const int hdc_nr_plan = 6;
    const double hdc_1_zpos = -40.656, hdc_2_zpos = 39.332;
    const double hdc_1_left = 40.0, hdc_1_right = -40.0, hdc_1y_offset = 0.0, hdc_1_top = -40.,
                 hdc_1_bot = 40., hdc_1x_offset = 0.0;
    const double hdc_2_left = 40.0, hdc_2_right = -40.0, hdc_2y_offset = 0.0, hdc_2_top = -40.,
                 hdc_2_bot = 40., hdc_2x_offset = 0.;
    const double hscin_1x_zpos = 52.1, hscin_1y_zpos = 61.7, hscin_2x_zpos = 271.4,
                 hscin_2y_zpos = 282.4;
    const double hscin_1x_thick = 1.000 * 1.067, hscin_1y_thick = 1.000 * 1.067,
                 hscin_2x_thick = 1.000 * 1.067, hscin_2y_thick = 1.000 * 1.067;
    const double hscin_1x_left = 50., hscin_1x_right = -50., hscin_1x_offset = 0.0;
    const double hscin_1y_top = -45., hscin_1y_bot = 45., hscin_1y_offset = 0.0;
    const double hscin_2x_left = 55., hscin_2x_right = -55., hscin_2x_offset = 0.;
    const double hscin_2y_top = -62.5, hscin_2y_bot = 62.5, hscin_2y_offset = 0;
    const double hcer_1_zentrance = -291.700, hcer_1_zmirror = -84.900, hcer_1_zexit = -61.700;
    const double hcer_2_zentrance = 72.600, hcer_2_zmirror = 179.400, hcer_2_zexit = 202.600;
    const double hcal_4ta_zpos = 341.0;
    const double hcal_left = 63.00, hcal_right = -63.00, hcal_top = -70.00, hcal_bottom = 70.00;
    const double nsig_max = 99.0e0;
    double sigma = S->wire_sigma > 0 ? S->wire_sigma : 0.020; /* hdc_sigma, hut.inc:125-126 */

    double hdc_del_plane = hdc_thick + hdc_wire_thick + hdc_cath_thick;
    double radw, drift, tmpran1, tmpran2;
    hut_hits h;
    *resmult = 1.0;
    for (int i = 0; i < 12; i++) { h.xdc[i] = 0.; h.ydc[i] = 0.; }

    radw = hfoil_exit_thick / hfoil_exit_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    radw = hcer_entr_thick / hcer_entr_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    drift = hcer_1_zmirror - hcer_1_zentrance - hcer_mirglass_thick / 2;
    radw = drift / hcer_1_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    radw = hcer_mirglass_thick / hcer_mirglass_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    drift = hcer_1_zexit - hcer_1_zmirror - hcer_mirglass_thick / 2;
    radw = drift / hcer_1_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    radw = hcer_exit_thick / hcer_exit_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);

    drift = (hdc_1_zpos - 0.5 * hdc_nr_plan * hdc_del_plane) - hcer_1_zexit;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);

    for (int jchamber = 1; jchamber <= 2; jchamber++) {
        if (jchamber == 2) {
            drift = hdc_2_zpos - hdc_1_zpos - hdc_nr_plan * hdc_del_plane;
            project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
            radw = drift / hair_radlen;
            if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
        }
        radw = hdc_entr_thick / hdc_entr_radlen;
        if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
        int npl_off = (jchamber - 1) * hdc_nr_plan;
        for (int iplane = 1; iplane <= hdc_nr_plan; iplane++) {
            radw = hdc_cath_thick / hdc_cath_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            drift = 0.5 * hdc_thick;
            radw = drift / hdc_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            drift = drift + hdc_cath_thick;
            project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
            radw = hdc_wire_thick / hdc_wire_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            if (wcs_flag) {
                tmpran1 = or_gauss1(hg, nsig_max);
                tmpran2 = or_gauss1(hg, nsig_max);
            } else {
                tmpran1 = 0.;
                tmpran2 = 0.;
            }
            h.xdc[npl_off + iplane - 1] = t->xs + sigma * tmpran1 * *resmult;
            h.ydc[npl_off + iplane - 1] = t->ys + sigma * tmpran2 * *resmult;
            if (iplane == 2 || iplane == 5) h.xdc[npl_off + iplane - 1] = 0.;
            else h.ydc[npl_off + iplane - 1] = 0.;
            drift = 0.5 * hdc_thick;
            radw = drift / hdc_radlen;
            if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
            drift = drift + hdc_wire_thick;
            project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
        }
        radw = hdc_exit_thick / hdc_exit_radlen;
        if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
        if (jchamber == 1) {
            if (t->xs > (hdc_1_bot - hdc_1x_offset) || t->xs < (hdc_1_top - hdc_1x_offset) ||
                t->ys > (hdc_1_left - hdc_1y_offset) || t->ys < (hdc_1_right - hdc_1y_offset)) {
                cnt[S_DC1]++;
                *stop_code = STOP_S_DC1;
                return 0;
            }
        } else {
            if (t->xs > (hdc_2_bot - hdc_2x_offset) || t->xs < (hdc_2_top - hdc_2x_offset) ||
                t->ys > (hdc_2_left - hdc_2y_offset) || t->ys < (hdc_2_right - hdc_2y_offset)) {
                cnt[S_DC2]++;
                *stop_code = STOP_S_DC2;
                return 0;
            }
        }
        radw = hdc_cath_thick / hdc_cath_radlen;
        if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    }

    drift = hscin_1x_zpos - hdc_2_zpos - 0.5 * hdc_nr_plan * hdc_del_plane;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    if (t->ys > (hscin_1x_left + hscin_1y_offset) || t->ys < (hscin_1x_right + hscin_1y_offset)) {
        cnt[S_S1]++;
        *stop_code = STOP_S_S1X;
        return 0;
    }
    radw = hscin_1x_thick / hscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    drift = hscin_1y_zpos - hscin_1x_zpos;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    if (t->xs > (hscin_1y_bot + hscin_1x_offset) || t->xs < (hscin_1y_top + hscin_1x_offset)) {
        cnt[S_S1]++;
        *stop_code = STOP_S_S1Y;
        return 0;
    }
    radw = hscin_1y_thick / hscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);

    drift = hcer_2_zentrance - hscin_1y_zpos;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    radw = hcer_2_entr_thick / hcer_2_entr_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    drift = hcer_2_zmirror - hcer_2_zentrance;
    radw = drift / hcer_2_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    radw = hcer_mir_thick / hcer_mir_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    drift = hcer_2_zexit - hcer_2_zmirror;
    radw = drift / hcer_2_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    radw = hcer_2_exit_thick / hcer_2_exit_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);

    drift = hscin_2x_zpos - hcer_2_zexit;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    if (t->ys > (hscin_2x_left + hscin_2y_offset) || t->ys < (hscin_2x_right + hscin_2y_offset)) {
        cnt[S_S3]++;
        *stop_code = STOP_S_S2X;
        return 0;
    }
    radw = hscin_2x_thick / hscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);
    drift = hscin_2y_zpos - hscin_2x_zpos;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    if (t->xs > (hscin_2y_bot + hscin_2x_offset) || t->xs < (hscin_2y_top + hscin_2x_offset)) {
        cnt[S_S2]++;
        *stop_code = STOP_S_S2Y;
        return 0;
    }
    radw = hscin_2y_thick / hscin_radlen;
    if (ms_flag) musc(hg, *m2, *p, radw, &t->dydzs, &t->dxdzs);

    drift = hcal_4ta_zpos - hscin_2y_zpos;
    radw = drift / hair_radlen;
    project(g, t, S, &t->xs, &t->ys, drift, decay_flag, dflag, m2, p, pathlen);
    if (ms_flag) musc_ext(hg, *m2, *p, radw, drift, &t->dydzs, &t->dxdzs, &t->ys, &t->xs);
    if (t->ys > hcal_left || t->ys < hcal_right || t->xs > hcal_bottom || t->xs < hcal_top) {
        cnt[S_CAL]++;
        *stop_code = STOP_S_CAL;
        return 0;
    }

    for (int jchamber = 1; jchamber <= 2; jchamber++) {
        int npl_off = (jchamber - 1) * hdc_nr_plan;
        for (int iplane = 1; iplane <= hdc_nr_plan; iplane++) {
            double z0 = jchamber == 1 ? hdc_1_zpos : hdc_2_zpos;
            h.zdc[npl_off + iplane - 1] = z0 + ((double)iplane - 0.5 - 0.5 * hdc_nr_plan) * hdc_del_plane;
        }
    }
    fit_planes(S, &h, x_fp, dx_fp, y_fp, dy_fp);
    *fit_done = 1;

    double xcal = *x_fp + *dx_fp * hcal_4ta_zpos;
    double ycal = *y_fp + *dy_fp * hcal_4ta_zpos;
    if (ycal > (hcal_left - 5.0) || ycal < (hcal_right + 5.0) || xcal > (hcal_bottom - 5.0) ||
        xcal < (hcal_top + 5.0)) {
        cnt[S_CAL]++;
        *stop_code = STOP_S_CALFID;
        return 0;
    }
    return 1;
}

/* mc_shms, shms/mc_shms.f:1-1108 (use_coll true, skip_hb / sieves false) */
void or_mc_shms(const or_maps *M, const or_settings *S, or_rng *g, or_track *t, int64_t *cnt,
                double p_spec, double *dpp, double *x, double *y, double *z, double *dxdz,
                double *dydz, double *x_fp, double *dx_fp, double *y_fp, double *dy_fp,
                double *m2, double *resmult, double fry, int *ok_spec, double *pathlen,
                int *stop_code, int *fit_done)
{
    const or_fwd_map *F = M->shms_fwd;
    const double h_entr = 8.5, v_entr = 12.5, h_exit = 8.65, v_exit = 12.85, x_off = +0.00,
                 y_off = +0.00;
    const double zd_hbin = 118.39, zd_hbmen = 17.61, zd_hbmex = 80.0, zd_hbout = 17.61;
    const double z_entr = 25.189, z_thick = 6.35;
    const double zd_q1in = 58.39, zd_q1men = 28.35, zd_q1mid = 93.65, zd_q1mex = 93.65,
                 zd_q1out = 28.35;
    const double zd_q2in = 25.55, zd_q2men = 39.1, zd_q2mid = 79.35, zd_q2mex = 79.35,
                 zd_q2out = 39.1;
    const double zd_q3in = 28.10, zd_q3men = 39.1, zd_q3mid = 79.35, zd_q3mex = 79.35,
                 zd_q3out = 39.1;
    const double zd_q3d1trans = 18.00, zd_d1flare = 30.10, zd_d1men = 39.47;
    const double zd_d1mid = 36.406263, zd_d1out = 60.68, zd_fp = 307.95;
    const double hcer_1_zentrance = -291.700;
    /* shms/apertures_shms.inc:9-28 */
    const double r_HBx = 11.2, r_HBfym = -4.13, r_HBfyp = 11.75, r_HBmenym = -5.45,
                 r_HBmenyp = 11.74, r_HBmexym = -10.25, r_HBmexyp = 11.71, r_HBbym = -11.71,
                 r_HBbyp = 11.70;
    const double r_Q1 = 20.00, r_Q2 = 30.00, r_Q3 = 30.00, r_D1 = 30.00;
    const int decay_flag = S->decay_flag, ms_flag = S->ms_flag, wcs_flag = S->wcs_flag;

    double p, xt, yt, zdrift;
    double dpp_recon, dth_recon, dph_recon, y_recon;
    int dflag = 0;
    *ok_spec = 0;
    *stop_code = STOP_OK;
    *fit_done = 0;
    cnt[S_TRIALS]++;
    if (S->sieve & 2) /* shms/mc_shms.f:276-290: in front of the bender */
        sieve_slit(g, 0, zd_hbin, 2.5, 1.64, 0.3, *x, *y, dxdz, dydz);
    if (S->sieve & 1) /* :293-310: at the collimator */
        sieve_slit(g, 0, zd_hbin + zd_hbmen + zd_hbmex + zd_hbout + z_entr, 2.5, 1.64, 0.3, *x, *y, dxdz, dydz);

    t->dpps = *dpp;
    p = p_spec * (1. + t->dpps / 100.);
    if (S->mom_offset != 0.0) t->dpps = 100. * ((1. + t->dpps / 100.) / (1. + S->mom_offset) - 1.);
    t->xs = *x;
    t->ys = *y;
    t->zs = *z;
    t->dxdzs = *dxdz;
    t->dydzs = *dydz;

#define STOPS(c, code) do { cnt[c]++; *stop_code = (code); return; } while (0)
    /* HB, :421-495 */
    zdrift = zd_hbin;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    xt = t->xs; yt = t->ys;
    or_rotate_vaxis(t, 1.5, &xt, &yt);
    yt = yt + 1.51;
    if ((xt * xt > r_HBx * r_HBx) || (yt > r_HBfyp) || (yt < r_HBfym)) STOPS(S_HB_IN, STOP_S_HB_IN);
    zdrift = zd_hbmen;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    xt = t->xs; yt = t->ys;
    or_rotate_vaxis(t, 1.5, &xt, &yt);
    yt = yt + 0.98;
    if ((xt * xt > r_HBx * r_HBx) || (yt > r_HBmenyp) || (yt < r_HBmenym)) STOPS(S_HB_MEN, STOP_S_HB_MEN);
    transp(F, g, t, S, 3, decay_flag, &dflag, m2, &p, zd_hbmex, pathlen);
    xt = t->xs; yt = t->ys;
    or_rotate_vaxis(t, -1.5, &xt, &yt);
    yt = yt + 0.98;
    if ((xt * xt > r_HBx * r_HBx) || (yt > r_HBmexyp) || (yt < r_HBmexym)) STOPS(S_HB_MEX, STOP_S_HB_MEX);
    zdrift = zd_hbout;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    xt = t->xs; yt = t->ys;
    or_rotate_vaxis(t, -1.5, &xt, &yt);
    yt = yt + 1.51;
    if ((xt * xt > r_HBx * r_HBx) || (yt > r_HBbyp) || (yt < r_HBbym)) STOPS(S_HB_OUT, STOP_S_HB_OUT);

    /* collimator, :500-568 */
    if (S->using_coll && (*m2 > 100.0 * 100.0) && (*m2 < 200.0 * 200.0)) {
        zdrift = z_entr;
        project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
        if (!coll_step(g, t, S, cnt, 5, m2, &p, p_spec, decay_flag, &dflag, pathlen))
            STOPS(S_COLL, STOP_COLL);
    } else {
        zdrift = z_entr;
        project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
        xt = t->xs; yt = t->ys;
        if (fabs(yt - y_off) > h_entr) STOPS(S_SLIT_HOR, STOP_SLIT_HOR);
        if (fabs(xt - x_off) > v_entr) STOPS(S_SLIT_VERT, STOP_SLIT_VERT);
        if (fabs(xt - x_off) > (-(v_entr / h_entr * fabs(yt - y_off)) + 3 * v_entr / 2))
            STOPS(S_SLIT_OCT, STOP_SLIT_OCT);
        zdrift = z_thick;
        project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
        xt = t->xs; yt = t->ys;
        if (fabs(yt - y_off) > (h_exit)) STOPS(S_SLIT_HOR, STOP_SLIT_HOR);
        if (fabs(xt - x_off) > (v_exit)) STOPS(S_SLIT_VERT, STOP_SLIT_VERT);
        if (fabs(xt - x_off) > ((-v_exit) / (h_exit)*fabs(yt - y_off) + 3 * (v_exit) / 2))
            STOPS(S_SLIT_OCT, STOP_SLIT_OCT);
    }

#define RQ(r) ((t->xs * t->xs + t->ys * t->ys) > (r) * (r))
    /* Q1, :571-642 */
    zdrift = zd_q1in - z_entr - z_thick;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_Q1)) STOPS(S_Q1_IN, STOP_S_Q1_IN);
    zdrift = zd_q1men;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_Q1)) STOPS(S_Q1_MEN, STOP_S_Q1_MEN);
    transp(F, g, t, S, 7, decay_flag, &dflag, m2, &p, zd_q1mid, pathlen);
    if (RQ(r_Q1)) STOPS(S_Q1_MID, STOP_S_Q1_MID);
    transp(F, g, t, S, 8, decay_flag, &dflag, m2, &p, zd_q1mex, pathlen);
    if (RQ(r_Q1)) STOPS(S_Q1_MEX, STOP_S_Q1_MEX);
    zdrift = zd_q1out;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_Q1)) STOPS(S_Q1_OUT, STOP_S_Q1_OUT);
    /* Q2, :644-712 */
    zdrift = zd_q2in;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_Q2)) STOPS(S_Q2_IN, STOP_S_Q2_IN);
    zdrift = zd_q2men;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_Q2)) STOPS(S_Q2_MEN, STOP_S_Q2_MEN);
    transp(F, g, t, S, 12, decay_flag, &dflag, m2, &p, zd_q2mid, pathlen);
    if (RQ(r_Q2)) STOPS(S_Q2_MID, STOP_S_Q2_MID);
    transp(F, g, t, S, 13, decay_flag, &dflag, m2, &p, zd_q2mex, pathlen);
    if (RQ(r_Q2)) STOPS(S_Q2_MEX, STOP_S_Q2_MEX);
    zdrift = zd_q2out;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_Q2)) STOPS(S_Q2_OUT, STOP_S_Q2_OUT);
    /* Q3, :714-781 */
    zdrift = zd_q3in;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_Q3)) STOPS(S_Q3_IN, STOP_S_Q3_IN);
    zdrift = zd_q3men;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_Q3)) STOPS(S_Q3_MEN, STOP_S_Q3_MEN);
    transp(F, g, t, S, 17, decay_flag, &dflag, m2, &p, zd_q3mid, pathlen);
    if (RQ(r_Q3)) STOPS(S_Q3_MID, STOP_S_Q3_MID);
    transp(F, g, t, S, 18, decay_flag, &dflag, m2, &p, zd_q3mex, pathlen);
    if (RQ(r_Q3)) STOPS(S_Q3_MEX, STOP_S_Q3_MEX);
    zdrift = zd_q3out;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_Q3)) STOPS(S_Q3_OUT, STOP_S_Q3_OUT);
    /* D1 mechanical entrance, :815-827 */
    zdrift = zd_q3d1trans;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    if (RQ(r_D1)) STOPS(S_D1_IN, STOP_S_D1_IN);
#undef RQ
#define RT(r) ((xt * xt + yt * yt) > (r) * (r))
    /* flare and magnetic entrance, :829-869 */
    zdrift = zd_d1flare;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    xt = t->xs; yt = t->ys;
    or_rotate_haxis(t, 9.200, &xt, &yt);
    xt = xt - 3.5;
    if (RT(r_D1)) STOPS(S_D1_FLR, STOP_S_D1_FLR);
    zdrift = zd_d1men;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    xt = t->xs; yt = t->ys;
    or_rotate_haxis(t, 9.200, &xt, &yt);
    xt = xt + 2.82;
    if (RT(r_D1)) STOPS(S_D1_MEN, STOP_S_D1_MEN);
    /* dipole interior, :871-990; magnetic exit :995-1011 */
    {
        static const double ang[8] = {6.9, 4.6, 2.3, 0.0, -2.3, -4.6, -6.9, -9.2};
        static const double off[8] = {8.05, 11.75, 13.96, 14.70, 13.96, 11.75, 8.05, 2.82};
        static const int ctr[8] = {S_D1_MID1, S_D1_MID2, S_D1_MID3, S_D1_MID4,
                                   S_D1_MID5, S_D1_MID6, S_D1_MID7, S_D1_MEX};
        for (int k = 0; k < 8; k++) {
            transp(F, g, t, S, 23 + k, decay_flag, &dflag, m2, &p, zd_d1mid, pathlen);
            xt = t->xs; yt = t->ys;
            or_rotate_haxis(t, ang[k], &xt, &yt);
            xt = xt + off[k];
            if (RT(r_D1)) STOPS(ctr[k], STOP_S_D1_MID1 + k);
        }
    }
    /* mechanical exit, :1015-1031 */
    zdrift = zd_d1out;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    xt = t->xs; yt = t->ys;
    or_rotate_haxis(t, -9.20, &xt, &yt);
    xt = xt - 6.88;
    if (RT(r_D1)) STOPS(S_D1_OUT, STOP_S_D1_OUT);
#undef RT
#undef STOPS

    /* to the Cherenkov entrance, :1036-1047 */
    zdrift = zd_fp + hcer_1_zentrance;
    project(g, t, S, &t->xs, &t->ys, zdrift, decay_flag, &dflag, m2, &p, pathlen);
    cnt[S_HUT]++;
    t->xs_fp = t->xs;
    t->ys_fp = t->ys;
    t->dxdzs_fp = t->dxdzs;
    t->dydzs_fp = t->dydzs;

    or_rng hut_store;
    or_rng *hg = or_hut_stream(g, &hut_store);
    int ok = shms_hut(g, hg, t, S, cnt, m2, &p, x_fp, dx_fp, y_fp, dy_fp, ms_flag, wcs_flag,
                      decay_flag, &dflag, resmult, pathlen, stop_code, fit_done);
    if (hg != g) g->used += hg->used;
    if (!ok) return;
    t->xs = *x_fp;
    t->ys = *y_fp;
    t->dxdzs = *dx_fp;
    t->dydzs = *dy_fp;
    recon(M->shms_rec, t, fry, &dpp_recon, &dth_recon, &dph_recon, &y_recon);
    *dpp = dpp_recon;
    *dxdz = dph_recon;
    *dydz = dth_recon;
    *y = y_recon;
    *ok_spec = 1;
    cnt[S_SUCCESSES]++;
}

/* ========================================================================= */
/* Batch driver                                                               */
/* ========================================================================= */

int or_run(const or_maps *M, int arm, const or_settings *S, int64_t n, const or_events_in *in,
           or_events_out *out, int rng_kind, const double *tape, int64_t stride,
           uint64_t seed, uint64_t event0, int nthreads, int64_t *counters)
{
    if (!M || !S || !in || !out || n < 0) return OR_E_ARG;
    if (arm < 1 || arm > 5) return OR_E_ARG;
    if (arm == 2 && !M->sos_fwd) return OR_E_ARG;
    if ((arm == 3 && !M->hrsr_fwd) || (arm == 4 && !M->hrsl_fwd)) return OR_E_ARG;
    if (rng_kind == 0 && (!tape || stride <= 0)) return OR_E_TAPE;
    int ncount = arm == 1 ? H_NCOUNT : arm == 2 ? SO_NCOUNT : arm == 5 ? S_NCOUNT : R_NCOUNT;
    or_ranlux *rl = NULL;
    if (rng_kind == 2) {
        rl = or_ranlux_new(3, (int)(seed & 0x7fffffff));
        nthreads = 1;
    }
    if (nthreads <= 0) nthreads = 1;
    int64_t tape_overflow = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads) reduction(+ : tape_overflow)
#endif
    {
        int64_t cnt[S_NCOUNT];
        memset(cnt, 0, sizeof cnt);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
        for (int64_t i = 0; i < n; i++) {
            or_rng g;
            memset(&g, 0, sizeof g);
            g.kind = rng_kind;
            if (rng_kind == 0) {
                g.tape = tape + i * stride;
                g.stride = stride;
            } else if (rng_kind == 1) {
                g.key[0] = (uint32_t)seed;
                g.key[1] = (uint32_t)(seed >> 32);
                g.event = event0 + (uint64_t)i;
            } else {
                g.rl = rl;
            }
            or_track t;
            memset(&t, 0, sizeof t);
            double dpp = in->dpp[i], x = in->x[i], y = in->y[i], z = in->z ? in->z[i] : 0.0;
            double dxdz = in->dxdz[i], dydz = in->dydz[i];
            double fry = in->fry ? in->fry[i] : x;
            double x_fp = out->x_fp[i], dx_fp = out->dx_fp[i], y_fp = out->y_fp[i],
                   dy_fp = out->dy_fp[i];
            double m2 = S->m2, resmult = out->resmult ? out->resmult[i] : 0.0;
            double pathlen = out->pathlen[i];
            int ok, stop, fit_done;
            if (arm == 1)
                or_mc_hms(M, S, &g, &t, cnt, S->p_spec, &dpp, &x, &y, &z, &dxdz, &dydz, &x_fp,
                          &dx_fp, &y_fp, &dy_fp, &m2, &resmult, fry, &ok, &pathlen, &stop,
                          &fit_done);
            else if (arm == 2)
                or_mc_sos(M, S, &g, &t, cnt, S->p_spec, &dpp, &x, &y, &z, &dxdz, &dydz, &x_fp,
                          &dx_fp, &y_fp, &dy_fp, &m2, &resmult, fry, &ok, &pathlen, &stop,
                          &fit_done);
            else if (arm == 3 || arm == 4)
                or_mc_hrs(M, arm, S, &g, &t, cnt, S->p_spec, &dpp, &x, &y, &z, &dxdz, &dydz,
                          &x_fp, &dx_fp, &y_fp, &dy_fp, &m2, &resmult, fry, &ok, &pathlen, &stop,
                          &fit_done);
            else
                or_mc_shms(M, S, &g, &t, cnt, S->p_spec, &dpp, &x, &y, &z, &dxdz, &dydz, &x_fp,
                           &dx_fp, &y_fp, &dy_fp, &m2, &resmult, fry, &ok, &pathlen, &stop,
                           &fit_done);
            if (g.overflow) {
                stop = STOP_TAPE_OVERFLOW;
                ok = 0;
                tape_overflow++;
            }
            out->dpp[i] = dpp;
            out->y[i] = y;
            out->dxdz[i] = dxdz;
            out->dydz[i] = dydz;
            out->x_fp[i] = x_fp;
            out->dx_fp[i] = dx_fp;
            out->y_fp[i] = y_fp;
            out->dy_fp[i] = dy_fp;
            out->pathlen[i] = pathlen;
            if (out->resmult) out->resmult[i] = resmult;
            if (out->m2) out->m2[i] = m2;
            out->ok[i] = (uint8_t)ok;
            out->stop[i] = (uint8_t)stop;
            if (out->used) out->used[i] = (int32_t)g.used;
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (int c = 0; c < ncount; c++) counters[c] += cnt[c];
        }
    }
    if (rl) free(rl);
    return tape_overflow ? OR_E_TAPE : OR_OK;
}

/* Reconstruction only (mc_hms_recon / mc_shms_recon) for n focal-plane tracks. */
int or_recon_run(const or_maps *M, int arm, int64_t n, const double *x_fp, const double *dx_fp,
                 const double *y_fp, const double *dy_fp, const double *fry, double *delta,
                 double *dxdz, double *dydz, double *ytar)
{
    if (!M || arm < 1 || arm > 5 || n < 0) return OR_E_ARG;
    const or_recon_map *R = arm == 1 ? M->hms_rec : arm == 2 ? M->sos_rec : arm == 3 ? M->hrsr_rec
                          : arm == 4 ? M->hrsl_rec : M->shms_rec;
    if (!R) return OR_E_ARG;
    for (int64_t i = 0; i < n; i++) {
        or_track t;
        memset(&t, 0, sizeof t);
        t.xs = x_fp[i];
        t.dxdzs = dx_fp[i];
        t.ys = y_fp[i];
        t.dydzs = dy_fp[i];
        double dp, dt, dph, yt;
        if (arm == 2) recon_sos(R, &t, fry ? fry[i] : 0.0, &dp, &dt, &dph, &yt);
        else recon(R, &t, fry ? fry[i] : 0.0, &dp, &dt, &dph, &yt);
        delta[i] = dp;
        dxdz[i] = dph;
        dydz[i] = dt;
        ytar[i] = yt;
    }
    return OR_OK;
}

int64_t or_philox_normals(uint64_t seed, uint64_t event, int n, double nsigmax, double *out)
{
    or_rng g;
    memset(&g, 0, sizeof g);
    g.kind = 1;
    g.key[0] = (uint32_t)seed;
    g.key[1] = (uint32_t)(seed >> 32);
    g.event = event;
    for (int i = 0; i < n; ++i) out[i] = or_gauss1(&g, nsigmax);
    return g.pos;
}

/* gauss1 on a uniform tape (unit tests): returns the draw, *used = uniforms read. */
double or_gauss1_tape(const double *tape, int64_t len, double nsigmax, int64_t *used)
{
    or_rng g;
    memset(&g, 0, sizeof g);
    g.kind = 0;
    g.tape = tape;
    g.stride = len;
    double v = or_gauss1(&g, nsigmax);
    if (used) *used = g.overflow ? -1 : g.pos;
    return v;
}

/* RANLUX state after RLUXGO(lux, seed, 0, 0) in RLUXUT form, then n values of
 * RANLUX(RVEC, n) (unit tests). */
void or_ranlux_sequence(int lux, int seed, int32_t isdext[25], double *rvec, int n)
{
    or_ranlux *r = or_ranlux_new(lux, seed);
    or_ranlux_save(r, isdext);
    or_ranlux_fill(r, rvec, n);
    free(r);
}

/* ========================================================================= */
/* Target step: the event preparation before an arm                          */
/* ========================================================================= */

/* Window and wall materials, target.inc:5-28 (X0_cm = X0 / rho there). */
#define OR_RHO_AL 2.70
#define OR_Z_AL 13.
#define OR_A_AL 26.98
#define OR_X0CM_AL (24.01 / 2.70)
#define OR_RHO_MYLAR 1.39
#define OR_Z_MYLAR 4.545
#define OR_A_MYLAR 8.735
#define OR_X0CM_MYLAR (39.95 / 1.39)
#define OR_RHO_KEVLAR 0.74
#define OR_Z_KEVLAR 2.67
#define OR_A_KEVLAR 4.67
#define OR_X0CM_KEVLAR (55.2 / 0.74)
#define OR_RHO_AIR 0.00121
#define OR_Z_AIR 7.2
#define OR_A_AIR 14.4
#define OR_X0CM_AIR (36.66 / 0.00121)
#define OR_INCH_CM 2.54
#define OR_TARGET_PI 3.14159265358979

/* Radiation length of the target nucleus in g/cm^2, init.f:18-42. */
double or_target_X0(double A, double Z)
{
    const double alpha = 1. / 137.0359895; /* constants.inc:44 */
    double L1 = log(184.15) - log(Z) / 3.0;
    double L2 = log(1194.) - 2. * log(Z) / 3.0;
    if (Z == 1) {
        L1 = 5.31;
        L2 = 6.144;
    }
    double za2 = (Z * alpha) * (Z * alpha);
    double fc = za2 * (1.202 + za2 * (-1.0369 + za2 * 1.008 / (za2 + 1)));
    long nA = lround(A);
    if (nA == 1) return 61.28;
    if (nA == 2) return 122.4;
    if (nA == 4) return 94.32;
    return 716.405 * A / Z / (Z * (L1 - fc) + L2);
}

/* trip_thru_target, target.f:1-246: path lengths through the target and the
 * windows for particle narm (1 beam, 2 scattered electron, 3 hadron) seen by
 * spectrometer `spec` (electron_arm / hadron_arm), the radiation lengths
 * crossed and the sampled energy loss (enerloss_new, draws in the order
 * target, Al, air, kevlar, mylar; a zero thickness draws nothing).
 * Returns 0, or -1 for a spectrometer code the reference does not define. */
int or_trip_thru_target(or_rng *g, const or_target *T, int narm, int spec, double zpos,
                        double energy, double theta, double mass, int typeflag, double *eloss,
                        double *radlen)
{
    const double X0_cm = T->X0 / T->rho;
    const int liquid = T->Z < 2.4;
    double s_Al = 0.0, s_air = 0.0, s_kevlar = 0.0, s_mylar = 0.0, s_target, forward_path;
    double e_t, e_al, e_air, e_kev, e_myl;
    if (narm == 1) { /* target.f:33-50 */
        s_target = (T->length / 2. + zpos) / fabs(cos(T->angle));
        if (liquid) {
            if (T->can == 1) s_Al = s_Al + 0.0028 * OR_INCH_CM;
            else if (T->can == 2) s_Al = s_Al + 0.0050 * OR_INCH_CM;
            else if (T->can == 3) s_Al = s_Al + 0.013;
        }
        *radlen = s_target / X0_cm + s_Al / OR_X0CM_AL;
        or_enerloss_new(g, s_target, T->rho, T->Z, T->A, energy, mass, typeflag, &e_t);
        or_enerloss_new(g, s_Al, OR_RHO_AL, OR_Z_AL, OR_A_AL, energy, mass, typeflag, &e_al);
        *eloss = e_t + e_al;
        return 0;
    }
    /* target.f:71-101 (electron) and :184-217 (hadron): the same tables */
    if (spec == 1) {
        s_Al = 0.016 * OR_INCH_CM; s_air = 15; s_kevlar = 0.015 * OR_INCH_CM;
        s_mylar = 0.005 * OR_INCH_CM;
        forward_path = (T->length / 2. - zpos) / fabs(cos(theta + T->angle));
    } else if (spec == 2) {
        s_Al = 0.008 * OR_INCH_CM; s_air = 15; s_kevlar = 0.005 * OR_INCH_CM;
        s_mylar = 0.003 * OR_INCH_CM;
        forward_path = (T->length / 2. - zpos) / fabs(cos(theta - T->angle));
    } else if (spec == 3 || spec == 4) {
        s_Al = 0.013 * OR_INCH_CM; s_air = 15; s_kevlar = 0. * OR_INCH_CM;
        s_mylar = 0.010 * OR_INCH_CM;
        forward_path = (T->length / 2. - zpos) / fabs(cos(theta - T->angle));
    } else if (spec == 5 || spec == 6) {
        s_Al = (0.02 + 0.01) * OR_INCH_CM; s_air = 57.27; s_kevlar = 0.0; s_mylar = 0.0;
        forward_path = (T->length / 2. - zpos) / fabs(cos(theta - T->angle));
    } else {
        return -1;
    }
    s_target = forward_path;
    if (liquid) {
        if (T->can == 1) { /* beer can, target.f:105-113 */
            double side_path = 1.325 * OR_INCH_CM / fabs(sin(theta));
            if (forward_path < side_path) {
                s_Al = s_Al + 0.005 * OR_INCH_CM / fabs(cos(theta));
            } else {
                s_target = side_path;
                s_Al = s_Al + 0.005 * OR_INCH_CM / fabs(sin(theta));
            }
        } else if (T->can == 2) { /* pudding can, target.f:114-139 */
            double tt = tan(theta);
            double t = tt * tt;
            double atmp = 1 + t;
            double btmp = -2 * zpos * t;
            double ctmp = zpos * zpos * t - (T->length / 2.) * (T->length / 2.);
            double z_can = (-btmp + sqrt(btmp * btmp - 4. * atmp * ctmp)) / 2. / atmp;
            double side_path = (z_can - zpos) / fabs(cos(theta));
            s_target = side_path;
            double costmp = z_can / (T->length / 2.);
            double th_can = 0.0;
            if (fabs(costmp) <= 1) th_can = acos(z_can / (T->length / 2.));
            /* |costmp - 1| <= 1e-6 gives 0; beyond that the reference keeps
             * the previous event's value (a saved local); 0 here */
            s_Al = s_Al + 0.0050 * OR_INCH_CM / fabs(sin(OR_TARGET_PI / 2 - (theta - th_can)));
        } else if (T->can == 3) { /* 2017 cryo cell, target.f:140-165 */
            double ecir = 1.315 * 2.54;
            double ecor = (1.315 + 0.0071) * 2.54;
            double entec = T->length - ecir;
            double twall = ecor - ecir;
            double tcm = zpos + T->length / 2.0;
            double tliquid, tal;
            if ((tcm + ecir / tan(theta)) < entec) {
                tliquid = ecir / sin(theta);
                tal = twall / sin(theta);
            } else {
                /* the compiled reference evaluates sin and cos of theta in
                 * this branch with one sincos call (the compiler merges them),
                 * whose sine can differ from sin() in the last bit */
                double sn, cs;
                sincos(theta, &sn, &cs);
                double d = T->length - ecir - tcm;
                double ds = d * sn;
                tliquid = (sqrt(ecir * ecir - ds * ds) + d * cs);
                tal = +(sqrt(ecor * ecor - ds * ds) - sqrt(ecir * ecir - ds * ds)) * twall /
                      (ecor - ecir);
            }
            s_Al = s_Al + tal;
            s_target = tliquid;
        }
    }
    *radlen = s_target / X0_cm + s_Al / OR_X0CM_AL + s_air / OR_X0CM_AIR +
              s_kevlar / OR_X0CM_KEVLAR + s_mylar / OR_X0CM_MYLAR;
    or_enerloss_new(g, s_target, T->rho, T->Z, T->A, energy, mass, typeflag, &e_t);
    or_enerloss_new(g, s_Al, OR_RHO_AL, OR_Z_AL, OR_A_AL, energy, mass, typeflag, &e_al);
    or_enerloss_new(g, s_air, OR_RHO_AIR, OR_Z_AIR, OR_A_AIR, energy, mass, typeflag, &e_air);
    or_enerloss_new(g, s_kevlar, OR_RHO_KEVLAR, OR_Z_KEVLAR, OR_A_KEVLAR, energy, mass, typeflag,
                    &e_kev);
    or_enerloss_new(g, s_mylar, OR_RHO_MYLAR, OR_Z_MYLAR, OR_A_MYLAR, energy, mass, typeflag,
                    &e_myl);
    *eloss = e_t + e_al + e_air + e_kev + e_myl;
    return 0;
}

/* target_musc, target.f:548-574: Lynch-Dahl planar angles, each truncated at
 * 3.5 sigma (gauss1). */
void or_target_musc(or_rng *g, double p, double beta, double teff, double dangles[2])
{
    const double Es = 13.6, epsilon = 0.088, nsig_max = 3.5;
    double theta_sigma = Es / p / beta * sqrt(teff) * (1 + epsilon * log10(teff / (beta * beta)));
    dangles[0] = theta_sigma * or_gauss1(g, nsig_max);
    dangles[1] = theta_sigma * or_gauss1(g, nsig_max);
}

/* One event of the target step for one particle:
 *   complete_ev, event.f:1031-1040 (trip_thru_target with typeflag 1 at the
 *     vertex energy, the loss zeroed without using_Eloss);
 *   montecarlo, simc.f:1376-1441 (hadron) / :1649-1706 (electron): delta,
 *     target_musc, the beam angles, the TRANSPORT transform, the offsets and
 *     the drift to z = 0; simc.f:1364-1369 for the beam (narm 1). */
static int prep_event(or_rng *g, const or_target *T, const or_prep *P, double cos_th,
                      double sin_th, double sin_phi, const or_vertex_in *in, int64_t i,
                      or_prep_out *out)
{
    const double tz = in->z[i];
    const double zpos = tz - T->zoffset;
    double eloss, teff, dang[2] = {0.0, 0.0};
    if (P->narm == 1) {
        if (or_trip_thru_target(g, T, 1, 0, zpos, P->beam_energy, 0.0, P->mass, 1, &eloss, &teff))
            return -1;
        if (!P->using_eloss) eloss = 0.0;
        if (P->mc_smear) or_target_musc(g, in->E[i], 1.0, teff, dang);
        if (out->eloss) out->eloss[i] = eloss;
        if (out->teff) out->teff[i] = teff;
        if (out->dang1) out->dang1[i] = dang[0];
        if (out->dang2) out->dang2[i] = dang[1];
        return 0;
    }
    const double E = in->E[i], p = in->p[i];
    const double Ev = in->E_vertex ? in->E_vertex[i] : E;
    if (or_trip_thru_target(g, T, P->narm, P->spectrometer, zpos, Ev, in->theta[i], P->mass, 1,
                            &eloss, &teff))
        return -1;
    if (!P->using_eloss) eloss = 0.0;
    double delta;
    if (P->narm == 3) {
        if (P->using_eloss)
            delta = (sqrt(fabs((E - eloss) * (E - eloss) - P->mass * P->mass)) - P->p_spec) /
                    P->p_spec * 100.;
        else
            delta = in->delta[i];
        if (P->mc_smear) or_target_musc(g, p, p / E, teff, dang);
    } else {
        delta = 100 * (E - eloss - P->coulomb - P->p_spec) / P->p_spec;
        if (P->mc_smear) or_target_musc(g, p, 1.0, teff, dang);
    }
    const double din1 = in->dang_yp ? in->dang_yp[i] : 0.0;
    const double din2 = in->dang_xp ? in->dang_xp[i] : 0.0;
    const double yptar = in->yptar[i] + dang[0] + din1;
    const double xptar = in->xptar[i] + dang[1] + din2 * cos_th;
    const double tx = in->x[i], ty = in->y[i];
    double x = -ty;
    double y = -tx * cos_th - tz * sin_th * sin_phi;
    double z = tz * cos_th + tx * sin_th * sin_phi;
    x = x - P->offset[0];
    y = y - P->offset[1];
    z = z - P->offset[2];
    const double dx = xptar - P->offset[3];
    const double dy = yptar - P->offset[4];
    x = x - z * dx;
    y = y - z * dy;
    if (out->dpp) out->dpp[i] = delta;
    if (out->x) out->x[i] = x;
    if (out->y) out->y[i] = y;
    if (out->z) out->z[i] = 0.0;
    if (out->dxdz) out->dxdz[i] = dx;
    if (out->dydz) out->dydz[i] = dy;
    if (out->fry) out->fry[i] = x;
    if (out->eloss) out->eloss[i] = eloss;
    if (out->teff) out->teff[i] = teff;
    if (out->dang1) out->dang1[i] = dang[0];
    if (out->dang2) out->dang2[i] = dang[1];
    return 0;
}

int or_target_prep(const or_target *T, const or_prep *P, int64_t n, const or_vertex_in *in,
                   or_prep_out *out, int rng_kind, const double *tape, int64_t stride,
                   uint64_t seed, uint64_t event0, int nthreads, int32_t *used)
{
    const double cos_th = cos(P->th_spec), sin_th = sin(P->th_spec), sin_phi = sin(P->phi);
    if (nthreads <= 0) nthreads = 1;
    int64_t bad = 0, overflow = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : bad, overflow)
#endif
    for (int64_t i = 0; i < n; i++) {
        or_rng g;
        memset(&g, 0, sizeof g);
        g.kind = rng_kind;
        if (rng_kind == 0) {
            g.tape = tape + i * stride;
            g.stride = stride;
        } else {
            g.key[0] = (uint32_t)seed;
            g.key[1] = (uint32_t)(seed >> 32);
            g.event = event0 + (uint64_t)i;
            g.tag = OR_TARGET_TAG;
        }
        if (prep_event(&g, T, P, cos_th, sin_th, sin_phi, in, i, out)) bad++;
        if (g.overflow) overflow++;
        if (used) used[i] = (int32_t)g.used;
    }
    if (bad) return -1;
    return overflow ? -2 : 0;
}

/* ========================================================================= */
/* Target-field tracking (trg_track.f)                                        */
/* ========================================================================= */

/* trgInit, trg_track.f:243-325: 51 x 51 (z, r) grid at 2 cm spacing; the
 * map file's lines run over z fastest, columns 3 and 4 are B_z and B_r. */
int or_trg_init(or_field *F, const char *map, double theta_e, double phi_e, double theta_p,
                double phi_p)
{
    const double pi180 = 3.141592653 / 180.;
    (void)phi_e;
    (void)phi_p; /* "for performance reasons B_phi is always treated 0" (:255-256, :375) */
    F->stht[0] = sin(theta_e * pi180);
    F->ctht[0] = cos(theta_e * pi180);
    F->stht[1] = sin(theta_p * pi180);
    F->ctht[1] = cos(theta_p * pi180);
    for (int ir = 0; ir < OR_TRG_N; ir++) {
        F->rr[ir] = 2. * (double)ir;
        F->zz[ir] = 2. * (double)ir;
    }
    if (map && strcmp(map, "0") == 0) { /* :290-299 */
        memset(F->bz, 0, sizeof F->bz);
        memset(F->br, 0, sizeof F->br);
    } else if (map) { /* :300-310 */
        FILE *f = fopen(map, "r");
        if (!f) return -1;
        char line[512];
        for (int ir = 0; ir < OR_TRG_N; ir++)
            for (int iz = 0; iz < OR_TRG_N; iz++) {
                double v[7];
                if (!fgets(line, sizeof line, f) ||
                    sscanf(line, "%lf %lf %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3], &v[4],
                           &v[5], &v[6]) != 7) {
                    fclose(f);
                    return -1;
                }
                F->bz[ir][iz] = v[2];
                F->br[ir][iz] = v[3];
            }
        fclose(f);
    } else { /* uniform 5 T over 26 cm in z and 16 cm in r, :311-323 */
        for (int ir = 0; ir < OR_TRG_N; ir++)
            for (int iz = 0; iz < OR_TRG_N; iz++) {
                F->br[ir][iz] = 0.;
                F->bz[ir][iz] = (F->rr[ir] <= 16. && F->zz[iz] <= 26.) ? 5.0 : 0.0;
            }
    }
    return 0;
}

/* trgField, trg_track.f:328-418: rotate into the field frame (z' along the
 * field), bilinear interpolation in (|z'|, r), rotate back. */
void or_trg_field(const or_field *F, const double x_[3], double B_[3], int spect)
{
    const int k = spect == -1 ? 0 : 1;
    const double stht = F->stht[k], ctht = F->ctht[k];
    double x[3], B[3], z, r, az, ar, a0, a1;
    x[0] = x_[0];
    x[1] = stht * x_[2] + ctht * x_[1];
    x[2] = ctht * x_[2] - stht * x_[1];
    z = fabs(x[2]);
    r = sqrt(x[0] * x[0] + x[1] * x[1]);
    const int i = (int)((z - F->zz[0]) / (F->zz[1] - F->zz[0])) + 1;
    const int j = (int)((r - F->rr[0]) / (F->rr[1] - F->rr[0])) + 1;
    if ((i + 1 > OR_TRG_N) || (i < 1) || (j + 1 > OR_TRG_N) || (j < 1)) {
        B_[0] = 0.;
        B_[1] = 0.;
        B_[2] = 0.;
        return;
    }
    const int i0 = i - 1, j0 = j - 1; /* B(i, j) = b[j0][i0] */
    az = ((z - F->zz[i0]) / (F->zz[1] - F->zz[0]));
    ar = ((r - F->rr[j0]) / (F->rr[1] - F->rr[0]));
    a0 = az * (F->bz[j0][i0 + 1] - F->bz[j0][i0]) + F->bz[j0][i0];
    a1 = az * (F->bz[j0 + 1][i0 + 1] - F->bz[j0 + 1][i0]) + F->bz[j0 + 1][i0];
    B[2] = (ar * (a1 - a0) + a0);
    if (r > 0.) {
        a0 = az * (F->br[j0][i0 + 1] - F->br[j0][i0]) + F->br[j0][i0];
        a1 = az * (F->br[j0 + 1][i0 + 1] - F->br[j0 + 1][i0]) + F->br[j0 + 1][i0];
        B[1] = (ar * (a1 - a0) + a0) / r;
        if (x[2] < 0.) B[1] = -B[1];
        B[0] = B[1] * x[0];
        B[1] = B[1] * x[1];
        B_[0] = B[0];
        B_[1] = -stht * B[2] + ctht * B[1];
        B_[2] = ctht * B[2] + stht * B[1];
    } else {
        B_[0] = 0.;
        B_[1] = -stht * B[2];
        B_[2] = ctht * B[2];
    }
}

/* trgDeriv, trg_track.f:423-459 (the u(7..9) terms are not used by trgRK4) */
static void trg_deriv(const or_field *F, const double u[9], double dudt[9], int spect,
                      double factor)
{
    double B[3];
    or_trg_field(F, u, B, spect);
    dudt[0] = u[3];
    dudt[1] = u[4];
    dudt[2] = u[5];
    dudt[6] = u[4] * B[2] - u[5] * B[1];
    dudt[7] = u[5] * B[0] - u[3] * B[2];
    dudt[8] = u[3] * B[1] - u[4] * B[0];
    dudt[3] = dudt[6] * factor;
    dudt[4] = dudt[7] * factor;
    dudt[5] = dudt[8] * factor;
}

/* trgRK4, trg_track.f:461-498 (Numerical Recipes rk4 on u(1..6)) */
void or_trg_rk4(const or_field *F, const double u0[9], double u1[9], double h, int spect,
                double factor)
{
    double ut[9], dudt[9], dut[9], dum[9], hh, h6;
    hh = h * 0.5;
    h6 = h / 6.;
    trg_deriv(F, u0, dudt, spect, factor);
    for (int i = 0; i < 6; i++) ut[i] = u0[i] + hh * dudt[i];
    trg_deriv(F, ut, dut, spect, factor);
    for (int i = 0; i < 6; i++) ut[i] = u0[i] + hh * dut[i];
    trg_deriv(F, ut, dum, spect, factor);
    for (int i = 0; i < 6; i++) {
        ut[i] = u0[i] + h * dum[i];
        dum[i] = dut[i] + dum[i];
    }
    trg_deriv(F, ut, dut, spect, factor);
    for (int i = 0; i < 6; i++) u1[i] = u0[i] + h6 * (dudt[i] + dut[i] + 2. * dum[i]);
}

/* trgTrackToPlane, trg_track.f:154-238: step until the plane
 * a x + b y + c z + d = 0 is crossed, then interpolate linearly.  The first
 * step only chooses the direction; with no crossing in it the loop restarts
 * from it and compares against the starting distance (:211-224). */
void or_trg_track_to_plane(const or_field *F, double u[9], double E, double dl, double a,
                           double b, double c, double d, int *ok, int spect)
{
    double ts, n, an, bn, cn, dn, maxdist, dist0, dist1, u0[9], u1[9];
    int steps, max_steps;
    if (!*ok) return;
    n = 1 / sqrt(a * a + b * b + c * c);
    an = a * n;
    bn = b * n;
    cn = c * n;
    dn = d * n;
    const double factor = 90. / E;
    ts = -dl / sqrt(u[3] * u[3] + u[4] * u[4] + u[5] * u[5]);
    dist0 = u[0] * an + u[1] * bn + u[2] * cn + dn;
    maxdist = fmax(fabs(dist0) * 4., 1.0);
    or_trg_rk4(F, u, u1, ts, spect, factor);
    dist1 = u1[0] * an + u1[1] * bn + u1[2] * cn + dn;
    if ((copysign(1., dist0) == copysign(1., dist1)) && (fabs(dist0) < fabs(dist1))) ts = -ts;
    steps = 0;
    max_steps = (int)(fmax(dist0, 10. * dl) / dl) * 10;
    if (copysign(1., dist0) == copysign(1., dist1)) {
        dist1 = dist0;
        while ((copysign(1., dist0) == copysign(1., dist1)) && *ok) {
            or_trg_rk4(F, u1, u0, ts, spect, factor);
            dist0 = u0[0] * an + u0[1] * bn + u0[2] * cn + dn;
            if (copysign(1., dist0) == copysign(1., dist1)) {
                or_trg_rk4(F, u0, u1, ts, spect, factor);
                dist1 = u1[0] * an + u1[1] * bn + u1[2] * cn + dn;
            }
            *ok = (fabs(dist1) < maxdist) && steps < max_steps;
            steps = steps + 1;
        }
    } else {
        for (int i = 0; i < 6; i++) u0[i] = u[i];
    }
    if (*ok)
        for (int i = 0; i < 6; i++) u[i] = u0[i] + (u1[i] - u0[i]) * dist0 / (dist0 - dist1);
}

/* track_from_tgt, trg_track.f:591-658: from the vertex to the plane z = 0,
 * then to z = 100 cm (field free); the first call's status is not kept. */
void or_track_from_tgt(const or_field *F, double *x, double *y, double *z, double *dx,
                       double *dy, double mom, double mass, int spect, int *ok)
{
    const double cc = 29.9792458;
    double vT[9] = {0};
    const double vel = fabs(mom) / sqrt(mom * mom + mass * mass) * cc;
    const double eng = copysign(1., mom) * sqrt(mom * mom + mass * mass);
    vT[0] = *x;
    vT[1] = *y;
    vT[2] = *z;
    vT[5] = vel / sqrt(1 + *dx * *dx + *dy * *dy);
    vT[3] = *dx * vT[5];
    vT[4] = *dy * vT[5];
    *ok = 1;
    or_trg_track_to_plane(F, vT, eng, 1., 0., 0., 1., 0., ok, spect);
    *ok = 1;
    or_trg_track_to_plane(F, vT, eng, 1., 0., 0., 1., -100., ok, spect);
    *x = vT[0];
    *y = vT[1];
    *z = vT[2];
    *dx = vT[3] / vT[5];
    *dy = vT[4] / vT[5];
}

/* track_to_tgt, trg_track.f:738-876: iterate field tracking from z = 100 cm
 * back to the beam plane and the arm's reconstruction with the corrected
 * vertical position, at most 10 times. */
void or_track_to_tgt(const or_field *F, const or_maps *M, int arm, const double track[4],
                     double *delta, double *y, double *dx, double *dy, double frx, double fry,
                     double mom, double mass, double ctheta, double stheta, int spect, int *ok)
{
    const double cc = 29.9792458;
    double vT[9] = {0}, vTx[9] = {0}, xx, delx, xxd, vel, eng, mom_0;
    int n;
    xx = -fry;
    vel = fabs(mom) / sqrt(mom * mom + mass * mass) * cc;
    eng = copysign(1., mom) * sqrt(mom * mom + mass * mass);
    mom_0 = mom / (1.e0 + *delta / 100.e0);
    vT[0] = -fry + 100. * *dx;
    vT[1] = *y + 100. * *dy;
    vT[2] = 100.;
    vT[5] = vel / sqrt(1 + *dy * *dy + *dx * *dx);
    vT[3] = *dx * vT[5];
    vT[4] = *dy * vT[5];
    or_trg_track_to_plane(F, vT, eng, 1., 0., -ctheta, stheta, frx, ok, spect);
    n = 0;
    delx = 1.;
    while ((delx > .0001) && (n < 10) && *ok) {
        delx = fabs(-fry - vT[0]);
        vTx[0] = -fry;
        for (int i = 1; i < 6; i++) vTx[i] = vT[i];
        or_trg_track_to_plane(F, vT, eng, 1., 0., 0., 1., 0., ok, spect);
        or_trg_track_to_plane(F, vTx, eng, 1., 0., 0., 1., 0., ok, spect);
        xx = xx + fmin(1., fmax(-1., (vTx[0] - vT[0])));
        xxd = xx;
        /* mc_*_recon(delta, dy, dx, y, xxd) on /track/ (:836-846) */
        double dl_, dxdz_, dydz_, yt_;
        or_recon_run(M, arm, 1, &track[0], &track[1], &track[2], &track[3], &xxd, &dl_, &dxdz_,
                     &dydz_, &yt_);
        *delta = dl_;
        *dy = dydz_;
        *dx = dxdz_;
        *y = yt_;
        mom = mom_0 * (1.e0 + *delta / 100.e0);
        vel = fabs(mom) / sqrt(mom * mom + mass * mass) * cc;
        eng = copysign(1., mom) * sqrt(mom * mom + mass * mass);
        vT[0] = xx + 100. * *dx;
        vT[1] = *y + 100. * *dy;
        vT[2] = 100.;
        vT[5] = vel / sqrt(1 + *dy * *dy + *dx * *dx);
        vT[3] = *dx * vT[5];
        vT[4] = *dy * vT[5];
        or_trg_track_to_plane(F, vT, eng, 1., 0., -ctheta, stheta, frx, ok, spect);
        n = n + 1;
    }
    if (delx > .2) *ok = 0;
    *dy = vT[4] / vT[5];
    *dx = vT[3] / vT[5];
    *y = vT[1];
}

void or_track_from_tgt_run(const or_field *F, int64_t n, double *x, double *y, double *z,
                           double *dx, double *dy, const double *mom, double mass, int spect,
                           uint8_t *ok, int nthreads)
{
    if (nthreads <= 0) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
#endif
    for (int64_t i = 0; i < n; i++) {
        int o = 0;
        or_track_from_tgt(F, &x[i], &y[i], &z[i], &dx[i], &dy[i], mom[i], mass, spect, &o);
        ok[i] = (uint8_t)o;
    }
}

void or_track_to_tgt_run(const or_field *F, const or_maps *M, int arm, int64_t n,
                         const double *x_fp, const double *dx_fp, const double *y_fp,
                         const double *dy_fp, double *delta, double *y, double *dx, double *dy,
                         const double *frx, const double *fry, const double *mom, double mass,
                         double ctheta, double stheta, int spect, uint8_t *ok, int nthreads)
{
    if (nthreads <= 0) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
#endif
    for (int64_t i = 0; i < n; i++) {
        const double tr[4] = {x_fp[i], dx_fp[i], y_fp[i], dy_fp[i]};
        int o = 1;
        or_track_to_tgt(F, M, arm, tr, &delta[i], &y[i], &dx[i], &dy[i], frx[i], fry[i], mom[i],
                        mass, ctheta, stheta, spect, &o);
        ok[i] = (uint8_t)o;
    }
}

/* ========================================================================= */
/* Event generator: H(e,e'p) elastic, radiation off                          */
/* ========================================================================= */

/* constants.inc:17-44 */
#define OR_C_ME 0.51099906
#define OR_C_MP 938.27231
#define OR_C_HBARC 197.327053
#define OR_C_PI 3.141592653589793

/* physics_angles, event.f:1572-1616: spectrometer slopes (dx = xptar,
 * dy = yptar) about the central angles to the physics theta and phi. */
void or_physics_angles(double theta0, double phi0, double dx, double dy, double *theta,
                       double *phi)
{
    const double costh = cos(theta0), sinth = sin(theta0), sinph = sin(phi0);
    const double r = sqrt(1. + dx * dx + dy * dy);
    *theta = acos((costh - dy * sinth * sinph) / r);
    if (dx != 0.0) {
        double p = atan((dy * costh + sinth * sinph) / dx);
        if (p <= 0) p = p + OR_C_PI;
        if (sinph < 0.) p = p + OR_C_PI;
        *phi = p;
    } else {
        *phi = phi0;
    }
}

/* spectrometer_angles, event.f:1618-1647: the inverse of physics_angles */
void or_spectrometer_angles(double theta0, double phi0, double *dx, double *dy, double theta,
                            double phi)
{
    const double x = sin(theta) * cos(phi), y = sin(theta) * sin(phi), z = cos(theta);
    const double x0 = sin(theta0) * cos(phi0), y0 = sin(theta0) * sin(phi0), z0 = cos(theta0);
    const double cos_dtheta = x * x0 + y * y0 + z * z0;
    *dx = x / cos_dtheta;
    double d = sqrt(1 / (cos_dtheta * cos_dtheta) - 1. - *dx * *dx);
    const double y_event = y / cos_dtheta;
    if (y_event < y0) d = -d;
    *dy = d;
}

/* fofa_best_fit, physics_proton.f:137-171 (Bosted's fit, Eqs. 4 and 5).
 * The reference's compiler evaluates x**2. and x**3. as products (x*x*x)
 * and x**4., x**5. as pow calls; the same here. */
void or_fofa_best_fit(double qsquar, double *GE, double *GM)
{
    const double mu_p = 2.793;
    const double Q2 = -qsquar * (OR_C_HBARC * OR_C_HBARC) * 1.e-6;
    const double Q = sqrt(fmax(Q2, 0.e0));
    const double Q3 = Q * Q * Q, Q4 = pow(Q, 4.), Q5 = pow(Q, 5.);
    double denom = 1. + 0.62 * Q + 0.68 * Q2 + 2.8 * Q3 + 0.83 * Q4;
    *GE = 1. / denom;
    denom = 1. + 0.35 * Q + 2.44 * Q2 + 0.5 * Q3 + 1.04 * Q4 + 0.34 * Q5;
    *GM = mu_p / denom;
}

/* sigMott, physics_proton.f:176-190: microbarn/sr */
double or_sigmott(double e0, double theta, double Q2)
{
    const double alpha = 1. / 137.0359895;
    const double t = 2. * alpha * OR_C_HBARC * e0 * cos(theta / 2.) / Q2;
    const double sig = t * t;
    return sig * 1.e4;
}

/* sigep, physics_proton.f:1-23 */
double or_sigep(double Q2, double E, double theta, double Ein)
{
    const double Mp2 = OR_C_MP * OR_C_MP;
    double GE, GM;
    or_fofa_best_fit(-Q2 / (OR_C_HBARC * OR_C_HBARC), &GE, &GM);
    const double qmu4mp = Q2 / 4. / Mp2;
    const double W1p = GM * GM * qmu4mp;
    const double W2p = (GE * GE + GM * GM * qmu4mp) / (1.0 + qmu4mp);
    const double t = tan(theta / 2.);
    const double Wp = W2p + 2. * W1p * (t * t);
    return or_sigmott(E, theta, Q2) * E / Ein * Wp;
}

/* One event of generate (event.f:157-212, :275-277, :322-325, :381-404) with
 * complete_ev's hydrogen-elastic branch (:497-563, :1013-1014) and
 * complete_main's weight (:1397-1398, :1450-1452, :1558). */
static void gen_hep_event(or_rng *g, const or_target *T, const or_hep *H, int64_t i,
                          or_hep_out *o)
{
    const double nsig_max = 3.0;
    double x = or_gauss1(g, nsig_max) * H->xwid + H->xoffset;
    double y = or_gauss1(g, nsig_max) * H->ywid + H->yoffset;
    double t3, t4, t5 = 0.0, t6 = 0.0;
    if (H->fr_pattern == 1) {
        t3 = or_grnd(g) * OR_C_PI;
        t4 = or_grnd(g) * OR_C_PI;
        t5 = cos(t3) * H->fr1;
        t6 = cos(t4) * H->fr2;
    } else if (H->fr_pattern == 2) {
        t3 = or_grnd(g) * 2. * OR_C_PI;
        t4 = sqrt(or_grnd(g)) * (H->fr2 - H->fr1) + H->fr1;
        t5 = cos(t3) * t4;
        t6 = sin(t3) * t4;
    } else if (H->fr_pattern == 3) {
        t3 = 2. * or_grnd(g) - 1.0;
        t4 = 2. * or_grnd(g) - 1.0;
        t5 = H->fr1 * t3;
        t6 = H->fr2 * t4;
    }
    x = x + t5;
    y = y + t6;
    const double z = (0.5 - or_grnd(g)) * T->length + T->zoffset;
    double eloss, teff;
    or_trip_thru_target(g, T, 1, 0, z - T->zoffset, H->Ebeam, 0.0, OR_C_ME, 1, &eloss, &teff);
    if (!H->using_eloss) eloss = 0.0;
    const double coul = H->using_coulomb ? H->coulomb : 0.0;
    const double Ein = H->Ebeam + (or_grnd(g) - 0.5) * H->dEbeam + coul - eloss;
    const double e_yptar = H->e_yptar[0] + or_grnd(g) * (H->e_yptar[1] - H->e_yptar[0]);
    const double e_xptar = H->e_xptar[0] + or_grnd(g) * (H->e_xptar[1] - H->e_xptar[0]);
    double e_theta, e_phi;
    or_physics_angles(H->e_theta, H->e_phi, e_xptar, e_yptar, &e_theta, &e_phi);
    /* complete_ev */
    const double uex = sin(e_theta) * cos(e_phi), uey = sin(e_theta) * sin(e_phi);
    const double uez = cos(e_theta);
    const double E = Ein * H->Mh / (H->Mh + Ein * (1. - uez));
    const int ok = !(E > Ein);
    const double e_delta = (E - H->e_P) * 100. / H->e_P;
    const double nu = Ein - E;
    const double Q2 = 2 * Ein * E * (1. - uez);
    const double q = sqrt(Q2 + nu * nu);
    const double uqx = -(E * uex / q), uqy = -(E * uey / q), uqz = (Ein - E * uez) / q;
    const double p_theta = acos(uqz);
    double p_phi = atan2(uqy, uqx);
    if (p_phi < 0.) p_phi = p_phi + 2. * OR_C_PI;
    double p_xptar, p_yptar;
    or_spectrometer_angles(H->p_theta, H->p_phi, &p_xptar, &p_yptar, p_theta, p_phi);
    const double p_E = sqrt(q * q + H->Mh * H->Mh);
    const double p_delta = (q - H->p_P) * 100. / H->p_P;
    const double r = sqrt(1. + e_yptar * e_yptar + e_xptar * e_xptar);
    const double jacobian = 1.0 / (r * r * r);
    const double sigcc = or_sigep(Q2, E, e_theta, Ein);
    const double weight = jacobian * sigcc;
#define OR_PUT(f, v) if (o->f) o->f[i] = (v)
    OR_PUT(x, x); OR_PUT(y, y); OR_PUT(z, z); OR_PUT(rasterx, t5); OR_PUT(rastery, t6);
    OR_PUT(Ein, Ein); OR_PUT(eloss, eloss); OR_PUT(teff, teff);
    OR_PUT(e_xptar, e_xptar); OR_PUT(e_yptar, e_yptar); OR_PUT(e_theta, e_theta);
    OR_PUT(e_phi, e_phi); OR_PUT(e_E, E); OR_PUT(e_delta, e_delta);
    OR_PUT(p_xptar, p_xptar); OR_PUT(p_yptar, p_yptar); OR_PUT(p_theta, p_theta);
    OR_PUT(p_phi, p_phi); OR_PUT(p_P, q); OR_PUT(p_E, p_E); OR_PUT(p_delta, p_delta);
    OR_PUT(Q2, Q2); OR_PUT(nu, nu); OR_PUT(q, q); OR_PUT(jacobian, jacobian);
    OR_PUT(sigcc, sigcc); OR_PUT(weight, weight); OR_PUT(ok, (uint8_t)ok);
#undef OR_PUT
}

int or_gen_hep(const or_target *T, const or_hep *H, int64_t n, or_hep_out *out, int rng_kind,
               const double *tape, int64_t stride, uint64_t seed, uint64_t event0, int nthreads,
               int32_t *used)
{
    if (nthreads <= 0) nthreads = 1;
    int64_t overflow = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : overflow)
#endif
    for (int64_t i = 0; i < n; i++) {
        or_rng g;
        memset(&g, 0, sizeof g);
        g.kind = rng_kind;
        if (rng_kind == 0) {
            g.tape = tape + i * stride;
            g.stride = stride;
        } else {
            g.key[0] = (uint32_t)seed;
            g.key[1] = (uint32_t)(seed >> 32);
            g.event = event0 + (uint64_t)i;
            g.tag = OR_GEN_TAG;
        }
        gen_hep_event(&g, T, H, i, out);
        if (g.overflow) overflow++;
        if (used) used[i] = (int32_t)g.used;
    }
    return overflow ? -2 : 0;
}

/* ========================================================================= */
/* Radiative tails: H(e,e'p) elastic with using_rad                           */
/* ========================================================================= */

/* OR_RAD_PERTURB (a conditioning probe for the tests, never the checker):
 * every elementary function of the radiative code returns its result moved
 * by one ulp, to measure how far last-bit differences of a libm travel. */
#ifdef OR_RAD_PERTURB
static double or_ulp_up(double x) { return nextafter(x, INFINITY); }
#define log(x) or_ulp_up(log(x))
#define exp(x) or_ulp_up(exp(x))
#define pow(x, y) or_ulp_up(pow(x, y))
#define tgamma(x) or_ulp_up(tgamma(x))
#define sin(x) or_ulp_up(sin(x))
#define cos(x) or_ulp_up(cos(x))
#define asin(x) or_ulp_up(asin(x))
#define acos(x) or_ulp_up(acos(x))
#define atan2(x, y) or_ulp_up(atan2(x, y))
#endif

/* constants.inc:38-44 (alpi = alpha/pi of the rounded alpha) */
#define OR_C_AMU 931.49432
static const double or_c_alpi = (1. / 137.0359895) / 3.141592653589793;
static const double or_c_euler = 0.577215665;

/* gamma: the callers (basicrad_init_ev init.f:744-790, peaked_rad_weight
 * radc.f:532,640-642, extrad_phi radc.f:675,713) declare `real*8 gamma`
 * without EXTERNAL, so the Fortran 2008 intrinsic GAMMA is called (libm
 * tgamma in the compiled reference), not the polynomial function gamma of
 * radc.f:94-116, which no caller reaches. */
double or_rad_gamma(double x) { return tgamma(x); }

/* lambda_dave, radc.f:763-812: effective bt of internal radiation per tail
 * (extended peaking approximation with plus_flag = 1). */
double or_lambda_dave(int itail, int plus_flag, int doing_proton, double e1, double e2, double e3, double p3,
                      double th)
{
    double lam = 0.0, plus_term = 0.0;
    if (plus_flag == 1 && itail < 3) {
        plus_term = log((1. - cos(th)) / 2.);
        if (doing_proton) plus_term = plus_term + 2. * log(e1 / e2);
    }
    if (itail == 1) {
        lam = or_c_alpi * ((2. * log(2. * e1 / OR_C_ME) - 1.) + plus_term);
    } else if (itail == 2) {
        lam = or_c_alpi * ((2. * log(2. * e2 / OR_C_ME) - 1.) + plus_term);
    } else if (itail == 3) {
        if (doing_proton) {
            lam = or_c_alpi * ((e3 / p3) * log((e3 + p3) / (e3 - p3)) - 2.);
            if (lam < 0) lam = 0.0;
        } else {
            lam = 0.0;
        }
    }
    return lam;
}

/* spence, brem.f:287-322 (the reference's approximate form) */
static double rad_spence(double ax)
{
    const double bx = fabs(ax);
    if (bx <= 1) return 0.0;
    const double l = log(bx);
    return -0.5 * (l * l);
}

/* inter, brem.f:218-248 (calculate_spence = .true., init.f:648) */
static double rad_inter(double alpha, double ar1, double ar2, double e1, double e2, double de)
{
    const double de2 = e1 - e2;
    const double amult = -1. / (alpha * (ar1 - ar2));
    double r = log(fabs((e2 / de) + ar1 * (de2 / de))) * log(fabs((ar1 - 1.) / ar1)) -
               log(fabs((e2 / de) + ar2 * (de2 / de))) * log(fabs((ar2 - 1.) / ar2));
    const double arg1 = (de2 / (e2 + ar1 * de2)) * (ar1 - 1.);
    const double arg2 = (de2 / (e2 + ar1 * de2)) * (ar1);
    const double arg3 = (de2 / (e2 + ar2 * de2)) * (ar2 - 1.);
    const double arg4 = (de2 / (e2 + ar2 * de2)) * (ar2);
    r = (((r - rad_spence(arg1)) + rad_spence(arg2)) + rad_spence(arg3)) - rad_spence(arg4);
    return r * amult / (OR_C_PI);
}

/* inter_prime, brem.f:581-598 */
static double rad_inter_prime(double alpha, double ar1, double ar2, double de)
{
    const double amult = -1. / (alpha * (ar1 - ar2));
    return (-1. / de) * amult / OR_C_PI * (log(fabs((ar1 - 1.) / ar1)) - log(fabs((ar2 - 1.) / ar2)));
}

/* brem, brem.f:6-216 (include_hard = calculate_spence = .true.,
 * exponentiate = use_expon, init.f:641-648).  Energies in MeV; returns the
 * radiative factor, with bsoft, bhard and dbsoft (1/MeV).  The compiled
 * reference evaluates x**0.5 as a square root (no pow call in brem.o). */
double or_brem(const or_rad *R, double ein, double eout, double egamma, int radiate_proton, double *bsoft,
               double *bhard, double *dbsoft)
{
    const double pi = 3.141592653589793, am = .93827231, ame = .00051099906, e2 = 1. / 137.0359895;
    const double ame2 = ame * ame, ame4 = ame2 * ame2, am2 = am * am;
    const double ak = ein / 1000., akp = eout / 1000., de = egamma / 1000.;
    const double eang = 2. * asin(sqrt(am / (2. * ak) * (ak / akp - 1.)));
    const double se = sin(eang / 2.);
    const double q2 = 4. * ak * akp * (se * se);
    const double ape = am + ak - akp;
    const double ap = sqrt(ape * ape - am2);
    const double pang = acos((ak - akp * cos(eang)) / ap);
    double aprod = 1.e0;
    const double bei = aprod * (-1. / (2. * pi)) * log(ak / de);
    const double dbei = aprod * (1. / (2. * pi * de));
    const double bef = aprod * (-1. / (2. * pi)) * log(akp / de);
    const double dbef = aprod * (1. / (2. * pi * de));
    aprod = -1.e0;
    double adot = ak * akp * (1. - cos(eang));
    double alpha = 2. * ame2 - 2. * adot;
    double ar1 = 0.5 + sqrt(adot * adot - ame4) / alpha;
    double ar2 = 0.5 - sqrt(adot * adot - ame4) / alpha;
    const double bee = aprod * adot * rad_inter(alpha, ar1, ar2, ak, akp, de);
    const double dbee = aprod * adot / (pi * alpha * (ar1 - ar2) * de) *
                        (log((ar1 - 1) / ar1) - log((ar2 - 1) / ar2));
    double bpi = 0, bpf = 0, bpp = 0, bepii = 0, bepff = 0, bepif = 0, bepfi = 0;
    double dbpi = 0, dbpf = 0, dbpp = 0, dbepii = 0, dbepff = 0, dbepif = 0, dbepfi = 0;
    if (radiate_proton) {
        const double ame_am2 = (ame * am) * (ame * am);
        aprod = 1.e0;
        bpi = aprod * (-1. / (2. * pi)) * log(am / de);
        dbpi = aprod * (1. / (2. * pi * de));
        bpf = aprod * (-1. / (2. * pi)) * log(ape / de);
        dbpf = aprod * (1 / (2. * pi * de));
        aprod = -1.e0;
        adot = am * ape;
        alpha = 2. * am2 - 2. * adot;
        ar1 = 0.5 + sqrt(adot * adot - am2 * am2) / alpha;
        ar2 = 0.5 - sqrt(adot * adot - am2 * am2) / alpha;
        bpp = aprod * adot * rad_inter(alpha, ar1, ar2, am, ape, de);
        dbpp = aprod * adot / (pi * alpha * (ar1 - ar2) * de) * (log((ar1 - 1) / ar1) - log((ar2 - 1) / ar2));
        /* the four electron-proton interference terms share one form */
        const double c_adot[4] = {ak * am, akp * ape - akp * ap * cos(eang + pang), ak * ape - ak * ap * cos(pang),
                                  akp * am};
        const double c_e1[4] = {ak, akp, ak, akp}, c_e2[4] = {am, ape, ape, am};
        const double c_prod[4] = {-1.e0, -1.e0, 1.e0, 1.e0};
        double bb[4], dbb[4];
        for (int k = 0; k < 4; ++k) {
            aprod = c_prod[k];
            adot = c_adot[k];
            alpha = am2 + ame2 - 2. * adot;
            ar1 = (am2 - adot + sqrt(adot * adot - ame_am2)) / alpha;
            ar2 = (am2 - adot - sqrt(adot * adot - ame_am2)) / alpha;
            bb[k] = aprod * adot * rad_inter(alpha, ar1, ar2, c_e1[k], c_e2[k], de);
            dbb[k] = aprod * adot / (pi * alpha * (ar1 - ar2) * de) * (log((ar1 - 1) / ar1) - log((ar2 - 1) / ar2));
        }
        bepii = bb[0]; bepff = bb[1]; bepif = bb[2]; bepfi = bb[3];
        dbepii = dbb[0]; dbepff = dbb[1]; dbepif = dbb[2]; dbepfi = dbb[3];
    }
    const double b = 2. * e2 * (bei + bef + bee);
    double bzz = 0.0, bz = 0.0, dbzz = 0.0, dbz = 0.0;
    if (radiate_proton) {
        bzz = 2. * e2 * (bpi + bpf + bpp);
        bz = 2. * e2 * (bepii + bepff + bepif + bepfi);
    }
    *bsoft = b + bz + bzz;
    *bhard = -1. * (e2 / pi) * (-28 / 9. + 13. / 6. * log(q2 / ame2));
    const double db = 2. * e2 * (dbei + dbef + dbee);
    if (radiate_proton) {
        dbzz = 2. * e2 * (dbpi + dbpf + dbpp);
        dbz = 2. * e2 * (dbepii + dbepff + dbepif + dbepfi);
    }
    *dbsoft = db + dbz + dbzz;
    *dbsoft = *dbsoft / 1000.;
    double r = R->use_expon == 1 ? -*dbsoft / exp(*bsoft) : 1. - *dbsoft;
    return r * (1. - *bhard);
}

/* bremos, brem.f:344-578: the same with off-shell protons (four-vectors,
 * MeV in; the initial proton at rest with the proton mass). */
double or_bremos(const or_rad *R, double egamma, double k_ix, double k_iy, double k_iz, double k_fx,
                 double k_fy, double k_fz, double p_ix, double p_iy, double p_iz, double p_fx, double p_fy,
                 double p_fz, double p_fe, int radiate_proton, double *bsoft, double *bhard, double *dbsoft)
{
    const double pi = 3.141592653589793, twopi = 2. * 3.141592653589793, ame = .00051099906;
    const double e2 = 1. / 137.0359895, mp = .93827231;
    const double ame2 = ame * ame, ame4 = ame2 * ame2;
    const double de = egamma / 1000.;
    const double kix = k_ix / 1000., kiy = k_iy / 1000., kiz = k_iz / 1000.;
    const double kfx = k_fx / 1000., kfy = k_fy / 1000., kfz = k_fz / 1000.;
    const double pix = p_ix / 1000., piy = p_iy / 1000., piz = p_iz / 1000.;
    const double pfe = p_fe / 1000., pfx = p_fx / 1000., pfy = p_fy / 1000., pfz = p_fz / 1000.;
    const double kie = sqrt(kix * kix + kiy * kiy + kiz * kiz + ame2);
    const double kfe = sqrt(kfx * kfx + kfy * kfy + kfz * kfz + ame2);
    const double pie = mp;
    const double dke = kfe - kie, dkx = kfx - kix, dky = kfy - kiy, dkz = kfz - kiz;
    const double q2 = -1. * (dke * dke - dkx * dkx - dky * dky - dkz * dkz);
    const double ami = mp;
    const double amf = sqrt(pfe * pfe - pfx * pfx - pfy * pfy - pfz * pfz);
    double aprod = 1.e0;
    const double bei = aprod * (-1. / twopi) * log(kie / de);
    const double dbei = aprod * (-1. / twopi) * (-1. / de);
    const double bef = aprod * (-1. / twopi) * log(kfe / de);
    const double dbef = aprod * (-1. / twopi) * (-1. / de);
    aprod = -1.e0;
    double adot = kie * kfe - kix * kfx - kiy * kfy - kiz * kfz;
    double alpha = 2. * ame2 - 2. * adot;
    double ar1 = 0.5 + sqrt(4. * (adot * adot) - 4. * ame4) / (2. * alpha);
    double ar2 = 0.5 - sqrt(4. * (adot * adot) - 4. * ame4) / (2. * alpha);
    const double bee = aprod * adot * rad_inter(alpha, ar1, ar2, kie, kfe, de);
    const double dbee = aprod * adot * rad_inter_prime(alpha, ar1, ar2, de);
    double bpi = 0, bpf = 0, bpp = 0, bepii = 0, bepff = 0, bepif = 0, bepfi = 0;
    double dbpi = 0, dbpf = 0, dbpp = 0, dbepii = 0, dbepff = 0, dbepif = 0, dbepfi = 0;
    if (radiate_proton) {
        aprod = 1.e0;
        bpi = aprod * (-1. / twopi) * log(pie / de);
        dbpi = aprod * (-1. / twopi) * (-1. / de);
        bpf = aprod * (-1. / twopi) * log(pfe / de);
        dbpf = aprod * (-1. / twopi) * (-1. / de);
        aprod = -1.e0;
        adot = pie * pfe - pix * pfx - piy * pfy - piz * pfz;
        alpha = ami * ami + amf * amf - 2. * adot;
        {
            const double m = ami * amf;
            ar1 = (2. * (amf * amf) - 2. * adot + sqrt(4. * (adot * adot) - 4. * (m * m))) / (2. * alpha);
            ar2 = (2. * (amf * amf) - 2. * adot - sqrt(4. * (adot * adot) - 4. * (m * m))) / (2. * alpha);
        }
        bpp = aprod * adot * rad_inter(alpha, ar1, ar2, pie, pfe, de);
        dbpp = aprod * adot * rad_inter_prime(alpha, ar1, ar2, de);
        /* interference terms: (electron leg, proton leg, sign, proton mass) */
        const double c_adot[4] = {kie * pie - kix * pix - kiy * piy - kiz * piz,
                                  kfe * pfe - kfx * pfx - kfy * pfy - kfz * pfz,
                                  kie * pfe - kix * pfx - kiy * pfy - kiz * pfz,
                                  kfe * pie - kfx * pix - kfy * piy - kfz * piz};
        const double c_e1[4] = {kie, kfe, kie, kfe}, c_e2[4] = {pie, pfe, pfe, pie};
        const double c_m[4] = {ami, amf, amf, ami};
        const double c_prod[4] = {-1.e0, -1.e0, 1.e0, 1.e0};
        double bb[4], dbb[4];
        for (int k = 0; k < 4; ++k) {
            aprod = c_prod[k];
            adot = c_adot[k];
            const double M = c_m[k], m = ame * M;
            alpha = M * M + ame2 - 2. * adot;
            ar1 = (2. * (M * M) - 2. * adot + sqrt(4. * (adot * adot) - 4. * (m * m))) / (2. * alpha);
            ar2 = (2. * (M * M) - 2. * adot - sqrt(4. * (adot * adot) - 4. * (m * m))) / (2. * alpha);
            bb[k] = aprod * adot * rad_inter(alpha, ar1, ar2, c_e1[k], c_e2[k], de);
            dbb[k] = aprod * adot * rad_inter_prime(alpha, ar1, ar2, de);
        }
        bepii = bb[0]; bepff = bb[1]; bepif = bb[2]; bepfi = bb[3];
        dbepii = dbb[0]; dbepff = dbb[1]; dbepif = dbb[2]; dbepfi = dbb[3];
    }
    const double b = 2. * e2 * (bei + bef + bee);
    double bzz = 0.0, bz = 0.0, dbzz = 0.0, dbz = 0.0;
    if (radiate_proton) {
        bzz = 2. * e2 * (bpi + bpf + bpp);
        bz = 2. * e2 * (bepii + bepff + bepif + bepfi);
    }
    *bsoft = b + bz + bzz;
    *bhard = -1. * (e2 / pi) * (-28 / 9. + 13. / 6. * log(q2 / ame2));
    const double db = 2. * e2 * (dbei + dbef + dbee);
    if (radiate_proton) {
        dbzz = 2. * e2 * (dbpi + dbpf + dbpp);
        dbz = 2. * e2 * (dbepii + dbepff + dbepif + dbepfi);
    }
    *dbsoft = db + dbz + dbzz;
    *dbsoft = *dbsoft / 1000.;
    double r = R->use_expon == 1 ? -*dbsoft / exp(*bsoft) : 1. - *dbsoft;
    return r * (1. - *bhard);
}

/* spen, radc.f:747-761 */
static double rad_spen(double x)
{
    double y = 1.0, s = 0.0;
    int i = 0;
    while (i <= 100 && fabs(y) > fabs(s) * 1.e-4) {
        i = i + 1;
        y = x * y;
        s = s + y / (double)(i * i);
    }
    return s;
}

/* The event fields generate_rad and its helpers read and change (the
 * hydrogen-elastic members of type(event) and type(event_main)). */
typedef struct {
    double Ein, e_E, e_P, e_delta, e_theta, e_phi, e_xptar, e_yptar;
    double uex, uey, uez;
    double p_P, p_E, p_delta, p_theta, p_phi, p_xptar, p_yptar;
    double upx, upy, upz;
    double Q2, nu, q, Em, Trec;
    double jacobian, z;
    double eloss[3], teff[3];
} rad_event;

/* The per-event radiative constants of /radccom/ (radc.inc:5-14). */
typedef struct {
    double bt[2], lambda[3], g[5], c[5], c_int[4], c_ext[4], g_int, g_ext, frac[3];
    double hardcorfac, etta;
    int rad_proton_this_ev;
} rad_state;

/* schwinger, radc.f:721-745 (include_hard = .true.) */
static void rad_schwinger(const or_rad *R, const rad_state *S, const rad_event *v, double Ecutoff, double *dsoft,
                          double *dhard)
{
    const double lq = log(v->Q2 / (OR_C_ME * OR_C_ME)) - 1.0;
    const double sn = sin(v->e_theta / 2.);
    const double s2 = sn * sn;
    const double b = 1. + 2. * v->nu * s2 / (R->targ_A * OR_C_AMU);
    const double spence = rad_spen(1. - s2) - 2.5893784;
    *dsoft = or_c_alpi * lq * log(v->Ein / (S->etta * S->etta) * v->e_E * b / (Ecutoff * Ecutoff));
    const double l = log(v->Ein / v->e_E);
    *dhard = -or_c_alpi * (2.166666 * lq + spence - (l * l) / 2.0);
}

/* basicrad_init_ev, init.f:732-804 */
static void rad_basicrad_init_ev(rad_state *S, double e1, double e2, double e3)
{
    const double one = 1.;
    const double e[3] = {e1, e2, e3};
    const double *lam = S->lambda, *bt = S->bt;
    double *g = S->g, *c = S->c, *ci = S->c_int, *ce = S->c_ext;
    g[1] = lam[0] + bt[0];
    g[2] = lam[1] + bt[1];
    g[3] = lam[2];
    g[0] = g[1] + g[2] + g[3];
    ci[1] = lam[0] / pow(e[0] * e[1], lam[0] / 2.);
    ci[2] = lam[1] / pow(e[0] * e[1], lam[1] / 2.);
    ci[3] = lam[2] / pow(OR_C_MP * e[2], lam[2] / 2.);
    for (int i = 1; i <= 3; ++i) ci[i] = ci[i] * exp(-or_c_euler * lam[i - 1]) / tgamma(one + lam[i - 1]);
    S->g_int = lam[0] + lam[1] + lam[2];
    ci[0] = ci[1] * ci[2] * S->g_int / lam[0] / lam[1];
    if (lam[2] > 0) ci[0] = ci[0] * ci[3] / lam[2];
    ci[0] = ci[0] * tgamma(one + lam[0]) * tgamma(one + lam[1]) * tgamma(one + lam[2]) /
            tgamma(one + S->g_int);
    for (int i = 1; i <= 2; ++i) ce[i] = bt[i - 1] / pow(e[i - 1], bt[i - 1]) / tgamma(one + bt[i - 1]);
    ce[3] = 0.0;
    S->g_ext = bt[0] + bt[1];
    ce[0] = ce[1] * ce[2] * S->g_ext / bt[0] / bt[1];
    ce[0] = ce[0] * tgamma(one + bt[0]) * tgamma(one + bt[1]) / tgamma(one + S->g_ext);
    for (int i = 1; i <= 2; ++i)
        c[i] = ci[i] * ce[i] * g[i] / lam[i - 1] / bt[i - 1] * tgamma(one + lam[i - 1]) *
               tgamma(one + bt[i - 1]) / tgamma(one + g[i]);
    c[3] = ci[3];
    c[0] = c[1] * c[2] * g[0] / g[1] / g[2];
    if (g[3] > 0) c[0] = c[0] * c[3] / g[3];
    c[0] = c[0] * tgamma(one + g[1]) * tgamma(one + g[2]) * tgamma(one + g[3]) /
           tgamma(one + g[0]);
    c[4] = g[4] / pow(e1 * e2, g[4]) / tgamma(one + g[4]);
    if (g[3] > 0) c[4] = c[4] / pow(e3, g[4]);
}

/* radc_init_ev, init.f:655-720 */
static void rad_init_ev(const or_rad *R, rad_state *S, const rad_event *v)
{
    S->etta = 1.0;
    for (int i = 0; i < 2; ++i) S->bt[i] = R->etatzai * v->teff[i];
    for (int i = 1; i <= 3; ++i)
        S->lambda[i - 1] = or_lambda_dave(i, 1, R->doing_tail[2], v->Ein, v->e_E, v->p_E, v->p_P, v->e_theta);
    S->rad_proton_this_ev = S->lambda[2] > 0;
    const double Ecutoff = 450.;
    double dsoft, dhard, dsoft_prime = 0.0;  /* schwinger leaves dsoft_prime at its initial 0 */
    if (R->intcor_mode == 0) {
        rad_schwinger(R, S, v, Ecutoff, &dsoft, &dhard);
    } else if (!R->use_offshell_rad) {
        or_brem(R, v->Ein, v->e_E, Ecutoff, S->rad_proton_this_ev, &dsoft, &dhard, &dsoft_prime);
    } else {
        or_bremos(R, Ecutoff, 0.0, 0.0, v->Ein, v->e_P * v->uex, v->e_P * v->uey, v->e_P * v->uez, 0.0, 0.0, 0.0,
                  v->p_P * v->upx, v->p_P * v->upy, v->p_P * v->upz, v->p_E, S->rad_proton_this_ev, &dsoft,
                  &dhard, &dsoft_prime);
    }
    S->hardcorfac = 1. / (1. - dhard);
    S->g[4] = -dsoft_prime * Ecutoff + S->bt[0] + S->bt[1];
    rad_basicrad_init_ev(S, v->Ein, v->e_E, v->p_E);
    for (int i = 1; i <= 3; ++i) S->frac[i - 1] = S->g[i] / S->g[0];
}

/* basicrad, radc.f:3-91: Egamma from g Egamma^(g-1) / (hi^g - lo^g) over
 * [lo, hi] (one uniform); itail 0 is the combined tail (4). */
static void rad_basicrad(or_rng *gr, const rad_state *S, int itail, double lo, double hi, double *Egamma,
                         double *weight, double *val_reciprocal)
{
    *Egamma = 0.0;
    *weight = 0.0;
    *val_reciprocal = 0.0;
    if (itail == 0) itail = 4;
    const double g = S->g[itail];
    if (g <= 0) {
        *weight = 1.0;
        return;
    }
    if (hi <= lo || hi <= 0) return;
    const double power_hi = pow(hi, g);
    double power_lo = 0.0;
    if (lo > 0) power_lo = pow(lo, g);
    const double ymin = power_lo / power_hi;
    const double y = ymin + or_grnd(gr) * (1. - ymin);
    const double x = pow(y, 1. / g);
    *Egamma = x * hi;
    if (*Egamma > 0) *val_reciprocal = pow(*Egamma, 1. - g) * (power_hi - power_lo) / g;
    *weight = S->c[itail] / g * (power_hi - power_lo);
}

/* extrad_phi, radc.f:668-718 (extrad_flag 3 with itail 0 stops the
 * reference; not reached from supported settings) */
static double rad_extrad_phi(const or_rad *R, const rad_state *S, int extrad_flag, int itail, double E1, double E2,
                             double Egamma)
{
    const double E[2] = {E1, E2};
    double phi = 1.0;
    if (extrad_flag == 2) {
        if (itail == 0)
            phi = 1. - (S->bt[0] / E[0] + S->bt[1] / E[1]) / (S->g[1] + S->g[2]) * Egamma;
        else if (itail == 1 || itail == 2)
            phi = 1. - S->bt[itail - 1] / E[itail - 1] / S->g[itail] * Egamma;
    } else if (extrad_flag == 3) {
        if (itail == 1 || itail == 2) {
            const double x = Egamma / E[itail - 1];
            const double t = S->bt[itail - 1] / R->etatzai;
            phi = phi * (1. - x + (x * x) / R->etatzai) *
                  exp(t * ((R->etatzai - 0.5) - R->etatzai * x + (x * x) / 2.)) * tgamma(1. + S->bt[itail - 1]);
        }
    }
    return phi;
}

/* peaked_rad_weight, radc.f:523-650 (rad_flag 0 or 1) */
static double rad_peaked_weight(const or_rad *R, const rad_state *S, int extrad_flag, const rad_event *v,
                                double Egamma, double emin, double emax, double basicrad_weight)
{
    const double ein = v->Ein, eout = v->e_E, eul = 0.577215665;
    if (extrad_flag <= 2) {
        const double phi_ext = rad_extrad_phi(R, S, extrad_flag, 0, ein, eout, Egamma);
        if (R->rad_flag == 1) return basicrad_weight * phi_ext;
    }
    double dsoft_intmin = 1.0, dsoft_intmax, dhard, dpm, dpx;
    if (!R->use_offshell_rad) {
        if (emin > 0) or_brem(R, ein, eout, emin, S->rad_proton_this_ev, &dsoft_intmin, &dhard, &dpm);
        or_brem(R, ein, eout, emax, S->rad_proton_this_ev, &dsoft_intmax, &dhard, &dpx);
    } else {
        const double kx = v->e_E * v->uex, ky = v->e_E * v->uey, kz = v->e_E * v->uez;
        const double px = v->p_P * v->upx, py = v->p_P * v->upy, pz = v->p_P * v->upz;
        if (emin > 0)
            or_bremos(R, emin, 0.0, 0.0, ein, kx, ky, kz, 0.0, 0.0, 0.0, px, py, pz, v->p_E, S->rad_proton_this_ev,
                      &dsoft_intmin, &dhard, &dpm);
        or_bremos(R, emax, 0.0, 0.0, ein, kx, ky, kz, 0.0, 0.0, 0.0, px, py, pz, v->p_E, S->rad_proton_this_ev,
                  &dsoft_intmax, &dhard, &dpx);
    }
    double w;
    if (emin > 0)
        w = S->c_ext[0] / S->g_ext *
            (exp(-dsoft_intmax) * pow(emax, S->g_ext) - exp(-dsoft_intmin) * pow(emin, S->g_ext));
    else
        w = S->c_ext[0] / S->g_ext * (exp(-dsoft_intmax) * pow(emax, S->g_ext));
    const double *g = S->g, *bt = S->bt;
    w = w * exp(-eul * g[4]) / tgamma(1. + g[4]) * tgamma(1. + g[4] - bt[0] - bt[1]) *
        tgamma(1. + bt[0]) * tgamma(1. + bt[1]) / tgamma(1. + g[4]);
    if (w < 0) w = 0;
    return w;
}

double or_rad_etatzai(double Z)
{
    double L1 = log(184.15) - log(Z) / 3.0, L2 = log(1194.) - 2. * log(Z) / 3.0;  /* init.f:19-24 */
    if (Z == 1) {
        L1 = 5.31;
        L2 = 6.144;
    }
    return (12.0 + (Z + 1.) / (Z * L1 + L2)) / 9.0;  /* init.f:634 */
}

/* complete_ev for hydrogen elastic (event.f:497-563, :1013-1014) and the
 * scattered particles' target losses (:1031-1040), then radc_init_ev
 * (:1045).  Returns 0 when E' > Ein (complete_ev's failure). */
static int rad_complete_ev(or_rng *gr, const or_target *T, const or_hep *H, const or_rad *R, rad_state *S,
                           rad_event *v)
{
    v->jacobian = 1.0;
    v->uex = sin(v->e_theta) * cos(v->e_phi);
    v->uey = sin(v->e_theta) * sin(v->e_phi);
    v->uez = cos(v->e_theta);
    v->e_E = v->Ein * H->Mh / (H->Mh + v->Ein * (1. - v->uez));
    if (v->e_E > v->Ein) return 0;
    v->e_P = v->e_E;
    v->e_delta = (v->e_P - H->e_P) * 100. / H->e_P;
    v->nu = v->Ein - v->e_E;
    v->Q2 = 2 * v->Ein * v->e_E * (1. - v->uez);
    v->q = sqrt(v->Q2 + v->nu * v->nu);
    const double uqx = -v->e_P * v->uex / v->q, uqy = -v->e_P * v->uey / v->q;
    const double uqz = (v->Ein - v->e_P * v->uez) / v->q;
    v->Em = 0.0;
    v->upx = uqx;
    v->upy = uqy;
    v->upz = uqz;
    v->p_P = v->q;
    v->p_theta = acos(v->upz);
    v->p_phi = atan2(v->upy, v->upx);
    if (v->p_phi < 0.) v->p_phi = v->p_phi + 2. * OR_C_PI;
    or_spectrometer_angles(H->p_theta, H->p_phi, &v->p_xptar, &v->p_yptar, v->p_theta, v->p_phi);
    v->p_E = sqrt(v->p_P * v->p_P + H->Mh * H->Mh);
    v->p_delta = (v->p_P - H->p_P) * 100. / H->p_P;
    v->Trec = 0.0;
    const double r = sqrt(1. + v->e_yptar * v->e_yptar + v->e_xptar * v->e_xptar);
    v->jacobian = v->jacobian / (r * r * r);
    or_trip_thru_target(gr, T, 2, R->e_arm, v->z - T->zoffset, v->e_E, v->e_theta, OR_C_ME, 1, &v->eloss[1],
                        &v->teff[1]);
    or_trip_thru_target(gr, T, 3, R->p_arm, v->z - T->zoffset, v->p_E, v->p_theta, H->Mh, 1, &v->eloss[2],
                        &v->teff[2]);
    if (!H->using_eloss) {
        v->eloss[1] = 0.0;
        v->eloss[2] = 0.0;
    }
    rad_init_ev(R, S, v);
    return 1;
}

/* One event of generate with using_rad (event.f:157-212, :275-277,
 * :322-325, :381-404) and generate_rad (radc.f:120-517) for doing_hyd_elast,
 * then complete_main's weight (event.f:1397-1398, :1450-1452, :1558). */
static void gen_hep_rad_event(or_rng *gr, const or_target *T, const or_hep *H, const or_rad *R, int64_t i,
                              or_hep_out *o, or_rad_out *ro)
{
    const double nsig_max = 3.0;
    rad_event v;
    memset(&v, 0, sizeof v);
    rad_state S;
    memset(&S, 0, sizeof S);
    double x = or_gauss1(gr, nsig_max) * H->xwid + H->xoffset;
    double y = or_gauss1(gr, nsig_max) * H->ywid + H->yoffset;
    double t3, t4, t5 = 0.0, t6 = 0.0;
    if (H->fr_pattern == 1) {
        t3 = or_grnd(gr) * OR_C_PI;
        t4 = or_grnd(gr) * OR_C_PI;
        t5 = cos(t3) * H->fr1;
        t6 = cos(t4) * H->fr2;
    } else if (H->fr_pattern == 2) {
        t3 = or_grnd(gr) * 2. * OR_C_PI;
        t4 = sqrt(or_grnd(gr)) * (H->fr2 - H->fr1) + H->fr1;
        t5 = cos(t3) * t4;
        t6 = sin(t3) * t4;
    } else if (H->fr_pattern == 3) {
        t3 = 2. * or_grnd(gr) - 1.0;
        t4 = 2. * or_grnd(gr) - 1.0;
        t5 = H->fr1 * t3;
        t6 = H->fr2 * t4;
    }
    x = x + t5;
    y = y + t6;
    v.z = (0.5 - or_grnd(gr)) * T->length + T->zoffset;
    or_trip_thru_target(gr, T, 1, 0, v.z - T->zoffset, H->Ebeam, 0.0, OR_C_ME, 1, &v.eloss[0], &v.teff[0]);
    if (!H->using_eloss) v.eloss[0] = 0.0;
    const double coul = H->using_coulomb ? H->coulomb : 0.0;
    v.Ein = H->Ebeam + (or_grnd(gr) - 0.5) * H->dEbeam + coul - v.eloss[0];
    v.e_yptar = H->e_yptar[0] + or_grnd(gr) * (H->e_yptar[1] - H->e_yptar[0]);
    v.e_xptar = H->e_xptar[0] + or_grnd(gr) * (H->e_xptar[1] - H->e_xptar[0]);
    or_physics_angles(H->e_theta, H->e_phi, v.e_xptar, v.e_yptar, &v.e_theta, &v.e_phi);
    int ok = rad_complete_ev(gr, T, H, R, &S, &v);
    double gen_weight = 1.0, Eg[3] = {0.0, 0.0, 0.0};
    double orig_Ein = 0, orig_e_E = 0, orig_e_delta = 0, orig_p_E = 0, orig_p_P = 0, orig_p_delta = 0;
    int ntail = 0;
    if (ok) {
        /* generate_rad */
        ok = 0;
        const int extrad_flag = R->extrad_flag != 0 ? R->extrad_flag : (R->rad_flag == 0 ? 3 : 1);
        double rad_weight = 1;
        int peaked_basis_flag = 1;
        if (R->rad_flag <= 1) {
            peaked_basis_flag = 0;
            const double u = or_grnd(gr);
            if (u >= S.frac[0] + S.frac[1]) ntail = 3;
            else if (u >= S.frac[0]) ntail = 2;
            else ntail = 1;
        } else if (R->rad_flag == 2) {
            ntail = (int)(or_grnd(gr) * 3.) + 1;
            if (ntail == 4) ntail = 3;
        } else {
            ntail = 0;
        }
        const double Mp = OR_C_MP;
        const double max_delta_Trec = fmax(v.Trec - R->vertex_Trec[0], R->vertex_Trec[1] - v.Trec);
        const double Egamma1_max = R->doing_tail[0] ? R->Egamma_tot_max : 0.0;
        const int hardwired = R->Egamma_gen_max > 0.01;
        double Emin[3] = {0, 0, 0}, Emax[3] = {0, 0, 0}, bw = 0.0, bvr = 0.0;
        int fail = 0;
        if (R->doing_tail[0] && (ntail == 0 || ntail == 1)) {
            double ebeam_max = Mp * R->edge_e_E[1] / (Mp - R->edge_e_E[1] * (1. - v.uez));
            if (ebeam_max < 0) ebeam_max = 1.e10;
            const double ebeam_min = Mp * R->edge_e_E[0] / (Mp - R->edge_e_E[0] * (1. - v.uez));
            Emin[0] = v.Ein - ebeam_max;
            Emax[0] = v.Ein - ebeam_min;
            Emax[0] = fmin(Emax[0], R->edge_Em[1]);
            Emax[0] = fmin(Emax[0], Egamma1_max);
            Emin[0] = Emin[0] - R->dE_edge_test;
            Emax[0] = Emax[0] + R->dE_edge_test;
            if (hardwired) Emax[0] = R->Egamma_gen_max;
            rad_basicrad(gr, &S, 1 * peaked_basis_flag, Emin[0], Emax[0], &Eg[0], &bw, &bvr);
            if (bw <= 0) {
                fail = 1;
            } else {
                v.Ein = v.Ein - Eg[0];
                if (!rad_complete_ev(gr, T, H, R, &S, &v)) {
                    fail = 1;
                } else {
                    rad_weight = rad_weight * bw;
                    if (R->rad_flag <= 1)
                        rad_weight = rad_peaked_weight(R, &S, extrad_flag, &v, Eg[0], Emin[0], Emax[0], bw);
                    else
                        rad_weight = rad_weight * rad_extrad_phi(R, &S, extrad_flag, 1, v.Ein, v.e_E, Eg[0]);
                }
            }
        }
        if (!fail && R->doing_tail[1] && (ntail == 0 || ntail == 2)) {
            Emin[1] = v.e_E - R->edge_e_E[1];
            Emax[1] = v.e_E - R->edge_e_E[0];
            /* doing_eep */
            Emax[1] = fmin(Emax[1], (R->edge_Em[1] - v.Em) - Eg[0] + max_delta_Trec);
            if (ntail != 0 || !S.rad_proton_this_ev)
                Emin[1] = fmax(Emin[1], (R->edge_Em[0] - v.Em) - Eg[0] - max_delta_Trec);
            Emax[1] = fmin(Emax[1], R->Egamma_tot_max - Eg[0]);
            Emin[1] = Emin[1] - R->dE_edge_test;
            Emax[1] = Emax[1] + R->dE_edge_test;
            if (hardwired) Emax[1] = R->Egamma_gen_max;
            rad_basicrad(gr, &S, 2 * peaked_basis_flag, Emin[1], Emax[1], &Eg[1], &bw, &bvr);
            if (bw <= 0) {
                fail = 1;
            } else {
                rad_weight = rad_weight * bw;
                if (R->rad_flag <= 1)
                    rad_weight = rad_peaked_weight(R, &S, extrad_flag, &v, Eg[1], Emin[1], Emax[1], bw);
                else
                    rad_weight = rad_weight * rad_extrad_phi(R, &S, extrad_flag, 2, v.Ein, v.e_E, Eg[1]);
            }
        }
        if (!fail && S.rad_proton_this_ev && (ntail == 0 || ntail == 3)) {
            Emin[2] = v.p_E - R->edge_p_E[1];
            Emax[2] = v.p_E - R->edge_p_E[0];
            Emax[2] = fmin(Emax[2], (R->edge_Em[1] - v.Em) - Eg[0] - Eg[1] + max_delta_Trec);
            Emin[2] = fmax(Emin[2], (R->edge_Em[0] - v.Em) - Eg[0] - Eg[1] - max_delta_Trec);
            Emax[2] = fmin(Emax[2], R->Egamma_tot_max - Eg[0] - Eg[1]);
            Emin[2] = Emin[2] - R->dE_edge_test;
            Emax[2] = Emax[2] + R->dE_edge_test;
            if (hardwired) Emax[2] = R->Egamma_gen_max;
            rad_basicrad(gr, &S, 3 * peaked_basis_flag, Emin[2], Emax[2], &Eg[2], &bw, &bvr);
            if (bw <= 0) {
                fail = 1;
            } else {
                rad_weight = rad_weight * bw;
                if (R->rad_flag <= 1)
                    rad_weight = rad_peaked_weight(R, &S, extrad_flag, &v, Eg[2], Emin[2], Emax[2], bw);
                else
                    rad_weight = rad_weight * rad_extrad_phi(R, &S, extrad_flag, 3, v.Ein, v.e_E, Eg[2]);
            }
        }
        if (!fail) {
            orig_Ein = v.Ein + Eg[0];
            orig_e_E = v.e_E - Eg[1];
            if (orig_e_E > 0e0) {
                orig_e_delta = (orig_e_E - H->e_P) / H->e_P * 100.;
                orig_p_E = v.p_E - Eg[2];
                if (orig_p_E > H->Mh) {
                    orig_p_P = sqrt(orig_p_E * orig_p_E - H->Mh * H->Mh);
                    orig_p_delta = (orig_p_P - H->p_P) / H->p_P * 100.;
                    gen_weight = gen_weight * rad_weight / S.hardcorfac;
                    ok = 1;
                }
            }
        }
    }
    const double sigcc = or_sigep(v.Q2, v.e_E, v.e_theta, v.Ein);
    const double weight = v.jacobian * gen_weight * sigcc;
#define OR_PUT(f, val) if (o->f) o->f[i] = (val)
    OR_PUT(x, x); OR_PUT(y, y); OR_PUT(z, v.z); OR_PUT(rasterx, t5); OR_PUT(rastery, t6);
    OR_PUT(Ein, v.Ein); OR_PUT(eloss, v.eloss[0]); OR_PUT(teff, v.teff[0]);
    OR_PUT(e_xptar, v.e_xptar); OR_PUT(e_yptar, v.e_yptar); OR_PUT(e_theta, v.e_theta);
    OR_PUT(e_phi, v.e_phi); OR_PUT(e_E, v.e_E); OR_PUT(e_delta, v.e_delta);
    OR_PUT(p_xptar, v.p_xptar); OR_PUT(p_yptar, v.p_yptar); OR_PUT(p_theta, v.p_theta);
    OR_PUT(p_phi, v.p_phi); OR_PUT(p_P, v.p_P); OR_PUT(p_E, v.p_E); OR_PUT(p_delta, v.p_delta);
    OR_PUT(Q2, v.Q2); OR_PUT(nu, v.nu); OR_PUT(q, v.q); OR_PUT(jacobian, v.jacobian);
    OR_PUT(sigcc, sigcc); OR_PUT(weight, weight); OR_PUT(ok, (uint8_t)ok);
#undef OR_PUT
#define OR_PUT(f, val) if (ro->f) ro->f[i] = (val)
    OR_PUT(Egamma1, Eg[0]); OR_PUT(Egamma2, Eg[1]); OR_PUT(Egamma3, Eg[2]); OR_PUT(gen_weight, gen_weight);
    OR_PUT(orig_Ein, orig_Ein); OR_PUT(orig_e_E, orig_e_E); OR_PUT(orig_e_delta, orig_e_delta);
    OR_PUT(orig_p_E, orig_p_E); OR_PUT(orig_p_P, orig_p_P); OR_PUT(orig_p_delta, orig_p_delta);
    OR_PUT(eloss2, v.eloss[1]); OR_PUT(teff2, v.teff[1]); OR_PUT(eloss3, v.eloss[2]); OR_PUT(teff3, v.teff[2]);
    OR_PUT(hardcorfac, S.hardcorfac); OR_PUT(ntail, (int32_t)ntail);
#undef OR_PUT
}

int or_gen_hep_rad(const or_target *T, const or_hep *H, const or_rad *R, int64_t n, or_hep_out *out,
                   or_rad_out *rout, int rng_kind, const double *tape, int64_t stride, uint64_t seed,
                   uint64_t event0, int nthreads, int32_t *used)
{
    if (R->rad_flag < 0 || R->rad_flag > 3 || R->extrad_flag < 0 || R->extrad_flag > 3) return -3;
    /* rad_flag 1 with the Friedrich form reads stale saved locals in the
     * reference (radc.f:610-636); rad_flag >= 2 with extrad_flag 3 stops
     * there for the combined tail only, which the peaked basis alone uses */
    if (R->rad_flag == 1 && R->extrad_flag == 3) return -3;
    if (nthreads <= 0) nthreads = 1;
    int64_t overflow = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : overflow)
#endif
    for (int64_t i = 0; i < n; i++) {
        or_rng g;
        memset(&g, 0, sizeof g);
        g.kind = rng_kind;
        if (rng_kind == 0) {
            g.tape = tape + i * stride;
            g.stride = stride;
        } else {
            g.key[0] = (uint32_t)seed;
            g.key[1] = (uint32_t)(seed >> 32);
            g.event = event0 + (uint64_t)i;
            g.tag = OR_GEN_TAG;
        }
        gen_hep_rad_event(&g, T, H, R, i, out, rout);
        if (g.overflow) overflow++;
        if (used) used[i] = (int32_t)g.used;
    }
    return overflow ? -2 : 0;
}

#ifdef OR_RAD_PERTURB
#undef log
#undef exp
#undef pow
#undef tgamma
#undef sin
#undef cos
#undef asin
#undef acos
#undef atan2
#endif
