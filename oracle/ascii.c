/*
 * Oracle (test infrastructure): JPL binary ephemeris (.440/.441) loading and evaluation.
 * Restates reference src/ascii_ephem.c (reader :105-252, Chebyshev work :27-65,
 * body evaluation :275-384) and the format signature check of src/assist.c:66-106.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* column order, reference src/ascii_ephem.h:42-61 */
enum { A_MER, A_VEN, A_EMB, A_MAR, A_JUP, A_SAT, A_URA, A_NEP, A_PLU, A_LUN, A_SUN, A_NUT, A_LIB, A_MAN, A_TDB, A_N };

struct orc_ascii {
    double beg, end, inc, cau, cem;
    int32_t num, ver;
    int32_t off[A_N], ncf[A_N], niv[A_N], ncm[A_N];
    double mass[ORC_NPLANETS];
    double J2E, J3E, J4E, J2SUN, AU, RE, CLIGHT, ASUN;
    size_t len, rec; /* file and block sizes in bytes */
    size_t nblk;     /* data blocks in the file */
    double *words;   /* whole file */
    double *con;
    char (*str)[8];
};

static int rd(const unsigned char *buf, size_t n, size_t *pos, void *dst, size_t k) {
    if (*pos + k > n) return 0;
    memcpy(dst, buf + *pos, k);
    *pos += k;
    return 1;
}

/* reference src/ascii_ephem.c:69-78: first constant whose 6-character name matches */
static double get_constant(const orc_ascii *a, const char *name) {
    for (int p = 0; p < a->num; p++)
        if (strncmp(name, a->str[p], 6) == 0) return a->con[p];
    return 0.0;
}

/* reference src/ascii_ephem.c:80-102: name space-padded to 6 characters, exact match */
int orc_ascii_find_constant(const orc_ascii *a, const char *name, double *out) {
    if (out) *out = 0.0;
    if (!a || !name || !out) return 0;
    char key[6];
    for (int i = 0; i < 6; i++) key[i] = ' ';
    for (int i = 0; i < 6 && name[i]; i++) key[i] = name[i];
    for (int p = 0; p < a->num; p++) {
        if (memcmp(a->str[p], key, 6) == 0) { *out = a->con[p]; return 1; }
    }
    return 0;
}

/* reference src/assist.c:66-106 */
int orc_ascii_signature(const unsigned char *buf, size_t n) {
    if (n < 0x00FC + 18) return 0;
    int valid = 0;
    for (int i = 0; i < 3; i++) {
        const unsigned char *name = buf + 0x00FC + 6 * i;
        if (!((name[0] >= 'A' && name[0] <= 'Z') || (name[0] >= 'a' && name[0] <= 'z'))) continue;
        int alnum = 0, ok = 1;
        for (int j = 0; j < 6; j++) {
            unsigned char c = name[j];
            if ((c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9')) alnum++;
            else if (c != ' ' && c != 0) { ok = 0; break; }
        }
        if (ok && alnum >= 2) valid++;
    }
    return valid >= 2;
}

void orc_ascii_free(orc_ascii *a) {
    if (!a) return;
    free(a->words);
    free(a->con);
    free(a->str);
    free(a);
}

/* reference src/ascii_ephem.c:105-252 */
orc_ascii *orc_ascii_load(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (len < 0x0B28) { fclose(f); return NULL; }
    unsigned char *buf = calloc(1, ((size_t)len + 7) & ~(size_t)7);
    if (!buf || fread(buf, 1, (size_t)len, f) != (size_t)len) { fclose(f); free(buf); return NULL; }
    fclose(f);
    const size_t n = (size_t)len;
    orc_ascii *a = calloc(1, sizeof(orc_ascii));
    a->words = (double *)buf;
    a->len = n;
    size_t pos = 0x0A5C;
    int ok = rd(buf, n, &pos, &a->beg, 8) && rd(buf, n, &pos, &a->end, 8) && rd(buf, n, &pos, &a->inc, 8) &&
             rd(buf, n, &pos, &a->num, 4) && rd(buf, n, &pos, &a->cau, 8) && rd(buf, n, &pos, &a->cem, 8);
    for (int p = 0; p < A_N; p++) a->ncm[p] = 3;
    a->ncm[A_NUT] = 2;
    a->ncm[A_TDB] = 1;
    for (int p = 0; ok && p < 12; p++)
        ok = rd(buf, n, &pos, &a->off[p], 4) && rd(buf, n, &pos, &a->ncf[p], 4) && rd(buf, n, &pos, &a->niv[p], 4);
    ok = ok && rd(buf, n, &pos, &a->ver, 4) && rd(buf, n, &pos, &a->off[12], 4) && rd(buf, n, &pos, &a->ncf[12], 4) &&
         rd(buf, n, &pos, &a->niv[12], 4);
    if (!ok || a->num < 0 || a->num > 100000 || !(a->inc > 0)) { orc_ascii_free(a); return NULL; }
    /* names: the first 400 at 0x00FC, the rest at 0x0B28 */
    const int nstr = a->num > 400 ? a->num : 400;
    a->str = calloc((size_t)nstr, sizeof(*a->str));
    for (int p = 0; p < 400; p++) memcpy(a->str[p], buf + 0x00FC + 6 * p, 6);
    pos = 0x0B28;
    for (int p = 400; p < a->num; p++) ok = ok && rd(buf, n, &pos, a->str[p], 6);
    for (int p = 13; ok && p < 15; p++)
        ok = rd(buf, n, &pos, &a->off[p], 4) && rd(buf, n, &pos, &a->ncf[p], 4) && rd(buf, n, &pos, &a->niv[p], 4);
    if (!ok) { orc_ascii_free(a); return NULL; }
    for (int p = 0; p < A_N; p++) a->off[p] -= 1;
    a->rec = sizeof(double) * 2;
    for (int p = 0; p < A_N; p++) a->rec += sizeof(double) * (size_t)a->ncf[p] * a->niv[p] * a->ncm[p];
    if (a->rec < 16 || n < 2 * a->rec) { orc_ascii_free(a); return NULL; }
    a->nblk = n / a->rec - 2;
    /* constants: block 1 */
    a->con = calloc((size_t)(a->num ? a->num : 1), sizeof(double));
    if (a->rec + 8 * (size_t)a->num > n) { orc_ascii_free(a); return NULL; }
    memcpy(a->con, buf + a->rec, 8 * (size_t)a->num);
    /* masses in ASSIST body order (reference src/ascii_ephem.c:202-217) */
    a->mass[0] = get_constant(a, "GMS   ");
    a->mass[1] = get_constant(a, "GM1   ");
    a->mass[2] = get_constant(a, "GM2   ");
    const double emrat = get_constant(a, "EMRAT  ");
    const double gmb = get_constant(a, "GMB   ");
    a->mass[3] = (emrat / (1. + emrat)) * gmb;
    a->mass[4] = 1. / (1 + emrat) * gmb;
    a->mass[5] = get_constant(a, "GM4   ");
    a->mass[6] = get_constant(a, "GM5   ");
    a->mass[7] = get_constant(a, "GM6   ");
    a->mass[8] = get_constant(a, "GM7   ");
    a->mass[9] = get_constant(a, "GM8   ");
    a->mass[10] = get_constant(a, "GM9   ");
    a->J2E = get_constant(a, "J2E   ");
    a->J3E = get_constant(a, "J3E   ");
    a->J4E = get_constant(a, "J4E   ");
    a->J2SUN = get_constant(a, "J2SUN ");
    a->AU = get_constant(a, "AU    ");
    a->RE = get_constant(a, "RE    ");
    a->CLIGHT = get_constant(a, "CLIGHT");
    a->ASUN = get_constant(a, "ASUN  ");
    return a;
}

/* reference src/ascii_ephem.c:27-65.  `t0` is the fraction of the block, `t1` the block
 * length in days.  At the very end of coverage (t0 == 1) the reference indexes one
 * sub-interval past the column; the oracle evaluates the last sub-interval at z = 1. */
static void ascii_work(const double *P, int ncm, int ncf, int niv, double t0, double t1, double *u, double *v,
                       double *w) {
    double T[24], S[24], U[24];
    double t = t0 * (double)niv;
    t0 = 2.0 * fmod(t, 1.0) - 1.0;
    const double c = (double)(niv * 2) / t1 / 86400.0;
    int b = (int)t;
    if (b >= niv) { b = niv - 1; t0 = 1.0; }
    T[0] = 1.0; T[1] = t0;
    S[0] = 0.0; S[1] = 1.0;
    U[0] = 0.0; U[1] = 0.0; U[2] = 4.0;
    for (int p = 2; p < ncf; p++) {
        T[p] = 2.0 * t0 * T[p - 1] - T[p - 2];
        S[p] = 2.0 * t0 * S[p - 1] + 2.0 * T[p - 1] - S[p - 2];
    }
    for (int p = 3; p < ncf; p++) U[p] = 2.0 * t0 * U[p - 1] + 4.0 * S[p - 1] - U[p - 2];
    for (int m = 0; m < ncm; m++) {
        u[m] = v[m] = w[m] = 0.0;
        const int n = ncf * (m + b * ncm);
        for (int p = 0; p < ncf; p++) {
            u[m] += T[p] * P[n + p];
            v[m] += S[p] * P[n + p] * c;
            w[m] += U[p] * P[n + p] * c * c;
        }
    }
}

double orc_ascii_emrat(const orc_ascii *a) { return a->cem; }
void orc_ascii_bounds(const orc_ascii *a, double *beg, double *end) { *beg = a->beg; *end = a->end; }
double orc_ascii_constant(const orc_ascii *a, int which) {
    switch (which) {
        case 0: return a->J2E;
        case 1: return a->J3E;
        case 2: return a->J4E;
        case 3: return a->J2SUN;
        case 4: return a->AU;
        case 5: return a->RE;
        case 6: return a->CLIGHT;
        case 7: return a->ASUN;
        default: return 0.0;
    }
}

/* reference src/ascii_ephem.c:275-384: GM and barycentric (Moon: via the EMB) position,
 * velocity, acceleration of ASSIST body `body` in au, au/d, au/d^2. */
int orc_ascii_calc(const orc_ascii *a, double jd_ref, double jd_rel, int body, double *GM, double o[9]) {
    if (!a) return ORC_ERROR_EPHEM_FILE;
    if (body < 0 || body >= ORC_NPLANETS) return ORC_ERROR_NEPHEM;
    *GM = a->mass[body];
    if (jd_ref + jd_rel < a->beg || jd_ref + jd_rel > a->end) return ORC_ERROR_COVERAGE;
    uint32_t blk = (uint32_t)((jd_ref + jd_rel - a->beg) / a->inc);
    if (blk >= a->nblk) blk = (uint32_t)a->nblk - 1; /* reference reads past the data at t == end */
    const double *z = a->words + (blk + 2) * a->rec / sizeof(double);
    const double t = ((jd_ref - a->beg - (double)blk * a->inc) + jd_rel) / a->inc;
    static const int col[ORC_NPLANETS] = {A_SUN, A_MER, A_VEN, -1, -1, A_MAR, A_JUP, A_SAT, A_URA, A_NEP, A_PLU};
    double u[3], v[3], w[3];
    if (body == 3 || body == 4) {
        double eu[3], ev[3], ew[3], lu[3], lv[3], lw[3];
        ascii_work(&z[a->off[A_EMB]], a->ncm[A_EMB], a->ncf[A_EMB], a->niv[A_EMB], t, a->inc, eu, ev, ew);
        ascii_work(&z[a->off[A_LUN]], a->ncm[A_LUN], a->ncf[A_LUN], a->niv[A_LUN], t, a->inc, lu, lv, lw);
        const double f = body == 3 ? -1.0 / (1.0 + a->cem) : a->cem / (1.0 + a->cem);
        for (int i = 0; i < 3; i++) {
            u[i] = eu[i]; u[i] += lu[i] * f;
            v[i] = ev[i]; v[i] += lv[i] * f;
            w[i] = ew[i]; w[i] += lw[i] * f;
        }
    } else {
        const int c = col[body];
        ascii_work(&z[a->off[c]], a->ncm[c], a->ncf[c], a->niv[c], t, a->inc, u, v, w);
    }
    for (int i = 0; i < 3; i++) {
        o[i] = u[i] / a->cau;
        o[3 + i] = v[i] / (a->cau / 86400.);
        o[6 + i] = w[i] / (a->cau / (86400. * 86400.));
    }
    return ORC_SUCCESS;
}
