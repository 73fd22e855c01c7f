// JPL binary ephemeris (.440/.441) reader and host-side evaluation.
//
// File layout (reference src/ascii_ephem.c:105-252): block 0 holds three 84-character title
// lines, the first 400 six-character constant names from byte 0x00FC, the header values from
// 0x0A5C (begin/end JD, days per block, constant count, km per au, Earth/Moon mass ratio, then
// offset/coefficients/sub-intervals of columns 0-11, the ephemeris number and column 12), the
// remaining constant names from 0x0B28, and columns 13-14.  Block 1 holds the constant values;
// block k + 2 covers [beg + k inc, beg + (k + 1) inc].  Within a block a column's sub-interval
// b, component m starts at word off + ncf (m + b ncm).
//
// Beyond the reference's reader this one checks, before anything is evaluated, that every
// column lies inside a block and that the file holds whole blocks, and clamps the block and
// sub-interval indices at the very end of coverage (where the reference reads one block or
// one sub-interval past the data).
#include <fcntl.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "../../include/ascii_ephem.h"
#include "internal.h"

namespace {

constexpr off_t kNames0 = 0x00FC;   // first 400 constant names
constexpr off_t kHeader = 0x0A5C;   // header values
constexpr int kFirstNames = 400;
constexpr int kMaxCoef = 24;        // Chebyshev work arrays of the evaluation

// Sequential reads from a file offset.
struct Cursor {
    int fd;
    off_t pos;
    bool ok = true;
    template <typename T>
    void get(T* dst) {
        if (ok && pread(fd, dst, sizeof(T), pos) == (ssize_t)sizeof(T)) pos += sizeof(T);
        else ok = false;
    }
    void get_name(char* dst) {
        if (ok && pread(fd, dst, 6, pos) == 6) pos += 6;
        else ok = false;
    }
};

// Value of the first constant whose name starts with the 6 characters of `name6`; warns and
// returns 0 when absent (reference src/ascii_ephem.c:69-78).
double constant6(const ascii_s* a, const char* name6) {
    for (int p = 0; p < a->num; p++)
        if (strncmp(name6, a->str[p], 6) == 0) return a->con[p];
    fprintf(stderr, "WARNING: Constant [%s] not found in ephemeris file.\n", name6);
    return 0.0;
}

void release(ascii_s* a) {
    if (!a) return;
    if (a->device_cache) assist_gpu::free_device_ephem((assist_gpu::DeviceEphemCache*)a->device_cache);
    if (a->map) munmap(a->map, a->len);
    if (a->str) {
        const int n = a->num > kFirstNames ? a->num : kFirstNames;
        for (int p = 0; p < n; p++) free(a->str[p]);
        free(a->str);
    }
    free(a->con);
    free(a);
}

// Chebyshev sums of one column (km, km/s, km/s^2) at fraction `f` in [0, 1] of a block of
// `days` days (reference src/ascii_ephem.c:27-65).
void column_sums(const double* P, int ncm, int ncf, int niv, double f, double days, double* u, double* v,
                 double* w) {
    double T[kMaxCoef], S[kMaxCoef], U[kMaxCoef];
    const double s = f * (double)niv;
    double z = 2.0 * fmod(s, 1.0) - 1.0;
    const double c = (double)(niv * 2) / days / 86400.0;
    int b = (int)s;
    if (b >= niv) { b = niv - 1; z = 1.0; }  // f == 1: the end of the block's last sub-interval
    if (b < 0) b = 0;
    T[0] = 1.0; T[1] = z;
    S[0] = 0.0; S[1] = 1.0;
    U[0] = 0.0; U[1] = 0.0; U[2] = 4.0;
    for (int p = 2; p < ncf; p++) {
        T[p] = 2.0 * z * T[p - 1] - T[p - 2];
        S[p] = 2.0 * z * S[p - 1] + 2.0 * T[p - 1] - S[p - 2];
    }
    for (int p = 3; p < ncf; p++) U[p] = 2.0 * z * U[p - 1] + 4.0 * S[p - 1] - U[p - 2];
    for (int m = 0; m < ncm; m++) {
        const double* cf = P + ncf * (m + b * ncm);
        double su = 0.0, sv = 0.0, sw = 0.0;
        for (int p = 0; p < ncf; p++) {
            su += T[p] * cf[p];
            sv += S[p] * cf[p] * c;
            sw += U[p] * cf[p] * c * c;
        }
        u[m] = su; v[m] = sv; w[m] = sw;
    }
}

}  // namespace

extern "C" {

struct ascii_s* assist_ascii_init(char* path) {
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return nullptr;
    struct stat sb;
    if (fstat(fd, &sb) < 0) {
        close(fd);
        fprintf(stderr, "Error while trying to determine filesize.\n");
        return nullptr;
    }
    ascii_s* a = (ascii_s*)calloc(1, sizeof(ascii_s));
    Cursor cur{fd, kHeader};
    cur.get(&a->beg);
    cur.get(&a->end);
    cur.get(&a->inc);
    cur.get(&a->num);
    cur.get(&a->cau);
    cur.get(&a->cem);
    for (int p = 0; p < ASCII_N; p++) a->ncm[p] = 3;
    a->ncm[ASCII_NUT] = 2;
    a->ncm[ASCII_TDB] = 1;
    for (int p = 0; p < 12; p++) { cur.get(&a->off[p]); cur.get(&a->ncf[p]); cur.get(&a->niv[p]); }
    cur.get(&a->ver);
    cur.get(&a->off[12]); cur.get(&a->ncf[12]); cur.get(&a->niv[12]);
    if (!cur.ok || a->num < 0 || a->num > 100000 || !(a->inc > 0) || !(a->end >= a->beg)) {
        close(fd);
        free(a);
        return nullptr;
    }
    const int nstr = a->num > kFirstNames ? a->num : kFirstNames;
    a->str = (char**)calloc(nstr, sizeof(char*));
    for (int p = 0; p < nstr; p++) a->str[p] = (char*)calloc(8, 1);
    Cursor names{fd, kNames0};
    for (int p = 0; p < kFirstNames; p++) names.get_name(a->str[p]);
    for (int p = kFirstNames; p < a->num; p++) cur.get_name(a->str[p]);  // continue after the header
    for (int p = 13; p < 15; p++) { cur.get(&a->off[p]); cur.get(&a->ncf[p]); cur.get(&a->niv[p]); }
    bool ok = cur.ok && names.ok;
    for (int p = 0; p < ASCII_N; p++) a->off[p] -= 1;
    a->len = (size_t)sb.st_size;
    a->rec = sizeof(double) * 2;
    for (int p = 0; p < ASCII_N && ok; p++) {
        if (a->ncf[p] < 0 || a->niv[p] < 0 || a->ncf[p] > kMaxCoef || a->niv[p] > 64) ok = false;
        else a->rec += sizeof(double) * (size_t)a->ncf[p] * a->niv[p] * a->ncm[p];
    }
    // every column inside a block, and the file whole blocks covering [beg, end]
    for (int p = 0; p < ASCII_N && ok; p++) {
        const size_t words = (size_t)a->ncf[p] * a->niv[p] * a->ncm[p];
        if (words && (a->off[p] < 2 || (size_t)a->off[p] + words > a->rec / sizeof(double))) ok = false;
    }
    const double nblk_cover = ceil((a->end - a->beg) / a->inc);
    if (ok && (a->len < 2 * a->rec || (double)(a->len / a->rec - 2) < nblk_cover || 8 * (size_t)a->num > a->rec))
        ok = false;
    if (!ok) {
        close(fd);
        release(a);
        return nullptr;
    }
    void* map = mmap(nullptr, a->len, PROT_READ, MAP_SHARED, fd, 0);
    if (map == MAP_FAILED) {
        close(fd);
        release(a);
        fprintf(stderr, "Error while calling mmap().\n");
        return nullptr;
    }
    a->map = map;
    a->con = (double*)calloc(a->num > 0 ? a->num : 1, sizeof(double));
    memcpy(a->con, (const char*)map + a->rec, sizeof(double) * (size_t)a->num);
    // masses in ASSIST body order (Earth and Moon split from the EMB mass by EMRAT)
    a->mass[ASSIST_BODY_SUN] = constant6(a, "GMS   ");
    a->mass[ASSIST_BODY_MERCURY] = constant6(a, "GM1   ");
    a->mass[ASSIST_BODY_VENUS] = constant6(a, "GM2   ");
    const double emrat = constant6(a, "EMRAT ");
    const double gmb = constant6(a, "GMB   ");
    a->mass[ASSIST_BODY_EARTH] = (emrat / (1. + emrat)) * gmb;
    a->mass[ASSIST_BODY_MOON] = 1. / (1 + emrat) * gmb;
    a->mass[ASSIST_BODY_MARS] = constant6(a, "GM4   ");
    a->mass[ASSIST_BODY_JUPITER] = constant6(a, "GM5   ");
    a->mass[ASSIST_BODY_SATURN] = constant6(a, "GM6   ");
    a->mass[ASSIST_BODY_URANUS] = constant6(a, "GM7   ");
    a->mass[ASSIST_BODY_NEPTUNE] = constant6(a, "GM8   ");
    a->mass[ASSIST_BODY_PLUTO] = constant6(a, "GM9   ");
    a->J2E = constant6(a, "J2E   ");
    a->J3E = constant6(a, "J3E   ");
    a->J4E = constant6(a, "J4E   ");
    a->J2SUN = constant6(a, "J2SUN ");
    a->AU = constant6(a, "AU    ");
    a->RE = constant6(a, "RE    ");
    a->CLIGHT = constant6(a, "CLIGHT");
    a->ASUN = constant6(a, "ASUN  ");
    close(fd);
#if defined(MADV_RANDOM)
    madvise(a->map, a->len, MADV_RANDOM);
#endif
    return a;
}

void assist_ascii_free(struct ascii_s* ascii) { release(ascii); }

void assist_ascii_work(double* P, int ncm, int ncf, int niv, double t0, double t1, double* u, double* v, double* w) {
    if (!P || ncf < 0 || ncf > kMaxCoef || niv < 1 || ncm < 0 || ncm > 3) return;
    column_sums(P, ncm, ncf, niv, t0, t1, u, v, w);
}

enum ASSIST_STATUS assist_ascii_calc(struct ascii_s* a, double jd_ref, double jd_rel, int body, double* const GM,
                                     double* const x, double* const y, double* const z, double* const vx,
                                     double* const vy, double* const vz, double* const ax, double* const ay,
                                     double* const az) {
    if (!a || !a->map) return ASSIST_ERROR_EPHEM_FILE;
    if (body < 0 || body >= ASSIST_BODY_NPLANETS) return ASSIST_ERROR_NEPHEM;
    *GM = a->mass[body];
    if (jd_ref + jd_rel < a->beg || jd_ref + jd_rel > a->end) return ASSIST_ERROR_COVERAGE;
    const size_t nblk = a->len / a->rec - 2;
    size_t blk = (size_t)(uint32_t)((jd_ref + jd_rel - a->beg) / a->inc);
    if (blk >= nblk) blk = nblk - 1;
    const double* blkp = (const double*)a->map + (blk + 2) * (a->rec / sizeof(double));
    const double f = ((jd_ref - a->beg - (double)blk * a->inc) + jd_rel) / a->inc;
    static const int col[ASSIST_BODY_NPLANETS] = {ASCII_SUN, ASCII_MER, ASCII_VEN, -1, -1, ASCII_MAR,
                                                  ASCII_JUP, ASCII_SAT, ASCII_URA, ASCII_NEP, ASCII_PLU};
    double u[3], v[3], w[3];
    if (body == ASSIST_BODY_EARTH || body == ASSIST_BODY_MOON) {
        // barycentric Earth and Moon from the EMB and the geocentric Moon
        double eu[3], ev[3], ew[3], lu[3], lv[3], lw[3];
        column_sums(blkp + a->off[ASCII_EMB], a->ncm[ASCII_EMB], a->ncf[ASCII_EMB], a->niv[ASCII_EMB], f, a->inc, eu, ev, ew);
        column_sums(blkp + a->off[ASCII_LUN], a->ncm[ASCII_LUN], a->ncf[ASCII_LUN], a->niv[ASCII_LUN], f, a->inc, lu, lv, lw);
        const double k = body == ASSIST_BODY_EARTH ? -1.0 / (1.0 + a->cem) : a->cem / (1.0 + a->cem);
        for (int i = 0; i < 3; i++) {
            u[i] = eu[i] + lu[i] * k;
            v[i] = ev[i] + lv[i] * k;
            w[i] = ew[i] + lw[i] * k;
        }
    } else {
        const int c = col[body];
        if (a->ncf[c] < 1 || a->niv[c] < 1 || a->ncm[c] != 3) return ASSIST_ERROR_NEPHEM;
        column_sums(blkp + a->off[c], 3, a->ncf[c], a->niv[c], f, a->inc, u, v, w);
    }
    const double au = a->cau;
    *x = u[0] / au; *y = u[1] / au; *z = u[2] / au;
    *vx = v[0] / (au / 86400.); *vy = v[1] / (au / 86400.); *vz = v[2] / (au / 86400.);
    *ax = w[0] / (au / (86400. * 86400.)); *ay = w[1] / (au / (86400. * 86400.)); *az = w[2] / (au / (86400. * 86400.));
    return ASSIST_SUCCESS;
}

enum ASSIST_STATUS assist_ascii_calc_from_ephem(const struct assist_ephem* ephem, double jd_ref, double jd_rel,
                                                int body, double* const GM, double* const x, double* const y,
                                                double* const z, double* const vx, double* const vy,
                                                double* const vz, double* const ax, double* const ay,
                                                double* const az) {
    return assist_ascii_calc(ephem->ascii_planets, jd_ref, jd_rel, body, GM, x, y, z, vx, vy, vz, ax, ay, az);
}

int assist_ascii_find_constant(const struct ascii_s* ascii, const char* name, double* out_value) {
    if (out_value) *out_value = 0.0;
    if (!ascii || !ascii->str || !ascii->con || !name || !out_value) return 0;
    char key[6];
    memset(key, ' ', sizeof key);
    for (int i = 0; i < 6 && name[i]; i++) key[i] = name[i];
    for (int p = 0; p < ascii->num; p++) {
        if (memcmp(ascii->str[p], key, 6) == 0) {
            *out_value = ascii->con[p];
            return 1;
        }
    }
    return 0;
}

}  // extern "C"
