// DAF/SPK type-2 reader (host side).
//
// Follows the reference reader's conventions (reference src/spk.c:215-402):
// file record check ("DAF/SPK", ND=2, NI=6), summary records from FWARD along
// the NEXT chain, targets in first-seen order, each target's begin epoch and
// segment length taken from its first segment, the file memory-mapped.
// Constants come from the comment area after the line "Initial conditions and
// constants used for integration:" (reference src/spk.c:696-766), masses are
// joined by NAIF code / "MAnnnn" name (reference src/spk.c:137-210).
#include <fcntl.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdint>

#include "internal.h"

namespace {

constexpr int kRecord = RECORD_LENGTH;

inline double jul(double eph) { return 2451545.0 + eph / 86400.0; }

// Appends a segment's word addresses to target t, doubling its arrays' capacity as needed.
bool add_segment(spk_target& t, int one, int two) {
    const int next = t.ind + 1;
    if (next >= t.allocated_ind) {
        int cap = t.allocated_ind ? 2 * t.allocated_ind : 32;
        while (cap <= next) cap *= 2;
        int* o = (int*)realloc(t.one, (size_t)cap * sizeof(int));
        if (!o) return false;
        t.one = o;
        int* w = (int*)realloc(t.two, (size_t)cap * sizeof(int));
        if (!w) return false;
        t.two = w;
        t.allocated_ind = cap;
    }
    t.ind = next;
    t.one[next] = one;
    t.two[next] = two;
    return true;
}

// Index of the target with this code in s, appending a zeroed one (capacity grown in steps of 32,
// the spare entries zero) when it is new; -1 when memory runs out.
int target_slot(spk_s* s, int code) {
    for (int i = 0; i < s->num; i++)
        if (s->targets[i].code == code) return i;
    if (s->num >= s->allocated_num) {
        const int cap = s->allocated_num + 32;
        spk_target* t = (spk_target*)realloc(s->targets, (size_t)cap * sizeof(spk_target));
        if (!t) return -1;
        memset(t + s->allocated_num, 0, (size_t)(cap - s->allocated_num) * sizeof(spk_target));
        s->targets = t;
        s->allocated_num = cap;
    }
    spk_target& t = s->targets[s->num];
    memset(&t, 0, sizeof t);
    t.code = code;
    t.ind = -1;
    return s->num++;
}

}  // namespace

extern "C" {

int assist_spk_free(struct spk_s* pl) {
    if (!pl) return -1;
    if (pl->device_cache) assist_gpu::free_device_ephem((assist_gpu::DeviceEphemCache*)pl->device_cache);
    if (pl->targets) {
        for (int m = 0; m < pl->num; m++) {
            free(pl->targets[m].one);
            free(pl->targets[m].two);
        }
        free(pl->targets);
    }
    if (pl->map) munmap(pl->map, pl->len);
    memset(pl, 0, sizeof *pl);
    free(pl);
    return 0;
}

union record_t* assist_load_spk_file_record(int fd) {
    union record_t* r = (union record_t*)calloc(1, sizeof(union record_t));
    if (!r) return nullptr;
    if (pread(fd, r->buf, kRecord, 0) != kRecord || strncmp(r->file.locidw, "DAF/SPK", 7) != 0 ||
        8 * (r->file.nd + (r->file.ni + 1) / 2) != (int)sizeof(struct sum_s)) {
        free(r);
        return nullptr;
    }
    return r;
}

struct spk_s* assist_spk_init(const char* path) {
    int fd = open(path, O_RDONLY);
    if (fd < 0) return nullptr;
    union record_t* rec = assist_load_spk_file_record(fd);
    if (!rec || rec->file.fward < 2) {
        free(rec);
        close(fd);
        return nullptr;
    }
    // the first summary record (its word 2, PREV, is zero)
    if (pread(fd, rec->buf, kRecord, (off_t)(rec->file.fward - 1) * kRecord) != kRecord || rec->buf[8] != 0) {
        free(rec);
        close(fd);
        return nullptr;
    }
    spk_s* s = (spk_s*)calloc(1, sizeof(spk_s));
    bool ok = s != nullptr;
    for (int guard = 0; ok && guard < 1000000; guard++) {
        const int ns = (int)rec->summary.nsum;
        for (int b = 0; ok && b < ns && b < 25; b++) {
            const struct sum_s& sm = rec->summary.s[b];
            const int m = target_slot(s, sm.tar);
            if (m < 0) { ok = false; break; }
            spk_target& t = s->targets[m];
            if (t.ind < 0) {
                t.cen = sm.cen;
                t.beg = jul(sm.beg);
                t.res = jul(sm.end) - t.beg;
            }
            ok = add_segment(t, sm.one, sm.two);
            t.end = jul(sm.end);
        }
        const long nx = (long)rec->summary.next - 1;
        if (nx < 0) break;
        if (pread(fd, rec->buf, kRecord, (off_t)nx * kRecord) != kRecord) break;
    }
    free(rec);
    struct stat sb;
    if (!ok || fstat(fd, &sb) < 0) {
        close(fd);
        assist_spk_free(s);
        return nullptr;
    }
    s->len = (size_t)sb.st_size;
    void* map = mmap(nullptr, s->len, PROT_READ, MAP_SHARED, fd, 0);
    close(fd);
    if (map == MAP_FAILED) {
        s->len = 0;
        assist_spk_free(s);
        return nullptr;
    }
    s->map = map;
#if defined(MADV_RANDOM)
    (void)madvise(map, s->len, MADV_RANDOM);
#endif
    // every word address the evaluator can touch must lie inside the file
    const size_t nwords = s->len / 8;
    for (int i = 0; i < s->num; i++) {
        const spk_target& t = s->targets[i];
        for (int k = 0; k <= t.ind; k++) {
            if (t.one[k] < 1 || t.two[k] < 4 || (size_t)t.two[k] > nwords || t.one[k] > t.two[k]) {
                assist_spk_free(s);
                return nullptr;
            }
        }
    }
    return s;
}

struct spk_target* assist_spk_find_target(const struct spk_s* pl, int code) {
    if (!pl) return nullptr;
    for (int i = 0; i < pl->num; i++)
        if (pl->targets[i].code == code) return &pl->targets[i];
    return nullptr;
}

void parse_comments(int fd, int first_summary_record, char** comments) {
    *comments = nullptr;
    const int nrec = first_summary_record - 2;
    if (nrec <= 0) return;
    char* buf = (char*)malloc((size_t)nrec * kRecord + 1);
    if (!buf) return;
    size_t len = 0;
    for (int i = 0; i < nrec; i++) {
        if (pread(fd, buf + len, kRecord, (off_t)(1 + i) * kRecord) != kRecord) {
            free(buf);
            return;
        }
        // trailing NUL and EOT bytes pad a comment record
        int k = kRecord;
        while (k > 0 && (buf[len + k - 1] == 0 || buf[len + k - 1] == 4)) k--;
        len += (size_t)k;
    }
    for (size_t i = 0; i < len; i++)
        if (buf[i] == 0) buf[i] = '\n';  // a NUL ends a comment line
    buf[len] = 0;
    *comments = buf;
}

struct spk_constants_and_masses assist_load_spk_constants_and_masses(const char* path) {
    struct spk_constants_and_masses c;
    memset(&c, 0, sizeof c);
    int fd = open(path, O_RDONLY);
    if (fd < 0) return c;
    union record_t* rec = assist_load_spk_file_record(fd);
    char* text = nullptr;
    if (rec) parse_comments(fd, rec->file.fward, &text);
    free(rec);
    close(fd);
    if (!text) return c;
    bool in_sec = false;
    char* save = nullptr;
    for (char* line = strtok_r(text, "\n", &save); line; line = strtok_r(nullptr, "\n", &save)) {
        if (strstr(line, "Initial conditions and constants used for integration:")) in_sec = true;
        if (!in_sec) continue;
        for (char* p = line; *p; p++)
            if (*p == 'D' || *p == 'd') *p = 'e';  // Fortran exponents
        char key[64], vs[64];
        if (sscanf(line, "%63s %63s", key, vs) != 2) continue;
        const double v = strtod(vs, nullptr);
        if (!strcmp(key, "cau") || !strcmp(key, "AU")) c.AU = v;
        else if (!strcmp(key, "EMRAT")) c.EMRAT = v;
        else if (!strcmp(key, "J2E")) c.J2E = v;
        else if (!strcmp(key, "J3E")) c.J3E = v;
        else if (!strcmp(key, "J4E")) c.J4E = v;
        else if (!strcmp(key, "J2SUN")) c.J2SUN = v;
        else if (!strcmp(key, "RE")) c.RE = v;
        else if (!strcmp(key, "CLIGHT")) c.CLIGHT = v;
        else if (!strcmp(key, "ASUN")) c.ASUN = v;
        else if (!strncmp(key, "GM", 2) || !strncmp(key, "MA", 2)) {
            char** nn = (char**)realloc(c.masses.names, (c.masses.count + 1) * sizeof(char*));
            if (!nn) break;
            c.masses.names = nn;
            double* nv = (double*)realloc(c.masses.values, (c.masses.count + 1) * sizeof(double));
            if (!nv) break;
            c.masses.values = nv;
            c.masses.names[c.masses.count] = strdup(key);
            c.masses.values[c.masses.count] = v;
            c.masses.count++;
        }
    }
    free(text);
    return c;
}

void assist_apply_spk_constants(struct assist_ephem* ephem, const struct spk_constants_and_masses* d) {
    if (!ephem || !d) return;
    ephem->AU = d->AU;
    ephem->EMRAT = d->EMRAT;
    ephem->J2E = d->J2E;
    ephem->J3E = d->J3E;
    ephem->J4E = d->J4E;
    ephem->J2SUN = d->J2SUN;
    ephem->RE = d->RE;
    ephem->CLIGHT = d->CLIGHT;
    ephem->ASUN = d->ASUN;
    ephem->Re_eq = ephem->RE / ephem->AU;
    ephem->Rs_eq = ephem->ASUN / ephem->AU;
    ephem->c_AU_per_day = (ephem->CLIGHT / ephem->AU) * 86400;
    ephem->c_squared = ephem->c_AU_per_day * ephem->c_AU_per_day;
    ephem->over_c_squared = 1.0 / ephem->c_squared;
}

void assist_free_spk_constants_and_masses(struct spk_constants_and_masses* d) {
    if (!d || !d->masses.names) return;
    for (size_t i = 0; i < d->masses.count; i++) free(d->masses.names[i]);
    free(d->masses.names);
    free(d->masses.values);
    d->masses.names = nullptr;
    d->masses.values = nullptr;
    d->masses.count = 0;
}

void assist_spk_join_masses(struct spk_s* sp, const struct mass_data* masses, double emrat) {
    static const struct { const char* name; int code; } planet_codes[] = {
        {"GMS", 10}, {"GM1", 1}, {"GM2", 2}, {"GMB", 399}, {"GMB", 3}, {"GMB", 301},
        {"GM4", 4}, {"GM5", 5}, {"GM6", 6}, {"GM7", 7}, {"GM8", 8}, {"GM9", 9}};
    if (!sp || !masses || !masses->names) return;
    for (int m = 0; m < sp->num; m++) {
        spk_target& t = sp->targets[m];
        if (t.mass != 0) continue;
        char label[64] = {0};
        for (const auto& pc : planet_codes)
            if (t.code == pc.code) { snprintf(label, sizeof label, "%s", pc.name); break; }
        if (!label[0]) snprintf(label, sizeof label, "MA%04d", t.code - 2000000);
        for (size_t i = 0; i < masses->count; i++) {
            if (strcmp(masses->names[i], label) == 0) {
                if (t.code == 399) t.mass = masses->values[i] * (emrat / (1. + emrat));
                else if (t.code == 301) t.mass = masses->values[i] * (1. / (1. + emrat));
                else t.mass = masses->values[i];
                break;
            }
        }
        if (t.mass == 0 && t.code != 199 && t.code != 299) printf("Mass not found for target code: %d\n", t.code);
    }
}

}  // extern "C"

namespace assist_host {

int find_target(const spk_s* s, int code) {
    const spk_target* t = assist_spk_find_target(s, code);
    return t ? (int)(t - s->targets) : -1;
}

// Record of target t for time jd_rel (reference src/spk.c:430-445, 500-514).  Segment and
// record indices are clamped to the data actually present.
static const double* locate(const spk_s* s, const spk_target& t, double jd_ref, double jd_rel, int* P, double* z) {
    int n = (int)((jd_ref + jd_rel - t.beg) / t.res);
    const int nseg = segments(t);
    n = n < 0 ? 0 : (n >= nseg ? nseg - 1 : n);
    const double* base = (const double*)s->map;
    const double* val = base + t.two[n] - 1;
    const int R = (int)val[-1];
    *P = (R - 2) / 3;
    int b = (int)(((jd_ref - jul(val[-3])) + jd_rel) / (val[-2] / 86400.0));
    const int N = (int)val[0];
    b = b < 0 ? 0 : (b >= N ? N - 1 : b);
    val = base + (t.one[n] - 1) + (long)b * R;
    *z = ((jd_ref - jul(val[0])) + jd_rel) / (val[1] / 86400.0);
    return val;
}

void target_pva(const spk_s* s, const spk_target& t, double jd_ref, double jd_rel, double u[3], double v[3],
                double w[3]) {
    int P;
    double z;
    const double* val = locate(s, t, jd_ref, jd_rel, &P, &z);
    if (P > 32) P = 32;
    double T[32], S[32], U[32];
    const double c = 1.0 / val[1];
    T[0] = 1.0; T[1] = z;
    S[0] = 0.0; S[1] = 1.0;
    U[0] = 0.0; U[1] = 0.0; U[2] = 4.0;
    for (int p = 2; p < P; p++) {
        T[p] = 2.0 * z * T[p - 1] - T[p - 2];
        S[p] = 2.0 * z * S[p - 1] + 2.0 * T[p - 1] - S[p - 2];
    }
    for (int p = 3; p < P; p++) U[p] = 2.0 * z * U[p - 1] + 4.0 * S[p - 1] - U[p - 2];
    for (int n = 0; n < 3; n++) {
        const int b = 2 + n * P;
        u[n] = v[n] = w[n] = 0.0;
        for (int p = 0; p < P; p++) {
            const double coeff = val[b + p];
            u[n] += coeff * T[p];
            v[n] += coeff * S[p] * c;
            w[n] += coeff * U[p] * c * c;
        }
    }
}

int asteroid_pos(const spk_s* s, double jd_ref, double jd_rel, int m, double* GM, double out[3]) {
    if (!s) return ASSIST_ERROR_AST_FILE;
    if (m < 0 || m >= s->num) return ASSIST_ERROR_NAST;
    const spk_target& t = s->targets[m];
    if (jd_ref + jd_rel < t.beg || jd_ref + jd_rel > t.end) return ASSIST_ERROR_COVERAGE;
    *GM = t.mass;
    int P;
    double z;
    const double* val = locate(s, t, jd_ref, jd_rel, &P, &z);
    if (P > 32) P = 32;
    double T[32];
    T[0] = 1.0; T[1] = z;
    for (int p = 2; p < P; p++) T[p] = 2.0 * z * T[p - 1] - T[p - 2];
    for (int n = 0; n < 3; n++) {
        const int b = 2 + n * P;
        double u = 0.0;
        for (int p = 0; p < P; p++) u += val[b + p] * T[p];
        out[n] = u / 149597870.7;
    }
    return ASSIST_SUCCESS;
}

static const int kNaifByAssist[ASSIST_BODY_NPLANETS] = {10, 1, 2, 399, 301, 4, 5, 6, 7, 8, 9};

// Barycentric state of planet-file target t (NAIF `code`) in au, au/day, au/day^2 (reference
// src/spk.c:550-610): its own Chebyshev series, plus the EMB's for the Earth and the Moon.
static int calc_from_target(const assist_ephem* e, const spk_target* t, int code, double jd_ref, double jd_rel,
                            double* GM, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                            double* ax, double* ay, double* az) {
    const spk_s* pl = e->spk_planets;
    if (!pl || !t) return ASSIST_ERROR_NEPHEM;
    if (jd_ref + jd_rel < t->beg || jd_ref + jd_rel > t->end) return ASSIST_ERROR_COVERAGE;
    double u[3], v[3], w[3];
    target_pva(pl, *t, jd_ref, jd_rel, u, v, w);
    if (code == 301 || code == 399) {
        const int ei = e->spk_emb_index >= 0 ? e->spk_emb_index : find_target(pl, 3);
        if (ei < 0) return ASSIST_ERROR_NEPHEM;
        double eu[3], ev[3], ew[3];
        target_pva(pl, pl->targets[ei], jd_ref, jd_rel, eu, ev, ew);
        for (int i = 0; i < 3; i++) { u[i] += eu[i]; v[i] += ev[i]; w[i] += ew[i]; }
    }
    *GM = t->mass;
    const double au = e->AU, spd = 86400.;
    *x = u[0] / au; *y = u[1] / au; *z = u[2] / au;
    *vx = v[0] / (au / spd); *vy = v[1] / (au / spd); *vz = v[2] / (au / spd);
    *ax = w[0] / (au / (spd * spd)); *ay = w[1] / (au / (spd * spd)); *az = w[2] / (au / (spd * spd));
    return ASSIST_SUCCESS;
}

int planets_calc(const assist_ephem* e, double jd_ref, double jd_rel, int body, double* GM, double* x, double* y,
                 double* z, double* vx, double* vy, double* vz, double* ax, double* ay, double* az) {
    if (body < 0 || body >= ASSIST_BODY_NPLANETS) return ASSIST_ERROR_NEPHEM;
    const spk_s* pl = e->spk_planets;
    if (!pl) return ASSIST_ERROR_NEPHEM;
    const int code = kNaifByAssist[body];
    int idx = e->spk_target_index[body];
    if (idx < 0 || idx >= pl->num || pl->targets[idx].code != code) idx = find_target(pl, code);
    const int rc = calc_from_target(e, idx >= 0 ? &pl->targets[idx] : nullptr, code, jd_ref, jd_rel, GM, x, y, z, vx,
                                    vy, vz, ax, ay, az);
    if (rc == ASSIST_SUCCESS || body != ASSIST_BODY_EARTH) return rc;
    // Earth from the EMB and the Moon when 399 is absent (reference src/spk.c:668-691)
    if (e->spk_emb_index < 0 || e->spk_target_index[ASSIST_BODY_MOON] < 0) return rc;
    const spk_target& emb = pl->targets[e->spk_emb_index];
    double eu[3], ev[3], ew[3], mu[3], mv[3], mw[3], u[3], v[3], w[3];
    target_pva(pl, emb, jd_ref, jd_rel, eu, ev, ew);
    target_pva(pl, pl->targets[e->spk_target_index[ASSIST_BODY_MOON]], jd_ref, jd_rel, mu, mv, mw);
    const double frac = 1.0 / (1.0 + e->EMRAT);
    for (int i = 0; i < 3; i++) {
        u[i] = -frac * mu[i] + eu[i];
        v[i] = -frac * mv[i] + ev[i];
        w[i] = -frac * mw[i] + ew[i];
    }
    const double au = e->AU, spd = 86400.;
    *x = u[0] / au; *y = u[1] / au; *z = u[2] / au;
    *vx = v[0] / (au / spd); *vy = v[1] / (au / spd); *vz = v[2] / (au / spd);
    *ax = w[0] / (au / (spd * spd)); *ay = w[1] / (au / (spd * spd)); *az = w[2] / (au / (spd * spd));
    *GM = 0.0;
    return ASSIST_SUCCESS;
}

}  // namespace assist_host

extern "C" {

enum ASSIST_STATUS assist_spk_calc(const struct spk_s* pl, double jd_ref, double jd_rel, int m, double* GM, double* x,
                                   double* y, double* z) {
    double out[3];
    const int rc = assist_host::asteroid_pos(pl, jd_ref, jd_rel, m, GM, out);
    if (rc == ASSIST_SUCCESS) { *x = out[0]; *y = out[1]; *z = out[2]; }
    return (enum ASSIST_STATUS)rc;
}

enum ASSIST_STATUS assist_spk_calc_planets(const struct assist_ephem* ephem, double jd_ref, double jd_rel, int code,
                                           double* GM, double* x, double* y, double* z, double* vx, double* vy,
                                           double* vz, double* ax, double* ay, double* az) {
    if (!ephem) return ASSIST_ERROR_NEPHEM;
    const spk_target* t = assist_spk_find_target(ephem->spk_planets, code);
    if (!t) return ASSIST_ERROR_NEPHEM;
    return (enum ASSIST_STATUS)assist_host::calc_from_target(ephem, t, code, jd_ref, jd_rel, GM, x, y, z, vx, vy, vz,
                                                             ax, ay, az);
}

enum ASSIST_STATUS assist_spk_calc_planets_by_assist(const struct assist_ephem* ephem, double jd_ref, double jd_rel,
                                                     int assist_body, double* GM, double* out_x, double* out_y,
                                                     double* out_z, double* out_vx, double* out_vy, double* out_vz,
                                                     double* out_ax, double* out_ay, double* out_az) {
    if (!ephem) return ASSIST_ERROR_NEPHEM;
    return (enum ASSIST_STATUS)assist_host::planets_calc(ephem, jd_ref, jd_rel, assist_body, GM, out_x, out_y, out_z,
                                                         out_vx, out_vy, out_vz, out_ax, out_ay, out_az);
}

struct mpos_s assist_spk_target_pos(const struct spk_s* pl, const struct spk_target* target, double jd_ref,
                                    double jd_rel) {
    struct mpos_s pos;
    memset(&pos, 0, sizeof pos);
    if (pl && target) assist_host::target_pva(pl, *target, jd_ref, jd_rel, pos.u, pos.v, pos.w);
    return pos;
}

}  // extern "C"
