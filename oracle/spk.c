/*
 * Oracle (test infrastructure): DAF/SPK type-2 loading and Chebyshev evaluation.
 * Restates reference src/spk.c and the ephemeris front door of src/forces.c.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Bumped on every ephemeris load or free: empties the per-thread ephemeris caches (below). */
static volatile long orc_generation = 1;

#define REC 1024

/* reference src/spk.c:133-134 */
static double jul(double eph) { return 2451545.0 + eph / 86400.0; }

static unsigned char *read_file(const char *path, size_t *n) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (len < REC) { fclose(f); return NULL; }
    size_t padded = ((size_t)len + 7) & ~(size_t)7;
    unsigned char *buf = calloc(1, padded);
    if (!buf) { fclose(f); return NULL; }
    if (fread(buf, 1, (size_t)len, f) != (size_t)len) { fclose(f); free(buf); return NULL; }
    fclose(f);
    *n = (size_t)len;
    return buf;
}

static void spk_free(orc_spk *s) {
    if (!s) return;
    for (int i = 0; i < s->num; i++) { free(s->targets[i].one); free(s->targets[i].two); }
    free(s->targets);
    free(s->words);
    free(s);
}

/* reference src/spk.c:215-265 (file record) and :267-402 (summaries) */
static orc_spk *spk_load(const char *path) {
    size_t n = 0;
    unsigned char *buf = read_file(path, &n);
    if (!buf) return NULL;
    int32_t nd, ni, fward;
    if (strncmp((const char *)buf, "DAF/SPK", 7) != 0) { free(buf); return NULL; }
    memcpy(&nd, buf + 8, 4);
    memcpy(&ni, buf + 12, 4);
    memcpy(&fward, buf + 76, 4);
    if (8 * (nd + (ni + 1) / 2) != 40) { free(buf); return NULL; }
    size_t off = (size_t)(fward - 1) * REC;
    if (off + REC > n || buf[off + 8] != 0) { free(buf); return NULL; }
    orc_spk *s = calloc(1, sizeof(orc_spk));
    s->words = (double *)buf;
    s->nbytes = n;
    int cap = 0;
    for (;;) {
        double next, nsum;
        memcpy(&next, buf + off, 8);
        memcpy(&nsum, buf + off + 16, 8);
        for (int b = 0; b < (int)nsum; b++) {
            const unsigned char *sp = buf + off + 24 + 40 * b;
            double beg, end;
            int32_t iv[6];
            memcpy(&beg, sp, 8);
            memcpy(&end, sp + 8, 8);
            memcpy(iv, sp + 16, 24);
            int m = -1;
            for (int i = 0; i < s->num; i++)
                if (s->targets[i].code == iv[0]) { m = i; break; }
            if (m < 0) {
                if (s->num == cap) {
                    cap = cap ? 2 * cap : 32;
                    s->targets = realloc(s->targets, cap * sizeof(orc_target));
                }
                m = s->num++;
                orc_target *t = &s->targets[m];
                memset(t, 0, sizeof(*t));
                t->code = iv[0];
                t->cen = iv[1];
                t->beg = jul(beg);
                t->res = jul(end) - t->beg;
            }
            orc_target *t = &s->targets[m];
            t->one = realloc(t->one, (t->nseg + 1) * sizeof(int));
            t->two = realloc(t->two, (t->nseg + 1) * sizeof(int));
            t->one[t->nseg] = iv[4];
            t->two[t->nseg] = iv[5];
            t->nseg++;
            t->end = jul(end);
        }
        long nx = (long)next - 1;
        if (nx < 0) break;
        off = (size_t)nx * REC;
        if (off + REC > n) break;
    }
    return s;
}

static void parse_constants(orc_ephem *e, const char *path, char ***names, double **vals, int *count) {
    size_t n = 0;
    unsigned char *buf = read_file(path, &n);
    *names = NULL; *vals = NULL; *count = 0;
    if (!buf) return;
    int32_t fward;
    memcpy(&fward, buf + 76, 4);
    int nrec = fward - 2;
    if (nrec <= 0) { free(buf); return; }
    char *text = malloc((size_t)nrec * REC + 1);
    size_t tot = 0;
    /* reference src/spk.c:49-110: strip trailing NUL/EOT of each record, NUL -> newline */
    for (int i = 0; i < nrec; i++) {
        const unsigned char *r = buf + (size_t)(1 + i) * REC;
        long k = REC;
        while (k > 0 && (r[k - 1] == 0 || r[k - 1] == 4)) k--;
        memcpy(text + tot, r, (size_t)k);
        tot += (size_t)k;
    }
    text[tot] = 0;
    for (size_t i = 0; i < tot; i++) if (text[i] == 0) text[i] = '\n';
    free(buf);
    int in_sec = 0;
    char *save = NULL;
    for (char *line = strtok_r(text, "\n", &save); line; line = strtok_r(NULL, "\n", &save)) {
        if (strstr(line, "Initial conditions and constants used for integration:")) in_sec = 1;
        if (!in_sec) continue;
        for (char *p = line; *p; p++) if (*p == 'D' || *p == 'd') *p = 'e';
        char key[64], vs[64];
        if (sscanf(line, "%63s %63s", key, vs) != 2) continue;
        double v = strtod(vs, NULL);
        if (!strcmp(key, "cau") || !strcmp(key, "AU")) e->AU = v;
        else if (!strcmp(key, "EMRAT")) e->EMRAT = v;
        else if (!strcmp(key, "J2E")) e->J2E = v;
        else if (!strcmp(key, "J3E")) e->J3E = v;
        else if (!strcmp(key, "J4E")) e->J4E = v;
        else if (!strcmp(key, "J2SUN")) e->J2SUN = v;
        else if (!strcmp(key, "RE")) e->RE = v;
        else if (!strcmp(key, "CLIGHT")) e->CLIGHT = v;
        else if (!strcmp(key, "ASUN")) e->ASUN = v;
        else if (!strncmp(key, "GM", 2) || !strncmp(key, "MA", 2)) {
            *names = realloc(*names, (*count + 1) * sizeof(char *));
            *vals = realloc(*vals, (*count + 1) * sizeof(double));
            (*names)[*count] = strdup(key);
            (*vals)[*count] = v;
            (*count)++;
        }
    }
    free(text);
}

/* reference src/spk.c:137-210 */
static void join_masses(orc_spk *s, char **names, double *vals, int count, double emrat) {
    static const struct { const char *name; int code; } pc[] = {
        {"GMS", 10}, {"GM1", 1}, {"GM2", 2}, {"GMB", 399}, {"GMB", 3}, {"GMB", 301},
        {"GM4", 4}, {"GM5", 5}, {"GM6", 6}, {"GM7", 7}, {"GM8", 8}, {"GM9", 9}};
    if (!s) return;
    for (int m = 0; m < s->num; m++) {
        orc_target *t = &s->targets[m];
        if (t->mass != 0) continue;
        char label[64] = {0};
        for (size_t i = 0; i < sizeof(pc) / sizeof(pc[0]); i++)
            if (t->code == pc[i].code) { snprintf(label, sizeof label, "%s", pc[i].name); break; }
        if (!label[0]) snprintf(label, sizeof label, "MA%04d", t->code - 2000000);
        for (int i = 0; i < count; i++) {
            if (strcmp(names[i], label) == 0) {
                if (t->code == 399) t->mass = vals[i] * (emrat / (1. + emrat));
                else if (t->code == 301) t->mass = vals[i] * (1. / (1. + emrat));
                else t->mass = vals[i];
                break;
            }
        }
    }
}

static int find_target(const orc_spk *s, int code) {
    if (!s) return -1;
    for (int i = 0; i < s->num; i++) if (s->targets[i].code == code) return i;
    return -1;
}

static const int naif_by_assist[ORC_NPLANETS] = {10, 1, 2, 399, 301, 4, 5, 6, 7, 8, 9};

/* Planets from a JPL binary file: constants from its table, asteroid masses from its MAnnnn
 * constants (reference src/assist.c:291-327). */
static orc_ephem *ascii_ephem_load(orc_ephem *e, const char *planets_path, const char *asteroids_path, int *err) {
    e->ascii = orc_ascii_load(planets_path);
    if (!e->ascii) { *err = ORC_ERROR_EPHEM_FILE; free(e); return NULL; }
    if (asteroids_path) e->asteroids = spk_load(asteroids_path);
    for (int k = 0; k < ORC_NPLANETS; k++) e->target_index[k] = -1;
    e->emb_index = -1;
    e->J2E = orc_ascii_constant(e->ascii, 0);
    e->J3E = orc_ascii_constant(e->ascii, 1);
    e->J4E = orc_ascii_constant(e->ascii, 2);
    e->J2SUN = orc_ascii_constant(e->ascii, 3);
    e->AU = orc_ascii_constant(e->ascii, 4);
    e->RE = orc_ascii_constant(e->ascii, 5);
    e->CLIGHT = orc_ascii_constant(e->ascii, 6);
    e->ASUN = orc_ascii_constant(e->ascii, 7);
    e->EMRAT = orc_ascii_emrat(e->ascii);
    e->Re_eq = e->RE / e->AU;
    e->Rs_eq = e->ASUN / e->AU;
    e->c_AU_per_day = (e->CLIGHT / e->AU) * 86400.0;
    e->c_squared = e->c_AU_per_day * e->c_AU_per_day;
    e->over_c_squared = 1.0 / e->c_squared;
    if (e->asteroids) {
        for (int m = 0; m < e->asteroids->num; m++) {
            orc_target *t = &e->asteroids->targets[m];
            if (t->code >= 2000000) {
                char key[16]; /* only the first 6 characters take part in the match */
                snprintf(key, sizeof key, "MA%04d", t->code - 2000000);
                double gm = 0.0;
                if (orc_ascii_find_constant(e->ascii, key, &gm)) t->mass = gm;
            }
        }
    }
    return e;
}

orc_ephem *orc_ephem_load(const char *planets_path, const char *asteroids_path, int *err) {
    orc_generation++;
    orc_ephem *e = calloc(1, sizeof(orc_ephem));
    e->jd_ref = 2451545.0;
    *err = ORC_SUCCESS;
    {
        /* format detection, reference src/assist.c:108-129 */
        unsigned char head[0x00FC + 18];
        FILE *f = fopen(planets_path, "rb");
        size_t got = f ? fread(head, 1, sizeof head, f) : 0;
        if (f) fclose(f);
        if (!(got >= 8 && strncmp((const char *)head, "DAF/SPK ", 8) == 0) && orc_ascii_signature(head, got))
            return ascii_ephem_load(e, planets_path, asteroids_path, err);
    }
    e->planets = spk_load(planets_path);
    if (!e->planets) { *err = ORC_ERROR_EPHEM_FILE; free(e); return NULL; }
    if (asteroids_path) e->asteroids = spk_load(asteroids_path);
    for (int k = 0; k < ORC_NPLANETS; k++) e->target_index[k] = find_target(e->planets, naif_by_assist[k]);
    e->emb_index = find_target(e->planets, 3);
    char **names; double *vals; int count;
    parse_constants(e, planets_path, &names, &vals, &count);
    /* reference src/spk.c:769-789 */
    e->Re_eq = e->RE / e->AU;
    e->Rs_eq = e->ASUN / e->AU;
    e->c_AU_per_day = (e->CLIGHT / e->AU) * 86400;
    e->c_squared = e->c_AU_per_day * e->c_AU_per_day;
    e->over_c_squared = 1.0 / e->c_squared;
    join_masses(e->planets, names, vals, count, e->EMRAT);
    join_masses(e->asteroids, names, vals, count, e->EMRAT);
    for (int i = 0; i < count; i++) free(names[i]);
    free(names); free(vals);
    return e;
}

void orc_ephem_free(orc_ephem *e) {
    if (!e) return;
    orc_generation++;
    spk_free(e->planets);
    spk_free(e->asteroids);
    orc_ascii_free(e->ascii);
    free(e);
}

int orc_ephem_num_asteroids(const orc_ephem *e) { return e->asteroids ? e->asteroids->num : 0; }

/* reference src/assist.c:524-554 */
void orc_time_bounds(const orc_ephem *e, double *t_beg, double *t_end) {
    double beg = -INFINITY, end = INFINITY;
    if (e->ascii) orc_ascii_bounds(e->ascii, &beg, &end);
    const orc_spk *sets[2] = {e->planets, e->asteroids};
    for (int f = 0; f < 2; f++) {
        if (!sets[f]) continue;
        for (int i = 0; i < sets[f]->num; i++) {
            if (sets[f]->targets[i].beg > beg) beg = sets[f]->targets[i].beg;
            if (sets[f]->targets[i].end < end) end = sets[f]->targets[i].end;
        }
    }
    if (t_beg) *t_beg = beg - e->jd_ref;
    if (t_end) *t_end = end - e->jd_ref;
}

/* Locate the type-2 record for target t at (jd_ref, jd_rel); reference src/spk.c:430-445,500-514. */
static const double *locate(const orc_spk *s, const orc_target *t, double jd_ref, double jd_rel,
                            int *P, double *z) {
    int n = (int)((jd_ref + jd_rel - t->beg) / t->res);
    if (n >= t->nseg) n = t->nseg - 1; /* reference reads past its index array at t == end */
    const double *val = s->words + t->two[n] - 1;
    int R = (int)val[-1];
    *P = (R - 2) / 3;
    int b = (int)(((jd_ref - jul(val[-3])) + jd_rel) / (val[-2] / 86400.0));
    int N = (int)val[0];
    if (b >= N) b = N - 1; /* reference: no clamp (reads the directory at t == end) */
    val = s->words + (t->one[n] - 1) + (long)b * R;
    *z = ((jd_ref - jul(val[0])) + jd_rel) / (val[1] / 86400.0);
    return val;
}

/* Position, velocity, acceleration series; reference src/spk.c:492-547 (km, km/s, km/s^2). */
static void target_pva(const orc_spk *s, const orc_target *t, double jd_ref, double jd_rel,
                       double u[3], double v[3], double w[3]) {
    int P;
    double z;
    const double *val = locate(s, t, jd_ref, jd_rel, &P, &z);
    double T[32], S[32], U[32];
    double c = 1.0 / val[1];
    T[0] = 1.0; T[1] = z;
    S[0] = 0.0; S[1] = 1.0;
    U[0] = 0.0; U[1] = 0.0; U[2] = 4.0;
    for (int p = 2; p < P; p++) {
        T[p] = 2.0 * z * T[p - 1] - T[p - 2];
        S[p] = 2.0 * z * S[p - 1] + 2.0 * T[p - 1] - S[p - 2];
    }
    for (int p = 3; p < P; p++) U[p] = 2.0 * z * U[p - 1] + 4.0 * S[p - 1] - U[p - 2];
    for (int n = 0; n < 3; n++) {
        int b = 2 + n * P;
        u[n] = v[n] = w[n] = 0.0;
        for (int p = 0; p < P; p++) {
            double coeff = val[b + p];
            u[n] += coeff * T[p];
            v[n] += coeff * S[p] * c;
            w[n] += coeff * U[p] * c * c;
        }
    }
}

/* Asteroid (heliocentric) position in au; reference src/spk.c:405-481. */
static int asteroid_pos(const orc_spk *s, double jd_ref, double jd_rel, int m, double *GM, double x[3]) {
    if (!s) return ORC_ERROR_AST_FILE;
    if (m < 0 || m >= s->num) return ORC_ERROR_NAST;
    const orc_target *t = &s->targets[m];
    if (jd_ref + jd_rel < t->beg || jd_ref + jd_rel > t->end) return ORC_ERROR_COVERAGE;
    *GM = t->mass;
    int P;
    double z;
    const double *val = locate(s, t, jd_ref, jd_rel, &P, &z);
    double T[32];
    T[0] = 1.0; T[1] = z;
    for (int p = 2; p < P; p++) T[p] = 2.0 * z * T[p - 1] - T[p - 2];
    for (int n = 0; n < 3; n++) {
        int b = 2 + n * P;
        double u = 0.0;
        for (int p = 0; p < P; p++) u += val[b + p] * T[p];
        u /= 149597870.7;
        x[n] = u;
    }
    return ORC_SUCCESS;
}

/* reference src/spk.c:550-610 and :632-693 */
static int planet_calc(const orc_ephem *e, double jd_ref, double t, int body, double *GM, double o[9]) {
    if (body < 0 || body >= ORC_NPLANETS) return ORC_ERROR_NEPHEM;
    if (e->ascii) return orc_ascii_calc(e->ascii, jd_ref, t, body, GM, o);
    const orc_spk *pl = e->planets;
    int code = naif_by_assist[body];
    int idx = e->target_index[body];
    if (idx >= 0) {
        const orc_target *tg = &pl->targets[idx];
        if (jd_ref + t < tg->beg || jd_ref + t > tg->end) return ORC_ERROR_COVERAGE;
        *GM = tg->mass;
        double u[3], v[3], w[3];
        target_pva(pl, tg, jd_ref, t, u, v, w);
        if (code == 301 || code == 399) {
            if (e->emb_index < 0) return ORC_ERROR_NEPHEM;
            double eu[3], ev[3], ew[3];
            target_pva(pl, &pl->targets[e->emb_index], jd_ref, t, eu, ev, ew);
            for (int i = 0; i < 3; i++) { u[i] += eu[i]; v[i] += ev[i]; w[i] += ew[i]; }
        }
        const double au = e->AU, spd = 86400.;
        for (int i = 0; i < 3; i++) {
            o[i] = u[i] / au;
            o[3 + i] = v[i] / (au / spd);
            o[6 + i] = w[i] / (au / (spd * spd));
        }
        return ORC_SUCCESS;
    }
    if (body == 3 && e->emb_index >= 0 && e->target_index[4] >= 0) {
        /* Earth from EMB and Moon (reference src/spk.c:668-691) */
        const orc_target *emb = &pl->targets[e->emb_index], *moon = &pl->targets[e->target_index[4]];
        if (jd_ref + t < emb->beg || jd_ref + t > emb->end) return ORC_ERROR_COVERAGE;
        double eu[3], ev[3], ew[3], mu[3], mv[3], mw[3];
        target_pva(pl, emb, jd_ref, t, eu, ev, ew);
        target_pva(pl, moon, jd_ref, t, mu, mv, mw);
        double frac = 1.0 / (1.0 + e->EMRAT);
        const double au = e->AU, spd = 86400.;
        for (int i = 0; i < 3; i++) {
            o[i] = (-frac * mu[i] + eu[i]) / au;
            o[3 + i] = (-frac * mv[i] + ev[i]) / (au / spd);
            o[6 + i] = (-frac * mw[i] + ew[i]) / (au / (spd * spd));
        }
        *GM = 0.0;
        return ORC_SUCCESS;
    }
    return ORC_ERROR_NEPHEM;
}

/* reference src/forces.c:175-263 (without the transparent 7-slot cache) */
static int all_ephem_eval(const orc_ephem *e, int i, double t, double out[10]) {
    const double jd_ref = e->jd_ref;
    if (i < ORC_NPLANETS) {
        int f = planet_calc(e, jd_ref, t, i, &out[0], out + 1);
        if (f != ORC_SUCCESS) return f;
    } else {
        double x[3];
        int f = asteroid_pos(e->asteroids, jd_ref, t, i - ORC_NPLANETS, &out[0], x);
        if (f != ORC_SUCCESS) return f;
        double s[10];
        f = orc_all_ephem(e, 0, t, s);
        if (f != ORC_SUCCESS) return f;
        out[1] = x[0] + s[1];
        out[2] = x[1] + s[2];
        out[3] = x[2] + s[3];
        for (int k = 4; k < 10; k++) out[k] = NAN;
    }
    return ORC_SUCCESS;
}

/* Ephemeris cache, as reference src/forces.c:175-263: up to 7 results per body keyed by the
 * exact time, which IAS15's repeated substep times hit; a miss fills an empty entry or replaces
 * the one whose time is oldest in the direction of integration.  An entry holds exactly what the
 * evaluation returned, so results are unchanged.  The cache is per thread and is emptied when
 * any ephemeris is loaded or freed. */
#define ORC_CACHE_BODIES 64
typedef struct {
    long gen;
    const orc_ephem *e;
    int used[ORC_CACHE_BODIES];
    double t[ORC_CACHE_BODIES][7];
    double v[ORC_CACHE_BODIES][7][10];
} orc_cache_t;
static __thread orc_cache_t orc_tc;
static __thread int orc_tc_dir = 1;

void orc_cache_direction(int dir) { orc_tc_dir = dir >= 0 ? 1 : -1; }

int orc_all_ephem(const orc_ephem *e, int i, double t, double out[10]) {
    if (i < 0 || i >= ORC_CACHE_BODIES) return all_ephem_eval(e, i, t, out);
    orc_cache_t *c = &orc_tc;
    if (c->gen != orc_generation || c->e != e) {
        c->gen = orc_generation;
        c->e = e;
        memset(c->used, 0, sizeof c->used);
    }
    for (int k = 0; k < c->used[i]; k++)
        if (c->t[i][k] == t) {
            memcpy(out, c->v[i][k], 10 * sizeof(double));
            return ORC_SUCCESS;
        }
    const int f = all_ephem_eval(e, i, t, out);
    if (f != ORC_SUCCESS) return f;
    int slot = c->used[i];
    if (slot < 7) {
        c->used[i]++;
    } else {
        slot = 0;
        for (int k = 1; k < 7; k++)
            if ((orc_tc_dir > 0) ? (c->t[i][k] < c->t[i][slot]) : (c->t[i][k] > c->t[i][slot])) slot = k;
    }
    c->t[i][slot] = t;
    memcpy(c->v[i][slot], out, 10 * sizeof(double));
    return ORC_SUCCESS;
}

/* Raw target evaluation for tests: which_file 0 planets, 1 asteroids. out: u[3] v[3] w[3] in km units. */
int orc_spk_eval_target(const orc_ephem *e, int which_file, int target_idx, double t, double out[9]) {
    const orc_spk *s = which_file ? e->asteroids : e->planets;
    if (!s || target_idx < 0 || target_idx >= s->num) return ORC_ERROR_NEPHEM;
    const orc_target *tg = &s->targets[target_idx];
    if (e->jd_ref + t < tg->beg || e->jd_ref + t > tg->end) return ORC_ERROR_COVERAGE;
    target_pva(s, tg, e->jd_ref, t, out, out + 3, out + 6);
    return ORC_SUCCESS;
}
