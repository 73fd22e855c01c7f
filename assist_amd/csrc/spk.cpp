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

namespace assist_host {

static constexpr int kRecord = 1024;

static inline double jul(double eph) { return 2451545.0 + eph / 86400.0; }

spk_s* spk_init(const char* path) {
    int fd = open(path, O_RDONLY);
    if (fd < 0) return nullptr;
    unsigned char rec[kRecord];
    if (pread(fd, rec, kRecord, 0) != kRecord || strncmp((const char*)rec, "DAF/SPK", 7) != 0) {
        close(fd);
        return nullptr;
    }
    int32_t nd, ni, fward;
    memcpy(&nd, rec + 8, 4);
    memcpy(&ni, rec + 12, 4);
    memcpy(&fward, rec + 76, 4);
    if (8 * (nd + (ni + 1) / 2) != 40 || fward < 2) { close(fd); return nullptr; }
    off_t off = (off_t)(fward - 1) * kRecord;
    if (pread(fd, rec, kRecord, off) != kRecord || rec[8] != 0) { close(fd); return nullptr; }
    spk_s* s = new spk_s();
    for (int guard = 0; guard < 1000000; guard++) {
        double next, nsum;
        memcpy(&next, rec, 8);
        memcpy(&nsum, rec + 16, 8);
        const int ns = (int)nsum;
        for (int b = 0; b < ns && b < 25; b++) {
            const unsigned char* sp = rec + 24 + 40 * b;
            double beg, end;
            int32_t iv[6];
            memcpy(&beg, sp, 8);
            memcpy(&end, sp + 8, 8);
            memcpy(iv, sp + 16, 24);
            int m = -1;
            for (int i = 0; i < (int)s->targets.size(); i++)
                if (s->targets[i].code == iv[0]) { m = i; break; }
            if (m < 0) {
                spk_target t;
                t.code = iv[0];
                t.cen = iv[1];
                t.mass = 0;
                t.beg = jul(beg);
                t.res = jul(end) - t.beg;
                s->targets.push_back(t);
                m = (int)s->targets.size() - 1;
            }
            s->targets[m].one.push_back(iv[4]);
            s->targets[m].two.push_back(iv[5]);
            s->targets[m].end = jul(end);
        }
        long nx = (long)next - 1;
        if (nx < 0) break;
        if (pread(fd, rec, kRecord, (off_t)nx * kRecord) != kRecord) break;
    }
    s->num = (int)s->targets.size();
    struct stat sb;
    if (fstat(fd, &sb) < 0) { close(fd); delete s; return nullptr; }
    s->len = (size_t)sb.st_size;
    void* map = mmap(nullptr, s->len, PROT_READ, MAP_SHARED, fd, 0);
    close(fd);
    if (map == MAP_FAILED) { delete s; return nullptr; }
    s->map = map;
    // every word address the evaluator can touch must lie inside the file
    const size_t nwords = s->len / 8;
    for (auto& t : s->targets) {
        for (size_t k = 0; k < t.one.size(); k++) {
            if (t.one[k] < 1 || t.two[k] < 4 || (size_t)t.two[k] > nwords || t.one[k] > t.two[k]) {
                munmap(map, s->len);
                delete s;
                return nullptr;
            }
        }
    }
    return s;
}

void spk_free(spk_s* s) {
    if (!s) return;
    if (s->gpu) assist_gpu::free_device_ephem(s->gpu);
    if (s->map) munmap(s->map, s->len);
    delete s;
}

int find_target(const spk_s* s, int code) {
    if (!s) return -1;
    for (int i = 0; i < s->num; i++)
        if (s->targets[i].code == code) return i;
    return -1;
}

spk_constants load_constants(const char* path) {
    spk_constants c;
    int fd = open(path, O_RDONLY);
    if (fd < 0) return c;
    unsigned char rec[kRecord];
    if (pread(fd, rec, kRecord, 0) != kRecord || strncmp((const char*)rec, "DAF/SPK", 7) != 0) { close(fd); return c; }
    int32_t fward;
    memcpy(&fward, rec + 76, 4);
    std::string text;
    // comment records 2 .. fward-1; trailing NUL/EOT bytes of each record are dropped and NUL
    // ends a line (reference src/spk.c:49-110)
    for (int i = 0; i < fward - 2; i++) {
        if (pread(fd, rec, kRecord, (off_t)(1 + i) * kRecord) != kRecord) break;
        int k = kRecord;
        while (k > 0 && (rec[k - 1] == 0 || rec[k - 1] == 4)) k--;
        text.append((const char*)rec, (size_t)k);
    }
    close(fd);
    for (auto& ch : text)
        if (ch == 0) ch = '\n';
    bool in_sec = false;
    size_t pos = 0;
    while (pos < text.size()) {
        size_t nl = text.find('\n', pos);
        if (nl == std::string::npos) nl = text.size();
        std::string line = text.substr(pos, nl - pos);
        pos = nl + 1;
        if (line.empty()) continue;
        if (line.find("Initial conditions and constants used for integration:") != std::string::npos) in_sec = true;
        if (!in_sec) continue;
        for (auto& ch : line)
            if (ch == 'D' || ch == 'd') ch = 'e';
        char key[64], vs[64];
        if (sscanf(line.c_str(), "%63s %63s", key, vs) != 2) continue;
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
            c.names.push_back(key);
            c.values.push_back(v);
        }
    }
    return c;
}

void join_masses(spk_s* s, const spk_constants& c, double emrat) {
    static const struct { const char* name; int code; } planet_codes[] = {
        {"GMS", 10}, {"GM1", 1}, {"GM2", 2}, {"GMB", 399}, {"GMB", 3}, {"GMB", 301},
        {"GM4", 4}, {"GM5", 5}, {"GM6", 6}, {"GM7", 7}, {"GM8", 8}, {"GM9", 9}};
    if (!s) return;
    for (auto& t : s->targets) {
        if (t.mass != 0) continue;
        char label[64] = {0};
        for (auto& pc : planet_codes)
            if (t.code == pc.code) { snprintf(label, sizeof label, "%s", pc.name); break; }
        if (!label[0]) snprintf(label, sizeof label, "MA%04d", t.code - 2000000);
        for (size_t i = 0; i < c.names.size(); i++) {
            if (c.names[i] == label) {
                if (t.code == 399) t.mass = c.values[i] * (emrat / (1. + emrat));
                else if (t.code == 301) t.mass = c.values[i] * (1. / (1. + emrat));
                else t.mass = c.values[i];
                break;
            }
        }
        if (t.mass == 0 && t.code != 199 && t.code != 299) printf("Mass not found for target code: %d\n", t.code);
    }
}

// Record of target t for time jd_rel (reference src/spk.c:430-445, 500-514).  Segment and
// record indices are clamped to the data actually present.
static const double* locate(const spk_s* s, const spk_target& t, double jd_ref, double jd_rel, int* P, double* z) {
    int n = (int)((jd_ref + jd_rel - t.beg) / t.res);
    const int nseg = (int)t.two.size();
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

int planets_calc(const assist_ephem* e, double jd_ref, double jd_rel, int body, double* GM, double* x, double* y,
                 double* z, double* vx, double* vy, double* vz, double* ax, double* ay, double* az) {
    if (body < 0 || body >= ASSIST_BODY_NPLANETS) return ASSIST_ERROR_NEPHEM;
    const spk_s* pl = e->spk_planets;
    if (!pl) return ASSIST_ERROR_NEPHEM;
    const int code = kNaifByAssist[body];
    int idx = e->spk_target_index[body];
    if (idx < 0 || idx >= pl->num || pl->targets[idx].code != code) idx = find_target(pl, code);
    double u[3], v[3], w[3];
    const double au = e->AU, spd = 86400.;
    if (idx >= 0) {
        const spk_target& t = pl->targets[idx];
        if (jd_ref + jd_rel < t.beg || jd_ref + jd_rel > t.end) return ASSIST_ERROR_COVERAGE;
        *GM = t.mass;
        target_pva(pl, t, jd_ref, jd_rel, u, v, w);
        if (code == 301 || code == 399) {
            int ei = e->spk_emb_index >= 0 ? e->spk_emb_index : find_target(pl, 3);
            if (ei < 0) return ASSIST_ERROR_NEPHEM;
            double eu[3], ev[3], ew[3];
            target_pva(pl, pl->targets[ei], jd_ref, jd_rel, eu, ev, ew);
            for (int i = 0; i < 3; i++) { u[i] += eu[i]; v[i] += ev[i]; w[i] += ew[i]; }
        }
    } else if (body == ASSIST_BODY_EARTH && e->spk_emb_index >= 0 && e->spk_target_index[ASSIST_BODY_MOON] >= 0) {
        // Earth from EMB and Moon when 399 is absent (reference src/spk.c:668-691)
        const spk_target& emb = pl->targets[e->spk_emb_index];
        if (jd_ref + jd_rel < emb.beg || jd_ref + jd_rel > emb.end) return ASSIST_ERROR_COVERAGE;
        double eu[3], ev[3], ew[3], mu[3], mv[3], mw[3];
        target_pva(pl, emb, jd_ref, jd_rel, eu, ev, ew);
        target_pva(pl, pl->targets[e->spk_target_index[ASSIST_BODY_MOON]], jd_ref, jd_rel, mu, mv, mw);
        const double frac = 1.0 / (1.0 + e->EMRAT);
        for (int i = 0; i < 3; i++) {
            u[i] = -frac * mu[i] + eu[i];
            v[i] = -frac * mv[i] + ev[i];
            w[i] = -frac * mw[i] + ew[i];
        }
        *GM = 0.0;
    } else {
        return ASSIST_ERROR_NEPHEM;
    }
    *x = u[0] / au; *y = u[1] / au; *z = u[2] / au;
    *vx = v[0] / (au / spd); *vy = v[1] / (au / spd); *vz = v[2] / (au / spd);
    *ax = w[0] / (au / (spd * spd)); *ay = w[1] / (au / (spd * spd)); *az = w[2] / (au / (spd * spd));
    return ASSIST_SUCCESS;
}

}  // namespace assist_host
