// Batched GPU entry points of libassist, the device ephemeris, and the
// simulation hooks (force evaluation and integration) built on them.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/ascii_ephem.h"
#include "device_types.h"
#include "internal.h"
#include "kernels.h"

using namespace assist_gpu;

namespace assist_gpu {

struct DeviceEphemCache {
    int device = -1;
    double jd_ref = 0;
    const spk_s* asteroids = nullptr;  // the asteroid file the tables were built with
    DevEphem host;                     // host copy with device pointers
    DevEphem* d_ephem = nullptr;
    double* d_records = nullptr;
};

void free_device_ephem(DeviceEphemCache* c) {
    if (!c) return;
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && c->device >= 0) {
        (void)hipSetDevice(c->device);
        if (c->d_ephem) (void)hipFree(c->d_ephem);
        if (c->d_records) (void)hipFree(c->d_records);
        (void)hipSetDevice(cur);
    }
    delete c;
}

}  // namespace assist_gpu

namespace {

std::mutex g_mutex;

inline double jul(double eph) { return 2451545.0 + eph / 86400.0; }

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    fprintf(stderr, "(ASSIST) HIP error in %s: %s\n", what, hipGetErrorString(e));
    return false;
}

// Flattens one SPK target into records [jd_ref-jul(MID), 86400/RADIUS, then the coefficients
// interleaved by degree: x_0, y_0, z_0, x_1, ...], each coordinate zero-padded to kCoefPad.
// Returns false when the target's segments are not contiguous with equal record length.
bool flatten_target(const spk_s* s, const spk_target& t, double jd_ref, std::vector<double>& out, DevSlot& slot,
                    size_t& offset_words) {
    const double* base = (const double*)s->map;
    double intlen = 0, init0 = 0;
    int rsize = 0, nrec_total = 0;
    for (int k = 0; k < assist_host::segments(t); k++) {
        const double* dir = base + t.two[k] - 1;  // -> N
        const double init = dir[-3], len = dir[-2];
        const int R = (int)dir[-1], N = (int)dir[0];
        if (k == 0) { intlen = len; init0 = init; rsize = R; }
        if (len != intlen || R != rsize || N <= 0) return false;
        if (fabs(init - (init0 + (double)nrec_total * intlen)) > 1e-6 * intlen) return false;
        const long need = (long)t.one[k] - 1 + (long)N * R;
        if (need > (long)t.two[k]) return false;
        nrec_total += N;
    }
    const int P = (rsize - 2) / 3;
    if (P < 2 || P > kCoefPad) return false;
    offset_words = out.size();
    out.resize(out.size() + (size_t)nrec_total * kRecWords, 0.0);
    double* dst = out.data() + offset_words;
    for (int k = 0; k < assist_host::segments(t); k++) {
        const double* dir = base + t.two[k] - 1;
        const int N = (int)dir[0];
        const double* src = base + t.one[k] - 1;
        for (int r = 0; r < N; r++) {
            const double* rec = src + (size_t)r * rsize;
            dst[0] = jd_ref - jul(rec[0]);
            dst[1] = 1.0 / (rec[1] / 86400.0);
            for (int c = 0; c < 3; c++)
                for (int p = 0; p < P; p++) dst[2 + 3 * p + c] = rec[2 + c * P + p];
            dst += kRecWords;
        }
    }
    slot.rec = nullptr;
    slot.c_init = jd_ref - jul(init0);
    slot.inv_intlen = 1.0 / (intlen / 86400.0);
    slot.nrec = nrec_total;
    slot.P = P;
    return true;
}

// Re-lays one column of a JPL binary planets file as device records, one per sub-interval of
// `inc / niv` days: [jd_ref - JD(mid), 1/radius (1/day), then the coefficients * scale
// interleaved by degree as above], each coordinate zero-padded to kCoefPad.  The evaluation is then the SPK one: the record index
// (jd_ref - beg + t) niv / inc and z = (jd_ref - mid + t) / radius are the block and sub-interval
// selection of reference src/ascii_ephem.c:315-318 and :34-39, written over the whole file.
// `scale` folds a constant factor into the coefficients (the Earth's and the Moon's offsets from
// the EMB are -1/(1+EMRAT) and EMRAT/(1+EMRAT) times the geocentric Moon).
bool flatten_ascii_column(const ascii_s* a, int col, double scale, double jd_ref, std::vector<double>& out,
                          DevSlot& slot, size_t& offset_words) {
    const int ncf = a->ncf[col], niv = a->niv[col];
    if (a->ncm[col] != 3 || ncf < 2 || ncf > kCoefPad || niv < 1) return false;
    const size_t nblk = (size_t)ceil((a->end - a->beg) / a->inc);
    const size_t blk_words = a->rec / sizeof(double);
    if (nblk < 1 || (nblk + 2) * a->rec > a->len) return false;
    const double sub = a->inc / niv;
    const double* base = (const double*)a->map;
    offset_words = out.size();
    out.resize(out.size() + nblk * niv * kRecWords, 0.0);
    double* dst = out.data() + offset_words;
    for (size_t blk = 0; blk < nblk; blk++) {
        const double* block = base + (blk + 2) * blk_words + a->off[col];
        for (int k = 0; k < niv; k++) {
            const double* cf = block + (size_t)ncf * 3 * k;
            dst[0] = jd_ref - (a->beg + (double)blk * a->inc + (k + 0.5) * sub);
            dst[1] = 2.0 / sub;
            for (int m = 0; m < 3; m++)
                for (int p = 0; p < ncf; p++) dst[2 + 3 * p + m] = cf[m * ncf + p] * scale;
            dst += kRecWords;
        }
    }
    slot.rec = nullptr;
    slot.c_init = jd_ref - a->beg;
    slot.inv_intlen = 1.0 / sub;
    slot.nrec = (int)(nblk * niv);
    slot.P = ncf;
    return true;
}

// The device copy of the ephemeris on the current device if one is resident (nothing is built
// or allocated): ASSIST_BATCH_ERROR_UNSUPPORTED otherwise.
int lookup_device_ephem(const assist_ephem* e, const DevEphem** out, DevEphem* host_copy) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!e || (!e->spk_planets && !e->ascii_planets)) return ASSIST_ERROR_EPHEM_FILE;
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return ASSIST_BATCH_ERROR_DEVICE;
    const DeviceEphemCache* c =
        (const DeviceEphemCache*)(e->spk_planets ? e->spk_planets->device_cache : e->ascii_planets->device_cache);
    if (c && c->device == dev && c->jd_ref == e->jd_ref && c->asteroids == e->spk_asteroids) {
        *out = c->d_ephem;
        *host_copy = c->host;
        return ASSIST_SUCCESS;
    }
    return ASSIST_BATCH_ERROR_UNSUPPORTED;
}

// Builds (or reuses) the device copy of the ephemeris on the current device.
int get_device_ephem(const assist_ephem* e, const DevEphem** out, DevEphem* host_copy) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!e || (!e->spk_planets && !e->ascii_planets)) return ASSIST_ERROR_EPHEM_FILE;
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return ASSIST_BATCH_ERROR_DEVICE;
    spk_s* pl = e->spk_planets;
    ascii_s* jpl = pl ? nullptr : e->ascii_planets;
    DeviceEphemCache* c = (DeviceEphemCache*)(pl ? pl->device_cache : jpl->device_cache);
    if (c && c->device == dev && c->jd_ref == e->jd_ref && c->asteroids == e->spk_asteroids) {
        *out = c->d_ephem;
        if (host_copy) *host_copy = c->host;
        return ASSIST_SUCCESS;
    }
    if (c) {
        free_device_ephem(c);
        if (pl) pl->device_cache = nullptr;
        else jpl->device_cache = nullptr;
    }
    static const int planet_codes[12] = {10, 1, 2, 3, 399, 301, 4, 5, 6, 7, 8, 9};
    // slot -> JPL binary column, and the factor folded into its coefficients
    const int jpl_col[12] = {ASCII_SUN, ASCII_MER, ASCII_VEN, ASCII_EMB, ASCII_LUN, ASCII_LUN,
                             ASCII_MAR, ASCII_JUP, ASCII_SAT, ASCII_URA, ASCII_NEP, ASCII_PLU};
    const double cem = jpl ? jpl->cem : 0.0;
    const double jpl_scale[12] = {1, 1, 1, 1, -1.0 / (1.0 + cem), cem / (1.0 + cem), 1, 1, 1, 1, 1, 1};
    DevEphem H;
    memset(&H, 0, sizeof H);
    std::vector<double> words;
    size_t offs[kNumSlots];
    const int n_ast = e->spk_asteroids ? e->spk_asteroids->num : 0;
    if (n_ast > 16) return ASSIST_BATCH_ERROR_UNSUPPORTED;
    for (int s = 0; s < kNumSlots; s++) {
        if (s < kFirstAsteroidSlot && jpl) {
            if (!flatten_ascii_column(jpl, jpl_col[s], jpl_scale[s], e->jd_ref, words, H.slot[s], offs[s]))
                return ASSIST_BATCH_ERROR_UNSUPPORTED;
            H.slot[s].kind = 0;
            continue;
        }
        const spk_s* file = s < kFirstAsteroidSlot ? pl : e->spk_asteroids;
        int idx;
        if (s < kFirstAsteroidSlot) idx = assist_host::find_target(pl, planet_codes[s]);
        else idx = (s - kFirstAsteroidSlot) < n_ast ? s - kFirstAsteroidSlot : -1;
        if (idx < 0) {
            if (s < kFirstAsteroidSlot) return ASSIST_ERROR_NEPHEM;  // Earth/Moon/EMB fallbacks not on the GPU path
            // missing asteroid slot: point it at the Sun's data (evaluated but never used)
            H.slot[s] = H.slot[0];
            offs[s] = offs[0];
            H.slot[s].kind = 1;
            continue;
        }
        if (!flatten_target(file, file->targets[idx], e->jd_ref, words, H.slot[s], offs[s])) return ASSIST_BATCH_ERROR_UNSUPPORTED;
        H.slot[s].kind = s < kFirstAsteroidSlot ? 0 : 1;
    }
    static const int body_slot[kNumPlanets] = {0, 1, 2, 4, 5, 6, 7, 8, 9, 10, 11};
    for (int b = 0; b < kNumBodies; b++) {
        if (b < kNumPlanets) {
            if (jpl) {
                H.GM[b] = jpl->mass[b];
            } else {
                const int idx = assist_host::find_target(pl, planet_codes[body_slot[b]]);
                H.GM[b] = idx >= 0 ? pl->targets[idx].mass : 0.0;
            }
        } else {
            H.GM[b] = (b - kNumPlanets) < n_ast ? e->spk_asteroids->targets[b - kNumPlanets].mass : 0.0;
        }
    }
    // planets: km -> au with the file's own au (reference src/spk.c:590, src/ascii_ephem.c:369)
    const double au_km = jpl ? jpl->cau : e->AU;
    H.inv_au = 1.0 / au_km;
    H.inv_au_ast = 1.0 / 149597870.7;
    H.vel_scale = 86400.0 / au_km;
    // coverage windows: every planet-file target, and every asteroid target
    {
        double pb = -INFINITY, pe = INFINITY, ab = -INFINITY, ae = INFINITY;
        if (jpl) { pb = jpl->beg; pe = jpl->end; }
        else for (int i = 0; i < pl->num; i++) { pb = fmax(pb, pl->targets[i].beg); pe = fmin(pe, pl->targets[i].end); }
        for (int i = 0; i < n_ast; i++) {
            ab = fmax(ab, e->spk_asteroids->targets[i].beg);
            ae = fmin(ae, e->spk_asteroids->targets[i].end);
        }
        H.t_beg = pb - e->jd_ref;
        H.t_end = pe - e->jd_ref;
        H.ta_beg = n_ast > 0 ? ab - e->jd_ref : H.t_beg;
        H.ta_end = n_ast > 0 ? ae - e->jd_ref : H.t_end;
    }
    H.n_ast = n_ast;
    // first slot from which every slot can run the short series: its records have at most
    // kCoefShort coefficients, or it is a missing asteroid's (whose positions no term uses)
    H.short_from = kNumSlots;
    for (int s = kNumSlots - 1; s >= 0; s--) {
        const bool unused = s >= kFirstAsteroidSlot + n_ast;
        if (!unused && H.slot[s].P > kCoefShort) break;
        H.short_from = s;
    }
    H.J2E = e->J2E; H.J3E = e->J3E; H.J4E = e->J4E; H.J2SUN = e->J2SUN;
    H.Re_eq = e->Re_eq; H.Rs_eq = e->Rs_eq; H.c_squared = e->c_squared; H.over_c_squared = e->over_c_squared;
    const double d2r = M_PI / 180.;
    H.e_cosa = cos(0.0 * d2r); H.e_sina = sin(0.0 * d2r); H.e_cosd = cos(90.0 * d2r); H.e_sind = sin(90.0 * d2r);
    H.s_cosa = cos(286.13 * d2r); H.s_sina = sin(286.13 * d2r); H.s_cosd = cos(63.87 * d2r); H.s_sind = sin(63.87 * d2r);
    H.k_J2 = 3.0 * H.J2E * H.Re_eq * H.Re_eq / 2.0;
    H.k_J3 = 5.0 * H.J3E * H.Re_eq * H.Re_eq * H.Re_eq / 2.0;
    H.k_J4 = 5.0 * H.J4E * H.Re_eq * H.Re_eq * H.Re_eq * H.Re_eq / 8.0;
    H.k_sJ2 = 3.0 * H.J2SUN * H.Rs_eq * H.Rs_eq / 2.0;
    c = new DeviceEphemCache();
    c->device = dev;
    c->jd_ref = e->jd_ref;
    c->asteroids = e->spk_asteroids;
    if (!hip_ok(hipMalloc(&c->d_records, words.size() * sizeof(double)), "hipMalloc(ephemeris)") ||
        !hip_ok(hipMemcpy(c->d_records, words.data(), words.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy(ephemeris)")) {
        free_device_ephem(c);
        return ASSIST_BATCH_ERROR_DEVICE;
    }
    for (int s = 0; s < kNumSlots; s++) H.slot[s].rec = c->d_records + offs[s];
    if (!hip_ok(hipMalloc(&c->d_ephem, sizeof(DevEphem)), "hipMalloc(DevEphem)") ||
        !hip_ok(hipMemcpy(c->d_ephem, &H, sizeof(DevEphem), hipMemcpyHostToDevice), "hipMemcpy(DevEphem)")) {
        free_device_ephem(c);
        return ASSIST_BATCH_ERROR_DEVICE;
    }
    c->host = H;
    if (pl) pl->device_cache = c;
    else jpl->device_cache = c;
    *out = c->d_ephem;
    if (host_copy) *host_copy = H;
    return ASSIST_SUCCESS;
}

int make_force_opts(const assist_extras* x, ForceOpts* fo) {
    if (x) {
        // EIH sources beyond the 11 planet-file bodies would be asteroids, whose velocities the
        // reference does not have (NaN, reference src/forces.c:222-223)
        if ((x->forces & ASSIST_FORCE_GR_EIH) && (x->gr_eih_sources < 0 || x->gr_eih_sources > ASSIST_BODY_NPLANETS))
            return ASSIST_BATCH_ERROR_UNSUPPORTED;
        if (x->geocentric != 0 && x->geocentric != 1) return ASSIST_BATCH_ERROR_ARG;
        fo->forces = x->forces;
        fo->gr_eih_sources = x->gr_eih_sources;
        fo->alpha = x->alpha; fo->nk = x->nk; fo->nm = x->nm; fo->nn = x->nn; fo->r0 = x->r0;
        fo->geocentric = x->geocentric;
    } else {
        fo->forces = 0x7F;
        fo->gr_eih_sources = 1;
        fo->alpha = 1.0; fo->nk = 0.0; fo->nm = 2.0; fo->nn = 5.093; fo->r0 = 1.0;
        fo->geocentric = 0;
    }
    return ASSIST_SUCCESS;
}

// The kernel's predictor-corrector tolerance for a configuration (IntegratorOpts::pc_tol): the
// mixed form's FP32 terms carry round-off far above FP64's, so by default it stops once the real
// particle's b6 changes by less than epsilon / 100 of its acceleration -- a hundredth of the
// step's own truncation-error scale -- rather than at FP64 round-off over every vector.
double corrector_tolerance(const assist_batch_config& cfg) {
    if (cfg.corrector_tolerance > 0) return cfg.corrector_tolerance;
    if (cfg.corrector_tolerance < 0 || cfg.precision != ASSIST_PRECISION_MIXED || !(cfg.epsilon > 0)) return 0.0;
    return fmax(1e-16, 1e-2 * cfg.epsilon);
}

// The kernel's first-pass tolerance (IntegratorOpts::pc_tol1).  IAS15 predicts each step's b
// from the last one, so the first pass of an attempt usually changes b6 by about epsilon of the
// acceleration, and the next pass by a thousandth of that (CPU oracle, configuration-4 mix: the
// ratio of the second pass's change to the first's has median 6e-4, 95th percentile 9e-3).  By
// default the mixed form accepts the attempt after its first pass when that pass changed the
// real particle's b6 by less than 3 epsilon of its acceleration, tightened by pc_aref (below).
double first_pass_tolerance(const assist_batch_config& cfg) {
    if (cfg.first_pass_tolerance > 0) return cfg.first_pass_tolerance;
    if (cfg.first_pass_tolerance < 0 || cfg.precision != ASSIST_PRECISION_MIXED || !(cfg.epsilon > 0)) return 0.0;
    return 3.0 * cfg.epsilon;
}

// The first-pass test's acceleration scale (IntegratorOpts::pc_aref).  A single pass leaves a
// residual of the same sign step after step, which the orbit turns into an energy error; where
// the acceleration is strong (the perihelion passages of eccentric orbits, planetary
// approaches) that error weighs most, so there the tolerance shrinks in proportion to
// pc_aref / max|a|, down to a tenth (0.3 epsilon with the mixed form's default).  The default,
// 6e-5 au/day^2, is the Sun's pull at about 2.2 au: main-belt orbits keep the full tolerance.
// CPU oracle, 2 304 configuration-4 orbits over 10 years: 1.20x fewer evaluations than a flat
// 0.3 epsilon, positions within 15 cm of IAS15's own stop on every orbit whose IAS15 result
// moves by less than 1 m between epsilon = 1e-9 and 1e-11 (a flat 3 epsilon: 2.7 m on one
// eccentric NEO-like orbit); the floor keeps the longest high-eccentricity spans of the
// uneven-span workload from taking more evaluations than a flat 0.3 epsilon.
double first_pass_accel(const assist_batch_config& cfg) {
    if (cfg.first_pass_accel > 0) return cfg.first_pass_accel;
    if (cfg.first_pass_accel < 0) return 0.0;
    return 6e-5;
}

int lanes_for(int nvec, int requested) {
    int g = 1;
    while (g < nvec) g <<= 1;
    if (requested > 0) {
        if (requested != 1 && requested != 2 && requested != 4 && requested != 8) return -1;
        if (requested < g) return -1;
        return requested;
    }
    if (g > 8) return -1;
    // default: 8-lane groups for every vector count (they spread the 28 ephemeris targets over 8
    // lanes and keep the substep cache; 2- and 4-lane groups, which keep 7-14 slots per lane in
    // registers, were slower in every measured configuration); the autotuner may pick 1 lane for
    // large real-only batches
    return 8;
}

// Persistent workspace on the current device (grown as needed; one launch at a time).
// Captured launches of the integration (set_counter + ias15_kernel), keyed by everything the
// launch takes: the argument block (device pointers included) and the launch shape.
struct GraphEntry {
    IntegrateArgs args;
    int G, threads, waves, blocks;
    bool mixed;
    hipGraphExec_t exec;
};
constexpr size_t kMaxGraphs = 8;

struct Workspace {
    int device = -1;
    double* scratch = nullptr;
    size_t scratch_words = 0;
    int* queue = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_in = nullptr;
    hipStream_t graph_stream = nullptr;  // default priority, non-blocking
    std::vector<GraphEntry> graphs;      // most recently used last
    // queue order of uneven batches (keys, sorted keys, indices, order, sort temporaries)
    unsigned long long* okeys = nullptr;
    unsigned long long* okeys_alt = nullptr;
    int* ovals = nullptr;
    int* order = nullptr;
    void* otemp = nullptr;
    size_t order_n = 0, otemp_bytes = 0;
    // learned start: rows [n][8], features [n][16], network outputs [n][32]
    double* prows = nullptr;
    float* pX = nullptr;
    float* pY = nullptr;
    size_t pred_n = 0;
};
Workspace g_ws;

void clear_graphs() {
    for (auto& e : g_ws.graphs) (void)hipGraphExecDestroy(e.exec);
    g_ws.graphs.clear();
}

int ensure_workspace(size_t scratch_words) {
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return ASSIST_BATCH_ERROR_DEVICE;
    if (g_ws.device != dev) {
        g_ws.graphs.clear();  // (those of another device are not destroyed from this one)
        g_ws = Workspace();
        g_ws.device = dev;
        if (!hip_ok(hipMalloc(&g_ws.queue, sizeof(int)), "hipMalloc(queue)")) return ASSIST_BATCH_ERROR_DEVICE;
        if (!hip_ok(hipEventCreate(&g_ws.ev0), "hipEventCreate") || !hip_ok(hipEventCreate(&g_ws.ev1), "hipEventCreate") ||
            !hip_ok(hipEventCreateWithFlags(&g_ws.ev_in, hipEventDisableTiming), "hipEventCreate") ||
            !hip_ok(hipStreamCreateWithFlags(&g_ws.graph_stream, hipStreamNonBlocking), "hipStreamCreate"))
            return ASSIST_BATCH_ERROR_DEVICE;
    }
    if (scratch_words > g_ws.scratch_words) {
        clear_graphs();  // captured launches hold the old scratch pointer
        if (g_ws.scratch) (void)hipFree(g_ws.scratch);
        g_ws.scratch = nullptr;
        g_ws.scratch_words = 0;
        if (!hip_ok(hipMalloc(&g_ws.scratch, scratch_words * sizeof(double)), "hipMalloc(scratch)")) return ASSIST_BATCH_ERROR_DEVICE;
        g_ws.scratch_words = scratch_words;
    }
    return ASSIST_SUCCESS;
}

// Buffers of the span sort for n systems (the workspace of the current device must exist).
int ensure_order(int n) {
    if ((size_t)n <= g_ws.order_n) return ASSIST_SUCCESS;
    clear_graphs();  // captured launches hold the old order pointer
    (void)hipFree(g_ws.okeys); (void)hipFree(g_ws.okeys_alt); (void)hipFree(g_ws.ovals); (void)hipFree(g_ws.order);
    (void)hipFree(g_ws.otemp);
    g_ws.okeys = g_ws.okeys_alt = nullptr;
    g_ws.ovals = g_ws.order = nullptr;
    g_ws.otemp = nullptr;
    g_ws.order_n = g_ws.otemp_bytes = 0;
    const size_t tb = span_order_temp_bytes(n);
    if (!hip_ok(hipMalloc(&g_ws.okeys, (size_t)n * sizeof(unsigned long long)), "hipMalloc(order)") ||
        !hip_ok(hipMalloc(&g_ws.okeys_alt, (size_t)n * sizeof(unsigned long long)), "hipMalloc(order)") ||
        !hip_ok(hipMalloc(&g_ws.ovals, (size_t)n * sizeof(int)), "hipMalloc(order)") ||
        !hip_ok(hipMalloc(&g_ws.order, (size_t)n * sizeof(int)), "hipMalloc(order)") ||
        !hip_ok(hipMalloc(&g_ws.otemp, tb > 0 ? tb : 1), "hipMalloc(order)"))
        return ASSIST_BATCH_ERROR_DEVICE;
    g_ws.order_n = (size_t)n;
    g_ws.otemp_bytes = tb;
    return ASSIST_SUCCESS;
}

// Buffers of the learned start for n systems (the workspace of the current device must exist).
int ensure_predict(size_t n) {
    if (n <= g_ws.pred_n) return ASSIST_SUCCESS;
    (void)hipFree(g_ws.prows); (void)hipFree(g_ws.pX); (void)hipFree(g_ws.pY);
    g_ws.prows = nullptr;
    g_ws.pX = g_ws.pY = nullptr;
    g_ws.pred_n = 0;
    if (!hip_ok(hipMalloc(&g_ws.prows, n * 8 * sizeof(double)), "hipMalloc(predict)") ||
        !hip_ok(hipMalloc(&g_ws.pX, n * ASSIST_MLP_IN * sizeof(float)), "hipMalloc(predict)") ||
        !hip_ok(hipMalloc(&g_ws.pY, n * ASSIST_MLP_OUT * sizeof(float)), "hipMalloc(predict)"))
        return ASSIST_BATCH_ERROR_DEVICE;
    g_ws.pred_n = n;
    return ASSIST_SUCCESS;
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    bool alloc(size_t n) { return n == 0 || hip_ok(hipMalloc(&p, n * sizeof(T)), "hipMalloc"); }
};

}  // namespace

// ---------------------------------------------------------------------------------------------
// simulation hooks: every real particle with its variational particles is one system
// ---------------------------------------------------------------------------------------------
namespace {

struct SimLayout {
    int n_real = 0, nvmax = 0;
    std::vector<std::vector<int>> vars;  // per real particle: var-config indices
};

SimLayout layout_of(const reb_simulation* sim) {
    SimLayout L;
    L.n_real = sim->N - sim->N_var;
    L.vars.assign(L.n_real > 0 ? L.n_real : 0, {});
    for (int v = 0; v < sim->N_var_config; v++) {
        const int tp = sim->var_config[v].testparticle;
        if (tp >= 0 && tp < L.n_real) L.vars[tp].push_back(v);
    }
    for (auto& v : L.vars) L.nvmax = (int)v.size() > L.nvmax ? (int)v.size() : L.nvmax;
    return L;
}

void pack(const reb_simulation* sim, const assist_extras* ax, const SimLayout& L, std::vector<double>& states,
          std::vector<double>& params, bool& has_params) {
    const int nv = 1 + L.nvmax;
    states.assign((size_t)L.n_real * nv * 6, 0.0);
    has_params = ax && ax->particle_params && (ax->forces & ASSIST_FORCE_NON_GRAVITATIONAL);
    params.assign(has_params ? (size_t)L.n_real * nv * 3 : 0, 0.0);
    for (int j = 0; j < L.n_real; j++) {
        const reb_particle& p = sim->particles[j];
        double* s = &states[(size_t)j * nv * 6];
        s[0] = p.x; s[1] = p.y; s[2] = p.z; s[3] = p.vx; s[4] = p.vy; s[5] = p.vz;
        if (has_params) for (int c = 0; c < 3; c++) params[(size_t)j * nv * 3 + c] = ax->particle_params[3 * j + c];
        for (size_t k = 0; k < L.vars[j].size(); k++) {
            const int vc = L.vars[j][k];
            const reb_particle& q = sim->particles[sim->var_config[vc].index];
            double* d = s + 6 * (k + 1);
            d[0] = q.x; d[1] = q.y; d[2] = q.z; d[3] = q.vx; d[4] = q.vy; d[5] = q.vz;
            if (has_params)
                for (int c = 0; c < 3; c++) params[((size_t)j * nv + k + 1) * 3 + c] = ax->particle_params[3 * (L.n_real + vc) + c];
        }
    }
}

}  // namespace

// ---- launch of the integration kernel and its autotuner ----------------------------------
namespace {

// The workspace, the graph cache and the tuner's choices are process-wide: the integration entry
// points hold this lock from the shape choice to the kernel's completion, so that calls from
// several host threads run one after another instead of sharing one work queue and scratch.
std::mutex g_launch_mutex;

struct LaunchShape {
    int G = 8;         // lanes per system
    int threads = 256; // block size
    int waves = 1;     // kernel variant (register budget for 1 or 2 waves per SIMD)
};

constexpr int kTuneMinSystems = 16384;  // smaller batches take the default shape
// Systems per trial launch: many resident rounds even at one lane per system (65,536 lanes fill
// the GPU), so that every shape is timed at throughput rather than by its slowest systems; with
// a learned first step a system covers the trial span in a few steps, and 4 systems per lane
// made the one-lane kernel look slower than 8-lane groups while it is 1.5x faster in full.
constexpr int kTuneSample = 1048576;
constexpr double kTuneSpan = 30.0;      // days integrated per trial system

// Replays (capturing it first when new) the launch set_counter + ias15_kernel for `a` on the
// library's graph stream, after the work already queued on `s`; *ms is the graph's HIP-event time.
int run_ias15_graph(const IntegrateArgs& a, bool mixed, const LaunchShape& sh, int blocks, int total_groups,
                    hipStream_t s, float* ms) {
    GraphEntry* hit = nullptr;
    for (auto& e : g_ws.graphs) {
        if (e.G == sh.G && e.threads == sh.threads && e.waves == sh.waves && e.blocks == blocks && e.mixed == mixed &&
            memcmp(&e.args, &a, sizeof a) == 0) {
            hit = &e;
            break;
        }
    }
    hipStream_t gs = g_ws.graph_stream;
    if (!hip_ok(hipEventRecord(g_ws.ev_in, s), "hipEventRecord") || !hip_ok(hipStreamWaitEvent(gs, g_ws.ev_in, 0), "hipStreamWaitEvent"))
        return ASSIST_BATCH_ERROR_DEVICE;
    if (!hit) {
        if (g_ws.graphs.size() >= kMaxGraphs) {
            (void)hipGraphExecDestroy(g_ws.graphs.front().exec);
            g_ws.graphs.erase(g_ws.graphs.begin());
        }
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (!hip_ok(hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
            return ASSIST_BATCH_ERROR_DEVICE;
        const bool launched = hip_ok(launch_set_counter(g_ws.queue, total_groups, gs), "launch(set_counter)") &&
                              hip_ok(launch_ias15(a, sh.G, mixed, sh.waves, blocks, sh.threads, gs, false), "launch(ias15)");
        const bool captured = hip_ok(hipStreamEndCapture(gs, &graph), "hipStreamEndCapture");
        if (!launched || !captured || !hip_ok(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate")) {
            if (graph) (void)hipGraphDestroy(graph);
            return ASSIST_BATCH_ERROR_DEVICE;
        }
        (void)hipGraphDestroy(graph);
        GraphEntry e;
        memset(&e, 0, sizeof e);
        e.args = a;
        e.G = sh.G; e.threads = sh.threads; e.waves = sh.waves; e.blocks = blocks; e.mixed = mixed;
        e.exec = exec;
        g_ws.graphs.push_back(e);
        hit = &g_ws.graphs.back();
    }
    if (!hip_ok(hipEventRecord(g_ws.ev0, gs), "hipEventRecord") || !hip_ok(hipGraphLaunch(hit->exec, gs), "hipGraphLaunch") ||
        !hip_ok(hipEventRecord(g_ws.ev1, gs), "hipEventRecord") || !hip_ok(hipEventSynchronize(g_ws.ev1), "ias15 graph"))
        return ASSIST_BATCH_ERROR_DEVICE;
    float t = 0.f;
    (void)hipEventElapsedTime(&t, g_ws.ev0, g_ws.ev1);
    if (ms) *ms = t;
    return ASSIST_SUCCESS;
}

// Launches set_counter + ias15_kernel on `s` and waits; *ms is the kernel's HIP-event time.
int run_ias15(IntegrateArgs a, bool mixed, const LaunchShape& sh, int blocks_per_cu_cap, hipStream_t s, float* ms,
              int* blocks_out, bool use_graph = false, bool use_cache = true) {
    int dev = 0, ncu = 0, per_cu = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice") ||
        !hip_ok(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute") ||
        !hip_ok(ias15_occupancy(sh.G, mixed, sh.waves, ext_model(a.fo), ias15_snap(a), a.nvec, sh.threads,
                                &per_cu),
                "occupancy"))
        return ASSIST_BATCH_ERROR_DEVICE;
    if (per_cu < 1) per_cu = 1;
    if (blocks_per_cu_cap > 0 && blocks_per_cu_cap < per_cu) per_cu = blocks_per_cu_cap;
    const long long need_blocks = ((long long)a.nsys * sh.G + sh.threads - 1) / sh.threads;
    const long long max_blocks = (long long)per_cu * ncu;
    // 8-lane groups fill the grid even for batches smaller than it: the kernel spreads the systems
    // over every wave, and the groups without one take Chebyshev items of their wave
    const bool spread = ias15_cache_words(sh.G, ext_model(a.fo)) > 0;
    const int blocks = (int)((need_blocks < max_blocks && !spread) ? need_blocks : max_blocks);
    const int total_groups = blocks * sh.threads / sh.G;
    // scratch: the b/e records of every lane (42 doubles), then the substep ephemeris cache
    const int cw = use_cache ? ias15_cache_words(sh.G, ext_model(a.fo)) : 0;
    int rc = ensure_workspace((size_t)blocks * sh.threads * (42 + cw));
    if (rc) return rc;
    a.scratch = g_ws.scratch;
    a.ecache = cw ? g_ws.scratch + (size_t)blocks * sh.threads * 42 : nullptr;
    a.queue = g_ws.queue;
    if (use_graph) {
        rc = run_ias15_graph(a, mixed, sh, blocks, total_groups, s, ms);
        if (rc) return rc;
        if (blocks_out) *blocks_out = blocks;
        return ASSIST_SUCCESS;
    }
    if (!hip_ok(launch_set_counter(g_ws.queue, total_groups, s), "launch(set_counter)") ||
        !hip_ok(hipEventRecord(g_ws.ev0, s), "hipEventRecord") ||
        !hip_ok(launch_ias15(a, sh.G, mixed, sh.waves, blocks, sh.threads, s), "launch(ias15)") ||
        !hip_ok(hipEventRecord(g_ws.ev1, s), "hipEventRecord") || !hip_ok(hipEventSynchronize(g_ws.ev1), "ias15"))
        return ASSIST_BATCH_ERROR_DEVICE;
    float t = 0.f;
    (void)hipEventElapsedTime(&t, g_ws.ev0, g_ws.ev1);
    if (ms) *ms = t;
    if (blocks_out) *blocks_out = blocks;
    return ASSIST_SUCCESS;
}

struct TuneKey {
    int dev, nvec, mixed, forces, has_params, cache;
    bool operator<(const TuneKey& o) const {
        return std::tie(dev, nvec, mixed, forces, has_params, cache) <
               std::tie(o.dev, o.nvec, o.mixed, o.forces, o.has_params, o.cache);
    }
};
std::map<TuneKey, LaunchShape> g_tuned;

// Picks lanes per system, block size and kernel variant for this batch's configuration by
// timing each candidate on copies of the first kTuneSample systems over at most kTuneSpan days
// (the caller's buffers are not touched), with the substep cache as the launch will run it; the
// choice is cached per device and configuration.
int tuned_shape(const IntegrateArgs& base, bool mixed, bool use_cache, hipStream_t s, LaunchShape* out) {
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return ASSIST_BATCH_ERROR_DEVICE;
    const TuneKey key{dev, base.nvec, mixed ? 1 : 0, base.fo.forces, base.params ? 1 : 0, use_cache ? 1 : 0};
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
        *out = it->second;
        return ASSIST_SUCCESS;
    }
    const int m = base.nsys < kTuneSample ? base.nsys : kTuneSample;
    const size_t sw = (size_t)m * base.nvec * 6;
    std::vector<double> h0(m), h1(m);
    DevBuf<double> st, t1;
    DevBuf<int> status;
    if (!st.alloc(sw) || !t1.alloc(m) || !status.alloc(m)) return ASSIST_BATCH_ERROR_DEVICE;
    if (!hip_ok(hipMemcpy(h0.data(), base.t0, m * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy") ||
        !hip_ok(hipMemcpy(h1.data(), base.t1, m * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy"))
        return ASSIST_BATCH_ERROR_DEVICE;
    for (int i = 0; i < m; i++) {
        const double d = h1[i] - h0[i];
        h1[i] = h0[i] + (d > kTuneSpan ? kTuneSpan : (d < -kTuneSpan ? -kTuneSpan : d));
    }
    if (!hip_ok(hipMemcpy(t1.p, h1.data(), m * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy"))
        return ASSIST_BATCH_ERROR_DEVICE;
    IntegrateArgs a = base;
    a.nsys = m;
    a.states = st.p;
    a.t1 = t1.p;
    a.status = status.p;
    a.steps = a.evals = a.rejected = nullptr;
    a.t_out = nullptr;
    a.trace = nullptr;
    a.trace_count = nullptr;
    a.snap_in = nullptr;
    a.snap_out = nullptr;
    a.t_pause = nullptr;
    a.archive = nullptr;
    a.archive_count = nullptr;
    a.archive_cap = 0;
    LaunchShape best;
    best.G = lanes_for(base.nvec, 0);
    float best_ms = 1e30f;
    // candidates: one lane per system (real-only systems) and 8-lane groups, 1 wave per SIMD;
    // 2- and 4-lane groups and the 2-wave variant were slower in every measured configuration
    // (they stay available through assist_batch_config)
    for (int G : {1, 8}) {
        if (G < base.nvec) continue;
        for (int threads : {256, 128}) {
            LaunchShape sh;
            sh.G = G;
            sh.threads = threads;
            sh.waves = 1;
            // each shape runs twice and keeps the faster time: a kernel's first launch can
            // include loading its code, which the events around it would count
            float ms = 1e30f;
            for (int rep = 0; rep < 2; rep++) {
                if (!hip_ok(hipMemcpyAsync(st.p, base.states, sw * sizeof(double), hipMemcpyDeviceToDevice, s),
                            "hipMemcpyAsync"))
                    return ASSIST_BATCH_ERROR_DEVICE;
                float t_ms = 0.f;
                const int rc = run_ias15(a, mixed, sh, 0, s, &t_ms, nullptr, false, use_cache);
                if (rc) return rc;
                if (t_ms < ms) ms = t_ms;
            }
            if (ms < best_ms) {
                best_ms = ms;
                best = sh;
            }
        }
    }
    g_tuned[key] = best;
    *out = best;
    return ASSIST_SUCCESS;
}

}  // namespace

static int integrate_device_impl(const struct assist_ephem* ephem, const struct assist_extras* opts,
                                 const struct assist_batch_config* cfg_in, int n, int n_var, const double* d_t0,
                                 const double* d_t1, double* d_states, const double* d_params, int* d_status,
                                 long long* d_steps, long long* d_evals, long long* d_rejected, void* stream,
                                 struct assist_batch_stats* stats, const struct assist_batch_extra* ex);

extern "C" {

void assist_batch_config_default(struct assist_batch_config* cfg) {
    memset(cfg, 0, sizeof *cfg);  // every field, including those added later, starts defined
    cfg->size = sizeof *cfg;
    cfg->epsilon = 1e-9;
    cfg->min_dt = 0.0;
    cfg->dt_initial = 0.001;
    cfg->max_steps = 1000000LL;
    cfg->precision = ASSIST_PRECISION_FP64;
    cfg->lanes_per_system = 0;
    cfg->block_size = 0;
    cfg->blocks_per_cu = 0;
}

void assist_batch_extra_init(struct assist_batch_extra* ex) {
    memset(ex, 0, sizeof *ex);
    ex->size = sizeof *ex;
    ex->trace_every = 1;
}

const char* assist_batch_error_string(int code) {
    if (code >= 0 && code < ASSIST_ERROR_N) return assist_error_messages[code];
    switch (code) {
    case ASSIST_BATCH_ERROR_ARG: return "Invalid argument.";
    case ASSIST_BATCH_ERROR_DEVICE: return "GPU runtime error (no device, allocation or launch failure).";
    case ASSIST_BATCH_ERROR_UNSUPPORTED: return "Option not supported by the GPU path.";
    default: return "Unknown error.";
    }
}

int assist_gpu_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void* assist_device_malloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    return p;
}
void assist_device_free(void* p) {
    if (p) (void)hipFree(p);
}
int assist_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    return hip_ok(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy") ? 0 : ASSIST_BATCH_ERROR_DEVICE;
}
int assist_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    return hip_ok(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy") ? 0 : ASSIST_BATCH_ERROR_DEVICE;
}
int assist_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    return hip_ok(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy") ? 0 : ASSIST_BATCH_ERROR_DEVICE;
}
int assist_device_synchronize(void) {
    return hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize") ? 0 : ASSIST_BATCH_ERROR_DEVICE;
}

int assist_batch_forces(const struct assist_ephem* ephem, const struct assist_extras* opts, int n, int n_var,
                        const double* t, const double* states, const double* params, double* acc, int* status,
                        int precision) {
    if (n < 0 || n_var < 0 || n_var > 7 || !t || !states || !acc) return ASSIST_BATCH_ERROR_ARG;
    if (n == 0) return ASSIST_SUCCESS;
    if (assist_gpu_device_count() <= 0) return ASSIST_BATCH_ERROR_DEVICE;
    ForcesArgs a;
    memset(&a, 0, sizeof a);
    int rc = make_force_opts(opts, &a.fo);
    if (rc) return rc;
    if (ext_model(a.fo) && precision == ASSIST_PRECISION_MIXED) return ASSIST_BATCH_ERROR_UNSUPPORTED;
    DevEphem H;
    rc = get_device_ephem(ephem, &a.E, &H);
    if (rc) return rc;
    coverage_window(H, a.fo);
    const int nv = 1 + n_var;
    const int G = lanes_for(nv, 0);
    if (G < 0) return ASSIST_BATCH_ERROR_ARG;
    DevBuf<double> dt, ds, dp, da;
    DevBuf<int> dst;
    const size_t ns = (size_t)n;
    if (!dt.alloc(ns) || !ds.alloc(ns * nv * 6) || !da.alloc(ns * nv * 3) || !dst.alloc(ns) ||
        (params && !dp.alloc(ns * nv * 3)))
        return ASSIST_BATCH_ERROR_DEVICE;
    if (!hip_ok(hipMemcpy(dt.p, t, ns * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy") ||
        !hip_ok(hipMemcpy(ds.p, states, ns * nv * 6 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy") ||
        (params && !hip_ok(hipMemcpy(dp.p, params, ns * nv * 3 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy")))
        return ASSIST_BATCH_ERROR_DEVICE;
    a.nsys = n;
    a.nvec = nv;
    a.t = dt.p;
    a.states = ds.p;
    a.params = params ? dp.p : nullptr;
    a.acc = da.p;
    a.status = dst.p;
    if (!hip_ok(launch_forces(a, G, precision == ASSIST_PRECISION_MIXED, nullptr), "launch(batch_forces)") ||
        !hip_ok(hipDeviceSynchronize(), "batch_forces"))
        return ASSIST_BATCH_ERROR_DEVICE;
    if (!hip_ok(hipMemcpy(acc, da.p, ns * nv * 3 * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy") ||
        (status && !hip_ok(hipMemcpy(status, dst.p, ns * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy")))
        return ASSIST_BATCH_ERROR_DEVICE;
    return ASSIST_SUCCESS;
}

int assist_batch_integrate_device(const struct assist_ephem* ephem, const struct assist_extras* opts,
                                  const struct assist_batch_config* cfg_in, int n, int n_var, const double* d_t0,
                                  const double* d_t1, double* d_states, const double* d_params, int* d_status,
                                  long long* d_steps, long long* d_evals, void* stream,
                                  struct assist_batch_stats* stats) {
    return assist_batch_integrate_device_ex(ephem, opts, cfg_in, n, n_var, d_t0, d_t1, d_states, d_params, d_status,
                                            d_steps, d_evals, stream, stats, nullptr);
}

int assist_batch_integrate_device_ex(const struct assist_ephem* ephem, const struct assist_extras* opts,
                                     const struct assist_batch_config* cfg_in, int n, int n_var, const double* d_t0,
                                     const double* d_t1, double* d_states, const double* d_params, int* d_status,
                                     long long* d_steps, long long* d_evals, void* stream,
                                     struct assist_batch_stats* stats, const struct assist_batch_extra* ex) {
    return integrate_device_impl(ephem, opts, cfg_in, n, n_var, d_t0, d_t1, d_states, d_params, d_status, d_steps,
                                 d_evals, nullptr, stream, stats, ex);
}

}  // extern "C"

static int integrate_device_locked(const struct assist_ephem* ephem, const struct assist_extras* opts,
                                   const struct assist_batch_config& cfg, int n, int n_var, const double* d_t0,
                                   const double* d_t1, double* d_states, const double* d_params, int* d_status,
                                   long long* d_steps, long long* d_evals, long long* d_rejected, void* stream,
                                   struct assist_batch_stats* stats, const struct assist_batch_extra* ex);

// The device entry point with a per-system rejection count (d_rejected, or nullptr).
static int integrate_device_impl(const struct assist_ephem* ephem, const struct assist_extras* opts,
                                 const struct assist_batch_config* cfg_in, int n, int n_var, const double* d_t0,
                                 const double* d_t1, double* d_states, const double* d_params, int* d_status,
                                 long long* d_steps, long long* d_evals, long long* d_rejected, void* stream,
                                 struct assist_batch_stats* stats, const struct assist_batch_extra* ex) {
    const bool dense = ex && ex->out_t;
    if (cfg_in && cfg_in->size != sizeof(struct assist_batch_config)) return ASSIST_BATCH_ERROR_ARG;
    if (ex && ex->size != sizeof(struct assist_batch_extra)) return ASSIST_BATCH_ERROR_ARG;
    if (n < 0 || n_var < 0 || n_var > 7 || !d_t0 || (!d_t1 && !dense) || !d_states || !d_status)
        return ASSIST_BATCH_ERROR_ARG;
    if (ex && ex->trace && (!ex->trace_count || ex->trace_capacity < 0 || ex->trace_every < 1 || ex->trace_skip < 0))
        return ASSIST_BATCH_ERROR_ARG;
    if (dense && (!ex->out_offset || !ex->out_states)) return ASSIST_BATCH_ERROR_ARG;
    if (ex && ex->archive && (!ex->archive_count || ex->archive_capacity < 1)) return ASSIST_BATCH_ERROR_ARG;
    if (dense && ex->snapshot_in) return ASSIST_BATCH_ERROR_UNSUPPORTED;
    struct assist_batch_config cfg;
    if (cfg_in) cfg = *cfg_in;
    else assist_batch_config_default(&cfg);
    if (n == 0) return ASSIST_SUCCESS;
    std::lock_guard<std::mutex> lock(g_launch_mutex);
    return integrate_device_locked(ephem, opts, cfg, n, n_var, d_t0, d_t1, d_states, d_params, d_status, d_steps,
                                   d_evals, d_rejected, stream, stats, ex);
}

static int integrate_device_locked(const struct assist_ephem* ephem, const struct assist_extras* opts,
                                   const struct assist_batch_config& cfg, int n, int n_var, const double* d_t0,
                                   const double* d_t1, double* d_states, const double* d_params, int* d_status,
                                   long long* d_steps, long long* d_evals, long long* d_rejected, void* stream,
                                   struct assist_batch_stats* stats, const struct assist_batch_extra* ex) {
    const bool dense = ex && ex->out_t;
    IntegrateArgs a;
    memset(&a, 0, sizeof a);
    int rc = make_force_opts(opts, &a.fo);
    if (rc) return rc;
    DevEphem H;
    rc = get_device_ephem(ephem, &a.E, &H);
    if (rc) return rc;
    coverage_window(H, a.fo);
    const int nv = 1 + n_var;
    const bool mixed = cfg.precision == ASSIST_PRECISION_MIXED;
    a.io.epsilon = cfg.epsilon;
    a.io.min_dt = cfg.min_dt;
    a.io.dt_initial = cfg.dt_initial;
    a.io.max_steps = cfg.max_steps;
    a.io.pc_tol = corrector_tolerance(cfg);
    a.io.pc_tol1 = first_pass_tolerance(cfg);
    a.io.pc_aref = first_pass_accel(cfg);
    a.nsys = n;
    a.nvec = nv;
    a.t0 = d_t0;
    a.t1 = d_t1;
    a.states = d_states;
    a.params = d_params;
    a.status = d_status;
    a.steps = d_steps;
    a.evals = d_evals;
    a.rejected = d_rejected;
    if (ex) {
        a.dt0 = ex->dt0;
        a.bw = ex->bw;
        a.trace = ex->trace;
        a.trace_count = ex->trace_count;
        a.trace_cap = ex->trace_capacity;
        a.trace_every = ex->trace_every;
        a.trace_skip = ex->trace_skip;
        a.out_off = ex->out_offset;
        a.out_t = ex->out_t;
        a.out_states = ex->out_states;
        a.snap_in = ex->snapshot_in;
        a.snap_out = ex->snapshot_out;
        a.t_pause = ex->t_pause;
        a.archive = ex->archive;
        a.archive_count = ex->archive ? ex->archive_count : nullptr;
        a.archive_cap = ex->archive ? ex->archive_capacity : 0;
    }
    hipStream_t s = (hipStream_t)stream;
    LaunchShape shape;
    int lanes_req = cfg.lanes_per_system;
    if (a.snap_in && lanes_req == 0) {
        // a resumed batch continues with the lanes per system its snapshots were taken with (the
        // one-lane and group kernels agree to rounding, not bit for bit)
        double g = 0.0;
        if (!hip_ok(hipMemcpy(&g, a.snap_in + 7, sizeof g, hipMemcpyDeviceToHost), "hipMemcpy(snapshot)"))
            return ASSIST_BATCH_ERROR_DEVICE;
        if (g == 1.0 || g == 2.0 || g == 4.0 || g == 8.0) lanes_req = (int)g;
    }
    shape.G = lanes_for(nv, lanes_req);
    if (shape.G < 0) return ASSIST_BATCH_ERROR_ARG;
    shape.threads = cfg.block_size > 0 ? cfg.block_size : kIas15Block;
    if (cfg.block_size == 0 && shape.G > 0) {
        // untuned batches too small to give every CU a 256-thread block take 128-thread blocks,
        // so that their waves spread over all CUs (configuration 2: 3% faster)
        int dev = 0, ncu = 0;
        if (hip_ok(hipGetDevice(&dev), "hipGetDevice") &&
            hip_ok(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute") &&
            (long long)n * shape.G < (long long)kIas15Block * ncu)
            shape.threads = 128;
    }
    if (shape.threads % 64 != 0 || shape.threads > kIas15Block || shape.threads % shape.G != 0) return ASSIST_BATCH_ERROR_ARG;
    shape.waves = cfg.waves_per_simd > 0 ? cfg.waves_per_simd : 1;
    if (shape.waves != 1 && shape.waves != 2) return ASSIST_BATCH_ERROR_ARG;
    // the extended force model (geocentric, several EIH sources) runs in FP64 with the 1-wave kernel
    const bool ext = ext_model(a.fo);
    if (ext && (mixed || shape.waves != 1)) return ASSIST_BATCH_ERROR_UNSUPPORTED;
    const bool tune = lanes_req == 0 && cfg.block_size == 0 && cfg.waves_per_simd == 0 &&
                      n >= kTuneMinSystems && !dense && !ext;
    // the substep cache pays when attempts take several passes: by default it is off with the
    // first-pass stop (configuration 4 mixed+learned: 158.3 ms without it, 161.9 ms with it)
    const bool cache = cfg.substep_cache_off < 0 || (cfg.substep_cache_off == 0 && a.io.pc_tol1 <= 0.0);
    if (tune) {
        rc = tuned_shape(a, mixed, cache, s, &shape);
        if (rc) return rc;
    }
    // large batches: hand out the systems of highest expected cost first (longest spans, smallest
    // perihelia), so that the launch does not end on an expensive system handed out last
    if (n >= kTuneMinSystems && !dense && d_t1 && !a.snap_in) {
        rc = ensure_workspace(0);
        if (rc) return rc;
        rc = ensure_order(n);
        if (rc) return rc;
        if (!hip_ok(launch_span_order(a.E, n, a.nvec, d_t0, d_t1, a.states, g_ws.okeys, g_ws.okeys_alt, g_ws.ovals,
                                      g_ws.order, g_ws.otemp, g_ws.otemp_bytes, s),
                    "launch(queue order)"))
            return ASSIST_BATCH_ERROR_DEVICE;
        a.order = g_ws.order;
    }
    float ms = 0.f;
    int blocks = 0;
    rc = run_ias15(a, mixed, shape, cfg.blocks_per_cu, s, &ms, &blocks, cfg.use_graph != 0, cache);
    if (rc) return rc;
    const int G = shape.G, threads = shape.threads, waves = shape.waves;
    if (stats) {
        stats->kernel_ms = ms;
        stats->lanes_per_system = G;
        stats->block_size = threads;
        stats->grid_blocks = blocks;
        stats->waves_per_simd = waves;
    }
    return ASSIST_SUCCESS;
}

extern "C" {

int assist_batch_integrate_device_noalloc(const struct assist_ephem* ephem, const struct assist_extras* opts,
                                          const struct assist_batch_config* cfg_in, int n, int n_var,
                                          const double* d_t0, const double* d_t1, double* d_states,
                                          const double* d_params, int* d_status, long long* d_steps,
                                          long long* d_evals, void* stream) {
    if (cfg_in && cfg_in->size != sizeof(struct assist_batch_config)) return ASSIST_BATCH_ERROR_ARG;
    if (n < 0 || n_var < 0 || n_var > 7 || !d_t0 || !d_t1 || !d_states || !d_status) return ASSIST_BATCH_ERROR_ARG;
    struct assist_batch_config cfg;
    if (cfg_in) cfg = *cfg_in;
    else assist_batch_config_default(&cfg);
    if (cfg.use_graph) return ASSIST_BATCH_ERROR_UNSUPPORTED;  // capturing a graph allocates
    if (n == 0) return ASSIST_SUCCESS;
    IntegrateArgs a;
    memset(&a, 0, sizeof a);
    int rc = make_force_opts(opts, &a.fo);
    if (rc) return rc;
    DevEphem H;
    rc = lookup_device_ephem(ephem, &a.E, &H);
    if (rc == ASSIST_BATCH_ERROR_UNSUPPORTED)
        fprintf(stderr, "(ASSIST) assist_batch_integrate_device_noalloc: the ephemeris is not resident on this GPU; "
                        "make it resident with any other batched call first.\n");
    if (rc) return rc;
    coverage_window(H, a.fo);
    const int nv = 1 + n_var;
    const bool mixed = cfg.precision == ASSIST_PRECISION_MIXED;
    const bool ext = ext_model(a.fo);
    if (ext && mixed) return ASSIST_BATCH_ERROR_UNSUPPORTED;
    const int G = lanes_for(nv, cfg.lanes_per_system);
    const int threads = cfg.block_size > 0 ? cfg.block_size : kIas15Block;
    if (G < 0 || threads % 64 != 0 || threads > kIas15Block || threads % G != 0) return ASSIST_BATCH_ERROR_ARG;
    if (cfg.waves_per_simd != 0 && cfg.waves_per_simd != 1) return ASSIST_BATCH_ERROR_UNSUPPORTED;
    a.io.epsilon = cfg.epsilon;
    a.io.min_dt = cfg.min_dt;
    a.io.dt_initial = cfg.dt_initial;
    a.io.max_steps = cfg.max_steps;
    a.io.pc_tol = corrector_tolerance(cfg);
    a.io.pc_tol1 = first_pass_tolerance(cfg);
    a.io.pc_aref = first_pass_accel(cfg);
    a.nsys = n;
    a.nvec = nv;
    a.t0 = d_t0;
    a.t1 = d_t1;
    a.states = d_states;
    a.params = d_params;
    a.status = d_status;
    a.steps = d_steps;
    a.evals = d_evals;
    std::lock_guard<std::mutex> lock(g_launch_mutex);
    double* scr = nullptr;
    int dev = 0, ncu = 0, per_cu = 0;
    if (!hip_ok(noalloc_scratch(&scr), "hipGetSymbolAddress(scratch)") || !hip_ok(hipGetDevice(&dev), "hipGetDevice") ||
        !hip_ok(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute") ||
        !hip_ok(ias15_occupancy(G, mixed, 1, ext, false, nv, threads, &per_cu), "occupancy"))
        return ASSIST_BATCH_ERROR_DEVICE;
    if (per_cu < 1) per_cu = 1;
    if (cfg.blocks_per_cu > 0 && cfg.blocks_per_cu < per_cu) per_cu = cfg.blocks_per_cu;
    // the grid: one pass of the persistent kernel over at most kNoAllocLanes resident lanes (the
    // b/e records of the static scratch); the work queue hands every system to them
    long long blocks = ((long long)n * G + threads - 1) / threads;
    if (blocks > (long long)per_cu * ncu) blocks = (long long)per_cu * ncu;
    if (blocks > kNoAllocLanes / threads) blocks = kNoAllocLanes / threads;
    if (blocks < 1) return ASSIST_BATCH_ERROR_ARG;
    a.queue = (int*)scr;
    a.scratch = scr + kNoAllocHeadWords;
    a.ecache = nullptr;  // no substep cache: same results, computed again in each iteration
    const int total_groups = (int)blocks * threads / G;
    hipStream_t s = (hipStream_t)stream;
    if (!hip_ok(launch_set_counter(a.queue, total_groups, s), "launch(set_counter)") ||
        !hip_ok(launch_ias15(a, G, mixed, 1, (int)blocks, threads, s), "launch(ias15)") ||
        !hip_ok(hipStreamSynchronize(s), "ias15"))
        return ASSIST_BATCH_ERROR_DEVICE;
    return ASSIST_SUCCESS;
}

int assist_batch_get_particles_device(const struct assist_ephem* ephem, int n, const int* d_body, const double* d_t,
                                      double* d_out, int* d_status, void* stream) {
    if (n < 0 || (n > 0 && (!d_body || !d_t || !d_out))) return ASSIST_BATCH_ERROR_ARG;
    if (n == 0) return ASSIST_SUCCESS;
    const DevEphem* E = nullptr;
    int rc = get_device_ephem(ephem, &E, nullptr);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (!hip_ok(launch_ephem_query(E, n, d_body, d_t, d_out, d_status, s), "launch(ephem_query)") ||
        !hip_ok(hipStreamSynchronize(s), "ephem_query"))
        return ASSIST_BATCH_ERROR_DEVICE;
    return ASSIST_SUCCESS;
}

int assist_batch_archive_interpolate_device(int n, int n_var, const double* d_archive, int capacity,
                                            const int* d_count, int nq, const int* d_q_sys, const double* d_q_t,
                                            double* d_out_states, int* d_status, void* stream) {
    if (n < 0 || n_var < 0 || n_var > 7 || capacity < 1 || nq < 0) return ASSIST_BATCH_ERROR_ARG;
    if (nq > 0 && (!d_archive || !d_count || !d_q_sys || !d_q_t || !d_out_states)) return ASSIST_BATCH_ERROR_ARG;
    if (nq == 0) return ASSIST_SUCCESS;
    hipStream_t s = (hipStream_t)stream;
    if (!hip_ok(launch_archive_interp(n, nq, 1 + n_var, capacity, d_archive, d_count, d_q_sys, d_q_t, d_out_states,
                                      d_status, s),
                "launch(archive_interp)") ||
        !hip_ok(hipStreamSynchronize(s), "archive_interp"))
        return ASSIST_BATCH_ERROR_DEVICE;
    return ASSIST_SUCCESS;
}

int assist_batch_archive_interpolate(int n, int n_var, const double* archive, int capacity, const int* count, int nq,
                                     const int* q_sys, const double* q_t, double* out_states, int* status) {
    if (n < 0 || n_var < 0 || n_var > 7 || capacity < 1 || nq < 0) return ASSIST_BATCH_ERROR_ARG;
    if (nq > 0 && (!archive || !count || !q_sys || !q_t || !out_states)) return ASSIST_BATCH_ERROR_ARG;
    if (nq == 0) return ASSIST_SUCCESS;
    if (assist_gpu_device_count() <= 0) return ASSIST_BATCH_ERROR_DEVICE;
    const int nv = 1 + n_var;
    const size_t words = (size_t)n * capacity * ASSIST_SNAPSHOT_WORDS(nv), nqs = (size_t)nq;
    DevBuf<double> da, dt, dout;
    DevBuf<int> dc, dq, dst;
    if (!da.alloc(words) || !dc.alloc((size_t)n) || !dq.alloc(nqs) || !dt.alloc(nqs) || !dout.alloc(nqs * nv * 6) ||
        !dst.alloc(nqs))
        return ASSIST_BATCH_ERROR_DEVICE;
    if (!hip_ok(hipMemcpy(da.p, archive, words * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy") ||
        !hip_ok(hipMemcpy(dc.p, count, (size_t)n * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy") ||
        !hip_ok(hipMemcpy(dq.p, q_sys, nqs * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy") ||
        !hip_ok(hipMemcpy(dt.p, q_t, nqs * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy"))
        return ASSIST_BATCH_ERROR_DEVICE;
    const int rc = assist_batch_archive_interpolate_device(n, n_var, da.p, capacity, dc.p, nq, dq.p, dt.p, dout.p,
                                                           dst.p, nullptr);
    if (rc) return rc;
    if (!hip_ok(hipMemcpy(out_states, dout.p, nqs * nv * 6 * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy") ||
        (status && !hip_ok(hipMemcpy(status, dst.p, nqs * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy")))
        return ASSIST_BATCH_ERROR_DEVICE;
    return ASSIST_SUCCESS;
}

int assist_batch_get_particles(const struct assist_ephem* ephem, int n, const int* body, const double* t, double* out,
                               int* status) {
    if (n < 0 || (n > 0 && (!body || !t || !out))) return ASSIST_BATCH_ERROR_ARG;
    if (n == 0) return ASSIST_SUCCESS;
    if (assist_gpu_device_count() <= 0) return ASSIST_BATCH_ERROR_DEVICE;
    const size_t ns = (size_t)n;
    DevBuf<int> db, dst;
    DevBuf<double> dt, dout;
    if (!db.alloc(ns) || !dt.alloc(ns) || !dout.alloc(ns * 10) || !dst.alloc(ns)) return ASSIST_BATCH_ERROR_DEVICE;
    if (!hip_ok(hipMemcpy(db.p, body, ns * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy") ||
        !hip_ok(hipMemcpy(dt.p, t, ns * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy"))
        return ASSIST_BATCH_ERROR_DEVICE;
    const int rc = assist_batch_get_particles_device(ephem, n, db.p, dt.p, dout.p, dst.p, nullptr);
    if (rc) return rc;
    if (!hip_ok(hipMemcpy(out, dout.p, ns * 10 * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy") ||
        (status && !hip_ok(hipMemcpy(status, dst.p, ns * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy")))
        return ASSIST_BATCH_ERROR_DEVICE;
    return ASSIST_SUCCESS;
}

int assist_batch_integrate(const struct assist_ephem* ephem, const struct assist_extras* opts,
                           const struct assist_batch_config* cfg, int n, int n_var, const double* t0,
                           const double* t1, double* states, const double* params, int* status, long long* steps,
                           struct assist_batch_stats* stats) {
    if (cfg && cfg->size != sizeof(struct assist_batch_config)) return ASSIST_BATCH_ERROR_ARG;
    if (n < 0 || n_var < 0 || n_var > 7 || !t0 || !t1 || !states) return ASSIST_BATCH_ERROR_ARG;
    if (n == 0) return ASSIST_SUCCESS;
    if (assist_gpu_device_count() <= 0) return ASSIST_BATCH_ERROR_DEVICE;
    const int nv = 1 + n_var;
    const size_t ns = (size_t)n;
    DevBuf<double> d0, d1, ds, dp;
    DevBuf<int> dst;
    DevBuf<long long> dsteps, devals, drej;
    if (!d0.alloc(ns) || !d1.alloc(ns) || !ds.alloc(ns * nv * 6) || !dst.alloc(ns) || !dsteps.alloc(ns) ||
        !devals.alloc(ns) || !drej.alloc(ns) || (params && !dp.alloc(ns * nv * 3)))
        return ASSIST_BATCH_ERROR_DEVICE;
    if (!hip_ok(hipMemcpy(d0.p, t0, ns * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy") ||
        !hip_ok(hipMemcpy(d1.p, t1, ns * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy") ||
        !hip_ok(hipMemcpy(ds.p, states, ns * nv * 6 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy") ||
        (params && !hip_ok(hipMemcpy(dp.p, params, ns * nv * 3 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy")))
        return ASSIST_BATCH_ERROR_DEVICE;
    struct assist_batch_stats st;
    memset(&st, 0, sizeof st);
    int rc = integrate_device_impl(ephem, opts, cfg, n, n_var, d0.p, d1.p, ds.p, params ? dp.p : nullptr, dst.p,
                                   dsteps.p, devals.p, drej.p, nullptr, &st, nullptr);
    if (rc) return rc;
    std::vector<long long> hsteps(ns), hevals(ns), hrej(ns);
    if (!hip_ok(hipMemcpy(states, ds.p, ns * nv * 6 * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy") ||
        !hip_ok(hipMemcpy(hsteps.data(), dsteps.p, ns * sizeof(long long), hipMemcpyDeviceToHost), "hipMemcpy") ||
        !hip_ok(hipMemcpy(hevals.data(), devals.p, ns * sizeof(long long), hipMemcpyDeviceToHost), "hipMemcpy") ||
        !hip_ok(hipMemcpy(hrej.data(), drej.p, ns * sizeof(long long), hipMemcpyDeviceToHost), "hipMemcpy") ||
        (status && !hip_ok(hipMemcpy(status, dst.p, ns * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy")))
        return ASSIST_BATCH_ERROR_DEVICE;
    if (steps) memcpy(steps, hsteps.data(), ns * sizeof(long long));
    for (size_t i = 0; i < ns; i++) { st.steps += hsteps[i]; st.evals += hevals[i]; st.rejected += hrej[i]; }
    if (stats) *stats = st;
    return ASSIST_SUCCESS;
}

void assist_additional_forces(struct reb_simulation* sim) {
    auto* ax = (struct assist_extras*)sim->extras;
    if (!ax || !ax->ephem) return;
    SimLayout L = layout_of(sim);
    if (L.n_real <= 0) return;
    std::vector<double> states, params, acc;
    bool hp = false;
    pack(sim, ax, L, states, params, hp);
    const int nv = 1 + L.nvmax;
    std::vector<double> t(L.n_real, sim->t);
    std::vector<int> st(L.n_real, 0);
    acc.assign((size_t)L.n_real * nv * 3, 0.0);
    const int rc = assist_batch_forces(ax->ephem, ax, L.n_real, L.nvmax, t.data(), states.data(), hp ? params.data() : nullptr,
                                       acc.data(), st.data(), ASSIST_PRECISION_FP64);
    if (rc) {
        reb_simulation_error(sim, assist_batch_error_string(rc));
        return;
    }
    for (int j = 0; j < L.n_real; j++) {
        if (st[j]) {
            reb_simulation_error(sim, assist_error_messages[st[j] < ASSIST_ERROR_N ? st[j] : 0]);
            return;
        }
        reb_particle& p = sim->particles[j];
        const double* a = &acc[(size_t)j * nv * 3];
        p.ax = a[0]; p.ay = a[1]; p.az = a[2];
        for (size_t k = 0; k < L.vars[j].size(); k++) {
            reb_particle& q = sim->particles[sim->var_config[L.vars[j][k]].index];
            q.ax = a[3 * (k + 1)]; q.ay = a[3 * (k + 1) + 1]; q.az = a[3 * (k + 1) + 2];
        }
    }
}

// Integrates every real particle of `sim` (with its variational particles) as its own system
// to tmax; with `dense`, the steps are not shortened to land on tmax and the state at tmax is
// interpolated inside the step that covers it.
static void sim_integrate(struct reb_simulation* sim, double tmax, bool dense) {
    auto* ax = (struct assist_extras*)sim->extras;
    if (!ax || !ax->ephem) {
        reb_simulation_error(sim, "No ASSIST extras attached to the simulation.");
        return;
    }
    SimLayout L = layout_of(sim);
    if (L.n_real <= 0) { sim->t = tmax; return; }
    std::vector<double> states, params;
    bool hp = false;
    pack(sim, ax, L, states, params, hp);
    const int nv = 1 + L.nvmax;
    std::vector<double> t0(L.n_real, sim->t), t1(L.n_real, tmax);
    std::vector<int> st(L.n_real, 0);
    std::vector<long long> steps(L.n_real, 0);
    struct assist_batch_config cfg;
    assist_batch_config_default(&cfg);
    cfg.epsilon = sim->ri_ias15.epsilon;
    cfg.min_dt = sim->ri_ias15.min_dt;
    if (sim->dt != 0) cfg.dt_initial = fabs(sim->dt);
    int rc;
    std::vector<double> out;
    if (dense) {
        std::vector<int> off(L.n_real + 1);
        for (int j = 0; j <= L.n_real; j++) off[j] = j;
        out.assign((size_t)L.n_real * nv * 6, 0.0);
        rc = assist_batch_integrate_dense(ax->ephem, ax, &cfg, L.n_real, L.nvmax, t0.data(), states.data(),
                                          hp ? params.data() : nullptr, off.data(), t1.data(), out.data(), st.data(),
                                          steps.data(), nullptr);
    } else {
        rc = assist_batch_integrate(ax->ephem, ax, &cfg, L.n_real, L.nvmax, t0.data(), t1.data(), states.data(),
                                    hp ? params.data() : nullptr, st.data(), steps.data(), nullptr);
    }
    if (rc) {
        reb_simulation_error(sim, assist_batch_error_string(rc));
        return;
    }
    const std::vector<double>& res = dense ? out : states;
    for (int j = 0; j < L.n_real; j++) {
        const double* s = &res[(size_t)j * nv * 6];
        reb_particle& p = sim->particles[j];
        p.x = s[0]; p.y = s[1]; p.z = s[2]; p.vx = s[3]; p.vy = s[4]; p.vz = s[5];
        for (size_t k = 0; k < L.vars[j].size(); k++) {
            reb_particle& q = sim->particles[sim->var_config[L.vars[j][k]].index];
            const double* d = s + 6 * (k + 1);
            q.x = d[0]; q.y = d[1]; q.z = d[2]; q.vx = d[3]; q.vy = d[4]; q.vz = d[5];
        }
        ax->steps_done += (int)steps[j];
        if (st[j] != 0) {
            reb_simulation_error(sim, st[j] > 0 && st[j] < ASSIST_ERROR_N ? assist_error_messages[st[j]]
                                                                        : "Integration of a particle did not finish.");
        }
    }
    if (!sim->status) sim->t = tmax;
}

void reb_simulation_integrate(struct reb_simulation* sim, double tmax) { sim_integrate(sim, tmax, false); }

void assist_integrate_or_interpolate(struct assist_extras* ax, double t) {
    if (!ax || !ax->sim) return;
    sim_integrate(ax->sim, t, true);
}

int assist_batch_integrate_dense(const struct assist_ephem* ephem, const struct assist_extras* opts,
                                 const struct assist_batch_config* cfg, int n, int n_var, const double* t0,
                                 double* states, const double* params, const int* out_offset, const double* out_t,
                                 double* out_states, int* status, long long* steps, struct assist_batch_stats* stats) {
    if (cfg && cfg->size != sizeof(struct assist_batch_config)) return ASSIST_BATCH_ERROR_ARG;
    if (n < 0 || n_var < 0 || n_var > 7 || !t0 || !states || !out_offset || !out_t || !out_states)
        return ASSIST_BATCH_ERROR_ARG;
    if (n == 0) return ASSIST_SUCCESS;
    if (out_offset[0] != 0) return ASSIST_BATCH_ERROR_ARG;
    for (int i = 0; i < n; i++) {
        const int a0 = out_offset[i], a1 = out_offset[i + 1];
        if (a1 < a0) return ASSIST_BATCH_ERROR_ARG;
        if (a1 == a0) continue;
        const double dir = (out_t[a1 - 1] >= t0[i]) ? 1.0 : -1.0;
        double prev = t0[i];
        for (int j = a0; j < a1; j++) {
            if (!((out_t[j] - prev) * dir >= 0.0)) return ASSIST_BATCH_ERROR_ARG;
            prev = out_t[j];
        }
    }
    if (assist_gpu_device_count() <= 0) return ASSIST_BATCH_ERROR_DEVICE;
    const int nv = 1 + n_var;
    const size_t ns = (size_t)n, m = (size_t)out_offset[n];
    DevBuf<double> d0, ds, dp, dt, dout;
    DevBuf<int> doff, dst;
    DevBuf<long long> dsteps;
    if (!d0.alloc(ns) || !ds.alloc(ns * nv * 6) || !dst.alloc(ns) || !dsteps.alloc(ns) || !doff.alloc(ns + 1) ||
        !dt.alloc(m ? m : 1) || !dout.alloc(m ? m * nv * 6 : 1) || (params && !dp.alloc(ns * nv * 3)))
        return ASSIST_BATCH_ERROR_DEVICE;
    if (!hip_ok(hipMemcpy(d0.p, t0, ns * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy") ||
        !hip_ok(hipMemcpy(ds.p, states, ns * nv * 6 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy") ||
        !hip_ok(hipMemcpy(doff.p, out_offset, (ns + 1) * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy") ||
        (m && !hip_ok(hipMemcpy(dt.p, out_t, m * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy")) ||
        (params && !hip_ok(hipMemcpy(dp.p, params, ns * nv * 3 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy")))
        return ASSIST_BATCH_ERROR_DEVICE;
    struct assist_batch_extra ex;
    assist_batch_extra_init(&ex);
    ex.out_offset = doff.p;
    ex.out_t = dt.p;
    ex.out_states = dout.p;
    const int rc = assist_batch_integrate_device_ex(ephem, opts, cfg, n, n_var, d0.p, nullptr, ds.p, params ? dp.p : nullptr,
                                                    dst.p, dsteps.p, nullptr, nullptr, stats, &ex);
    if (rc) return rc;
    if (!hip_ok(hipMemcpy(states, ds.p, ns * nv * 6 * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy") ||
        (m && !hip_ok(hipMemcpy(out_states, dout.p, m * nv * 6 * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy")) ||
        (status && !hip_ok(hipMemcpy(status, dst.p, ns * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy")) ||
        (steps && !hip_ok(hipMemcpy(steps, dsteps.p, ns * sizeof(long long), hipMemcpyDeviceToHost), "hipMemcpy")))
        return ASSIST_BATCH_ERROR_DEVICE;
    return ASSIST_SUCCESS;
}

int assist_step_features(const struct assist_ephem* ephem, int n, const double* d_rows, int stride, float* d_X,
                         double* d_omega) {
    if (!ephem || n < 0 || !d_rows || stride < 8 || !d_X) return ASSIST_BATCH_ERROR_ARG;
    if (n == 0) return ASSIST_SUCCESS;
    const DevEphem* E = nullptr;
    int rc = get_device_ephem(ephem, &E, nullptr);
    if (rc) return rc;
    if (!hip_ok(launch_features(E, n, d_rows, stride, d_X, d_omega, nullptr), "launch(features)") ||
        !hip_ok(hipDeviceSynchronize(), "features"))
        return ASSIST_BATCH_ERROR_DEVICE;
    return ASSIST_SUCCESS;
}

int assist_batch_predict_start(const struct assist_ephem* ephem, struct assist_mlp* net, int n, int n_var,
                               const double* d_t0, const double* d_t1, const double* d_states, double min_dt,
                               int iterations, double* d_dt0, double* d_bw) {
    if (!ephem || !net || n < 0 || n_var < 0 || n_var > 7 || !d_t0 || !d_t1 || !d_states || !d_dt0 || iterations < 1)
        return ASSIST_BATCH_ERROR_ARG;
    if (n == 0) return ASSIST_SUCCESS;
    const DevEphem* E = nullptr;
    int rc = get_device_ephem(ephem, &E, nullptr);
    if (rc) return rc;
    // (the buffers are the library's persistent workspace, shared with the integration entry points)
    std::lock_guard<std::mutex> lock(g_launch_mutex);
    rc = ensure_workspace(0);
    if (rc) return rc;
    rc = ensure_predict((size_t)n);
    if (rc) return rc;
    // the first pass sees the last step as 1 day; later passes see the previous prediction
    if (!hip_ok(launch_start_rows(n, 1 + n_var, d_t0, d_states, 1.0, g_ws.prows, nullptr), "launch(start_rows)"))
        return ASSIST_BATCH_ERROR_DEVICE;
    for (int it = 0; it < iterations; it++) {
        if (!hip_ok(launch_features(E, n, g_ws.prows, 8, g_ws.pX, nullptr, nullptr), "launch(features)"))
            return ASSIST_BATCH_ERROR_DEVICE;
        rc = assist_mlp_predict(net, g_ws.pX, n, g_ws.pY, nullptr);
        if (rc) return rc;
        const bool last = it == iterations - 1;
        if (!hip_ok(launch_start_post(n, g_ws.pY, d_t0, d_t1, min_dt, g_ws.prows, d_dt0, (last && d_bw) ? d_bw : nullptr,
                                      nullptr),
                    "launch(start_post)"))
            return ASSIST_BATCH_ERROR_DEVICE;
    }
    if (!hip_ok(hipDeviceSynchronize(), "predict_start")) return ASSIST_BATCH_ERROR_DEVICE;
    return ASSIST_SUCCESS;
}

}  // extern "C"
