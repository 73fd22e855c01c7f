// Launch interface of the HIP kernels (internal to libassist).
#pragma once
#include <hip/hip_runtime.h>

#include "device_types.h"

namespace assist_gpu {

// Largest block of the integration kernel.
constexpr int kIas15Block = 256;
// Doubles per LDS column of the integration kernel (b and g, 42 words, padded to 43).
constexpr int kLdsColWords = 43;
// LDS columns per lane group: one per lane; the 2-wave variant packs the 7 vectors of 8-lane
// groups into 7 columns (lanes that own no vector share one spare column) so that two blocks fit
// a CU's LDS.
__host__ __device__ constexpr int ias15_cols_per_group(int G, int nvec, int waves) {
    return (waves == 2 && G == 8 && nvec <= 7) ? 7 : G;
}
// IAS15 tables staged in LDS ahead of the columns (doubles): spacings h (8), predictor weights
// (9 x 8 and 9 x 7), 1/(h_i - h_j) (28), Newton-to-power weights (21), and the b-update weights
// of substep n as a full row (7 x 7: the weights, then 1 at k = n - 1 and 0 after it), padded
// to 256 words so that the columns keep their LDS bank alignment.
constexpr int kTabH = 0, kTabPX = 8, kTabPV = 80, kTabRRI = 143, kTabC = 171, kTabW = 192, kTabWords = 256;

struct ForcesArgs {
    const DevEphem* E;
    ForceOpts fo;
    int nsys;
    int nvec;               // 1 + number of variational particles per system
    const double* t;        // [nsys]
    const double* states;   // [nsys][nvec][6]
    const double* params;   // [nsys][nvec][3] or nullptr
    double* acc;            // [nsys][nvec][3]
    int* status;            // [nsys] or nullptr
};

struct IntegrateArgs {
    const DevEphem* E;
    ForceOpts fo;
    IntegratorOpts io;
    int nsys;
    int nvec;
    const double* t0;       // [nsys]
    const double* t1;       // [nsys]
    double* states;         // [nsys][nvec][6], updated in place
    const double* params;   // [nsys][nvec][3] or nullptr
    int* status;            // [nsys]
    long long* steps;       // [nsys] or nullptr
    long long* evals;       // [nsys] or nullptr
    long long* rejected;    // [nsys] or nullptr
    double* t_out;          // [nsys] or nullptr
    double* scratch;        // [threads][42]: br, er of each lane's last accepted step
    int* queue;             // next system to hand out
    // learned start (optional): first step and b warm start per system
    const double* dt0;      // [nsys] first trial step (days, sign ignored) or nullptr
    const double* bw;       // [nsys][21] b_k,c = bw * |a0| * dt^(k+1) for the real particle, or nullptr
    // step traces (optional): one record of kTraceWords doubles per sampled accepted step
    double* trace;
    int* trace_count;
    int trace_cap;
    int trace_every;        // record steps s >= trace_skip with s % trace_every == 0
    int trace_skip;
    int pad_;
    // dense output (optional): system s writes its state at out_t[out_off[s] .. out_off[s+1]-1]
    // (sorted in the direction of integration) to out_states[i][nvec][6] by interpolating the
    // accepted step that contains each time; steps are then not trimmed to land on t1 (t1 unused)
    const int* out_off;
    const double* out_t;
    double* out_states;
    // queue order (optional): the q-th system handed out is order[q] (longest span first)
    const int* order;
    // substep ephemeris cache of 8-lane groups (optional): [waves][8][7][kEcWords] doubles
    double* ecache;
    // checkpoint and resume (optional), snapshots of snap_words(nvec) doubles per system:
    // snap_in resumes each system from its snapshot (t0 and dt0 are then not read), snap_out
    // receives each system's snapshot when it stops, and a system with t_pause stops (status
    // kPaused) at the first step boundary at or past t_pause[s] in its direction
    const double* snap_in;
    double* snap_out;
    const double* t_pause;
    // step archive (optional): system s writes its snapshot at its start and after every
    // accepted step to archive[s][k] (k < archive_cap; later ones are counted, not written),
    // and the number of snapshots it took to archive_count[s]
    double* archive;
    int* archive_count;
    int archive_cap;
    int pad2_;
};

// Doubles of the substep ephemeris cache per launched thread (0 when the kernel keeps none):
// 8-lane groups outside the extended model keep 88 words per substep and group.  (One lane per
// system keeps none: 64 systems per wave need their slots evaluated on nearly every trip, and
// the lanes that would load instead wait on memory for all 28 slots.)
__host__ __device__ constexpr int ias15_cache_words(int G, bool ext) {
    return (G == 8 && !ext) ? 7 * 88 / 8 : 0;
}
// Doubles of LDS per wave for the exchange of the wave-wide Chebyshev rounds (8-lane groups
// outside the extended model): 8 groups x 29 rows (28 slots, the Sun's velocity) x 3.
__host__ __device__ constexpr int ias15_xch_words(int G, bool ext) {
    return (G == 8 && !ext) ? 8 * 29 * 3 : 0;
}

// Trace record: t, x0[3], v0[3] at the start of the step, the previous step, the accepted
// step, IAS15's proposal for the next step before the growth cap, a0[3], b[7][3] (k-major).
constexpr int kTraceWords = 34;

// Features of the learned start: one row per (t, x[3], v[3], dt_last) record with `stride`
// doubles between rows; writes X[n][16] (float) and, if omega is given, the heliocentric
// mean motion sqrt(GM_sun / r^3) per row.
hipError_t launch_features(const DevEphem* E, int n, const double* rows, int stride, float* X, double* omega,
                           hipStream_t s);
// Rows [t0, x, v, dt_last] of the real particles of a batch.
hipError_t launch_start_rows(int n, int nvec, const double* t0, const double* states, double dt_last, double* rows,
                             hipStream_t s);
// From network outputs Y[n][32]: dt0 = 10^Y0 (clamped to [dt_min, |t1 - t0|]) written to dt0 and to
// the rows' dt_last; bw[n][21] = Y[1 + 3k + c] (c / dt0)^(k+1) / (k+1)! when bw is given.
hipError_t launch_start_post(int n, const float* Y, const double* t0, const double* t1,
                             double dt_min, double* rows, double* dt0, double* bw, hipStream_t s);

hipError_t launch_forces(const ForcesArgs& a, int G, bool mixed, hipStream_t s);
// waves: the kernel variant compiled for 1 or 2 waves per SIMD (register budget 512 or 256).
hipError_t launch_ias15(const IntegrateArgs& a, int G, bool mixed, int waves, int blocks, int threads, hipStream_t s,
                        bool set_attributes = true);
hipError_t ias15_occupancy(int G, bool mixed, int waves, bool ext, bool snap, int nvec, int threads,
                           int* blocks_per_cu);
// Whether a launch uses the checkpoint, pause or step-archive features (kernels without them
// exist for the hot shapes).
inline bool ias15_snap(const IntegrateArgs& a) { return a.snap_in || a.snap_out || a.t_pause || a.archive; }
// The extended force model (geocentric frame or several EIH sources) has kernels of its own.
inline bool ext_model(const ForceOpts& f) { return f.geocentric != 0 || ((f.forces & 0x40) && f.gr_eih_sources > 1); }
size_t ias15_lds_bytes(int G, int nvec, int threads, int waves, bool ext);
hipError_t launch_set_counter(int* q, int v, hipStream_t s);
// Scratch of the allocation-free entry point: a static device array of the library (22 MB per
// GPU): kNoAllocHeadWords doubles for the work-queue counter, then the b/e records of up to
// kNoAllocLanes resident lanes (42 doubles each).  *ptr receives its device address on the
// current device (looked up once per device).
constexpr int kNoAllocLanes = 65536;
constexpr int kNoAllocHeadWords = 32;
constexpr int kNoAllocMaxDevices = 64;
hipError_t noalloc_scratch(double** ptr);
// Queue order of large batches: order[0..n) = system indices sorted by span |t1 - t0|, longest
// first, and among equal spans by the real particle's perihelion distance, smallest first (stable,
// so equal keys keep their index order).  Buffers: keys/keys_alt n words, vals n ints.
size_t span_order_temp_bytes(int n);
hipError_t launch_span_order(const DevEphem* E, int n, int nvec, const double* t0, const double* t1,
                             const double* states, unsigned long long* keys, unsigned long long* keys_alt, int* vals,
                             int* order, void* temp, size_t temp_bytes, hipStream_t s);
// States at query times from step archives (IntegrateArgs::archive layout): query q asks for
// system qsys[q] (0 <= qsys[q] < nsys, else status kBadSystem) at qt[q]; out[q][nvec][6];
// status kCoverage outside the archived steps.
hipError_t launch_archive_interp(int nsys, int nq, int nvec, int cap, const double* archive, const int* count,
                                 const int* qsys, const double* qt, double* out, int* status, hipStream_t s);
hipError_t launch_ephem_query(const DevEphem* E, int n, const int* body, const double* t, double* out, int* status,
                              hipStream_t s);

}  // namespace assist_gpu
