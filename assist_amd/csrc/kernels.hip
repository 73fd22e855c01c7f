// Batched ASSIST kernels for gfx950.
//
//  * batch_forces_kernel: one force evaluation per system (parity and the
//    assist_additional_forces hook).
//  * ias15_kernel: a persistent integrator.  Each group of G lanes owns one
//    system at a time; lane k of the group holds the IAS15 state of vector k
//    (k = 0 the real particle, k >= 1 its variational particles) in registers.
//    Every trip round the loop performs exactly one force evaluation for every
//    active group, whatever phase of its step the group is in (start-of-step
//    acceleration, predictor-corrector substep n, after a rejection), so the
//    expensive part runs with the whole wave converged.  Finished groups take
//    the next system from an atomic work queue.
//
// IAS15 follows Rein & Spiegel (2015), MNRAS 446, 1424, as configured by ASSIST
// (global step-size criterion, reference src/assist.c:440-446).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <type_traits>

#include "device_types.h"
#include "force_model.h"
#include "ias15_tables.h"
#include "kernels.h"

#include <hipcub/hipcub.hpp>

namespace assist_gpu {

__constant__ double cH[8] = IAS15_H_INIT;
__constant__ double cRRI[28] = IAS15_RRINV_INIT;
__constant__ double cC[21] = IAS15_C_INIT;
__constant__ double cD[21] = IAS15_D_INIT;
__constant__ double cPX[9 * 8] = IAS15_PX_INIT;
__constant__ double cPV[9 * 7] = IAS15_PV_INIT;

// ------------------------------------------------------------------------------------------
// batched force evaluation
// ------------------------------------------------------------------------------------------
template <int G, bool Mixed, bool Ext = false>
__global__ void __launch_bounds__(256) batch_forces_kernel(ForcesArgs a) {
    using TS = typename TermScalar<double, Mixed>::type;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (G - 1);
    const int sys = tid / G;
    if (sys >= a.nsys) return;  // whole groups exit together (nsys*G threads, G | blockDim)
    const int nv = a.nvec;
    const double* st = a.states + (size_t)sys * nv * 6;
    double xr[3], vr[3], A[3] = {0.0, 0.0, 0.0}, dA[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 3; c++) { xr[c] = st[c]; vr[c] = st[3 + c]; }
    bool ng = false;
    if (a.params) {
        const double* pp = a.params + (size_t)sys * nv * 3;
        A[0] = pp[0]; A[1] = pp[1]; A[2] = pp[2];
        ng = !(A[0] == 0.0 && A[1] == 0.0 && A[2] == 0.0);
        if (lane < nv) { dA[0] = pp[3 * lane]; dA[1] = pp[3 * lane + 1]; dA[2] = pp[3 * lane + 2]; }
    }
    EvalResult<double> R;
    Jac<TS> J;
    force_eval<G, Mixed, double, Ext>(*a.E, a.E->slot, a.E->GM, a.fo, lane, a.t[sys], xr, vr, A, ng, nv > 1, R, J);
    // each lane handles vectors lane, lane+G, ... (G may be smaller than nvec for parity runs)
    for (int v = lane; v < nv; v += G) {
        double x[3], vv[3], out[3];
#pragma unroll
        for (int c = 0; c < 3; c++) { x[c] = st[6 * v + c]; vv[c] = st[6 * v + 3 + c]; }
        double dAv[3] = {0.0, 0.0, 0.0};
        if (a.params) {
            const double* pp = a.params + ((size_t)sys * nv + v) * 3;
            dAv[0] = pp[0]; dAv[1] = pp[1]; dAv[2] = pp[2];
        }
        own_accel<double, TS>(v == 0 ? 0 : 1, R, J, x, vv, dAv, ng, out);
#pragma unroll
        for (int c = 0; c < 3; c++) a.acc[((size_t)sys * nv + v) * 3 + c] = out[c];
    }
    if (lane == 0 && a.status) a.status[sys] = R.status;
}

// ------------------------------------------------------------------------------------------
// IAS15
// ------------------------------------------------------------------------------------------
// Kernel-argument pointer (constant address space, so loads through it are scalar loads; the
// host compilation pass only type-checks the kernel body, with a plain pointer).
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(4))) const IntegrateArgs* KArgs;
#else
typedef const IntegrateArgs* KArgs;
#endif
__device__ __forceinline__ KArgs launder(KArgs p) {
    asm volatile("" : "+s"(p));
    return p;
}

// A 7 x 3 array in a lane's LDS column: element (k, c) at p[3k + c].  A column holds b (words
// 0-20) and g (21-41) and is padded to kLdsColWords doubles so that the 64 lanes of a wave,
// reading the same element of their columns, spread over the LDS banks.
struct LdsCol {
    double* p;
    __device__ double& operator[](int c) const { return p[c]; }
};
struct LdsArr {
    double* p;
    __device__ LdsCol operator[](int k) const { return LdsCol{p + 3 * k}; }
};

__device__ __forceinline__ void add_cs(double& p, double& cs, double in) {
    const double y = in - cs;
    const double t = p + y;
    cs = (t - p) - y;
    p = t;
}

// Re-expansion of the last accepted step's b (br) for a step of ratio q:
// e = Q(q) br (Taylor re-centring at h = 1, scaled by q^k); b = e + (br - er).
__device__ __forceinline__ void predict_e(const double br[7], double q, double e[7]) {
    const double q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3, q6 = q3 * q3, q7 = q3 * q4;
    e[0] = q * (br[6] * 7.0 + br[5] * 6.0 + br[4] * 5.0 + br[3] * 4.0 + br[2] * 3.0 + br[1] * 2.0 + br[0]);
    e[1] = q2 * (br[6] * 21.0 + br[5] * 15.0 + br[4] * 10.0 + br[3] * 6.0 + br[2] * 3.0 + br[1]);
    e[2] = q3 * (br[6] * 35.0 + br[5] * 20.0 + br[4] * 10.0 + br[3] * 4.0 + br[2]);
    e[3] = q4 * (br[6] * 35.0 + br[5] * 15.0 + br[4] * 5.0 + br[3]);
    e[4] = q5 * (br[6] * 21.0 + br[5] * 6.0 + br[4]);
    e[5] = q6 * (br[6] * 7.0 + br[5]);
    e[6] = q7 * br[6];
}

// Snap: the kernel carries the checkpoint, pause and step-archive features (the hot shapes are
// also compiled without them, for launches that use none).
template <int G, bool Mixed, int W, bool Ext = false, bool Snap = true>
__global__ void __launch_bounds__(kIas15Block) __attribute__((amdgpu_waves_per_eu(W, W)))
ias15_kernel(IntegrateArgs args) {
    // The arguments are read through the kernel-argument segment pointer, which is
    // re-laundered every loop trip: this keeps them out of SGPRs across the loop (where
    // they would otherwise be spilled to VGPR lanes) at the price of a few scalar loads.
    (void)args;
    KArgs ka = (KArgs)__builtin_amdgcn_kernarg_segment_ptr();
    using TS = typename TermScalar<double, Mixed>::type;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (G - 1);
    const int grp = tid / G;
    const int nv = ka->nvec;
    const bool own = lane < nv;
    // br[7][3], er[7][3] of the last accepted step, per lane, in the scratch buffer
    double* scr = ka->scratch + ((size_t)grp * G + lane) * 42;

    double x0[3] = {2.5, 0.0, 0.0}, v0[3] = {0.0, 0.01, 0.0}, csx[3], csv[3], a0[3];
    // b and g (7 x 3 each per vector) live in LDS, one column per owning lane, to relieve
    // register pressure; lanes that own no vector share one spare column (their values are
    // never used), which is what lets 7-vector systems on 8-lane groups fit two blocks per CU.
    extern __shared__ double s_lds[];
    double* const tab = s_lds;
    double* const s_cols = s_lds + kTabWords;
    for (int i = threadIdx.x; i < kTabWords; i += blockDim.x) {
        double v;
        if (i < kTabPX) v = cH[i - kTabH];
        else if (i < kTabPV) v = cPX[i - kTabPX];
        else if (i < kTabRRI) v = cPV[i - kTabPV];
        else if (i < kTabC) v = cRRI[i - kTabRRI];
        else if (i < kTabW) v = cC[i - kTabC];
        else if (i < kTabW + 49) {
            const int n = (i - kTabW) / 7 + 1, k = (i - kTabW) % 7;
            v = k < n - 1 ? cC[(n - 1) * (n - 2) / 2 + k] : (k == n - 1 ? 1.0 : 0.0);
        } else {
            v = 0.0;  // (padding)
        }
        tab[i] = v;
    }
    // the ephemeris descriptor (slot descriptors, masses and force-model constants), read by
    // every force evaluation, is staged in LDS as well: the slot descriptors are read with a
    // lane-dependent index in front of every record fetch, and the constants' global loads would
    // each wait on the memory latency in the middle of a trip
    static_assert(sizeof(DevEphem) % sizeof(unsigned long long) == 0, "DevEphem is copied in 8-byte words");
    __shared__ DevEphem s_E;
    {
        const unsigned long long* src = (const unsigned long long*)ka->E;
        unsigned long long* dst = (unsigned long long*)&s_E;
        for (int i = threadIdx.x; i < (int)(sizeof(DevEphem) / sizeof(unsigned long long)); i += blockDim.x) dst[i] = src[i];
    }
    const DevSlot* const s_slots = s_E.slot;
    const double* const s_gm = s_E.GM;
    __syncthreads();
    const int cols_per_group = ias15_cols_per_group(G, nv, W);
    const int col = (own || cols_per_group == G) ? (int)(threadIdx.x / G) * cols_per_group + lane
                                                 : (int)(blockDim.x / G) * cols_per_group;
    const LdsArr b{s_cols + (size_t)col * kLdsColWords}, g{s_cols + (size_t)col * kLdsColWords + 21};
    // this wave's exchange area of the wave-wide Chebyshev rounds, after the columns
    double* const s_xch = s_cols + ((size_t)(blockDim.x / G) * cols_per_group + 1) * kLdsColWords +
                          (size_t)(threadIdx.x >> 6) * ias15_xch_words(G, Ext);
    double A[3], dA[3];
    double t = 0, t_end = 0, dt = 0, dts = 0, dt_last = 0, pc_err_last = 0;
    // (phase: false at a step's first evaluation, true in its predictor-corrector substeps)
    bool phase = false;
    int n = 1, it = 0, last = 0, status = 0, attempts = 0;
    long long steps = 0, evals = 0, rejected = 0;
    bool ng = false;
    // the queue hands out positions q; order[q] (when given) is the system at that position.
    // 8-lane groups take their first positions slot-major over the grid's waves (group k of every
    // wave, then group k + 1): a batch smaller than the grid spreads over every wave, and the
    // groups left without a system help their wave's Chebyshev rounds from the first trip
    int q0 = grp;
    if constexpr (ias15_cache_words(G, Ext) > 0) {
        const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
        q0 = (grp & 7) * nwaves + (grp >> 3);
    }
    bool active = q0 < ka->nsys;
    int sys = active ? (ka->order ? ka->order[q0] : q0) : 0;
    int cur = 0, cend = 0;  // dense output: next and end index of this system's output times
    int nrec = 0;           // step archive: snapshots this system has taken
    constexpr bool kWaveRounds = ias15_cache_words(G, Ext) > 0;  // 8-lane groups
    SubstepCache scache{};
    // the next trip is a later iteration at substep nn: load its cached positions now
    auto prefetch = [&](int nn) {
        if constexpr (kWaveRounds) {
            if (ka->ecache) {
                scache.base = ka->ecache + (size_t)(tid >> 6) * (64 * ias15_cache_words(G, Ext));
                ec_prefetch(scache, lane, nn);
            }
        }
    };

    // Snapshot of the current system at a step boundary (layout in device_types.h): each lane
    // writes its own vector's block, lane 0 the header.  Before the first step the b and e
    // blocks are written as zeros (the lane's scratch then holds another system's).
    auto write_snap = [&](double* base) {
        if (own) {
            double* w = base + kSnapHeader + kSnapVector * lane;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                w[c] = x0[c]; w[3 + c] = v0[c]; w[6 + c] = csx[c]; w[9 + c] = csv[c]; w[12 + c] = a0[c];
            }
            const bool stepped = dt_last != 0.0;
            for (int k = 0; k < 42; k++) w[15 + k] = stepped ? scr[k] : 0.0;
        }
        if (lane == 0) {
            base[0] = t; base[1] = dt; base[2] = dt_last;
            base[3] = (double)steps; base[4] = (double)evals; base[5] = (double)rejected;
            base[6] = (double)status; base[7] = (double)G;
        }
    };
    auto archive_snap = [&]() {
        if (Snap && ka->archive) {
            if (nrec < ka->archive_cap)
                write_snap(ka->archive + ((size_t)sys * ka->archive_cap + nrec) * snap_words(nv));
            nrec++;
        }
    };

    auto load = [&](int s) {
        // the system's inputs are loaded first, all together, so that they wait on memory once
        const double* st = ka->states + ((size_t)s * nv + (own ? lane : 0)) * 6;
        double sv[6];
#pragma unroll
        for (int k = 0; k < 6; k++) sv[k] = st[k];
        double pa[3] = {0.0, 0.0, 0.0}, pd[3] = {0.0, 0.0, 0.0};
        if (ka->params) {
            const double* pp = ka->params + (size_t)s * nv * 3;
            const double* pl = pp + 3 * (own ? lane : 0);
#pragma unroll
            for (int c = 0; c < 3; c++) { pa[c] = pp[c]; pd[c] = pl[c]; }
        }
        // (the optional arrays are read through a pointer that falls back on t0[s], which is
        // always there, so that no load sits behind a branch that would wait for it)
        const double t0s = ka->t0[s];
        const bool dense = ka->out_t != nullptr;
        const int* po = dense ? ka->out_off + s : (const int*)(ka->t0 + s);
        const int off0 = po[0], off1 = po[dense ? 1 : 0];
        const double t1s = (dense ? ka->t0 : ka->t1)[s];
        const double dtl = (ka->dt0 ? ka->dt0 : ka->t0)[s];
        const double dt0s = ka->dt0 ? fabs(dtl) : ka->io.dt_initial;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            x0[c] = own ? sv[c] : 0.0;
            v0[c] = own ? sv[3 + c] : 0.0;
            csx[c] = csv[c] = a0[c] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < 7; k++)
#pragma unroll
            for (int c = 0; c < 3; c++) b[k][c] = g[k][c] = 0.0;
        A[0] = A[1] = A[2] = dA[0] = dA[1] = dA[2] = 0.0;
        ng = false;
        if (ka->params) {
            A[0] = pa[0]; A[1] = pa[1]; A[2] = pa[2];
            ng = !(A[0] == 0.0 && A[1] == 0.0 && A[2] == 0.0);
            if (own) { dA[0] = pd[0]; dA[1] = pd[1]; dA[2] = pd[2]; }
        }
        t = t0s;
        if (ka->out_t) {
            cur = off0;
            cend = off1;
            t_end = (cend > cur) ? ka->out_t[cend - 1] : t;
        } else {
            t_end = t1s;
        }
        dt = dt0s;
        if (dt * (t_end - t) < 0) dt = -dt;
        dt_last = 0.0;
        phase = false;
        status = 0;
        steps = evals = rejected = 0;
        last = 0;
        if (Snap && ka->snap_in) {
            // resume: the state, steps and counts of the snapshot, and the last step's b and e
            // for the prediction of the next step (the first evaluation recomputes a0)
            const double* sn = ka->snap_in + (size_t)s * snap_words(nv);
            t = sn[0];
            dt = sn[1];
            dt_last = sn[2];
            steps = (long long)sn[3];
            evals = (long long)sn[4];
            rejected = (long long)sn[5];
            if (dt * (t_end - t) < 0) dt = -dt;
            if (own) {
                const double* w = sn + kSnapHeader + kSnapVector * lane;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    x0[c] = w[c]; v0[c] = w[3 + c]; csx[c] = w[6 + c]; csv[c] = w[9 + c]; a0[c] = w[12 + c];
                }
                for (int k = 0; k < 42; k++) scr[k] = w[15 + k];
            }
        }
        nrec = 0;
        archive_snap();
    };
    auto finish = [&]() {
        if (own) {
            double* st = ka->states + ((size_t)sys * nv + lane) * 6;
#pragma unroll
            for (int c = 0; c < 3; c++) { st[c] = x0[c]; st[3 + c] = v0[c]; }
        }
        if (Snap && ka->snap_out) write_snap(ka->snap_out + (size_t)sys * snap_words(nv));
        if (Snap && lane == 0 && ka->archive_count) ka->archive_count[sys] = nrec;
        if (lane == 0) {
            if (ka->status) ka->status[sys] = status;
            if (ka->steps) ka->steps[sys] = steps;
            if (ka->evals) ka->evals[sys] = evals;
            if (ka->rejected) ka->rejected[sys] = rejected;
            if (ka->t_out) ka->t_out[sys] = t;
        }
        int next = 0;
        if (lane == 0) next = atomicAdd(ka->queue, 1);
        next = gbcast_i<G, 0>(next);
        active = next < ka->nsys;
        if (active) {
            sys = ka->order ? ka->order[next] : next;
            load(sys);
        } else {
            // a helper group evaluates at a fixed main-belt point (its results are discarded): a
            // common point keeps it on the branches the active groups take, and no Marsden branch
            ng = false;
            phase = false;
#pragma unroll
            for (int c = 0; c < 3; c++) { x0[c] = (c == 0) ? 2.5 : 0.0; v0[c] = (c == 1) ? 0.01 : 0.0; }
        }
    };
    // Prepare an attempt of the current step: trim to t_end, predict b from the last
    // accepted step, convert to g.  Returns false when the step cannot be taken.
    auto setup_attempt = [&](bool new_step) -> bool {
        const double dir = (t_end >= t) ? 1.0 : -1.0;
        if (new_step) {
            dts = dt;
            last = 0;
            if (!ka->out_t && (t + dts - t_end) * dir >= 0) { dts = t_end - t; last = 1; }
            attempts = 0;
        }
        if (!(fabs(dts) > 0.0) || !(fabs(dts) < 1e300) || ++attempts > 64) { status = kDtUnderflow; return false; }
        if (dt_last != 0.0 && own) {
            const double q = dts / dt_last;
            if (q > 20.0) {
#pragma unroll
                for (int k = 0; k < 7; k++)
#pragma unroll
                    for (int c = 0; c < 3; c++) b[k][c] = 0.0;
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    double br[7], er[7], e[7];
#pragma unroll
                    for (int k = 0; k < 7; k++) { br[k] = scr[3 * k + c]; er[k] = scr[21 + 3 * k + c]; }
                    predict_e(br, q, e);
#pragma unroll
                    for (int k = 0; k < 7; k++) b[k][c] = e[k] + (br[k] - er[k]);
                }
            }
        } else if (ka->bw && lane == 0) {
            // learned warm start of the first step: b_k = bw_k |a0| dt^(k+1)
            const double* w = ka->bw + (size_t)sys * 21;
            const double amag = sqrt(a0[0] * a0[0] + a0[1] * a0[1] + a0[2] * a0[2]);
            double hp = dts;
#pragma unroll
            for (int k = 0; k < 7; k++) {
#pragma unroll
                for (int c = 0; c < 3; c++) b[k][c] = w[3 * k + c] * amag * hp;
                hp *= dts;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 7; k++)
#pragma unroll
                for (int c = 0; c < 3; c++) b[k][c] = 0.0;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double b0 = b[0][c], b1 = b[1][c], b2 = b[2][c], b3 = b[3][c], b4 = b[4][c], b5 = b[5][c], b6 = b[6][c];
            g[0][c] = b6 * cD[15] + b5 * cD[10] + b4 * cD[6] + b3 * cD[3] + b2 * cD[1] + b1 * cD[0] + b0;
            g[1][c] = b6 * cD[16] + b5 * cD[11] + b4 * cD[7] + b3 * cD[4] + b2 * cD[2] + b1;
            g[2][c] = b6 * cD[17] + b5 * cD[12] + b4 * cD[8] + b3 * cD[5] + b2;
            g[3][c] = b6 * cD[18] + b5 * cD[13] + b4 * cD[9] + b3;
            g[4][c] = b6 * cD[19] + b5 * cD[14] + b4;
            g[5][c] = b6 * cD[20] + b5;
            g[6][c] = b6;
        }
        it = 1;
        pc_err_last = 1e300;
        n = 1;
        phase = true;
        return true;
    };

    if (active) load(sys);
    // 8-lane groups: a group whose queue ran dry stays in the loop as a helper while any group of
    // its wave is active.  It joins each trip's force evaluation up to the wave-wide Chebyshev
    // rounds, where its lanes take items of the groups that need fresh positions (the rounds of
    // a thinned-out wave, e.g. the last long-span systems of an uneven batch, drop from 4 to 1),
    // and leaves the trip right after the evaluation without touching any state or output.
    while (kWaveRounds ? (__ballot(active) != 0ULL) : active) {
        ka = launder(ka);
        const DevEphem& E = s_E;
        // zero-length span (dense output: no output times left)
        if (active && !phase && (ka->out_t ? cur >= cend : t == t_end)) { finish(); continue; }
        // ---- evaluation point of this trip ----
        double xe[3], ve[3], te;
        if (!phase) {
#pragma unroll
            for (int c = 0; c < 3; c++) { xe[c] = x0[c]; ve[c] = v0[c]; }
            te = t;
        } else {
            // predictor weights at spacing h_n (tables: the position and velocity
            // polynomials' coefficients times powers of h_n)
            const double* px = tab + kTabPX + 8 * n;
            const double* pv = tab + kTabPV + 7 * n;
            const double dt2 = dts * dts;
            double s[9];
            s[0] = dts * tab[kTabH + n];
#pragma unroll
            for (int k = 1; k < 9; k++) s[k] = dt2 * px[k - 1];
#pragma unroll
            for (int c = 0; c < 3; c++)
                xe[c] = -csx[c] + ((s[8] * b[6][c] + s[7] * b[5][c] + s[6] * b[4][c] + s[5] * b[3][c] + s[4] * b[2][c] +
                                    s[3] * b[1][c] + s[2] * b[0][c] + s[1] * a0[c] + s[0] * v0[c]) + x0[c]);
#pragma unroll
            for (int k = 1; k < 8; k++) s[k] = dts * pv[k - 1];
#pragma unroll
            for (int c = 0; c < 3; c++)
                ve[c] = -csv[c] + ((s[7] * b[6][c] + s[6] * b[5][c] + s[5] * b[4][c] + s[4] * b[3][c] + s[3] * b[2][c] +
                                    s[2] * b[1][c] + s[1] * b[0][c] + s[0] * a0[c]) + v0[c]);
            te = t + s[0];
        }
        // ---- force evaluation (whole group) ----
        double xr[3], vr[3];
#pragma unroll
        for (int c = 0; c < 3; c++) { xr[c] = gbcast<G, 0>(xe[c]); vr[c] = gbcast<G, 0>(ve[c]); }
        EvalResult<double> R;
        Jac<TS> J;
        const ForceOpts fo = ka->fo;
        // 8-lane groups: wave-wide Chebyshev rounds and the substep ephemeris cache (the first
        // iteration of an attempt stores the positions of its substeps, the later iterations,
        // at the same substep times, use the ones ec_prefetch loaded at the end of the last trip)
        const SubstepCache* scp = nullptr;
        if constexpr (kWaveRounds) {
            scache.base = ka->ecache ? ka->ecache + (size_t)(tid >> 6) * (64 * ias15_cache_words(G, Ext)) : nullptr;
            scache.xch = s_xch;
            scache.n = n;
            scache.mode = !active ? kEcUse : ((!ka->ecache || !phase) ? kEcEval : (it == 1 ? kEcFill : kEcUse));
            scp = &scache;
        }
        force_eval<G, Mixed, double, Ext>(E, s_slots, s_gm, fo, lane, te, xr, vr, A, ng, nv > 1, R, J, scp);
        if (kWaveRounds && !active) continue;
        double at[3];
        own_accel<double, TS>(lane, R, J, xe, ve, dA, ng, at);
        evals++;
        if (R.status != kOk) { status = R.status; finish(); continue; }

        if (!phase) {
#pragma unroll
            for (int c = 0; c < 3; c++) a0[c] = at[c];
            if (!setup_attempt(true)) { finish(); continue; }
            continue;
        }
        // ---- corrector: improve g_{n-1} and b_0..b_{n-1}.  Straight-line code for every
        //      substep (the groups of a wave are at different substeps): each row of b and g is
        //      read from LDS at once and the substep-dependent steps are selected or weighted,
        //      not branched, so the LDS latency is paid once per trip. ----
        double maxak = 0.0, maxdb6 = 0.0;
        {
            const double* rri = tab + kTabRRI + n * (n - 1) / 2;  // 1 / (h_n - h_j), j = 0..n-1
            const double* wn = tab + kTabW + 7 * (n - 1);          // b_k += wn[k] dg_{n-1}
            double r[6], w[7];
#pragma unroll
            for (int k = 0; k < 6; k++) r[k] = rri[k < n ? k : 0];
#pragma unroll
            for (int k = 0; k < 7; k++) w[k] = wn[k];
            const double rl = rri[n - 1];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double gv[6], bv[7];
#pragma unroll
                for (int k = 0; k < 6; k++) gv[k] = g[k][c];
#pragma unroll
                for (int k = 0; k < 7; k++) bv[k] = b[k][c];
                const double gold = g[n - 1][c];
                // ((gk r_0 - g_0) r_1 - g_1) r_2 ..., each step one fused multiply-add
                double val = at[c] - a0[c];
#pragma unroll
                for (int k = 0; k < 6; k++) val = (k < n - 1) ? fma(val, r[k], -gv[k]) : val;
                val *= rl;
                const double tmp = val - gold;
                g[n - 1][c] = val;
                // b_k + w_k tmp: the weights below k = n - 1, then b_{n-1} + tmp (weight 1), and
                // the rows above unchanged (weight 0)
#pragma unroll
                for (int k = 0; k < 7; k++) b[k][c] = fma(tmp, w[k], bv[k]);
                if (n == 7 && own) {
                    const double ak = fabs(at[c]);
                    if (is_normal(ak) && ak > maxak) maxak = ak;
                    const double db = fabs(tmp);
                    if (is_normal(db) && db > maxdb6) maxdb6 = db;
                }
            }
        }
        if (n < 7) {
            n++;
            if (it >= 2) prefetch(n);
            continue;
        }
        // ---- end of a predictor-corrector iteration ----
        // IAS15's test takes every vector of the system to FP64 round-off; with a tolerance the
        // real particle's own change decides (the variational vectors take the same iterations),
        // and with a first-pass tolerance an attempt whose first pass changed b6 by less ends there
        // (the tolerance shrinks by pc_aref / max|a|, down to a tenth, where the acceleration
        // exceeds pc_aref)
        const double pc_tol = ka->io.pc_tol;
        const double pc_mak = pc_tol > 0.0 ? gbcast<G, 0>(maxak) : gmax<G>(maxak);
        const double pc_err = pc_tol > 0.0 ? gbcast<G, 0>(maxdb6 / maxak) : gmax<G>(maxdb6) / pc_mak;
        const double pc_aref = ka->io.pc_aref;
        const double pc_tol1 = (pc_aref > 0.0 && pc_mak > pc_aref) ? ka->io.pc_tol1 * fmax(0.1, pc_aref / pc_mak) : ka->io.pc_tol1;
        if (!(pc_err < (pc_tol > 0.0 ? pc_tol : 1e-16)) && !(it == 1 && pc_err < pc_tol1) &&
            !(it > 2 && pc_err_last <= pc_err) && it < 12) {
            pc_err_last = pc_err;
            it++;
            n = 1;
            prefetch(1);
            continue;
        }
        // ---- step-size control (real particle only) ----
        double dt_new = dts, dt_prop = dts;
        if (ka->io.epsilon > 0) {
            double mak = 0.0, mb6 = 0.0;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double ak = fabs(at[c]);
                if (is_normal(ak) && ak > mak) mak = ak;
                const double bk = fabs(b[6][c]);
                if (is_normal(bk) && bk > mb6) mb6 = bk;
            }
            const double err = gbcast<G, 0>(mb6 / mak);
            if (is_normal(err)) dt_new = pow(ka->io.epsilon / err, 1.0 / 7.0) * dts;
            else dt_new = dts / 0.25;
            if (fabs(dt_new) < ka->io.min_dt) dt_new = copysign(ka->io.min_dt, dt_new);
            if (fabs(dt_new / dts) < 0.25) {
                rejected++;
                dts = dt_new;
                last = 0;
                if (!setup_attempt(false)) finish();
                continue;
            }
            dt_prop = dt_new;
            if (dt_new / dts > 1. / 0.25) dt_new = dts / 0.25;
        }
        // ---- accept: advance x0, v0 (Kahan) and store br, er ----
        if (ka->trace && lane == 0 && steps >= ka->trace_skip && steps % ka->trace_every == 0) {
            const int idx = atomicAdd(ka->trace_count, 1);
            if (idx < ka->trace_cap) {
                double* rec = ka->trace + (size_t)idx * kTraceWords;
                rec[0] = t;
#pragma unroll
                for (int c = 0; c < 3; c++) { rec[1 + c] = x0[c]; rec[4 + c] = v0[c]; rec[10 + c] = a0[c]; }
                rec[7] = dt_last;
                rec[8] = dts;
                rec[9] = dt_prop;
#pragma unroll
                for (int k = 0; k < 7; k++)
#pragma unroll
                    for (int c = 0; c < 3; c++) rec[13 + 3 * k + c] = b[k][c];
            }
        }
        if (ka->out_t) {
            // dense output: every output time in [t, t + dts) from this step's polynomial, as
            // reference src/assist.c:556-597 (assist_interpolate)
            while (cur < cend) {
                double h = (ka->out_t[cur] - t) / dts;
                if (!(h < 1.0)) break;
                if (h < 0.0) h = 0.0;
                if (own) {
                    double s[9], sv[8];
                    s[0] = dts * h;
                    s[1] = s[0] * s[0] / 2.;
                    s[2] = s[1] * h / 3.;
                    s[3] = s[2] * h / 2.;
                    s[4] = 3. * s[3] * h / 5.;
                    s[5] = 2. * s[4] * h / 3.;
                    s[6] = 5. * s[5] * h / 7.;
                    s[7] = 3. * s[6] * h / 4.;
                    s[8] = 7. * s[7] * h / 9.;
                    sv[0] = dts * h;
                    sv[1] = sv[0] * h / 2.;
                    sv[2] = 2. * sv[1] * h / 3.;
                    sv[3] = 3. * sv[2] * h / 4.;
                    sv[4] = 4. * sv[3] * h / 5.;
                    sv[5] = 5. * sv[4] * h / 6.;
                    sv[6] = 6. * sv[5] * h / 7.;
                    sv[7] = 7. * sv[6] * h / 8.;
                    double* o = ka->out_states + ((size_t)cur * nv + lane) * 6;
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        o[c] = x0[c] + (s[8] * b[6][c] + s[7] * b[5][c] + s[6] * b[4][c] + s[5] * b[3][c] + s[4] * b[2][c] +
                                        s[3] * b[1][c] + s[2] * b[0][c] + s[1] * a0[c] + s[0] * v0[c]);
                        o[3 + c] = v0[c] + sv[7] * b[6][c] + sv[6] * b[5][c] + sv[5] * b[4][c] + sv[4] * b[3][c] +
                                   sv[3] * b[2][c] + sv[2] * b[1][c] + sv[1] * b[0][c] + sv[0] * a0[c];
                    }
                }
                cur++;
            }
        }
        // end-of-step weights: the predictor at h = 1 (row 8 of the tables)
        double wx[8], wv[7];
        {
            const double dt2 = dts * dts;
#pragma unroll
            for (int k = 0; k < 8; k++) wx[k] = cPX[64 + k] * dt2;
#pragma unroll
            for (int k = 0; k < 7; k++) wv[k] = cPV[56 + k] * dts;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            add_cs(x0[c], csx[c], b[6][c] * wx[7]);
            add_cs(x0[c], csx[c], b[5][c] * wx[6]);
            add_cs(x0[c], csx[c], b[4][c] * wx[5]);
            add_cs(x0[c], csx[c], b[3][c] * wx[4]);
            add_cs(x0[c], csx[c], b[2][c] * wx[3]);
            add_cs(x0[c], csx[c], b[1][c] * wx[2]);
            add_cs(x0[c], csx[c], b[0][c] * wx[1]);
            add_cs(x0[c], csx[c], a0[c] * wx[0]);
            add_cs(x0[c], csx[c], v0[c] * dts);
            add_cs(v0[c], csv[c], b[6][c] * wv[6]);
            add_cs(v0[c], csv[c], b[5][c] * wv[5]);
            add_cs(v0[c], csv[c], b[4][c] * wv[4]);
            add_cs(v0[c], csv[c], b[3][c] * wv[3]);
            add_cs(v0[c], csv[c], b[2][c] * wv[2]);
            add_cs(v0[c], csv[c], b[1][c] * wv[1]);
            add_cs(v0[c], csv[c], b[0][c] * wv[0]);
            add_cs(v0[c], csv[c], a0[c] * dts);
        }
        if (own) {
            const double q = (dt_last != 0.0) ? dts / dt_last : 0.0;
            // the last step's b, loaded at once (one wait instead of one per component: the
            // stores below would otherwise keep the later components' loads behind them)
            double bl[21];
#pragma unroll
            for (int k = 0; k < 21; k++) bl[k] = scr[k];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double br[7], e[7];
#pragma unroll
                for (int k = 0; k < 7; k++) br[k] = bl[3 * k + c];
                if (dt_last != 0.0 && q <= 20.0) predict_e(br, q, e);
                else
#pragma unroll
                    for (int k = 0; k < 7; k++) e[k] = 0.0;
#pragma unroll
                for (int k = 0; k < 7; k++) { scr[21 + 3 * k + c] = e[k]; scr[3 * k + c] = b[k][c]; }
            }
        }
        t = last ? t_end : t + dts;
        steps++;
        dt_last = dts;
        if (!(last && ka->io.epsilon <= 0)) dt = dt_new;
        int bad = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) bad |= !(fabs(x0[c]) < 1e300) || !(fabs(v0[c]) < 1e300);
        bad = gor<G>(own ? bad : 0);
        if (bad) { status = kNonFinite; finish(); continue; }
        archive_snap();
        if (last || (ka->out_t && cur >= cend)) { finish(); continue; }
        if (Snap && ka->t_pause && (t - ka->t_pause[sys]) * dt_last >= 0) { status = kPaused; finish(); continue; }
        if (steps >= ka->io.max_steps) { status = kMaxSteps; finish(); continue; }
        phase = false;
    }
}

// ------------------------------------------------------------------------------------------
// learned start: features and post-processing
// ------------------------------------------------------------------------------------------
// Row r: t = rows[r*stride], x = rows[+1..3], v = rows[+4..6], dt_last = rows[+7].
// Features: heliocentric position / 3 au, heliocentric velocity * 40 d/au, log10 of the three
// smallest distances to the 11 planetary bodies (Sun, planets, Moon, Pluto), log10 dt_last.
__global__ void features_kernel(const DevEphem* __restrict__ Ep, int n, const double* __restrict__ rows, int stride,
                                float* __restrict__ X, double* __restrict__ omega) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const DevEphem& E = *Ep;
    const double* row = rows + (size_t)r * stride;
    const double t = row[0];
    double pos[12][3];
    double sunv[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
    for (int s = 0; s < kFirstAsteroidSlot; s++) {
        const ChebOut<double> c = (s == 0) ? cheb_eval<double, true>(E.slot[s], t) : cheb_eval<double, false>(E.slot[s], t);
        pos[s][0] = c.x; pos[s][1] = c.y; pos[s][2] = c.z;
        if (s == 0) { sunv[0] = c.vx * E.inv_au; sunv[1] = c.vy * E.inv_au; sunv[2] = c.vz * E.inv_au; }
    }
    double d[3] = {1e300, 1e300, 1e300};
    double xs[3];
#pragma unroll 1
    for (int s = 0; s < kFirstAsteroidSlot; s++) {
        if (s == kSlotEMB) continue;
        double p[3];
        const bool em = (s == kSlotEarth || s == kSlotMoon);
        for (int c = 0; c < 3; c++) p[c] = (em ? pos[s][c] + pos[kSlotEMB][c] : pos[s][c]) * E.inv_au;
        if (s == 0) for (int c = 0; c < 3; c++) xs[c] = p[c];
        const double dx = row[1] - p[0], dy = row[2] - p[1], dz = row[3] - p[2];
        const double dist = sqrt(dx * dx + dy * dy + dz * dz);
        if (dist < d[0]) { d[2] = d[1]; d[1] = d[0]; d[0] = dist; }
        else if (dist < d[1]) { d[2] = d[1]; d[1] = dist; }
        else if (dist < d[2]) d[2] = dist;
    }
    float* x = X + (size_t)r * 16;
    double h[3];
    for (int c = 0; c < 3; c++) {
        h[c] = row[1 + c] - xs[c];
        x[c] = (float)(h[c] / 3.0);
        x[3 + c] = (float)((row[4 + c] - sunv[c]) * 40.0);
    }
    for (int k = 0; k < 3; k++) x[6 + k] = (float)log10(fmax(d[k], 1e-12));
    x[9] = (float)log10(fmax(fabs(row[7]), 1e-6));
    for (int k = 10; k < 16; k++) x[k] = 0.f;
    if (omega) {
        const double rh = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
        omega[r] = sqrt(E.GM[0] / (rh * rh * rh));
    }
}

__global__ void start_rows_kernel(int n, int nvec, const double* __restrict__ t0, const double* __restrict__ states,
                                  double dt_last, double* __restrict__ rows) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double* row = rows + (size_t)r * 8;
    const double* st = states + (size_t)r * nvec * 6;
    row[0] = t0[r];
    for (int c = 0; c < 6; c++) row[1 + c] = st[c];
    row[7] = dt_last;
}

constexpr double kWarmScale = 0.175;

__global__ void start_post_kernel(int n, const float* __restrict__ Y, const double* __restrict__ t0, const double* __restrict__ t1, double dt_min,
                                  double* __restrict__ rows, double* __restrict__ dt0, double* __restrict__ bw) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float* y = Y + (size_t)r * 32;
    double dt = pow(10.0, (double)fmin(fmax(y[0], -4.f), 3.f));
    const double span = fabs(t1[r] - t0[r]);
    if (span > 0.0 && dt > span) dt = span;
    if (dt < dt_min) dt = dt_min;
    dt0[r] = dt;
    rows[(size_t)r * 8 + 7] = dt;
    if (bw) {
        // b_k = y_k |a0| (c dt_k / dt)^(k+1) / (k+1)! for a first step dt_k of either sign, with
        // c = (7! eps)^(1/7) for eps = 1e-9 (the scale at which IAS15 settles; see learned.py)
        const double w = kWarmScale / dt;
        double wp = w, fact = 1.0;
        for (int k = 0; k < 7; k++) {
            fact *= (double)(k + 1);
            for (int c = 0; c < 3; c++) bw[(size_t)r * 21 + 3 * k + c] = (double)y[1 + 3 * k + c] * wp / fact;
            wp *= w;
        }
    }
}

hipError_t launch_features(const DevEphem* E, int n, const double* rows, int stride, float* X, double* omega,
                           hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(features_kernel, dim3((n + 127) / 128), dim3(128), 0, s, E, n, rows, stride, X, omega);
    return hipGetLastError();
}

hipError_t launch_start_rows(int n, int nvec, const double* t0, const double* states, double dt_last, double* rows,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(start_rows_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, nvec, t0, states, dt_last, rows);
    return hipGetLastError();
}

hipError_t launch_start_post(int n, const float* Y, const double* t0, const double* t1,
                             double dt_min, double* rows, double* dt0, double* bw, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(start_post_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, Y, t0, t1, dt_min, rows,
                       dt0, bw);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Batched ephemeris queries (reference src/forces.c:175-263, assist_all_ephem, without its cache):
// query q asks for ASSIST body body[q] at time t[q] (days from jd_ref) and receives
// GM, x, y, z, vx, vy, vz, ax, ay, az in au, au/d, au/d^2; asteroids get the barycentric
// position (heliocentric + the Sun's) and NaN velocity and acceleration, as in the reference.
// One thread per query.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) ephem_query_kernel(const DevEphem* __restrict__ Ep, int n,
                                                          const int* __restrict__ body, const double* __restrict__ t,
                                                          double* __restrict__ out, int* __restrict__ status) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const DevEphem& E = *Ep;
    const int b = body[q];
    const double tq = t[q];
    double* o = out + (size_t)q * 10;
    int st = kOk;
    // ASSIST_ERROR_NEPHEM (4) for a negative body, ASSIST_ERROR_AST_FILE (2) for an asteroid without
    // an asteroid file, ASSIST_ERROR_NAST (3) past the file's asteroids
    if (b < 0) st = 4;
    else if (b >= kNumPlanets && E.n_ast == 0) st = 2;
    else if (b >= kNumPlanets + E.n_ast) st = 3;
    // a planet needs its own target's coverage; an asteroid its own and the Sun's, which makes
    // its position barycentric (reference src/forces.c:213-217, src/spk.c:417-419)
    else if (!(tq >= E.t_beg && tq <= E.t_end)) st = kCoverage;
    else if (b >= kNumPlanets && !(tq >= E.ta_beg && tq <= E.ta_end)) st = kCoverage;
    if (st != kOk) {
        for (int k = 0; k < 10; k++) o[k] = 0.0;
        if (status) status[q] = st;
        return;
    }
    double r[9];
    if (b < kNumPlanets) {
        const int s = b < 3 ? b : b + 1;  // body -> slot (the EMB's slot 3 is skipped)
        const ChebOut<double> c = cheb_eval<double, true, true>(E.slot[s], tq);
        r[0] = c.x; r[1] = c.y; r[2] = c.z; r[3] = c.vx; r[4] = c.vy; r[5] = c.vz; r[6] = c.ax; r[7] = c.ay; r[8] = c.az;
        if (s == kSlotEarth || s == kSlotMoon) {
            const ChebOut<double> e = cheb_eval<double, true, true>(E.slot[kSlotEMB], tq);
            r[0] += e.x; r[1] += e.y; r[2] += e.z; r[3] += e.vx; r[4] += e.vy; r[5] += e.vz;
            r[6] += e.ax; r[7] += e.ay; r[8] += e.az;
        }
        for (int k = 0; k < 9; k++) o[1 + k] = r[k] * E.inv_au;
    } else {
        const ChebOut<double> c = cheb_eval<double, false>(E.slot[kFirstAsteroidSlot + b - kNumPlanets], tq);
        const ChebOut<double> sun = cheb_eval<double, false>(E.slot[0], tq);
        o[1] = c.x * E.inv_au_ast + sun.x * E.inv_au;
        o[2] = c.y * E.inv_au_ast + sun.y * E.inv_au;
        o[3] = c.z * E.inv_au_ast + sun.z * E.inv_au;
        for (int k = 4; k < 10; k++) o[k] = __longlong_as_double(0x7ff8000000000000LL);
    }
    o[0] = E.GM[b];
    if (status) status[q] = kOk;
}

hipError_t launch_ephem_query(const DevEphem* E, int n, const int* body, const double* t, double* out, int* status,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ephem_query_kernel, dim3((n + 255) / 256), dim3(256), 0, s, E, n, body, t, out, status);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Interpolation in a step archive (the counterpart of reference src/assist.c:599-633, 682-752:
// the state at t between two saved step boundaries, from the earlier one's state and the later
// one's acceleration at the start of the step, b coefficients and step size).  One thread per
// (query, vector).
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) archive_interp_kernel(int nsys, int nq, int nvec, int cap, const double* __restrict__ archive,
                                                             const int* __restrict__ count, const int* __restrict__ qsys,
                                                             const double* __restrict__ qt, double* __restrict__ out,
                                                             int* __restrict__ status) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long long)nq * nvec) return;
    const int q = (int)(tid / nvec), v = (int)(tid % nvec);
    const int sw = snap_words(nvec);
    const int s = qsys[q];
    const double tq = qt[q];
    double* o = out + ((size_t)q * nvec + v) * 6;
    if (s < 0 || s >= nsys) {
        for (int c = 0; c < 6; c++) o[c] = 0.0;
        if (v == 0 && status) status[q] = kBadSystem;
        return;
    }
    int m = count[s];
    if (m > cap) m = cap;
    const double* rec0 = archive + (size_t)s * cap * sw;
    // the last record at or before tq in the direction of integration (records are in step order)
    int k = -1;
    if (m >= 1) {
        const double dir = (m >= 2 && rec0[(size_t)sw + 2] < 0.0) ? -1.0 : 1.0;
        if ((tq - rec0[0]) * dir >= 0.0) {
            int lo = 0, hi = m - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if ((tq - rec0[(size_t)mid * sw]) * dir >= 0.0) lo = mid;
                else hi = mid - 1;
            }
            k = lo;
        }
    }
    const double* a = k >= 0 ? rec0 + (size_t)k * sw : nullptr;
    int st = kOk;
    if (!a || (k == m - 1 && tq != a[0])) st = kCoverage;
    if (st != kOk) {
        for (int c = 0; c < 6; c++) o[c] = 0.0;
        if (v == 0 && status) status[q] = st;
        return;
    }
    const double* xa = a + kSnapHeader + kSnapVector * v;
    if (tq == a[0]) {
        for (int c = 0; c < 6; c++) o[c] = xa[c];
    } else {
        const double* bh = a + sw;  // the next record: the step from a to it
        const double* xb = bh + kSnapHeader + kSnapVector * v;
        const double dts = bh[2];
        const double h = (tq - a[0]) / dts;
        const double* a0 = xb + 12;
        const double* b = xb + 15;
        double s_[9], sv[8];
        s_[0] = dts * h;
        s_[1] = s_[0] * s_[0] / 2.;
        s_[2] = s_[1] * h / 3.;
        s_[3] = s_[2] * h / 2.;
        s_[4] = 3. * s_[3] * h / 5.;
        s_[5] = 2. * s_[4] * h / 3.;
        s_[6] = 5. * s_[5] * h / 7.;
        s_[7] = 3. * s_[6] * h / 4.;
        s_[8] = 7. * s_[7] * h / 9.;
        sv[0] = dts * h;
        sv[1] = sv[0] * h / 2.;
        sv[2] = 2. * sv[1] * h / 3.;
        sv[3] = 3. * sv[2] * h / 4.;
        sv[4] = 4. * sv[3] * h / 5.;
        sv[5] = 5. * sv[4] * h / 6.;
        sv[6] = 6. * sv[5] * h / 7.;
        sv[7] = 7. * sv[6] * h / 8.;
        for (int c = 0; c < 3; c++) {
            o[c] = xa[c] + (s_[8] * b[18 + c] + s_[7] * b[15 + c] + s_[6] * b[12 + c] + s_[5] * b[9 + c] +
                            s_[4] * b[6 + c] + s_[3] * b[3 + c] + s_[2] * b[c] + s_[1] * a0[c] + s_[0] * xa[3 + c]);
            o[3 + c] = xa[3 + c] + sv[7] * b[18 + c] + sv[6] * b[15 + c] + sv[5] * b[12 + c] + sv[4] * b[9 + c] +
                       sv[3] * b[6 + c] + sv[2] * b[3 + c] + sv[1] * b[c] + sv[0] * a0[c];
        }
    }
    if (v == 0 && status) status[q] = kOk;
}

hipError_t launch_archive_interp(int nsys, int nq, int nvec, int cap, const double* archive, const int* count,
                                 const int* qsys, const double* qt, double* out, int* status, hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    const long long threads = (long long)nq * nvec;
    hipLaunchKernelGGL(archive_interp_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, nsys, nq, nvec,
                       cap, archive, count, qsys, qt, out, status);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Queue order of large batches: systems sorted by expected length, longest first, so that the
// persistent kernel starts the long integrations early and its tail is not set by one started late.
// ------------------------------------------------------------------------------------------
// Queue keys: the span |t1 - t0| first, then 1/q, q the perihelion distance (au) of the real
// particle's osculating orbit about the solar-system barycentre with the Sun's GM (the Sun sits
// within 0.01 au of the barycentre).  IAS15 (eps 1e-9) takes about 7 + 0.15 |t1 - t0| / q steps on
// main-belt and NEO-like orbits (a fit over 800-10 000-day spans, 11% rms in log), but a close
// approach can multiply that, so the longest spans go first whatever their orbits, and among
// equal spans the smallest perihelia.  The key is the span's bit pattern with its low 20 bits
// (below 2^-32 of the span) replaced by min(1/q, 1000) x 1000.  Only the order of the queue
// depends on the key.
__global__ void span_keys_kernel(const DevEphem* __restrict__ E, int n, int nvec, const double* __restrict__ t0,
                                 const double* __restrict__ t1, const double* __restrict__ states,
                                 unsigned long long* __restrict__ keys, int* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double span = fabs(t1[i] - t0[i]);
    const double* st = states + (size_t)i * nvec * 6;
    const double mu = E->GM[0];
    const double x = st[0], y = st[1], z = st[2], vx = st[3], vy = st[4], vz = st[5];
    const double r = sqrt(x * x + y * y + z * z);
    const double v2 = vx * vx + vy * vy + vz * vz;
    const double hx = y * vz - z * vy, hy = z * vx - x * vz, hz = x * vy - y * vx;
    const double h2 = hx * hx + hy * hy + hz * hz;
    const double inv_a = 2.0 / r - v2 / mu;
    double q = r;
    if (inv_a > 0.0) {
        const double e2 = fmax(0.0, 1.0 - h2 * inv_a / mu);
        q = (1.0 - sqrt(e2)) / inv_a;
    }
    const double rate = (q == q) ? fmin(1.0 / fmax(q, 1e-3), 1000.0) : 0.0;
    const unsigned long long low = (unsigned long long)(rate * 1000.0);  // < 2^20
    // non-negative doubles sort like their bit patterns; NaN sorts last
    keys[i] = (span == span) ? (((unsigned long long)__double_as_longlong(span) & ~0xFFFFFull) | low) : 0ull;
    vals[i] = i;
}

size_t span_order_temp_bytes(int n) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, (const unsigned long long*)nullptr,
                                                       (unsigned long long*)nullptr, (const int*)nullptr,
                                                       (int*)nullptr, n, 0, 64);
    return bytes;
}

hipError_t launch_span_order(const DevEphem* E, int n, int nvec, const double* t0, const double* t1,
                             const double* states, unsigned long long* keys, unsigned long long* keys_alt, int* vals,
                             int* order, void* temp, size_t temp_bytes, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(span_keys_kernel, dim3((n + 255) / 256), dim3(256), 0, s, E, n, nvec, t0, t1, states, keys,
                       vals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, keys, keys_alt, vals, order, n, 0, 64, s);
}

// ------------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------------
template <int G, bool M, bool X = false>
static hipError_t launch_forces_t(const ForcesArgs& a, hipStream_t s) {
    const int threads = 256;
    const long long total = (long long)a.nsys * G;
    const int blocks = (int)((total + threads - 1) / threads);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((batch_forces_kernel<G, M, X>), dim3(blocks), dim3(threads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_forces(const ForcesArgs& a, int G, bool mixed, hipStream_t s) {
    if (ext_model(a.fo)) {
        if (mixed) return hipErrorInvalidValue;
        switch (G) {
        case 1: return launch_forces_t<1, false, true>(a, s);
        case 2: return launch_forces_t<2, false, true>(a, s);
        case 4: return launch_forces_t<4, false, true>(a, s);
        case 8: return launch_forces_t<8, false, true>(a, s);
        default: return hipErrorInvalidValue;
        }
    }
    switch (G * 2 + (mixed ? 1 : 0)) {
    case 2: return launch_forces_t<1, false>(a, s);
    case 3: return launch_forces_t<1, true>(a, s);
    case 4: return launch_forces_t<2, false>(a, s);
    case 5: return launch_forces_t<2, true>(a, s);
    case 8: return launch_forces_t<4, false>(a, s);
    case 9: return launch_forces_t<4, true>(a, s);
    case 16: return launch_forces_t<8, false>(a, s);
    case 17: return launch_forces_t<8, true>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template <int G, bool M, int W, bool X = false, bool S = true>
static const void* ias15_fn() {
    return (const void*)ias15_kernel<G, M, W, X, S>;
}

static const void* ias15_lookup(int G, bool mixed, int waves, bool ext, bool snap) {
    if (!snap && !ext && waves == 1) {  // the hot shapes, without the checkpoint features
        switch (G * 2 + (mixed ? 1 : 0)) {
        case 2: return ias15_fn<1, false, 1, false, false>();
        case 3: return ias15_fn<1, true, 1, false, false>();
        case 16: return ias15_fn<8, false, 1, false, false>();
        case 17: return ias15_fn<8, true, 1, false, false>();
        default: break;
        }
    }
    if (ext) {  // extended force model: FP64, 1-wave variant
        if (mixed || waves != 1) return nullptr;
        switch (G) {
        case 1: return ias15_fn<1, false, 1, true>();
        case 2: return ias15_fn<2, false, 1, true>();
        case 4: return ias15_fn<4, false, 1, true>();
        case 8: return ias15_fn<8, false, 1, true>();
        default: return nullptr;
        }
    }
    const int key = (G * 2 + (mixed ? 1 : 0)) * 4 + waves;
    switch (key) {
#define CASE(g, m, w) case (g * 2 + m) * 4 + w: return ias15_fn<g, (bool)m, w>();
    CASE(1, 0, 1) CASE(1, 1, 1) CASE(2, 0, 1) CASE(2, 1, 1) CASE(4, 0, 1) CASE(4, 1, 1) CASE(8, 0, 1) CASE(8, 1, 1)
    CASE(1, 0, 2) CASE(1, 1, 2) CASE(2, 0, 2) CASE(2, 1, 2) CASE(4, 0, 2) CASE(4, 1, 2) CASE(8, 0, 2) CASE(8, 1, 2)
#undef CASE
    default: return nullptr;
    }
}

size_t ias15_lds_bytes(int G, int nvec, int threads, int waves, bool ext) {
    return (kTabWords + ((size_t)(threads / G) * ias15_cols_per_group(G, nvec, waves) + 1) * kLdsColWords +
            (size_t)(threads / 64) * ias15_xch_words(G, ext)) *
           sizeof(double);
}

hipError_t launch_ias15(const IntegrateArgs& a, int G, bool mixed, int waves, int blocks, int threads, hipStream_t s,
                        bool set_attributes) {
    const void* f = ias15_lookup(G, mixed, waves, ext_model(a.fo), ias15_snap(a));
    if (!f) return hipErrorInvalidValue;
    const size_t lds = ias15_lds_bytes(G, a.nvec, threads, waves, ext_model(a.fo));
    hipError_t e = hipSuccess;
    if (set_attributes) {  // (set beforehand by ias15_occupancy when the launch is being captured)
        e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    void* params[] = {(void*)&a};
    e = hipLaunchKernel(f, dim3(blocks), dim3(threads), params, lds, s);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t ias15_occupancy(int G, bool mixed, int waves, bool ext, bool snap, int nvec, int threads,
                           int* blocks_per_cu) {
    const void* f = ias15_lookup(G, mixed, waves, ext, snap);
    if (!f) return hipErrorInvalidValue;
    const size_t lds = ias15_lds_bytes(G, nvec, threads, waves, ext);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, threads, lds);
}

// ---- scratch of the allocation-free integration entry point -------------------------------
// A file-scope device array: every GPU that runs this library holds its own copy until the
// process ends.  The work-queue counter sits in its first 256 bytes, then the b/e records (42
// doubles per lane) of at most kNoAllocLanes resident lanes.  Host code never passes the array's
// name to a kernel (on the host it names a stand-in, not device memory): it passes the device
// address that hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocScratchWords = kNoAllocHeadWords + (size_t)kNoAllocLanes * 42;
static __device__ __align__(256) double scratch[kNoAllocScratchWords];

hipError_t noalloc_scratch(double** ptr) {
    // one address per device, looked up once with that device current
    static std::mutex m;
    static double* addr[kNoAllocMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(m);
    if (!addr[dev]) {
        double* p = nullptr;
        e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    return hipSuccess;
}

// Sets the work-queue counter (a kernel, so the sequence can be captured in a graph).
__global__ void set_counter_kernel(int* q, int v) { *q = v; }

hipError_t launch_set_counter(int* q, int v, hipStream_t s) {
    hipLaunchKernelGGL(set_counter_kernel, dim3(1), dim3(1), 0, s, q, v);
    return hipGetLastError();
}

}  // namespace assist_gpu
