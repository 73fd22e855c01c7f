// The ASSIST force model on the GPU, evaluated cooperatively by a group of G
// lanes (G = 1, 2, 4 or 8) that all work on one system (one real particle and
// its variational particles) at one time t.
//
//   1. Chebyshev evaluation of the 28 ephemeris targets, slots spread over the
//      G lanes (reference src/spk.c:405-547, 550-610).
//   2. Composition: Earth and Moon = EMB + offset, asteroids = heliocentric +
//      Sun (reference src/spk.c:572-587, src/forces.c:213-219).
//   3. Per-body work on the lane that evaluated the body: point-mass
//      acceleration, its Jacobian, and the EIH sums over the planets
//      (reference src/forces.c:266-433, 1319-1501).
//   4. A butterfly sum over the group, after which every lane holds identical
//      totals.
//   5. The single-particle terms (non-gravitational, Earth J2/J3/J4, solar J2,
//      EIH, simple/potential GR), evaluated redundantly on every lane (the
//      lanes of a wave execute one instruction stream, so splitting them over
//      lanes would serialise anyway) (reference src/forces.c:435-1983).
//
// The variational equations are linear in the variational state, so instead
// of applying each term's partials to each variational particle the way the
// reference does, the terms add their partials into one 3x9 Jacobian
// [d a/d x | d a/d v | d a/d A] that the variational lanes apply once.  The
// real particle's acceleration is summed in the reference's term order
// (NG, Earth harmonics, solar J2, EIH, GR potential, GR simple, direct).
//
// Mixed precision (Mixed = true): Chebyshev evaluation, the Sun/planet/Earth/
// Moon point masses and everything in IAS15 stay in FP64; the 16 asteroid
// point masses, J3/J4, GR, non-gravitational forces and every Jacobian run in
// FP32.
#pragma once
#include <math.h>

#include <type_traits>

#include "device_types.h"

// Host-and-device qualifier: the kernels use these functions on the GPU; the
// operation-count tool instantiates them on the host with a counting scalar.
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AHD __host__ __device__ __forceinline__
#else
#define AHD inline
#endif

namespace assist_gpu {

enum : int {
    F_SUN = 0x01, F_PLANETS = 0x02, F_ASTEROIDS = 0x04, F_NONGRAV = 0x08, F_EARTH_H = 0x10,
    F_SUN_H = 0x20, F_GR_EIH = 0x40, F_GR_SIMPLE = 0x80, F_GR_POTENTIAL = 0x100
};

// ---- lane-group collectives ----------------------------------------------------------
// Butterfly collectives over the G lanes of a group (G = 1: identity), on DPP lane
// permutations within a row: quad_perm [1,0,3,2] pairs lanes i and i^1, quad_perm [2,3,0,1]
// pairs i and i^2, row_half_mirror pairs lane i of an 8-lane half-row with lane 7-i (after the
// first two steps every lane of a quad holds the quad total, so this adds the other quad's).
// Floating-point addition is commutative, so every lane ends with bit-identical totals.
#if defined(__HIP_DEVICE_COMPILE__)
enum : int { kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141 };
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL>((int)(b & 0xffffffffLL));
    const int hi = dpp_i<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
#endif

template <int G, typename T>
AHD T gsum(T v) {
    static_assert(G == 1 || G == 2 || G == 4 || G == 8, "group size");
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (G >= 2) v += dpp<kDppXor1>(v);
    if constexpr (G >= 4) v += dpp<kDppXor2>(v);
    if constexpr (G >= 8) v += dpp<kDppHalfMirror>(v);
#endif
    return v;
}
template <int G>
AHD double gmax(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (G >= 2) v = fmax(v, dpp<kDppXor1>(v));
    if constexpr (G >= 4) v = fmax(v, dpp<kDppXor2>(v));
    if constexpr (G >= 8) v = fmax(v, dpp<kDppHalfMirror>(v));
#endif
    return v;
}
// Broadcast from lane SRC of each group: a quad_perm that gives every lane of a quad that quad's
// lane SRC % 4, then (8-lane groups) row_half_mirror to fetch the other quad's copy for the
// lanes whose quad does not hold SRC.
#if defined(__HIP_DEVICE_COMPILE__)
constexpr int dpp_quad(int a, int b, int c, int d) { return a | (b << 2) | (c << 4) | (d << 6); }
template <int G, int SRC, typename T>
__device__ __forceinline__ T dpp_bcast(T v) {
    static_assert(SRC >= 0 && SRC < G, "source lane");
    if constexpr (G == 2) {
        return dpp<dpp_quad(SRC, SRC, 2 + SRC, 2 + SRC)>(v);
    } else if constexpr (G == 4) {
        return dpp<dpp_quad(SRC, SRC, SRC, SRC)>(v);
    } else {
        constexpr int q = SRC & 3;
        const T v1 = dpp<dpp_quad(q, q, q, q)>(v);
        const T v2 = dpp<kDppHalfMirror>(v1);
        const bool upper = (threadIdx.x & 4) != 0;
        return (upper == (SRC >= 4)) ? v1 : v2;
    }
}
#endif

template <int G, int SRC, typename T>
AHD T gbcast(T v) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (G > 1) return dpp_bcast<G, SRC>(v);
#endif
    return v;
}
template <int G, int SRC>
AHD int gbcast_i(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (G > 1) return __float_as_int(dpp_bcast<G, SRC>(__int_as_float(v)));
#endif
    return v;
}
// Value of lane `src` of each group, with a run-time source lane (ds_bpermute).
template <int G>
AHD double gshfl(double v, int src) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (G > 1) return __shfl(v, src, G);
#endif
    (void)src;
    return v;
}
template <int G>
AHD int gor(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (G >= 2) v |= dpp_i<kDppXor1>(v);
    if constexpr (G >= 4) v |= dpp_i<kDppXor2>(v);
    if constexpr (G >= 8) v |= dpp_i<kDppHalfMirror>(v);
#endif
    return v;
}

// Single-precision partner of a double-precision scalar type (float for double).
template <typename D>
struct SingleOf { using type = float; };

AHD double value_of(double x) { return x; }

// 1/sqrt(x): the hardware reciprocal square root with one refinement step on the GPU
// (within a few ulp), replacing a square root and a division.
AHD double inv_sqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return rsqrt(x);
#else
    return 1.0 / sqrt(x);
#endif
}
AHD float inv_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return rsqrtf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}
AHD float value_of(float x) { return x; }

AHD bool is_normal(double x) {
    const double a = fabs(x);
    return a >= 2.2250738585072014e-308 && a <= 1.7976931348623157e308;
}

// ---- Chebyshev evaluation -------------------------------------------------------------
// Record selection follows reference src/spk.c:440-445 (integer truncation of the
// record offset, then z in [-1, 1]); indices are clamped so a time at the very end of
// coverage reads the last record instead of the directory words.
template <typename D>
struct ChebOut {
    D x, y, z;     // km
    D vx, vy, vz;  // km/day (only when requested)
    D ax, ay, az;  // km/day^2 (only when requested)
};

struct alignas(16) Pair { double x, y; };

// The device evaluates the first NP terms of the zero-padded record (NP = kCoefPad, or
// kCoefShort where every record of the slot has at most that many) with every load issued up
// front; the operation-count instantiation (D != double) counts the file's P terms.  `rec`
// points at the record (in global memory on the device, in host memory otherwise).
template <typename D, bool Vel, bool Acc, int NP, typename Ptr>
AHD ChebOut<D> cheb_series(const DevSlot& S, Ptr rec, D t) {
    static_assert(NP >= 2 && NP <= kCoefPad, "series length");
    constexpr int NW = 2 + 3 * NP;  // words used (an odd count loads one more, inside the record)
    double c[kRecWords];
#pragma unroll
    for (int k = 0; k < NW; k += 2) {
        const Pair v = {rec[k], rec[k + 1]};
        c[k] = v.x;
        c[k + 1] = v.y;
    }
    const int P = std::is_same<D, double>::value ? NP : S.P;
    const D z = (D(c[0]) + t) * D(c[1]);
    // coefficient p of x, y and z at words 2 + 3p, 3 + 3p, 4 + 3p: the series consumes the
    // record in load order, so its first terms start while the later loads are in flight
    const double* cx = c + 2;
    const double* cy = c + 3;
    const double* cz = c + 4;
    const D z2 = D(2.0) * z;
    D T0 = D(1.0), T1 = z;
    D x = fma(D(cx[3]), z, D(cx[0]));
    D y = fma(D(cy[3]), z, D(cy[0]));
    D w = fma(D(cz[3]), z, D(cz[0]));
    ChebOut<D> o;
    D S0 = D(0.0), S1 = D(1.0);
    D vx = D(cx[3]), vy = D(cy[3]), vz = D(cz[3]);
    // second derivatives U_p = d^2 T_p / dz^2: U_0 = U_1 = 0 (reference src/spk.c:520-530)
    D U0 = D(0.0), U1 = D(0.0);
    D ax = D(0.0), ay = D(0.0), az = D(0.0);
#pragma unroll
    for (int p = 2; p < NP; p++) {
        if (p < P) {
            const D T2 = fma(z2, T1, -T0);
            x = fma(D(cx[3 * p]), T2, x);
            y = fma(D(cy[3 * p]), T2, y);
            w = fma(D(cz[3 * p]), T2, w);
            if (Vel || Acc) {
                const D S2 = fma(z2, S1, fma(D(2.0), T1, -S0));
                vx = fma(D(cx[3 * p]), S2, vx);
                vy = fma(D(cy[3 * p]), S2, vy);
                vz = fma(D(cz[3 * p]), S2, vz);
                if (Acc) {
                    const D U2 = fma(z2, U1, fma(D(4.0), S1, -U0));
                    ax = fma(D(cx[3 * p]), U2, ax);
                    ay = fma(D(cy[3 * p]), U2, ay);
                    az = fma(D(cz[3 * p]), U2, az);
                    U0 = U1;
                    U1 = U2;
                }
                S0 = S1;
                S1 = S2;
            }
            T0 = T1;
            T1 = T2;
        }
    }
    if (Vel || Acc) {
        o.vx = vx * D(c[1]);
        o.vy = vy * D(c[1]);
        o.vz = vz * D(c[1]);
    } else {
        o.vx = o.vy = o.vz = D(0.0);
    }
    if (Acc) {
        const D c2 = D(c[1]) * D(c[1]);
        o.ax = ax * c2;
        o.ay = ay * c2;
        o.az = az * c2;
    } else {
        o.ax = o.ay = o.az = D(0.0);
    }
    o.x = x;
    o.y = y;
    o.z = w;
    return o;
}

// Record index of slot S at time t (clamped to the records present).
template <typename D>
AHD int record_index(const DevSlot& S, D t) {
    const D fi = (D(S.c_init) + t) * D(S.inv_intlen);
    const int idx = (int)value_of(fi);
    return idx < 0 ? 0 : (idx >= S.nrec ? S.nrec - 1 : idx);
}

template <typename D, bool Vel, bool Acc = false, int NP = kCoefPad>
AHD ChebOut<D> cheb_eval(const DevSlot& S, D t) {
    const int idx = record_index(S, t);
#if defined(__HIP_DEVICE_COMPILE__)
    // the records live in global memory: address them as such (global, not flat, loads)
    typedef __attribute__((address_space(1))) const double GlobalDouble;
    GlobalDouble* rec = (GlobalDouble*)S.rec + (size_t)idx * kRecWords;
#else
    const double* rec = S.rec + (size_t)idx * kRecWords;
#endif
    return cheb_series<D, Vel, Acc, NP>(S, rec, t);
}

// ---- substep ephemeris cache and wave-wide Chebyshev rounds (8-lane groups) ---------------
// The reference keeps up to 7 ephemeris results per body keyed by the exact time (reference
// src/forces.c:175-263), which IAS15's substep times hit: every predictor-corrector iteration of
// a step evaluates the forces at the same 7 times.  The integration kernel does the same per
// lane group: the first iteration of a step attempt stores the raw Chebyshev positions of each
// substep in the group's cache (kEcFill) and the later iterations load them (kEcUse).  On a trip
// where some groups of a wave need fresh positions (kEcEval: the step's first evaluation, or
// kEcFill), their 28 evaluations each are spread slot-major over all active lanes of the wave, so
// the wave runs ceil(28 k / A) rounds for k such groups and A active lanes instead of 4, and the
// results reach the owning lanes through the wave's LDS exchange area.  Every value comes from the
// same code and inputs as in the per-group evaluation, so results are unchanged bit for bit.
enum : int { kEcEval = 0, kEcFill = 1, kEcUse = 2 };
constexpr int kEcWords = 88;  // per group and substep: raw positions (km) of the 28 slots, then
                              // the Sun's raw velocity (words 84-86), one pad word
// Rows of a group in a wave's LDS exchange area: the 28 slots' positions, then the Sun's velocity.
constexpr int kXchRows = kNumSlots + 1;
struct SubstepCache {
    double* base;             // this wave's region: [8 groups][7 substeps][kEcWords] doubles, or
                              // nullptr (no cache: every trip evaluates, mode kEcEval)
    double* xch;              // this wave's LDS exchange area: [8 groups][kXchRows][3] doubles
    int n;                    // substep 1..7 (kEcFill)
    int mode;
    double px[4][3], sv[3];   // kEcUse: this lane's positions and the Sun's velocity, loaded by
                              // ec_prefetch at the end of the previous trip
};

#if defined(__HIP_DEVICE_COMPILE__)
// Index of the i-th group (0..7) whose leader bit (bit 8 g) is set in `leaders` (uniform).
__device__ __forceinline__ int nth_group(unsigned long long leaders, int i) {
    int g = 0, cnt = 0;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const bool b = ((leaders >> (8 * t)) & 1ULL) != 0;
        g = (b && cnt == i) ? t : g;
        cnt += b ? 1 : 0;
    }
    return g;
}

// Loads this lane's cached positions of substep n (and the Sun's velocity) into sc for the next
// trip, so that their latency overlaps the rest of the current trip.
__device__ __forceinline__ void ec_prefetch(SubstepCache& sc, int lane, int n) {
    // the cache stores of the wave's earlier trips complete before these loads
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    const double* p = sc.base + ((size_t)(__lane_id() >> 3) * 7 + (n - 1)) * kEcWords;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int s = r * 8 + lane;
        const int sl = s < kNumSlots ? s : 0;  // (lanes past the last slot read slot 0)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double v = p[3 * sl + c];
            sc.px[r][c] = (s < kNumSlots) ? v : 0.0;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) sc.sv[c] = p[84 + c];
}

// Raw positions (km) of the slots r * 8 + lane (r < 4) of this lane's group and, on the group's
// lane 0, the Sun's raw velocity, for 8-lane groups (see the cache description above).  All
// active lanes of the wave call this together; whole groups are active or not.  The evaluating
// lanes write each result to the wave's LDS exchange area, from which the owners read it.
__device__ __forceinline__ void wave_positions(const DevSlot* slots, int short_from, int lane, double t, bool want_vel,
                                               const SubstepCache& sc, double px[4][3], double sv[3]) {
    constexpr unsigned long long kLeaders = 0x0101010101010101ULL;
    const int wl = __lane_id();
    const int grp = wl >> 3;
    const unsigned long long below = (1ULL << (8 * grp)) - 1ULL;
    const unsigned long long act = __ballot(1) & kLeaders;        // active groups
    const bool need = sc.mode != kEcUse;
    const unsigned long long nd = __ballot(need) & kLeaders;      // groups to evaluate
    const bool any_vel = __ballot(need && want_vel) != 0ULL;
    const int k = __popcll(nd);
    const int A = 8 * __popcll(act);                               // active lanes
    const int a = 8 * __popcll(act & below) + lane;                // this lane's active rank
    // rounds: item j = s k + i (slot s of the i-th group to evaluate) on the lane of active rank
    // j mod A in round j / A; the Sun's items (s = 0) fall in round 0, which evaluates velocities
    const int items = kNumSlots * k;
    const int R = (items + A - 1) / A;
    const float inv_k = k > 0 ? 1.0f / (float)k : 0.0f;
    double* const x = sc.xch;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (q >= R) break;
        const int j = q * A + a;
        const int s = (int)(((float)j + 0.5f) * inv_k);
        const int i = j - s * k;
        const int gi = nth_group(nd, i < k ? i : 0);
        const double ti = __shfl(t, 8 * gi, 64);
        const int nm = __shfl(sc.n * 4 + sc.mode, 8 * gi, 64);
        // a round whose slots all have at most kCoefShort coefficients runs the short series
        const bool short_round = __ballot(j < items && s < short_from) == 0ULL;
        if (j < items) {
            ChebOut<double> c;
            if (q == 0 && any_vel) c = cheb_eval<double, true>(slots[s], ti);
            else if (short_round) c = cheb_eval<double, false, false, kCoefShort>(slots[s], ti);
            else c = cheb_eval<double, false>(slots[s], ti);
            double* e = x + (gi * kXchRows + s) * 3;
            e[0] = c.x; e[1] = c.y; e[2] = c.z;
            if (s == 0) {
                double* ev = x + (gi * kXchRows + kNumSlots) * 3;
                ev[0] = c.vx; ev[1] = c.vy; ev[2] = c.vz;
            }
            if ((nm & 3) == kEcFill) {
                double* p = sc.base + ((size_t)gi * 7 + ((nm >> 2) - 1)) * kEcWords;
                p[3 * s] = c.x; p[3 * s + 1] = c.y; p[3 * s + 2] = c.z;
                if (s == 0) { p[84] = c.vx; p[85] = c.vy; p[86] = c.vz; }
            }
        }
    }
    // (LDS operations of one wave complete in order; the fence keeps the compiler's order too)
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (need) {
        const double* e = x + grp * kXchRows * 3;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int s = r * 8 + lane;
            const int sl = s < kNumSlots ? s : 0;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double v = e[3 * sl + c];
                px[r][c] = (s < kNumSlots) ? v : 0.0;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) sv[c] = e[3 * (kXchRows - 1) + c];
    } else {
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) px[r][c] = sc.px[r][c];
#pragma unroll
        for (int c = 0; c < 3; c++) sv[c] = sc.sv[c];
    }
}
#elif defined(__HIPCC__)
// (the host pass of hipcc only type-checks the kernels that call it)
__device__ void ec_prefetch(SubstepCache& sc, int lane, int n);
#endif

// ---- the 3x9 Jacobian of the real acceleration --------------------------------------------
template <typename T>
struct Jac {
    T x[3][3];  // d a_i / d x_j
    T v[3][3];  // d a_i / d v_j
    T A[3][3];  // d a_i / d A_j (Marsden parameters)
};

template <typename T>
AHD void jac_zero(Jac<T>& J) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) J.x[i][j] = J.v[i][j] = J.A[i][j] = T(0);
}

// Rotation into a body's equatorial frame (rows of R), written out as in
// reference src/forces.c:503-505 / 542-544.
template <typename T>
struct Rot {
    T r[3][3];
    AHD Rot(double cosa, double sina, double cosd, double sind) {
        r[0][0] = T(-sina); r[0][1] = T(cosa); r[0][2] = T(0);
        r[1][0] = T(-cosa * sind); r[1][1] = T(-sina * sind); r[1][2] = T(cosd);
        r[2][0] = T(cosa * cosd); r[2][1] = T(sina * cosd); r[2][2] = T(sind);
    }
};

// J += R^T S R for a symmetric S given as (xx, yy, zz, xy, yz, xz)
template <typename T>
AHD void add_rotated_sym(T J[3][3], const Rot<T>& R, T sxx, T syy, T szz, T sxy, T syz, T sxz) {
    const T S[3][3] = {{sxx, sxy, sxz}, {sxy, syy, syz}, {sxz, syz, szz}};
    T SR[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) SR[i][j] = S[i][0] * R.r[0][j] + S[i][1] * R.r[1][j] + S[i][2] * R.r[2][j];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) J[i][j] += R.r[0][i] * SR[0][j] + R.r[1][i] * SR[1][j] + R.r[2][i] * SR[2][j];
}

// ---- single-particle terms ---------------------------------------------------------------

// Earth J2/J3/J4 (reference src/forces.c:435-642).  J2 in D; J3/J4 and the partials in TS.
// The Earth's pole is fixed at RA 0, Dec 90 deg (reference src/forces.c:484-490): the rotation
// has cos RA = 1, sin RA = 0, sin Dec = 1 exactly and keeps cos Dec = cos(pi/2) = 6.1e-17,
// whose residue the reference carries, so it is written out for that form (rows (0, 1, 0),
// (-1, 0, c), (c, 0, 1) with c = cos Dec).  Powers of 1/r replace the reference's divisions.
template <typename D, typename TS>
AHD void term_earth_J2J4(const DevEphem& E, const D xr[3], const D xe[3], D GM, bool need_jac, D acc[3],
                         Jac<TS>& Jt) {
    const D c = D(E.e_cosd);
    const D ex = xr[0] - xe[0], ey = xr[1] - xe[1], ez = xr[2] - xe[2];
    const D r2 = ex * ex + ey * ey + ez * ez;
    const D ir = inv_sqrt(r2);
    const D ir2 = ir * ir;
    const D dx = ey, dy = -ex + ez * c, dz = ex * c + ez;
    const D costheta2 = dz * dz * ir2;
    const D J2p = D(E.k_J2) * (ir2 * ir2 * ir);  // 3 J2 Re^2 / (2 r^5)
    const D J2f = D(5.) * costheta2 - D(1.);
    D rx = GM * J2p * J2f * dx;
    D ry = GM * J2p * J2f * dy;
    D rz = GM * J2p * (J2f - D(2.)) * dz;
    const TS ct2 = TS(costheta2), tdx = TS(dx), tdy = TS(dy), tdz = TS(dz), tGM = TS(GM);
    const TS tir = TS(ir), tir2 = TS(ir2);
    const TS tir5 = tir2 * tir2 * tir;
    const TS J3p = TS(E.k_J3) * tir5;            // 5 J3 Re^3 / (2 r^5)
    const TS J3f = TS(3.) - TS(7.) * ct2;
    const TS J4p = TS(E.k_J4) * (tir5 * tir2);   // 5 J4 Re^4 / (8 r^7)
    const TS J4f = TS(63.) * ct2 * ct2 - TS(42.) * ct2 + TS(3.);
    const TS j3x = -tGM * J3p * tir2 * J3f * tdx * tdz;
    const TS j3y = -tGM * J3p * tir2 * J3f * tdy * tdz;
    const TS j3z = -tGM * J3p * (TS(6.) * ct2 - TS(7.) * ct2 * ct2 - TS(0.6));
    const TS j4x = tGM * J4p * J4f * tdx;
    const TS j4y = tGM * J4p * J4f * tdy;
    const TS j4z = tGM * J4p * (J4f + TS(12.) - TS(28.) * ct2) * tdz;
    rx += D(j3x); ry += D(j3y); rz += D(j3z);
    rx += D(j4x); ry += D(j4y); rz += D(j4z);
    acc[0] += -ry + rz * c;
    acc[1] += rx;
    acc[2] += ry * c + rz;
    if (!need_jac) return;
    const TS J2pT = TS(J2p), J2fT = TS(J2f);
    const TS f2 = TS(7.) * ct2 - TS(1.);
    const TS f3 = TS(35.) * ct2 * ct2 - TS(30.) * ct2 + TS(3.);
    TS sxx = tGM * J2pT * (J2fT - TS(5.) * f2 * tdx * tdx * tir2);
    TS syy = tGM * J2pT * (J2fT - TS(5.) * f2 * tdy * tdy * tir2);
    TS szz = tGM * J2pT * TS(-1.) * f3;
    TS sxy = tGM * J2pT * TS(-5.) * f2 * tdx * tdy * tir2;
    TS syz = tGM * J2pT * TS(-5.) * (f2 - TS(2.)) * tdy * tdz * tir2;
    TS sxz = tGM * J2pT * TS(-5.) * (f2 - TS(2.)) * tdx * tdz * tir2;
    const TS ct = tdz * tir;
    const TS g2 = TS(21) * (TS(-3.) * ct2 + TS(1.)) * tir2;
    const TS g3 = TS(3) * (TS(-21.) * ct2 * ct2 + TS(14.) * ct2 - TS(1.)) * tir2;
    const TS g4 = (TS(-63.) * ct2 * ct2 + TS(70.) * ct2 - TS(15.)) * ct * tir;
    sxx += tGM * J3p * ct * (g2 * tdx * tdx - J3f) * tir;
    syy += tGM * J3p * ct * (g2 * tdy * tdy - J3f) * tir;
    szz += tGM * J3p * g4;
    sxy += tGM * J3p * g2 * ct * tdx * tdy * tir;
    syz += tGM * J3p * g3 * tdy;
    sxz += tGM * J3p * g3 * tdx;
    const TS h2 = TS(33.) * ct2 * ct2 - TS(18.) * ct2 + TS(1.);
    const TS h3 = TS(33.) * ct2 * ct2 - TS(30.) * ct2 + TS(5.);
    const TS h4 = TS(231.) * ct2 * ct2 * ct2 - TS(315.) * ct2 * ct2 + TS(105.) * ct2 - TS(5.);
    sxx += tGM * J4p * (J4f - TS(21.) * h2 * tdx * tdx * tir2);
    syy += tGM * J4p * (J4f - TS(21.) * h2 * tdy * tdy * tir2);
    szz += tGM * J4p * TS(-3.) * h4;
    sxy += tGM * J4p * TS(-21.) * h2 * tdx * tdy * tir2;
    syz += tGM * J4p * TS(-21.) * h3 * tdy * tdz * tir2;
    sxz += tGM * J4p * TS(-21.) * h3 * tdx * tdz * tir2;
    // J += R^T S R; R's columns are u0 = (0, -1, c), u1 = (1, 0, 0), u2 = (0, c, 1) and
    // (R^T S R)_ij = u_i . (S u_j)
    const TS tc = TS(c);
    const TS w[3][3] = {{-sxy + tc * sxz, -syy + tc * syz, -syz + tc * szz},
                        {sxx, sxy, sxz},
                        {tc * sxy + sxz, tc * syy + syz, tc * syz + szz}};
#pragma unroll
    for (int j = 0; j < 3; j++) {
        Jt.x[0][j] += -w[j][1] + tc * w[j][2];
        Jt.x[1][j] += w[j][0];
        Jt.x[2][j] += tc * w[j][1] + w[j][2];
    }
}

// Solar J2 (reference src/forces.c:644-772), in D; partials in TS.
template <typename D, typename TS>
AHD void term_solar_J2(const DevEphem& E, const D xr[3], const D xs[3], D GMs, bool need_jac, D acc[3], Jac<TS>& Jt) {
    const D cosa = D(E.s_cosa), sina = D(E.s_sina), cosd = D(E.s_cosd), sind = D(E.s_sind);
    D dx = xr[0] - xs[0], dy = xr[1] - xs[1], dz = xr[2] - xs[2];
    const D r2 = dx * dx + dy * dy + dz * dz;
    const D ir = inv_sqrt(r2);
    const D ir2 = ir * ir;
    const D dxp = -dx * sina + dy * cosa;
    const D dyp = -dx * cosa * sind - dy * sina * sind + dz * cosd;
    const D dzp = dx * cosa * cosd + dy * sina * cosd + dz * sind;
    dx = dxp; dy = dyp; dz = dzp;
    const D costheta2 = dz * dz * ir2;
    const D pf = D(E.k_sJ2) * (ir2 * ir2 * ir);  // 3 J2sun Rs^2 / (2 r^5)
    const D f1 = D(5.) * costheta2 - D(1.);
    const D rx = GMs * pf * f1 * dx;
    const D ry = GMs * pf * f1 * dy;
    const D rz = GMs * pf * (f1 - D(2.)) * dz;
    acc[0] += -rx * sina - ry * cosa * sind + rz * cosa * cosd;
    acc[1] += rx * cosa - ry * sina * sind + rz * sina * cosd;
    acc[2] += +ry * cosd + rz * sind;
    if (!need_jac) return;
    const TS t2 = TS(costheta2), tdx = TS(dx), tdy = TS(dy), tdz = TS(dz), tir2 = TS(ir2);
    const TS gp = TS(GMs * pf), tf1 = TS(f1);
    const TS f2 = TS(7.) * t2 - TS(1.);
    const TS f3 = TS(35.) * t2 * t2 - TS(30.) * t2 + TS(3.);
    const TS sxx = gp * (tf1 - TS(5.) * f2 * tdx * tdx * tir2);
    const TS syy = gp * (tf1 - TS(5.) * f2 * tdy * tdy * tir2);
    const TS szz = gp * TS(-1.) * f3;
    const TS sxy = gp * TS(-5.) * f2 * tdx * tdy * tir2;
    const TS syz = gp * TS(-5.) * (f2 - TS(2.)) * tdy * tdz * tir2;
    const TS sxz = gp * TS(-5.) * (f2 - TS(2.)) * tdx * tdz * tir2;
    Rot<TS> R(E.s_cosa, E.s_sina, E.s_cosd, E.s_sind);
    add_rotated_sym(Jt.x, R, sxx, syy, szz, sxy, syz, sxz);
}

// Marsden non-gravitational acceleration (reference src/forces.c:774-1057), in TS.
template <typename D, typename TS>
AHD void term_nongrav(const ForceOpts& O, const D xr[3], const D vr[3], const D xs[3], const D vs[3], const double A[3],
                      bool need_jac, D acc[3], Jac<TS>& J) {
    const TS dx = TS(xr[0] - xs[0]), dy = TS(xr[1] - xs[1]), dz = TS(xr[2] - xs[2]);
    const TS dvx = TS(vr[0] - vs[0]), dvy = TS(vr[1] - vs[1]), dvz = TS(vr[2] - vs[2]);
    const TS A1 = TS(A[0]), A2 = TS(A[1]), A3 = TS(A[2]);
    const TS alpha = TS(O.alpha), nk = TS(O.nk), nm = TS(O.nm), nn = TS(O.nn), r0 = TS(O.r0);
    const TS r2 = dx * dx + dy * dy + dz * dz;
    const TS r = sqrt(r2);
    const TS x = r / r0;
    // x^-nm: for the default model's nm = 2 as 1 / x^2, within an ulp of pow and without its cost
    const TS xm = (nm == TS(2)) ? TS(1) / (x * x) : pow(x, -nm);
    // nk = 0 (the reference's default, src/assist.c:435): the factor (1 + x^nn)^-nk is pow(., -0),
    // exactly 1 for every argument, so leaving it out gives g bit for bit.
    const bool nk0 = nk == TS(0);
    const TS g = nk0 ? alpha * xm : alpha * xm * pow(TS(1.0) + pow(x, nn), -nk);
    const TS hx = dy * dvz - dz * dvy, hy = dz * dvx - dx * dvz, hz = dx * dvy - dy * dvx;
    const TS h = sqrt(hx * hx + hy * hy + hz * hz);
    const TS tx = hy * dz - hz * dy, ty = hz * dx - hx * dz, tz = hx * dy - hy * dx;
    const TS _t = sqrt(tx * tx + ty * ty + tz * tz);
    // (the radial, transverse and normal unit vectors through one reciprocal each)
    const TS gr = A1 * g / r, gt = A2 * g / _t, gh = A3 * g / h;
    acc[0] += D(gr * dx + gt * tx + gh * hx);
    acc[1] += D(gr * dy + gt * ty + gh * hy);
    acc[2] += D(gr * dz + gt * tz + gh * hz);
    if (!need_jac) return;
    const TS r3 = r * r * r;
    const TS v2 = dvx * dvx + dvy * dvy + dvz * dvz;
    const TS rdotv = dx * dvx + dy * dvy + dz * dvz;
    const TS vdott = dvx * tx + dvy * ty + dvz * tz;
    // dg/dr = (alpha/r0) (-nm x^(-nm-1) (1+x^nn)^-nk + x^-nm (-nk nn) x^(nn-1) (1+x^nn)^(-nk-1)).
    // With nk = +0 and nn >= +0 finite, -nk nn = -0; for 0 < x finite the other factors of the
    // second term are >= +0, so the term is -0 whenever x^-nm and x^(nn-1) are finite, and adding
    // -0 leaves the first term unchanged bit for bit.  x^(nn-1) is finite when |nn-1| (|e|+1)
    // stays below the overflow exponent, e the binary exponent of x (x in [2^(e-1), 2^e)); then
    // the three powers of the second term are skipped.
    int ex = 0;
    (void)frexp(x, &ex);
    const TS e_lim = sizeof(TS) == 8 ? TS(1000) : TS(120);
    const bool short_dg = nk0 && !signbit(nk) && nn >= TS(0) && !signbit(nn) && isfinite(nn) && x > TS(0) &&
                          isfinite(x) && isfinite(xm) && fabs(nn - TS(1)) * TS(abs(ex) + 1) < e_lim;
    const TS dgdr = short_dg ? (alpha / r0) * (-nm * pow(x, -nm - TS(1)))
                             : (alpha / r0) * (-nm * pow(x, -nm - TS(1)) * pow(TS(1.0) + pow(x, nn), -nk) +
                                               xm * (-nk * nn) * pow(x, nn - TS(1)) * pow(TS(1.0) + pow(x, nn), -nk - TS(1)));
    const TS dgx = dgdr * dx / r, dgy = dgdr * dy / r, dgz = dgdr * dz / r;
    const TS h3 = h * h * h, t3 = _t * _t * _t;
    const TS hxh3 = hx / h3, hyh3 = hy / h3, hzh3 = hz / h3;
    const TS txt3 = tx / t3, tyt3 = ty / t3, tzt3 = tz / t3;
    J.A[0][0] += g * dx / r; J.A[1][0] += g * dy / r; J.A[2][0] += g * dz / r;
    J.A[0][1] += g * tx / _t; J.A[1][1] += g * ty / _t; J.A[2][1] += g * tz / _t;
    J.A[0][2] += g * hx / h; J.A[1][2] += g * hy / h; J.A[2][2] += g * hz / h;
    const TS one = TS(1), two = TS(2);
    J.x[0][0] += A1 * (dgx * dx / r + g * (one / r - dx * dx / r3)) + A2 * (dgx * tx / _t + g * ((dx * dvx - rdotv) / _t - txt3 * (two * dx * vdott - rdotv * tx))) + A3 * (dgx * hx / h + g * (-hxh3) * (v2 * dx - rdotv * dvx));
    J.x[1][1] += A1 * (dgy * dy / r + g * (one / r - dy * dy / r3)) + A2 * (dgy * ty / _t + g * ((dy * dvy - rdotv) / _t - tyt3 * (two * dy * vdott - rdotv * ty))) + A3 * (dgy * hy / h + g * (-hyh3) * (v2 * dy - rdotv * dvy));
    J.x[2][2] += A1 * (dgz * dz / r + g * (one / r - dz * dz / r3)) + A2 * (dgz * tz / _t + g * ((dz * dvz - rdotv) / _t - tzt3 * (two * dz * vdott - rdotv * tz))) + A3 * (dgz * hz / h + g * (-hzh3) * (v2 * dz - rdotv * dvz));
    J.x[0][1] += A1 * (dgy * dx / r + g * (-dx * dy / r3)) + A2 * (dgy * tx / _t + g * ((two * dy * dvx - dx * dvy) / _t - txt3 * (two * dy * vdott - rdotv * ty))) + A3 * (dgy * hx / h + g * (dvz / h - hxh3 * (v2 * dy - rdotv * dvy)));
    J.x[1][0] += A1 * (dgx * dy / r + g * (-dy * dx / r3)) + A2 * (dgx * ty / _t + g * ((two * dx * dvy - dy * dvx) / _t - tyt3 * (two * dx * vdott - rdotv * tx))) + A3 * (dgx * hy / h + g * (-dvz / h - hyh3 * (v2 * dx - rdotv * dvx)));
    J.x[0][2] += A1 * (dgz * dx / r + g * (-dx * dz / r3)) + A2 * (dgz * tx / _t + g * ((two * dz * dvx - dx * dvz) / _t - txt3 * (two * dz * vdott - rdotv * tz))) + A3 * (dgz * hx / h + g * (-dvy / h - hxh3 * (v2 * dz - rdotv * dvz)));
    J.x[2][0] += A1 * (dgx * dz / r + g * (-dz * dx / r3)) + A2 * (dgx * tz / _t + g * ((two * dx * dvz - dz * dvx) / _t - tzt3 * (two * dx * vdott - rdotv * tx))) + A3 * (dgx * hz / h + g * (dvy / h - hzh3 * (v2 * dx - rdotv * dvx)));
    J.x[1][2] += A1 * (dgz * dy / r + g * (-dy * dz / r3)) + A2 * (dgz * ty / _t + g * ((two * dz * dvy - dy * dvz) / _t - tyt3 * (two * dz * vdott - rdotv * tz))) + A3 * (dgz * hy / h + g * (dvx / h - hyh3 * (v2 * dz - rdotv * dvz)));
    J.x[2][1] += A1 * (dgy * dz / r + g * (-dz * dy / r3)) + A2 * (dgy * tz / _t + g * ((two * dy * dvz - dz * dvy) / _t - tzt3 * (two * dy * vdott - rdotv * ty))) + A3 * (dgy * hz / h + g * (-dvx / h - hzh3 * (v2 * dy - rdotv * dvy)));
    J.v[0][0] += A2 * g * ((dy * dy + dz * dz) / _t - txt3 * r2 * tx) + A3 * g * (-hxh3 * (r2 * dvx - dx * rdotv));
    J.v[1][1] += A2 * g * ((dx * dx + dz * dz) / _t - tyt3 * r2 * ty) + A3 * g * (-hyh3 * (r2 * dvy - dy * rdotv));
    J.v[2][2] += A2 * g * ((dx * dx + dy * dy) / _t - tzt3 * r2 * tz) + A3 * g * (-hzh3 * (r2 * dvz - dz * rdotv));
    J.v[0][1] += A2 * g * (-dy * dx / _t - tyt3 * r2 * tx) + A3 * g * (-dz / h - hxh3 * (r2 * dvy - dy * rdotv));
    J.v[1][0] += A2 * g * (-dx * dy / _t - txt3 * r2 * ty) + A3 * g * (dz / h - hyh3 * (r2 * dvx - dx * rdotv));
    J.v[0][2] += A2 * g * (-dz * dx / _t - tzt3 * r2 * tx) + A3 * g * (dy / h - hxh3 * (r2 * dvz - dz * rdotv));
    J.v[2][0] += A2 * g * (-dx * dz / _t - txt3 * r2 * tz) + A3 * g * (-dy / h - hzh3 * (r2 * dvx - dx * rdotv));
    J.v[1][2] += A2 * g * (-dz * dy / _t - tzt3 * r2 * ty) + A3 * g * (-dx / h - hyh3 * (r2 * dvz - dz * rdotv));
    J.v[2][1] += A2 * g * (-dy * dz / _t - tyt3 * r2 * tz) + A3 * g * (dx / h - hzh3 * (r2 * dvy - dy * rdotv));
}

// EIH GR with the Sun as the only source (reference src/forces.c:1288-1983, beta = gamma = 1),
// given the group-summed planet sums: S0 = sum_k GM_k / r_ik, dS0 = -sum_k GM_k d_ik / r_ik^3
// (k over the 11 planetary bodies), S1 = sum_{k != Sun} GM_k / r_sk and the Sun's Newtonian
// acceleration aj from the same bodies.  In TS.
template <typename D, typename TS>
AHD void term_eih_sun(const DevEphem& E, const D xr[3], const D vr[3], const D xs[3], const D vs[3], D GMs, D S0,
                      const D dS0[3], D S1, const D aj[3], bool need_jac, D acc[3], Jac<TS>& J) {
    const TS oc2 = TS(E.over_c_squared);
    const TS dxij = TS(xr[0] - xs[0]), dyij = TS(xr[1] - xs[1]), dzij = TS(xr[2] - xs[2]);
    const TS vx = TS(vr[0]), vy = TS(vr[1]), vz = TS(vr[2]);
    const TS vxj = TS(vs[0]), vyj = TS(vs[1]), vzj = TS(vs[2]);
    const TS GMj = TS(GMs);
    const TS rij2 = dxij * dxij + dyij * dyij + dzij * dzij;
    const TS ir = inv_sqrt(rij2);
    const TS ir2 = ir * ir;
    const TS ir3 = ir2 * ir;
    const TS prefacij = GMj * ir3;
    const TS vi2 = vx * vx + vy * vy + vz * vz;
    const TS term2 = oc2 * vi2;
    const TS vj2 = vxj * vxj + vyj * vyj + vzj * vzj;
    const TS term3 = TS(2) * oc2 * vj2;
    const TS vidotvj = vx * vxj + vy * vyj + vz * vzj;
    const TS term4 = TS(-4) * oc2 * vidotvj;
    const TS rijdotvj = dxij * vxj + dyij * vyj + dzij * vzj;
    const TS term5 = TS(-1.5) * oc2 * (rijdotvj * rijdotvj) * ir2;
    const TS fx = TS(4) * vx - TS(3) * vxj, fy = TS(4) * vy - TS(3) * vyj, fz = TS(4) * vz - TS(3) * vzj;
    const TS f = dxij * fx + dyij * fy + dzij * fz;
    const TS pf = prefacij * f;
    const TS rel[3] = {vx - vxj, vy - vyj, vz - vzj};
    const TS term0 = TS(S0) * (TS(-4) * oc2);
    const TS term1 = TS(S1) * (-oc2);
    const TS axj = TS(aj[0]), ayj = TS(aj[1]), azj = TS(aj[2]);
    const TS term6 = TS(-0.5) * oc2 * (dxij * axj + dyij * ayj + dzij * azj);
    const TS t8 = GMj * ir * TS(3.5);
    const TS factor = term0 + term1 + term2 + term3 + term4 + term5 + term6;
    const TS dij[3] = {dxij, dyij, dzij};
    const TS ajv[3] = {axj, ayj, azj};
#pragma unroll
    for (int c = 0; c < 3; c++) acc[c] += D(-prefacij * dij[c] * factor + (pf * rel[c] * oc2 + t8 * ajv[c] * oc2));
    if (!need_jac) return;
    const TS pv = GMj * ir3;
    const TS g5 = TS(-3.0) * GMj * (ir3 * ir2);
    const TS dpref[3] = {g5 * dxij, g5 * dyij, g5 * dzij};
    const TS t5f = TS(3.0) * oc2 * rijdotvj * ir;
    const TS vj[3] = {vxj, vyj, vzj};
    const TS vi[3] = {vx, vy, vz};
    const TS fv[3] = {fx, fy, fz};
    TS dfx[3], dfv[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const TS dt5 = -t5f * (vj[q] * ir - rijdotvj * dij[q] * ir3);
        const TS dt6 = TS(-0.5) * oc2 * ajv[q];
        dfx[q] = TS(dS0[q]) * (TS(-4) * oc2) + dt5 + dt6;
        dfv[q] = TS(2.0) * oc2 * vi[q] + TS(-4) * oc2 * vj[q];
    }
    const TS pf7 = pv * f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const TS d7x = dpref[q] * f * rel[c] + pv * fv[q] * rel[c];
            TS d7v = pv * (TS(4) * dij[q]) * rel[c];
            if (q == c) d7v += pf7;
            const TS d8 = -GMj * ajv[c] * ir3 * dij[q] * TS(3.5);
            TS v = -dpref[q] * dij[c] * factor - pv * dij[c] * dfx[q];
            if (q == c) v -= pv * factor;
            J.x[c][q] += v + (d7x * oc2 + d8 * oc2);
            J.v[c][q] += -pv * dij[c] * dfv[q] + d7v * oc2;
        }
    }
}

// One EIH source j of reference src/forces.c:1319-1499 (real particle) and :1508-1982
// (partials), beta = gamma = 1, in FP64, for the extended force model (several sources and/or
// the geocentric frame):
//   xs, vs   source position and velocity relative to the frame origin (xj - xo, vj - vo);
//   S0, dS0  sum_k GM_k / r_ik over the 11 planet-file bodies and its gradient (all sources);
//   S1, aj   sum_{k != j} GM_k / r_jk and the source's Newtonian acceleration from the same
//            bodies, from barycentric positions;
//   ao       the frame origin's acceleration (zero for barycentric): term 6 uses aj - ao, term 8
//            uses aj, as the reference does.
// Term 7 and 8 contributions go to t78 (the reference adds their sums over the sources after the
// source loop); the rest goes to acc.
AHD void term_eih_source(const DevEphem& E, const double xr[3], const double vr[3], const double xs[3],
                         const double vs[3], double GMj, double S0, const double dS0[3], double S1,
                         const double aj[3], const double ao[3], bool need_jac, double acc[3], double t78[3],
                         Jac<double>& J) {
    const double oc2 = E.over_c_squared;
    const double dij[3] = {xr[0] - xs[0], xr[1] - xs[1], xr[2] - xs[2]};
    const double rij2 = dij[0] * dij[0] + dij[1] * dij[1] + dij[2] * dij[2];
    const double ir = inv_sqrt(rij2);
    const double ir2 = ir * ir;
    const double ir3 = ir2 * ir;
    const double prefacij = GMj * ir3;
    const double vi2 = vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2];
    const double term2 = oc2 * vi2;
    const double vj2 = vs[0] * vs[0] + vs[1] * vs[1] + vs[2] * vs[2];
    const double term3 = 2.0 * oc2 * vj2;
    const double vidotvj = vr[0] * vs[0] + vr[1] * vs[1] + vr[2] * vs[2];
    const double term4 = -4.0 * oc2 * vidotvj;
    const double rijdotvj = dij[0] * vs[0] + dij[1] * vs[1] + dij[2] * vs[2];
    const double term5 = -1.5 * oc2 * (rijdotvj * rijdotvj) * ir2;
    double fv[3], rel[3], am[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        fv[c] = 4.0 * vr[c] - 3.0 * vs[c];
        rel[c] = vr[c] - vs[c];
        am[c] = aj[c] - ao[c];
    }
    const double f = dij[0] * fv[0] + dij[1] * fv[1] + dij[2] * fv[2];
    const double pf = prefacij * f;
    const double term0 = S0 * (-4.0 * oc2);
    const double term1 = S1 * (-oc2);
    const double term6 = -0.5 * oc2 * (dij[0] * am[0] + dij[1] * am[1] + dij[2] * am[2]);
    const double t8 = GMj * ir * 3.5;
    const double factor = term0 + term1 + term2 + term3 + term4 + term5 + term6;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        acc[c] += -prefacij * dij[c] * factor;
        t78[c] += pf * rel[c] + t8 * aj[c];
    }
    if (!need_jac) return;
    const double g5 = -3.0 * GMj * (ir3 * ir2);
    const double dpref[3] = {g5 * dij[0], g5 * dij[1], g5 * dij[2]};
    const double t5f = 3.0 * oc2 * rijdotvj * ir;
    double dfx[3], dfv[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        dfx[q] = dS0[q] * (-4.0 * oc2) - t5f * (vs[q] * ir - rijdotvj * dij[q] * ir3) - 0.5 * oc2 * am[q];
        dfv[q] = 2.0 * oc2 * vr[q] - 4.0 * oc2 * vs[q];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double d7x = dpref[q] * f * rel[c] + prefacij * fv[q] * rel[c];
            double d7v = prefacij * (4.0 * dij[q]) * rel[c];
            if (q == c) d7v += pf;
            const double d8 = -GMj * aj[c] * ir3 * dij[q] * 3.5;
            double v = -dpref[q] * dij[c] * factor - prefacij * dij[c] * dfx[q];
            if (q == c) v -= prefacij * factor;
            J.x[c][q] += v + (d7x * oc2 + d8 * oc2);
            J.v[c][q] += -prefacij * dij[c] * dfv[q] + d7v * oc2;
        }
    }
}

// GR potential form (reference src/forces.c:1059-1161), Sun only, in TS.
template <typename D, typename TS>
AHD void term_gr_potential(const DevEphem& E, const D xr[3], const D xs[3], D GMs, bool need_jac, D acc[3], Jac<TS>& J) {
    const TS px = TS(xr[0] - xs[0]), py = TS(xr[1] - xs[1]), pz = TS(xr[2] - xs[2]);
    const TS r2 = px * px + py * py + pz * pz;
    const TS r = sqrt(r2);
    const TS G = TS(GMs);
    const TS prefac = TS(-6.0) * G * G / (TS(E.c_squared) * r2 * r2);
    acc[0] += D(prefac * px);
    acc[1] += D(prefac * py);
    acc[2] += D(prefac * pz);
    if (!need_jac) return;
    const TS u[3] = {px / r, py / r, pz / r};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) J.x[i][j] += (i == j ? prefac : TS(0)) + TS(-4.0) * prefac * u[i] * u[j];
}

// GR simple form (reference src/forces.c:1163-1286), Sun only, in TS.
template <typename D, typename TS>
AHD void term_gr_simple(const DevEphem& E, const D xr[3], const D vr[3], const D xs[3], const D vs[3], D GMs,
                        bool need_jac, D acc[3], Jac<TS>& J) {
    const TS p[3] = {TS(xr[0] - xs[0]), TS(xr[1] - xs[1]), TS(xr[2] - xs[2])};
    const TS pv[3] = {TS(vr[0] - vs[0]), TS(vr[1] - vs[1]), TS(vr[2] - vs[2])};
    const TS G = TS(GMs);
    const TS v2 = pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2];
    const TS r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const TS A = TS(4.0) * G / r - v2;
    const TS B = TS(4.0) * (p[0] * pv[0] + p[1] * pv[1] + p[2] * pv[2]);
    const TS prefac = G / (r * r * r * TS(E.c_squared));
#pragma unroll
    for (int c = 0; c < 3; c++) acc[c] += D(prefac * (A * p[c] + B * pv[c]));
    if (!need_jac) return;
    const TS dpdr = TS(-3.0) * prefac / r;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const TS Fi = A * p[i] + B * pv[i];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            J.x[i][j] += dpdr * p[j] / r * Fi + prefac * ((i == j ? A : TS(0)) - p[i] * (p[j] / r) * TS(4.0) * G / (r * r) + TS(4.0) * pv[j] * pv[i]);
            J.v[i][j] += prefac * (TS(-2.0) * pv[j] * p[i] + TS(4.0) * p[j] * pv[i] + (i == j ? B : TS(0)));
        }
    }
}

// ---- the full evaluation ---------------------------------------------------------------------
template <typename D>
struct EvalResult {
    D acc[3];    // real-particle acceleration
    int status;
};

// Distance [au] inside which the mixed form evaluates an asteroid's point mass in FP64.  The
// FP32 round-off of the term (~6e-8 of it) is amplified ~5e3 times in IAS15's b6 error
// estimate; beyond 0.2 au even the largest asteroid's term is < 1e-6 of the solar
// acceleration, so the amplified noise stays ~1e-10, below eps = 1e-9.
constexpr double kFp32MinDist = 0.2;

template <typename D, bool Mixed>
struct TermScalar { using type = typename std::conditional<Mixed, typename SingleOf<D>::type, D>::type; };

// Evaluates the force model for the real particle at (t, xr, vr) cooperatively over the G lanes
// of a group; every lane receives the real acceleration and the Jacobian J.  `slots` and `gm` are
// the ephemeris slot descriptors and masses (E.slot and E.GM, or copies of them in LDS).
// The EIH sources of the extended model: sources j < gr_eih_sources are the planet-file bodies
// in ASSIST order (Sun, Mercury, Venus, Earth, Moon, Mars, ..., Pluto).  For each source the
// lanes that own the other bodies add their GM/r_jk and Newtonian pulls (barycentric
// positions, as reference src/forces.c:1408-1432), the group sums them, and every lane then
// evaluates the source's terms with term_eih_source.  px/pv hold this lane's barycentric slot
// positions and velocities in au, au/d.
template <int G, int NR>
AHD void eih_sources(const DevEphem& E, const ForceOpts& O, const double* gm, int lane, const double xr[3],
                     const double vr[3], const double (&px)[NR][3], const double (&pv)[NR][3], const double xo[3],
                     const double vo[3], const double ao[3], double S0, const double dS0[3], bool need_jac,
                     double acc[3], Jac<double>& J) {
    const int nsrc = O.gr_eih_sources < kNumPlanets ? O.gr_eih_sources : kNumPlanets;
    double t78[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < kNumPlanets; j++) {
        if (j >= nsrc) break;
        const int sj = j < 3 ? j : j + 1;  // body -> slot (the EMB's slot 3 is skipped)
        double xj[3], vj[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            xj[c] = gshfl<G>(px[sj / G][c], sj % G);
            vj[c] = gshfl<G>(pv[sj / G][c], sj % G);
        }
        double s1 = 0.0, aj[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int s = r * G + lane;
            if (s >= kFirstAsteroidSlot || s == kSlotEMB) continue;
            const int b = slot_body(s);
            if (b == j) continue;
            const double dx = xj[0] - px[r][0], dy = xj[1] - px[r][1], dz = xj[2] - px[r][2];
            const double q2 = dx * dx + dy * dy + dz * dz;
            const double iq = inv_sqrt(q2);
            s1 += gm[b] * iq;
            const double f = gm[b] * (iq * iq * iq);
            aj[0] -= f * dx; aj[1] -= f * dy; aj[2] -= f * dz;
        }
        if constexpr (G > 1) {
            s1 = gsum<G>(s1);
#pragma unroll
            for (int c = 0; c < 3; c++) aj[c] = gsum<G>(aj[c]);
        }
        const double xs[3] = {xj[0] - xo[0], xj[1] - xo[1], xj[2] - xo[2]};
        const double vs[3] = {vj[0] - vo[0], vj[1] - vo[1], vj[2] - vo[2]};
        term_eih_source(E, xr, vr, xs, vs, gm[j], S0, dS0, s1, aj, ao, need_jac, acc, t78, J);
    }
    const double oc2 = E.over_c_squared;
#pragma unroll
    for (int c = 0; c < 3; c++) acc[c] += t78[c] * oc2;
}

//
// Ext selects the extended model (FP64 only): several EIH sources (gr_eih_sources up to 11) and
// the geocentric frame.  Every planet-file slot is then evaluated with velocity and acceleration;
// body positions are shifted by the Earth's (xk - xo), so every particle-body difference is the
// reference's x + (xo - xk) bit for bit; each source's planet sums come from barycentric
// positions, split over the lanes that own the bodies; and the Earth's acceleration is removed
// from the real particle's at the end (the indirect term, reference src/forces.c:157-172).
template <int G, bool Mixed, typename D, bool Ext = false>
AHD void force_eval(const DevEphem& E, const DevSlot* slots, const double* gm, const ForceOpts& O, int lane, D t, const D xr[3],
                    const D vr[3], const double A[3], bool ng, bool need_jac, EvalResult<D>& out,
                    Jac<typename TermScalar<D, Mixed>::type>& J, const SubstepCache* sc = nullptr) {
    using TS = typename TermScalar<D, Mixed>::type;
    static_assert(!Ext || (!Mixed && std::is_same<D, double>::value), "the extended model is FP64 only");
    constexpr int NR = (kNumSlots + G - 1) / G;
    constexpr int NRV = Ext ? NR : 1;  // rounds with velocities and accelerations kept
    const int forces = O.forces;
    const double tv = value_of(t);
    // (the window is coverage_window's, precomputed with the ephemeris)
    out.status = (tv >= O.t_lo && tv <= O.t_hi) ? kOk : kCoverage;
    const bool eih = (forces & F_GR_EIH) && O.gr_eih_sources > 0;
    const bool want_sun_vel = eih || ((forces & F_NONGRAV) && ng) || (forces & F_GR_SIMPLE);

    // kStream (one lane per system): the slots are evaluated one at a time in step 3 rather than
    // all first, which would keep 28 positions live through the per-body work
    constexpr bool kStream = (G == 1 && !Ext && std::is_same<D, double>::value);
    D sun[3], sunv[3], emb[3], earth[3];
    const D inv_au = D(E.inv_au);
    D px[kStream ? 1 : NR][3];
    D pv[NRV][3], pa[NRV][3];
    if constexpr (kStream) {
        const ChebOut<D> cs = cheb_eval<D, true>(slots[0], t);
        sun[0] = cs.x * inv_au; sun[1] = cs.y * inv_au; sun[2] = cs.z * inv_au;
        sunv[0] = cs.vx * inv_au; sunv[1] = cs.vy * inv_au; sunv[2] = cs.vz * inv_au;
        const ChebOut<D> ce = cheb_eval<D, false>(slots[kSlotEMB], t);
        emb[0] = ce.x; emb[1] = ce.y; emb[2] = ce.z;
        earth[0] = earth[1] = earth[2] = D(0.0);  // set in step 3
    } else {
    // 1. Chebyshev evaluations of the slots owned by this lane; round 0 (which holds the Sun,
    //    slot 0, on lane 0) also evaluates velocities when the Sun's velocity is needed.  With a
    //    substep cache (8-lane groups), the wave-wide rounds or the cache provide them instead.
    D sunv_l[3] = {D(0.0), D(0.0), D(0.0)};
    bool from_wave = false;
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (G == 8 && !Ext && std::is_same<D, double>::value) {
        if (sc) {
            wave_positions(slots, E.short_from, lane, t, want_sun_vel, *sc, px, sunv_l);
            from_wave = true;
        }
    }
#endif
#pragma unroll
    for (int r = 0; r < NR; r++) {
        if (from_wave) break;
        const int s = r * G + lane;
        if (s < kNumSlots) {
            ChebOut<D> c;
            if (Ext && r * G < kFirstAsteroidSlot) {
                c = cheb_eval<D, true, Ext>(slots[s], t);
                if (r == 0) { sunv_l[0] = c.vx; sunv_l[1] = c.vy; sunv_l[2] = c.vz; }
            } else if (r == 0 && want_sun_vel) {
                c = cheb_eval<D, true>(slots[s], t);
                sunv_l[0] = c.vx; sunv_l[1] = c.vy; sunv_l[2] = c.vz;
            } else {
                c = cheb_eval<D, false>(slots[s], t);
            }
            px[r][0] = c.x; px[r][1] = c.y; px[r][2] = c.z;
            if constexpr (Ext) {
                pv[r][0] = c.vx; pv[r][1] = c.vy; pv[r][2] = c.vz;
                pa[r][0] = c.ax; pa[r][1] = c.ay; pa[r][2] = c.az;
            }
        } else {
            px[r][0] = px[r][1] = px[r][2] = D(0.0);
            if constexpr (Ext) {
                pv[r][0] = pv[r][1] = pv[r][2] = D(0.0);
                pa[r][0] = pa[r][1] = pa[r][2] = D(0.0);
            }
        }
    }
    // 2. exchange Sun (au) and EMB (km); compose Earth, Moon and asteroids
#pragma unroll
    for (int c = 0; c < 3; c++) {
        sun[c] = gbcast<G, 0>(px[0][c] * inv_au);
        sunv[c] = gbcast<G, 0>(sunv_l[c] * inv_au);
        emb[c] = gbcast<G, kSlotEMB % G>(px[kSlotEMB / G][c]);
    }
    D embv[3], emba[3];
    if constexpr (Ext) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            embv[c] = gbcast<G, kSlotEMB % G>(pv[kSlotEMB / G][c]);
            emba[c] = gbcast<G, kSlotEMB % G>(pa[kSlotEMB / G][c]);
        }
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int s = r * G + lane;
        if (s >= kNumSlots) continue;
        if (s < kFirstAsteroidSlot) {
            const bool em = (s == kSlotEarth || s == kSlotMoon);
#pragma unroll
            for (int c = 0; c < 3; c++) px[r][c] = (em ? px[r][c] + emb[c] : px[r][c]) * inv_au;
            if constexpr (Ext) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    pv[r][c] = (em ? pv[r][c] + embv[c] : pv[r][c]) * inv_au;
                    pa[r][c] = (em ? pa[r][c] + emba[c] : pa[r][c]) * inv_au;
                }
            }
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++) px[r][c] = fma(px[r][c], D(E.inv_au_ast), sun[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) earth[c] = gbcast<G, kSlotEarth % G>(px[kSlotEarth / G][c]);
    }
    // frame origin: the Earth's barycentric state when geocentric, else zero
    D xo[3] = {D(0.0), D(0.0), D(0.0)}, vo[3] = {D(0.0), D(0.0), D(0.0)}, ao[3] = {D(0.0), D(0.0), D(0.0)};
    if constexpr (Ext) {
        if (O.geocentric) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                xo[c] = earth[c];
                vo[c] = gbcast<G, kSlotEarth % G>(pv[kSlotEarth / G][c]);
                ao[c] = gbcast<G, kSlotEarth % G>(pa[kSlotEarth / G][c]);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            sun[c] = sun[c] - xo[c];
            sunv[c] = sunv[c] - vo[c];
            earth[c] = earth[c] - xo[c];
        }
    }

    // 3. per-body point masses, Jacobian and EIH planet sums on the owning lane
    D ad[3] = {D(0.0), D(0.0), D(0.0)};  // direct acceleration, FP64 bodies
    TS adt[3] = {TS(0), TS(0), TS(0)};   // direct acceleration, TS bodies (asteroids in mixed mode)
    TS jd[6] = {TS(0), TS(0), TS(0), TS(0), TS(0), TS(0)};  // xx yy zz xy yz xz
    D S0 = D(0.0), S1 = D(0.0), dS0[3] = {D(0.0), D(0.0), D(0.0)}, aj[3] = {D(0.0), D(0.0), D(0.0)};
    const int nbody = kNumPlanets + E.n_ast;
    // body b at barycentric position pk (au)
    auto body_term = [&](int b, const D pk[3]) {
        const D GM = D(gm[b]);
        const bool allowed = (b == 0) ? (forces & F_SUN) : (b < kNumPlanets ? (forces & F_PLANETS) : (forces & F_ASTEROIDS));
        // (with the extended model, xo is the frame origin: x - (xk - xo) == x + (xo - xk))
        const D dx = Ext ? xr[0] - (pk[0] - xo[0]) : xr[0] - pk[0];
        const D dy = Ext ? xr[1] - (pk[1] - xo[1]) : xr[1] - pk[1];
        const D dz = Ext ? xr[2] - (pk[2] - xo[2]) : xr[2] - pk[2];
        // Mixed form: an asteroid's point mass runs in FP32 unless the particle is within
        // kFp32MinDist of it, where the FP32 round-off of the term would reach IAS15's
        // error estimate (see kFp32MinDist).
        if (Mixed && b >= kNumPlanets && value_of(dx * dx + dy * dy + dz * dz) > kFp32MinDist * kFp32MinDist) {
            const TS fx = TS(dx), fy = TS(dy), fz = TS(dz);
            const TS r2 = fx * fx + fy * fy + fz * fz;
            const TS ir = inv_sqrt(r2);
            const TS ir2 = ir * ir;
            const TS r3inv = ir2 * ir;
            const TS pref = TS(GM) * r3inv;
            if (allowed) { adt[0] -= pref * fx; adt[1] -= pref * fy; adt[2] -= pref * fz; }
            if (need_jac) {
                const TS r5 = TS(3) * r3inv * ir2;
                const TS g = TS(GM);
                jd[0] += g * (fx * fx * r5 - r3inv); jd[1] += g * (fy * fy * r5 - r3inv); jd[2] += g * (fz * fz * r5 - r3inv);
                jd[3] += g * (fx * fy * r5); jd[4] += g * (fy * fz * r5); jd[5] += g * (fx * fz * r5);
            }
            return;
        }
        const D r2 = dx * dx + dy * dy + dz * dz;
        const D ir = inv_sqrt(r2);
        const D ir2 = ir * ir;
        const D r3inv = ir2 * ir;
        const D pref = GM * r3inv;
        const D ax = -pref * dx, ay = -pref * dy, az = -pref * dz;
        if (allowed) { ad[0] += ax; ad[1] += ay; ad[2] += az; }
        if (need_jac) {
            const TS r3t = TS(r3inv);
            const TS r5 = TS(D(3.0) * r3inv * ir2);
            const TS g = TS(GM), fx = TS(dx), fy = TS(dy), fz = TS(dz);
            jd[0] += g * (fx * fx * r5 - r3t); jd[1] += g * (fy * fy * r5 - r3t); jd[2] += g * (fz * fz * r5 - r3t);
            jd[3] += g * (fx * fy * r5); jd[4] += g * (fy * fz * r5); jd[5] += g * (fx * fz * r5);
        }
        if (eih && b < kNumPlanets) {
            S0 += GM * ir;
            dS0[0] += ax; dS0[1] += ay; dS0[2] += az;
            if (!Ext && b != 0) {
                const D sx = sun[0] - pk[0], sy = sun[1] - pk[1], sz = sun[2] - pk[2];
                const D q2 = sx * sx + sy * sy + sz * sz;
                const D iq = inv_sqrt(q2);
                S1 += GM * iq;
                const D f = GM * (iq * iq * iq);
                aj[0] -= f * sx; aj[1] -= f * sy; aj[2] -= f * sz;
            }
        }
    };
    if constexpr (kStream) {
        // one system per lane: each slot is evaluated and folded in at once (the Sun and the
        // EMB, which the others need, were evaluated above); four slots in flight (3% faster
        // than two on 1,048,576 real-only systems)
        auto fold = [&](int s, const ChebOut<D>& c) {
            D pk[3];
            if (s < kFirstAsteroidSlot) {
                const bool em = (s == kSlotEarth || s == kSlotMoon);
                pk[0] = (em ? c.x + emb[0] : c.x) * inv_au;
                pk[1] = (em ? c.y + emb[1] : c.y) * inv_au;
                pk[2] = (em ? c.z + emb[2] : c.z) * inv_au;
                if (s == kSlotEarth) { earth[0] = pk[0]; earth[1] = pk[1]; earth[2] = pk[2]; }
            } else {
                pk[0] = fma(c.x, D(E.inv_au_ast), sun[0]);
                pk[1] = fma(c.y, D(E.inv_au_ast), sun[1]);
                pk[2] = fma(c.z, D(E.inv_au_ast), sun[2]);
            }
            const int b = slot_body(s);
            if (b < nbody) body_term(b, pk);
        };
#pragma unroll 4
        for (int s = 1; s < kFirstAsteroidSlot; s++) {
            if (s == kSlotEMB) continue;
            fold(s, cheb_eval<D, false>(slots[s], t));
        }
        // the asteroids with the short series where every asteroid record allows it (the bound
        // is the same on every lane: a scalar branch)
#if defined(__HIP_DEVICE_COMPILE__)
        const int short_from = __builtin_amdgcn_readfirstlane(E.short_from);
#else
        const int short_from = E.short_from;
#endif
        if (short_from <= kFirstAsteroidSlot) {
#pragma unroll 4
            for (int s = kFirstAsteroidSlot; s < kNumSlots; s++) fold(s, cheb_eval<D, false, false, kCoefShort>(slots[s], t));
        } else {
#pragma unroll 4
            for (int s = kFirstAsteroidSlot; s < kNumSlots; s++) fold(s, cheb_eval<D, false>(slots[s], t));
        }
        body_term(0, sun);
    } else {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int s = r * G + lane;
            if (s >= kNumSlots || s == kSlotEMB) continue;
            const int b = slot_body(s);
            if (b >= nbody) continue;
            body_term(b, px[r]);
        }
    }
    // 4. group sums
    if constexpr (G > 1) {
#pragma unroll
        for (int c = 0; c < 3; c++) { ad[c] = gsum<G>(ad[c]); adt[c] = gsum<G>(adt[c]); }
        if (need_jac) {
#pragma unroll
            for (int k = 0; k < 6; k++) jd[k] = gsum<G>(jd[k]);
        }
        if (eih) {
            S0 = gsum<G>(S0); S1 = gsum<G>(S1);
#pragma unroll
            for (int c = 0; c < 3; c++) { dS0[c] = gsum<G>(dS0[c]); aj[c] = gsum<G>(aj[c]); }
        }
    }
    // 5. single-particle terms, reference order
    D acc[3] = {D(0.0), D(0.0), D(0.0)};
    jac_zero(J);
    const D GMs = D(E.GM[0]);
    if ((forces & F_NONGRAV) && ng) term_nongrav<D, TS>(O, xr, vr, sun, sunv, A, need_jac, acc, J);
    if (forces & F_EARTH_H) term_earth_J2J4<D, TS>(E, xr, earth, D(E.GM[3]), need_jac, acc, J);
    if (forces & F_SUN_H) term_solar_J2<D, TS>(E, xr, sun, GMs, need_jac, acc, J);
    if constexpr (Ext) {
        if (eih) eih_sources<G>(E, O, gm, lane, xr, vr, px, pv, xo, vo, ao, S0, dS0, need_jac, acc, J);
    } else {
        if (eih) term_eih_sun<D, TS>(E, xr, vr, sun, sunv, GMs, S0, dS0, S1, aj, need_jac, acc, J);
    }
    if (forces & F_GR_POTENTIAL) term_gr_potential<D, TS>(E, xr, sun, GMs, need_jac, acc, J);
    if (forces & F_GR_SIMPLE) term_gr_simple<D, TS>(E, xr, vr, sun, sunv, GMs, need_jac, acc, J);
    if (forces & (F_SUN | F_PLANETS | F_ASTEROIDS)) {
#pragma unroll
        for (int c = 0; c < 3; c++) acc[c] += ad[c] + D(adt[c]);
        if (need_jac) {
            J.x[0][0] += jd[0]; J.x[1][1] += jd[1]; J.x[2][2] += jd[2];
            J.x[0][1] += jd[3]; J.x[1][0] += jd[3];
            J.x[1][2] += jd[4]; J.x[2][1] += jd[4];
            J.x[0][2] += jd[5]; J.x[2][0] += jd[5];
        }
    }
    if constexpr (Ext) {
        // indirect term of the geocentric frame (real particle only; ao is zero otherwise)
#pragma unroll
        for (int c = 0; c < 3; c++) acc[c] -= ao[c];
    }
    out.acc[0] = acc[0]; out.acc[1] = acc[1]; out.acc[2] = acc[2];
}

// Acceleration of this lane's own vector: the real particle (lane 0) or a variational
// particle (J applied to its state and its Marsden-parameter variation).
template <typename D, typename TS>
AHD void own_accel(int lane, const EvalResult<D>& R, const Jac<TS>& J, const D x[3], const D v[3], const double dA[3],
                   bool ng, D out[3]) {
    if (lane == 0) {
        out[0] = R.acc[0]; out[1] = R.acc[1]; out[2] = R.acc[2];
        return;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        TS s = J.x[c][0] * TS(x[0]) + J.x[c][1] * TS(x[1]) + J.x[c][2] * TS(x[2]);
        s += J.v[c][0] * TS(v[0]) + J.v[c][1] * TS(v[1]) + J.v[c][2] * TS(v[2]);
        if (ng) s += J.A[c][0] * TS(dA[0]) + J.A[c][1] * TS(dA[1]) + J.A[c][2] * TS(dA[2]);
        out[c] = D(s);
    }
}

}  // namespace assist_gpu
