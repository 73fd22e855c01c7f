// Algorithmic operation count of the batched integrator.
//
// Instantiates the GPU force model (force_model.h) on the host with counting
// scalar types and one lane per system (G = 1), so every arithmetic operation
// the algorithm needs for one system is counted exactly once (the device
// replicates the single-particle terms across the lanes of a group; those
// copies are not algorithmic work).  Counting convention: +, -, *, /, sqrt and
// pow count 1, a fused multiply-add counts 2, sign changes, comparisons, fabs,
// frexp and class tests count 0.  The IAS15 arithmetic is replayed from the
// kernel's expressions.
//
// Output (JSON): FP64 and FP32 operations per force evaluation for the force
// configurations used by the benchmarks, the Chebyshev share of an evaluation,
// and the IAS15 operations per component and per system.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

static long long g_ops64 = 0, g_ops32 = 0;

struct CntF;
struct Cnt {
    double v;
    Cnt() : v(0) {}
    Cnt(double x) : v(x) {}
    explicit Cnt(const CntF& f);
};
struct CntF {
    float v;
    CntF() : v(0) {}
    CntF(double x) : v((float)x) {}
    explicit CntF(const Cnt& d) : v((float)d.v) {}
};
inline Cnt::Cnt(const CntF& f) : v(f.v) {}

#define BIN(T, C, op)                                                             \
    inline T operator op(T a, T b) { C++; return T(a.v op b.v); }                 \
    inline T& operator op##=(T& a, T b) { C++; a.v = a.v op b.v; return a; }
BIN(Cnt, g_ops64, +) BIN(Cnt, g_ops64, -) BIN(Cnt, g_ops64, *) BIN(Cnt, g_ops64, /)
BIN(CntF, g_ops32, +) BIN(CntF, g_ops32, -) BIN(CntF, g_ops32, *) BIN(CntF, g_ops32, /)
#undef BIN
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt operator+(Cnt a) { return a; }
inline CntF operator+(CntF a) { return a; }
inline CntF operator-(CntF a) { return CntF(-a.v); }
inline Cnt fma(Cnt a, Cnt b, Cnt c) { g_ops64 += 2; return Cnt(a.v * b.v + c.v); }
inline Cnt sqrt(Cnt a) { g_ops64++; return Cnt(::sqrt(a.v)); }
inline CntF sqrt(CntF a) { g_ops32++; return CntF(::sqrtf(a.v)); }
inline CntF pow(CntF a, CntF b) { g_ops32++; return CntF(::powf(a.v, b.v)); }
inline Cnt pow(Cnt a, Cnt b) { g_ops64++; return Cnt(::pow(a.v, b.v)); }
inline double value_of(Cnt a) { return a.v; }
#define CMP(T, op) inline bool operator op(T a, T b) { return a.v op b.v; }
CMP(Cnt, ==) CMP(Cnt, <) CMP(Cnt, >) CMP(Cnt, >=) CMP(CntF, ==) CMP(CntF, <) CMP(CntF, >) CMP(CntF, >=)
#undef CMP
inline Cnt fabs(Cnt a) { return Cnt(::fabs(a.v)); }
inline CntF fabs(CntF a) { return CntF(::fabsf(a.v)); }
inline Cnt frexp(Cnt a, int* e) { return Cnt(::frexp(a.v, e)); }
inline CntF frexp(CntF a, int* e) { return CntF(::frexpf(a.v, e)); }
inline bool signbit(Cnt a) { return std::signbit(a.v); }
inline bool signbit(CntF a) { return std::signbit(a.v); }
inline bool isfinite(Cnt a) { return std::isfinite(a.v); }
inline bool isfinite(CntF a) { return std::isfinite(a.v); }
inline Cnt inv_sqrt(Cnt a) { g_ops64++; return Cnt(1.0 / ::sqrt(a.v)); }
inline CntF inv_sqrt(CntF a) { g_ops32++; return CntF(1.0f / ::sqrtf(a.v)); }

#include "force_model.h"

namespace assist_gpu {
template <>
struct SingleOf<Cnt> { using type = CntF; };
}  // namespace assist_gpu

using namespace assist_gpu;

// A host DevEphem whose slots have the benchmark ephemeris' shapes (planets 14 coefficients,
// Earth/Moon 13, asteroids 11); the values only need to be finite.
static DevEphem fake_ephem(std::vector<double>& buf) {
    DevEphem E;
    memset(&E, 0, sizeof E);
    buf.assign(kRecWords * kNumSlots, 0.5);
    for (int s = 0; s < kNumSlots; s++) {
        const int P = (s == kSlotEarth || s == kSlotMoon) ? 13 : (s < kFirstAsteroidSlot ? 14 : 11);
        DevSlot& S = E.slot[s];
        S.rec = buf.data() + kRecWords * s;
        S.c_init = 0.0;
        S.inv_intlen = 0.0;
        S.nrec = 1;
        S.P = P;
    }
    for (int b = 0; b < kNumBodies; b++) E.GM[b] = 1e-10 * (b + 1);
    E.inv_au = 1.0 / 149597870.7;
    E.inv_au_ast = 1.0 / 149597870.7;
    E.t_beg = -1e9;
    E.t_end = 1e9;
    E.ta_beg = -1e9;
    E.ta_end = 1e9;
    E.n_ast = 16;
    E.J2E = 1e-3; E.J3E = -2e-6; E.J4E = -1e-6; E.J2SUN = 2e-7; E.Re_eq = 4e-5; E.Rs_eq = 4e-3;
    E.c_squared = 3e4; E.over_c_squared = 1.0 / 3e4;
    E.e_cosa = 1; E.e_sina = 0; E.e_cosd = 6e-17; E.e_sind = 1;
    E.s_cosa = 0.27; E.s_sina = -0.96; E.s_cosd = 0.44; E.s_sind = 0.9;
    E.k_J2 = 1.5 * E.J2E * E.Re_eq * E.Re_eq; E.k_J3 = 2.5 * E.J3E * E.Re_eq * E.Re_eq * E.Re_eq;
    E.k_J4 = 0.625 * E.J4E * E.Re_eq * E.Re_eq * E.Re_eq * E.Re_eq; E.k_sJ2 = 1.5 * E.J2SUN * E.Rs_eq * E.Rs_eq;
    return E;
}

struct Counts { long long f64, f32; };

template <bool Mixed>
static Counts count_eval(const DevEphem& E, int forces, int nvec, bool ng) {
    using TS = typename TermScalar<Cnt, Mixed>::type;
    ForceOpts O;
    O.forces = forces;
    O.gr_eih_sources = 1;
    O.alpha = 1.0; O.nk = 0.0; O.nm = 2.0; O.nn = 5.093; O.r0 = 1.0;
    coverage_window(E, O);
    const Cnt xr[3] = {Cnt(2.1), Cnt(-0.7), Cnt(0.3)}, vr[3] = {Cnt(0.002), Cnt(0.009), Cnt(0.001)};
    const double A[3] = {1e-12, 1e-13, 1e-14};
    g_ops64 = g_ops32 = 0;
    EvalResult<Cnt> R;
    Jac<TS> J;
    force_eval<1, Mixed, Cnt>(E, E.slot, E.GM, O, 0, Cnt(1.0), xr, vr, A, ng, nvec > 1, R, J);
    for (int v = 1; v < nvec; v++) {
        Cnt out[3];
        const double dA[3] = {0, 0, 0};
        own_accel<Cnt, TS>(v, R, J, xr, vr, dA, ng, out);
    }
    return Counts{g_ops64, g_ops32};
}

// Chebyshev operations of one force evaluation (the 28 slots' positions, and the Sun's velocity
// when a term needs it): the part an evaluation served by the substep ephemeris cache skips.
static long long count_cheb(const DevEphem& E, bool sun_vel) {
    g_ops64 = g_ops32 = 0;
    for (int s = 0; s < kNumSlots; s++) {
        if (s == 0 && sun_vel) (void)cheb_eval<Cnt, true>(E.slot[s], Cnt(1.0));
        else (void)cheb_eval<Cnt, false>(E.slot[s], Cnt(1.0));
    }
    return g_ops64;
}

// IAS15 arithmetic of the kernel (kernels.hip), replayed with the counting type.
struct IasCounts {
    long long substep_scalar;   // per system and substep: the predictor coefficients
    long long predict_comp;     // per component and substep: predicted x and v
    long long correct_comp[8];  // per component at substep n (1..7)
    long long setup_comp;       // per component per step attempt: b from br/er, g from b
    long long accept_comp;      // per component per accepted step: x0, v0 update and br/er
    long long accept_scalar;    // per accepted step: the end-of-step weights
    long long control;          // per step attempt: step-size control (system)
};

static IasCounts count_ias15() {
    IasCounts c;
    memset(&c, 0, sizeof c);
    Cnt dts(1.0), h(0.5), s[9], sv[8];
    const Cnt px(0.01), pv(0.02);  // table weights (constant memory)
    g_ops64 = 0;
    s[0] = dts * h;
    Cnt dt2p = dts * dts;
    for (int k = 1; k < 9; k++) s[k] = dt2p * px;
    for (int k = 1; k < 8; k++) sv[k] = dts * pv;
    Cnt te = Cnt(0.0) + s[0];
    (void)te;
    c.substep_scalar = g_ops64;
    Cnt b[7], g[7], x0(1), v0(1), a0(1), csx(0), csv(0);
    for (int k = 0; k < 7; k++) b[k] = g[k] = Cnt(0.1);
    g_ops64 = 0;
    Cnt xe = -csx + ((s[8] * b[6] + s[7] * b[5] + s[6] * b[4] + s[5] * b[3] + s[4] * b[2] + s[3] * b[1] + s[2] * b[0] +
                      s[1] * a0 + s[0] * v0) + x0);
    Cnt ve = -csv + ((sv[7] * b[6] + sv[6] * b[5] + sv[5] * b[4] + sv[4] * b[3] + sv[3] * b[2] + sv[2] * b[1] +
                      sv[1] * b[0] + s[0] * a0) + v0);
    (void)xe; (void)ve;
    c.predict_comp = g_ops64 - 2;  // the two sign changes of csx, csv are free
    const Cnt r(0.7), cc(0.3);
    for (int n = 1; n <= 7; n++) {
        g_ops64 = 0;
        Cnt at(2.0);
        Cnt gk = at - a0;
        Cnt val = gk * r;
        for (int k = 0; k < n - 1; k++) val = (val - g[k]) * r;
        Cnt tmp = val - g[n - 1];
        for (int k = 0; k < n - 1; k++) b[k] += tmp * cc;
        b[n - 1] += tmp;
        c.correct_comp[n] = g_ops64;
    }
    g_ops64 = 0;
    {
        Cnt q(1.1), br[7], er[7], e[7];
        for (int k = 0; k < 7; k++) br[k] = er[k] = Cnt(0.2);
        Cnt q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3, q6 = q3 * q3, q7 = q3 * q4;
        e[0] = q * (br[6] * Cnt(7.0) + br[5] * Cnt(6.0) + br[4] * Cnt(5.0) + br[3] * Cnt(4.0) + br[2] * Cnt(3.0) + br[1] * Cnt(2.0) + br[0]);
        e[1] = q2 * (br[6] * Cnt(21.0) + br[5] * Cnt(15.0) + br[4] * Cnt(10.0) + br[3] * Cnt(6.0) + br[2] * Cnt(3.0) + br[1]);
        e[2] = q3 * (br[6] * Cnt(35.0) + br[5] * Cnt(20.0) + br[4] * Cnt(10.0) + br[3] * Cnt(4.0) + br[2]);
        e[3] = q4 * (br[6] * Cnt(35.0) + br[5] * Cnt(15.0) + br[4] * Cnt(5.0) + br[3]);
        e[4] = q5 * (br[6] * Cnt(21.0) + br[5] * Cnt(6.0) + br[4]);
        e[5] = q6 * (br[6] * Cnt(7.0) + br[5]);
        e[6] = q7 * br[6];
        for (int k = 0; k < 7; k++) b[k] = e[k] + (br[k] - er[k]);
        Cnt d(0.01);
        g[0] = b[6] * d + b[5] * d + b[4] * d + b[3] * d + b[2] * d + b[1] * d + b[0];
        g[1] = b[6] * d + b[5] * d + b[4] * d + b[3] * d + b[2] * d + b[1];
        g[2] = b[6] * d + b[5] * d + b[4] * d + b[3] * d + b[2];
        g[3] = b[6] * d + b[5] * d + b[4] * d + b[3];
        g[4] = b[6] * d + b[5] * d + b[4];
        g[5] = b[6] * d + b[5];
        c.setup_comp = g_ops64;
    }
    g_ops64 = 0;
    {
        auto add_cs = [](Cnt& p, Cnt& cs, Cnt in) {
            Cnt y = in - cs;
            Cnt t = p + y;
            cs = (t - p) - y;
            p = t;
        };
        Cnt wx[8], wv[7];
        {
            const long long before = g_ops64;
            Cnt dt2 = dts * dts;
            for (int k = 0; k < 8; k++) wx[k] = px * dt2;
            for (int k = 0; k < 7; k++) wv[k] = pv * dts;
            c.accept_scalar = g_ops64 - before;
            g_ops64 = before;
        }
        for (int k = 6; k >= 0; k--) add_cs(x0, csx, b[k] * wx[k + 1]);
        add_cs(x0, csx, a0 * wx[0]);
        add_cs(x0, csx, v0 * dts);
        for (int k = 6; k >= 0; k--) add_cs(v0, csv, b[k] * wv[k]);
        add_cs(v0, csv, a0 * dts);
        // er = Q(q) br_old (same arithmetic as the setup's e)
        c.accept_comp = g_ops64 + (c.setup_comp - 7 * 3 - 21 - 21 + 6);  // e only: setup minus b = e+(br-er) and g = Db
    }
    g_ops64 = 0;
    {
        Cnt mb6(1e-10), mak(1e-4), eps(1e-9);
        Cnt err = mb6 / mak;
        Cnt dtn = pow(eps / err, Cnt(1.0 / 7.0)) * dts;
        Cnt ratio = dtn / dts;
        Cnt dq = dts / Cnt(1.0);
        (void)ratio; (void)dq;
        c.control = g_ops64 + 1;  // + the ratio of the next prediction
    }
    return c;
}

int main() {
    std::vector<double> buf;
    const DevEphem E = fake_ephem(buf);
    printf("{\n  \"convention\": \"+,-,*,/,sqrt,pow = 1; fma = 2; per system, counted once\",\n");
    printf("  \"eval\": {\n");
    const int cfgs[][2] = {{7, 0}, {7, 1}, {1, 0}, {1, 1}, {2, 0}};
    bool first = true;
    for (auto& cf : cfgs) {
        for (int mixed = 0; mixed < 2; mixed++) {
            const Counts k = mixed ? count_eval<true>(E, 0x7F, cf[0], cf[1]) : count_eval<false>(E, 0x7F, cf[0], cf[1]);
            printf("%s    \"nvec%d_ng%d_%s\": {\"fp64\": %lld, \"fp32\": %lld}", first ? "" : ",\n", cf[0], cf[1],
                   mixed ? "mixed" : "fp64", k.f64, k.f32);
            first = false;
        }
    }
    printf("\n  },\n");
    printf("  \"cheb\": {\"sun_vel\": %lld, \"no_sun_vel\": %lld},\n", count_cheb(E, true), count_cheb(E, false));
    const IasCounts c = count_ias15();
    long long corr = 0;
    for (int n = 1; n <= 7; n++) corr += c.correct_comp[n];
    printf("  \"ias15\": {\"substep_scalar\": %lld, \"predict_comp\": %lld, \"correct_comp_7substeps\": %lld, "
           "\"setup_comp\": %lld, \"accept_comp\": %lld, \"accept_scalar\": %lld, \"control\": %lld}\n}\n",
           c.substep_scalar, c.predict_comp, corr, c.setup_comp, c.accept_comp, c.accept_scalar, c.control);
    return 0;
}
