// Device-side layout of the ephemeris and per-launch arguments.
//
// The SPK type-2 data of the 28 Chebyshev targets the force model needs are
// flattened into one device buffer.  Each target ("slot") is a contiguous run
// of equal-length records; a record holds
//     [ jd_ref - jul(MID),  86400/RADIUS,  x_0, y_0, z_0,  x_1, y_1, z_1,  ... ]
// (the coefficients interleaved by degree, in the order the evaluation uses them)
// with every coordinate zero-padded to kCoefPad coefficients, so that all
// records have one size and one fully unrolled evaluation (a zero coefficient
// adds an exact zero).
// The first two words replace the SPK record's MID and RADIUS with the
// quantities the evaluation actually uses (reference src/spk.c:445,514:
// z = ((jd_ref - jul(MID)) + t) / (RADIUS/86400)).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define ASSIST_HD __host__ __device__
#else
#define ASSIST_HD
#endif

namespace assist_gpu {

constexpr int kNumSlots = 28;       // 12 planet-file targets + 16 asteroids
constexpr int kNumBodies = 27;      // ASSIST body indices 0..26
constexpr int kNumPlanets = 11;     // ASSIST_BODY_NPLANETS
constexpr int kCoefPad = 14;                 // coefficients per coordinate on the device
constexpr int kRecWords = 2 + 3 * kCoefPad;  // 44 doubles = 352 bytes per record
// Short form of the series for records of at most kCoefShort coefficients (the trailing padded
// terms, which add exact zeros, are skipped; 11 is the asteroid records' count of the ephemeris
// files used here, and DE440's outer planets have 6-8)
constexpr int kCoefShort = 11;
constexpr int kSlotEMB = 3;         // Earth-Moon barycentre (no force of its own)
constexpr int kSlotEarth = 4;       // Earth relative to the EMB
constexpr int kSlotMoon = 5;        // Moon relative to the EMB
constexpr int kFirstAsteroidSlot = 12;

// Slot order: Sun, Mercury, Venus, EMB, Earth/EMB, Moon/EMB, Mars, Jupiter,
// Saturn, Uranus, Neptune, Pluto, asteroids 0..15.  slot -> ASSIST body index
// (-1 for the EMB).
ASSIST_HD constexpr int slot_body(int s) {
    return s < 3 ? s : (s == 3 ? -1 : s - 1);
}

struct DevSlot {
    const double* rec;   // first record
    double c_init;       // jd_ref - jul(INIT) of the first record
    double inv_intlen;   // 86400 / INTLEN  (1/record length in days)
    int nrec;
    int P;               // coefficients per coordinate in the file (<= kCoefPad)
    int kind;            // 0 planet-file target, 1 asteroid
    int pad_;
};

struct DevEphem {
    DevSlot slot[kNumSlots];
    double GM[kNumBodies];
    double inv_au;        // 1 / AU [km]              (planets, reference src/spk.c:590)
    double inv_au_ast;    // 1 / 149597870.7          (asteroids, reference src/spk.c:470)
    double vel_scale;     // 86400 / AU               (km/s -> au/d)
    double t_beg, t_end;    // coverage of every planet-file target, days relative to jd_ref
    double ta_beg, ta_end;  // coverage of every asteroid target (the planet window when none)
    int n_ast;              // asteroids present (0 if no asteroid file)
    int short_from;         // first slot from which every slot has at most kCoefShort coefficients
    // force-model constants (reference src/spk.c:769-789)
    double J2E, J3E, J4E, J2SUN, Re_eq, Rs_eq, c_squared, over_c_squared;
    // pole rotations: Earth (RA 0, Dec 90 deg) and Sun (RA 286.13, Dec 63.87 deg),
    // cos/sin computed on the host (reference src/forces.c:484-490,670-676)
    double e_cosa, e_sina, e_cosd, e_sind;
    double s_cosa, s_sina, s_cosd, s_sind;
    // zonal-harmonic prefactors: 3 J2E Re^2 / 2, 5 J3E Re^3 / 2, 5 J4E Re^4 / 8, 3 J2SUN Rs^2 / 2
    double k_J2, k_J3, k_J4, k_sJ2;
};

struct ForceOpts {
    int forces;          // ASSIST_FORCE_* bitmask
    int gr_eih_sources;  // EIH sources: the first gr_eih_sources planet-file bodies (0..11)
    double alpha, nk, nm, nn, r0;  // Marsden g(r) shape
    int geocentric;      // 1: particle states relative to the Earth (reference src/forces.c:72-87)
    int pad_;
    // coverage window of a force evaluation (days from jd_ref): the planet file's, narrowed to the
    // asteroid file's when a direct term is on and asteroids are loaded (set with the ephemeris)
    double t_lo, t_hi;
};

// Coverage window of a force evaluation.  The reference reports coverage errors from its direct
// term (reference src/forces.c:315-321, 362-367), whose variational loop fetches the first asteroid
// whatever the force flags, and from the geocentric frame (:76-80): the asteroid window therefore
// applies when a direct term is on and an asteroid file is loaded.  The planet window always
// applies: outside it the reference's other terms would use unset ephemeris values.
ASSIST_HD inline void coverage_window(const DevEphem& E, ForceOpts& O) {
    O.t_lo = E.t_beg;
    O.t_hi = E.t_end;
    if ((O.forces & 0x07) && E.n_ast > 0) {  // ASSIST_FORCE_SUN | _PLANETS | _ASTEROIDS
        O.t_lo = O.t_lo > E.ta_beg ? O.t_lo : E.ta_beg;
        O.t_hi = O.t_hi < E.ta_end ? O.t_hi : E.ta_end;
    }
}

// IAS15 configuration for the batched integrator.
struct IntegratorOpts {
    double epsilon;
    double min_dt;
    double dt_initial;
    long long max_steps;
    // predictor-corrector stop: 0 = IAS15's own test, max|db6| / max|a| over every vector of the
    // system below 1e-16; > 0 = the real particle's max|db6| / max|a| below pc_tol
    double pc_tol;
    // > 0: an attempt also ends after its first pass when that pass's change (the same measure)
    // is below pc_tol1, multiplied by max(0.1, pc_aref / max|a|) where max|a| (the same
    // measure's denominator) exceeds pc_aref > 0
    double pc_tol1;
    double pc_aref;
};

enum SystemStatus : int {
    kOk = 0,
    kCoverage = 5,        // ASSIST_ERROR_COVERAGE
    kNonFinite = -1,
    kMaxSteps = -2,
    kDtUnderflow = -3,
    kPaused = -4,         // stopped at the first step boundary at or past its pause time
    kBadSystem = 100,     // a query named no system of the batch (ASSIST_BATCH_ERROR_ARG)
};

// Snapshot of one system's IAS15 state at a step boundary (doubles): a header, then one block
// per vector.  Header: t, the next trial step, the last accepted step (signed; 0 before the
// first), steps, evaluations, rejections, status, lanes per system of the kernel that took it.
// Vector block: x[3], v[3], their Kahan
// compensations csx[3], csv[3], the acceleration at the start of the last accepted step a0[3],
// that step's b[7][3] and the correction e[7][3] of its prediction (k-major: b[3k + c]).
constexpr int kSnapHeader = 8;
constexpr int kSnapVector = 57;
ASSIST_HD constexpr int snap_words(int nvec) { return kSnapHeader + kSnapVector * nvec; }

}  // namespace assist_gpu
