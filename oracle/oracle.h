/*
 * ASSIST CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's hot path, used as the checker by
 * tests/, by __graft_entry__.smoke() and as the CPU baseline leg of bench.py.
 * The product (assist_amd/libassist.so) never links or calls it.
 *
 *   spk.c    DAF/SPK loading and type-2 Chebyshev evaluation
 *            (reference src/spk.c:133-210, 267-547, 550-693, 696-789)
 *   ascii.c  JPL binary (.440/.441) planets: reader and evaluation
 *            (reference src/ascii_ephem.c:27-65, 105-252, 275-384)
 *   forces.c the force model, term by term, in the reference's order
 *            (reference src/forces.c:49-1983)
 *   ias15.c  IAS15 restated from Rein & Spiegel (2015), MNRAS 446, 1424
 *            (REBOUND is not in the reference tree; parity of the stepper is
 *            pinned by closed-form Kepler motion, time reversal and
 *            variational-vs-shadow tests, not by reference vectors)
 *   batch.c  one independent integrator per system, OpenMP over systems
 *
 * Build with -ffp-contract=off (see Makefile) so the arithmetic is the
 * expression-by-expression IEEE evaluation of the formulas.
 */
#ifndef ASSIST_ORACLE_H
#define ASSIST_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_SUCCESS = 0, ORC_ERROR_EPHEM_FILE = 1, ORC_ERROR_AST_FILE = 2, ORC_ERROR_NAST = 3,
       ORC_ERROR_NEPHEM = 4, ORC_ERROR_COVERAGE = 5 };

enum { ORC_FORCE_SUN = 0x01, ORC_FORCE_PLANETS = 0x02, ORC_FORCE_ASTEROIDS = 0x04,
       ORC_FORCE_NON_GRAVITATIONAL = 0x08, ORC_FORCE_EARTH_HARMONICS = 0x10,
       ORC_FORCE_SUN_HARMONICS = 0x20, ORC_FORCE_GR_EIH = 0x40, ORC_FORCE_GR_SIMPLE = 0x80,
       ORC_FORCE_GR_POTENTIAL = 0x100 };

#define ORC_NPLANETS 11

typedef struct {
    int code, cen;
    double mass, beg, end, res;
    int nseg;
    int *one, *two;
} orc_target;

typedef struct {
    orc_target *targets;
    int num;
    double *words; /* whole file, as 8-byte words */
    size_t nbytes;
} orc_spk;

typedef struct orc_ascii orc_ascii; /* JPL binary (.440/.441) planets file, ascii.c */

typedef struct {
    double jd_ref;
    orc_spk *planets, *asteroids;
    orc_ascii *ascii; /* planets from a JPL binary file (then planets == NULL) */
    int target_index[ORC_NPLANETS]; /* index into planets->targets, -1 if absent */
    int emb_index;
    double AU, EMRAT, J2E, J3E, J4E, J2SUN, RE, CLIGHT, ASUN;
    double Re_eq, Rs_eq, c_AU_per_day, c_squared, over_c_squared;
} orc_ephem;

typedef struct {
    int forces;
    int gr_eih_sources;
    int geocentric;
    double alpha, nk, nm, nn, r0;
} orc_opts;

/* One reference-style "simulation": n_real real particles followed by n_var
 * first-order variational particles; var particle v belongs to real particle
 * var_tp[v].  p holds x,y,z,vx,vy,vz,ax,ay,az per particle. */
typedef struct {
    int n_real, n_var;
    const int *var_tp;
    double t;
    double *p;
    const double *params; /* 3*(n_real+n_var) Marsden A1,A2,A3 / dA, or NULL */
    int status;
} orc_sys;

typedef struct {
    double epsilon, min_dt, dt; /* dt: initial (then proposed) step */
    int exact_finish;
    long max_steps;
    /* predictor-corrector stop: 0 = max|db6| / max|a| over every component below 1e-16;
     * > 0 = the real particles' max|db6| / max|a| below pc_tol (the product's option) */
    double pc_tol;
    /* > 0: also stop after the first pass of an attempt when its change (the same measure) is
     * below pc_tol1 (the product's first-pass option); 0 = no first-pass stop */
    double pc_tol1;
    /* > 0: where the pass's max|a| (the measure's denominator) exceeds pc_aref, the first-pass
     * tolerance is pc_tol1 * max(0.1, pc_aref / max|a|) (the product's first-pass acceleration
     * scale) */
    double pc_aref;
} orc_ias15_cfg;

typedef struct {
    long steps, rejected, iterations, evals, pc_maxed;
    int status; /* 0 ok, >0 ephemeris error, -1 nonfinite, -2 max steps, -3 dt underflow */
    double dt_last_done, dt_next;
} orc_stats;

/* ---- ephemeris ---- */
orc_ephem *orc_ephem_load(const char *planets_path, const char *asteroids_path, int *err);
void orc_ephem_free(orc_ephem *e);
int orc_ephem_num_asteroids(const orc_ephem *e);
/* Body i at t (days rel. jd_ref): out = GM, x,y,z, vx,vy,vz, ax,ay,az (au, au/d, au/d^2). */
int orc_all_ephem(const orc_ephem *e, int i, double t, double out[10]);
/* Direction of integration for the ephemeris cache's replacement rule (this thread). */
void orc_cache_direction(int dir);
void orc_time_bounds(const orc_ephem *e, double *t_beg, double *t_end);
int orc_spk_eval_target(const orc_ephem *e, int which_file, int target_idx, double t, double out[9]);

/* ---- JPL binary planets (ascii.c) ---- */
int orc_ascii_signature(const unsigned char *buf, size_t n);
orc_ascii *orc_ascii_load(const char *path);
void orc_ascii_free(orc_ascii *a);
int orc_ascii_find_constant(const orc_ascii *a, const char *name, double *out);
int orc_ascii_calc(const orc_ascii *a, double jd_ref, double jd_rel, int body, double *GM, double o[9]);
double orc_ascii_emrat(const orc_ascii *a);
void orc_ascii_bounds(const orc_ascii *a, double *beg, double *end);
/* 0 J2E, 1 J3E, 2 J4E, 3 J2SUN, 4 AU, 5 RE, 6 CLIGHT, 7 ASUN */
double orc_ascii_constant(const orc_ascii *a, int which);

/* ---- forces ---- */
void orc_opts_default(orc_opts *o);
/* Accelerations of all particles of s at s->t (zeroes a first, like gravity=NONE). */
void orc_additional_forces(const orc_ephem *e, const orc_opts *o, orc_sys *s);

/* ---- IAS15 ---- */
void orc_ias15_cfg_default(orc_ias15_cfg *c);
int orc_integrate(const orc_ephem *e, const orc_opts *o, orc_ias15_cfg *cfg, orc_sys *s,
                  double t_end, orc_stats *stats);
void orc_ias15_tables(double h[8], double rr[28], double c[21], double d[21]);

/* ---- batched drivers (one independent system per entry) ---- */
/* states: nsys x (1+n_var) x 6; params: nsys x (1+n_var) x 3 or NULL; acc out: nsys x (1+n_var) x 3 */
int orc_batch_forces(const orc_ephem *e, const orc_opts *o, int nsys, int n_var, const double *t,
                     const double *states, const double *params, double *acc, int *status);
int orc_batch_integrate(const orc_ephem *e, const orc_opts *o, const orc_ias15_cfg *cfg, int nsys,
                        int n_var, const double *t0, const double *t1, double *states,
                        const double *params, int nthreads, int *status, long *steps, long *evals,
                        long *rejected, double *t_out);

#ifdef __cplusplus
}
#endif
#endif
