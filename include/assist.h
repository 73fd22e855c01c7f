/*
 * assist-amd public C ABI (libassist.so).
 *
 * Part 1 keeps the reference's API surface (reference src/assist.h:36-304,
 * src/forces.h:30-36): the same enums, the same `struct assist_ephem` and
 * `struct assist_extras` field layouts (the Python wrapper mirrors them with
 * ctypes, reference assist/ephem.py:95-120 and assist/extras.py:84-100), the
 * same function names, argument meanings and error conventions.
 *
 * REBOUND, which hosts the reference as a plugin, is not part of this build;
 * `struct reb_simulation` / `struct reb_particle` below are a minimal
 * project-defined simulation carrying exactly the fields the force model and
 * the drop-in entry points use.
 *
 * Part 2 adds batched entry points: many independent systems (one real test
 * particle plus its first-order variational particles), each with its own
 * time, IAS15 step, span and direction, evaluated/integrated on the GPU.
 */
#ifndef ASSIST_AMD_ASSIST_H
#define ASSIST_AMD_ASSIST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------
 * Part 1: reference API surface
 * ---------------------------------------------------------------------------------------- */

extern const char* assist_build_str;    /* reference src/assist.h:36 */
extern const char* assist_version_str;  /* reference src/assist.h:37 */
extern const char* assist_githash_str;  /* reference src/assist.h:38 */
extern const char* assist_error_messages[];  /* reference src/assist.c:52-59 */
extern const int assist_error_messages_N;    /* reference src/assist.c:60 */

/* reference src/assist.h:52-57 */
typedef enum {
    FILE_FORMAT_VALID_BSP = 0,
    FILE_FORMAT_ASCII_BIN = 2,
    FILE_FORMAT_UNKNOWN = 3
} ephemeris_file_format_t;

/* reference src/assist.h:76-87 */
enum ASSIST_FORCES {
    ASSIST_FORCE_NONE = 0,
    ASSIST_FORCE_SUN = 0x01,
    ASSIST_FORCE_PLANETS = 0x02,
    ASSIST_FORCE_ASTEROIDS = 0x04,
    ASSIST_FORCE_NON_GRAVITATIONAL = 0x08,
    ASSIST_FORCE_EARTH_HARMONICS = 0x10,
    ASSIST_FORCE_SUN_HARMONICS = 0x20,
    ASSIST_FORCE_GR_EIH = 0x40,
    ASSIST_FORCE_GR_SIMPLE = 0x80,
    ASSIST_FORCE_GR_POTENTIAL = 0x100,
};

/* reference src/assist.h:90-98 */
enum ASSIST_STATUS {
    ASSIST_SUCCESS,
    ASSIST_ERROR_EPHEM_FILE,
    ASSIST_ERROR_AST_FILE,
    ASSIST_ERROR_NAST,
    ASSIST_ERROR_NEPHEM,
    ASSIST_ERROR_COVERAGE,
    ASSIST_ERROR_N,
};

/* reference src/assist.h:102-116 */
enum ASSIST_BODY {
    ASSIST_BODY_SUN = 0,
    ASSIST_BODY_MERCURY = 1,
    ASSIST_BODY_VENUS = 2,
    ASSIST_BODY_EARTH = 3,
    ASSIST_BODY_MOON = 4,
    ASSIST_BODY_MARS = 5,
    ASSIST_BODY_JUPITER = 6,
    ASSIST_BODY_SATURN = 7,
    ASSIST_BODY_URANUS = 8,
    ASSIST_BODY_NEPTUNE = 9,
    ASSIST_BODY_PLUTO = 10,
    ASSIST_BODY_NPLANETS = 11,
};

struct ascii_s;
struct spk_s;

/* Same field order as reference src/assist.h:125-155. */
struct assist_ephem {
    double jd_ref;
    struct spk_s* spk_planets;
    struct spk_s* spk_asteroids;
    struct ascii_s* ascii_planets;
    int planets_source;
    int (*planets_calc)(const struct assist_ephem*, double, double, int, double* const, double* const,
                        double* const, double* const, double* const, double* const, double* const,
                        double* const, double* const, double* const);
    int spk_target_index[ASSIST_BODY_NPLANETS];
    int spk_emb_index;
    double AU;
    double EMRAT;
    double J2E;
    double J3E;
    double J4E;
    double J2SUN;
    double RE;
    double CLIGHT;
    double ASUN;
    double Re_eq;
    double Rs_eq;
    double c_AU_per_day;
    double c_squared;
    double over_c_squared;
};

/* reference src/assist.h:157-175 */
struct assist_cache_item {
    double GM, x, y, z, vx, vy, vz, ax, ay, az;
};
struct assist_ephem_cache {
    double* t;
    double dt_sign;
    struct assist_cache_item* items;
};

/* Minimal simulation (REBOUND is not part of this build). */
struct reb_particle {
    double x, y, z, vx, vy, vz, ax, ay, az;
    double m;
};
struct reb_variational_configuration {
    int index;        /* particle index of the variational particle */
    int testparticle; /* index of the real particle it belongs to */
};
struct reb_simulation_integrator_ias15 {
    double epsilon;   /* default 1e-9; 0 selects a fixed step */
    double min_dt;
    int adaptive_mode; /* 1 (global criterion) is the only mode; set by assist_init */
};
struct reb_simulation {
    double t;
    double dt;
    double dt_last_done;
    int N;            /* real + variational particles */
    int N_var;
    int N_var_config;
    int allocated_N;
    struct reb_particle* particles;
    struct reb_variational_configuration* var_config;
    int status;       /* 0 running/success, >0 error (reference REB_STATUS_GENERIC_ERROR semantics) */
    int exact_finish_time;
    struct reb_simulation_integrator_ias15 ri_ias15;
    void* extras;
    void (*additional_forces)(struct reb_simulation*);
    char error_message[256];
};

/* Same field order as reference src/assist.h:177-195. */
struct assist_extras {
    struct reb_simulation* sim;
    struct assist_ephem* ephem;
    struct assist_ephem_cache* ephem_cache;
    int extras_should_free_ephem;
    int geocentric;
    struct reb_particle* last_state;
    struct reb_particle* current_state;
    double* particle_params;
    int steps_done;
    int forces;
    int gr_eih_sources;
    double alpha;
    double nk;
    double nm;
    double nn;
    double r0;
};

/* Ephemeris lifecycle and queries -- reference src/assist.h:271-300, src/assist.c:188-554 */
struct assist_ephem* assist_ephem_create(char* planets_file_name, char* asteroids_file_name);
int assist_ephem_init(struct assist_ephem* ephem, char* user_planets_path, char* user_asteroids_path);
void assist_ephem_free(struct assist_ephem* ephem);
void assist_ephem_free_pointers(struct assist_ephem* ephem);
void assist_ephem_time_bounds(const struct assist_ephem* ephem, double* t_beg, double* t_end);
struct reb_particle assist_get_particle(const struct assist_ephem* ephem, const int particle_id, const double t);
struct reb_particle assist_get_particle_with_error(const struct assist_ephem* ephem, const int particle_id,
                                                   const double t, int* error);
/* reference src/forces.h:32-36 (cache may be NULL) */
int assist_all_ephem(const struct assist_ephem* ephem, struct assist_ephem_cache* cache, const int i, const double t,
                     double* const GM, double* const x, double* const y, double* const z, double* const vx,
                     double* const vy, double* const vz, double* const ax, double* const ay, double* const az);

/* Format detection and discovery -- reference src/assist.h:63-70, src/assist.c:66-154 */
int assist_detect_ascii_bin_signature(int fd);
ephemeris_file_format_t assist_detect_ephemeris_file_format(int fd);
int assist_discover_planets_path(char* out_path, size_t out_path_size, const char* assist_dir);

/* Extras lifecycle -- reference src/assist.h:197-229, src/assist.c:379-501 */
struct assist_extras* assist_attach(struct reb_simulation* sim, struct assist_ephem* ephem);
void assist_init(struct assist_extras* assist, struct reb_simulation* sim, struct assist_ephem* ephem);
void assist_detach(struct reb_simulation* sim, struct assist_extras* assist);
void assist_free(struct assist_extras* assist);
void assist_free_pointers(struct assist_extras* assist);
void assist_error(struct assist_extras* assist, const char* const msg);

/* The force hook -- reference src/forces.h:30 (assist_additional_forces, installed as
 * REBOUND's additional_forces at src/assist.c:444).  Evaluated on the GPU. */
void assist_additional_forces(struct reb_simulation* sim);

/* Minimal simulation API (stands in for REBOUND's). */
struct reb_simulation* reb_simulation_create(void);
void reb_simulation_init(struct reb_simulation* sim);          /* for caller-allocated structs */
void reb_simulation_free(struct reb_simulation* sim);
void reb_simulation_free_pointers(struct reb_simulation* sim); /* frees particles, keeps the struct */
void reb_simulation_add(struct reb_simulation* sim, struct reb_particle p);
int reb_simulation_add_variation_1st_order(struct reb_simulation* sim, int testparticle);
/* Integrates with IAS15 on the GPU to tmax (exact finish). */
void reb_simulation_integrate(struct reb_simulation* sim, double tmax);
void reb_simulation_error(struct reb_simulation* sim, const char* msg);

/* ------------------------------------------------------------------------------------------
 * Part 2: batched GPU entry points (no counterpart in the reference)
 * ---------------------------------------------------------------------------------------- */

enum ASSIST_BATCH_ERROR {
    ASSIST_BATCH_OK = 0,
    ASSIST_BATCH_ERROR_ARG = 100,      /* invalid argument */
    ASSIST_BATCH_ERROR_DEVICE = 101,   /* HIP runtime error or no GPU */
    ASSIST_BATCH_ERROR_UNSUPPORTED = 102, /* option not supported on the GPU path */
};

/* Per-system status codes written by the batched integrator. */
enum ASSIST_SYSTEM_STATUS {
    ASSIST_SYSTEM_OK = 0,
    ASSIST_SYSTEM_COVERAGE = 5,      /* == ASSIST_ERROR_COVERAGE */
    ASSIST_SYSTEM_NONFINITE = -1,
    ASSIST_SYSTEM_MAX_STEPS = -2,
    ASSIST_SYSTEM_DT_UNDERFLOW = -3,
    ASSIST_SYSTEM_PAUSED = -4,       /* stopped at its pause time (assist_batch_extra.t_pause) */
};

/* Snapshot of one system's IAS15 state at a step boundary, in doubles (device memory):
 *   header[ASSIST_SNAPSHOT_HEADER]: t, next trial step, last accepted step (signed; 0 before the
 *     first step), steps, force evaluations, rejected attempts, status, lanes per system of the
 *     kernel that took it;
 *   then per vector (real particle first) ASSIST_SNAPSHOT_VECTOR doubles: x[3], v[3], the Kahan
 *     compensations csx[3], csv[3], the acceleration a0[3] at the start of the last accepted step,
 *     that step's b[7][3] and the correction e[7][3] of its prediction (k-major: b[3k + c]).
 * A checkpoint holds everything IAS15 carries from one step to the next, so that an integration
 * resumed from it takes the same steps as one that was never interrupted. */
#define ASSIST_SNAPSHOT_HEADER 8
#define ASSIST_SNAPSHOT_VECTOR 57
#define ASSIST_SNAPSHOT_WORDS(nvec) (ASSIST_SNAPSHOT_HEADER + ASSIST_SNAPSHOT_VECTOR * (nvec))

enum ASSIST_PRECISION { ASSIST_PRECISION_FP64 = 0, ASSIST_PRECISION_MIXED = 1 };

struct assist_batch_config {
    size_t size;          /* sizeof(struct assist_batch_config), set by assist_batch_config_default;
                             the entry points reject any other value (ASSIST_BATCH_ERROR_ARG) */
    double epsilon;       /* IAS15 epsilon (default 1e-9); 0 = fixed step */
    double min_dt;        /* default 0 */
    double dt_initial;    /* first trial step in days (default 0.001) */
    long long max_steps;  /* per system (default 10^6) */
    int precision;        /* ASSIST_PRECISION_* */
    int lanes_per_system; /* 0 = choose (autotuned), else 1, 2, 4 or 8 (>= 1 + n_var) */
    int block_size;       /* 0 = choose, else a multiple of 64 up to 256 */
    int blocks_per_cu;    /* 0 = occupancy limit */
    int waves_per_simd;   /* 0 = choose (autotuned), else 1 or 2: the kernel variant's register budget */
    int use_graph;        /* 1: launch the integration as a captured HIP graph, captured once per
                             argument set and replayed (the launches of a replay take the same
                             device pointers); 0 (default): direct launches */
    int substep_cache_off; /* whether 8-lane groups keep the ephemeris positions of each substep
                             for the later predictor-corrector iterations of the step (results
                             are the same bit for bit either way).  0 (default): keep them unless
                             the first-pass stop is on (the mixed form's default), where most
                             attempts take one pass and the stores cost more than the reuse
                             saves; 1: evaluate them every time; -1: keep them */
    int reserved_;
    double corrector_tolerance; /* when the predictor-corrector iterations stop.  0 (default): by
                             precision -- FP64 takes IAS15's own test (every vector's
                             max|db6| / max|a| below 1e-16, as REBOUND does); the mixed form
                             stops when the real particle's max|db6| / max|a| is below
                             epsilon / 100 (1e-16 when epsilon is 0).  > 0: that tolerance on the
                             real particle, in either precision.  < 0: IAS15's own test. */
    double first_pass_tolerance; /* a step attempt also ends after its first predictor-corrector
                             pass when that pass changed the real particle's b6 by less than
                             this fraction of its acceleration (the same measure), tightened by
                             first_pass_accel below.  0 (default): by precision -- the mixed form
                             3 epsilon, FP64 none.  > 0: that tolerance, in either precision.
                             < 0: none. */
    double first_pass_accel; /* acceleration scale of the first-pass test (au/day^2): where the
                             real particle's max|a| exceeds it, the first-pass tolerance is
                             multiplied by first_pass_accel / max|a|, down to a tenth -- close
                             to the Sun and near planets the test tightens.  0 (default): 6e-5
                             (the Sun's pull at about 2.2 au).  > 0: that scale.  < 0: none (a
                             purely relative test). */
};

struct assist_batch_stats {
    long long steps, rejected, evals;  /* totals over the batch, filled by assist_batch_integrate;
                                          the entry points on device buffers leave them 0 (their
                                          per-system counts are in d_steps / d_evals) */
    double kernel_ms;     /* device time of the integration kernel (HIP events) */
    int lanes_per_system, block_size, grid_blocks, waves_per_simd;
};

void assist_batch_config_default(struct assist_batch_config* cfg);
struct assist_batch_extra;
/* Zeroes every field of `ex`, sets its size and trace_every = 1. */
void assist_batch_extra_init(struct assist_batch_extra* ex);
const char* assist_batch_error_string(int code);
/* Number of visible GPUs (0 when none / runtime unavailable). */
int assist_gpu_device_count(void);

/* Force model for n systems at their own times t[i].  `opts` carries forces, gr_eih_sources,
 * geocentric and the Marsden shape (alpha, nk, nm, nn, r0); NULL selects the defaults of
 * assist_init.  states: n x (1+n_var) x 6 (x,y,z,vx,vy,vz; real particle first), params:
 * n x (1+n_var) x 3 (A1,A2,A3 for the real particle, their variations for the variational
 * ones) or NULL.  acc: n x (1+n_var) x 3.  status: per system (ASSIST_ERROR_COVERAGE etc.).
 * geocentric = 1 (states relative to the Earth) and gr_eih_sources from 2 to 11 run in FP64
 * only: with ASSIST_PRECISION_MIXED they return ASSIST_BATCH_ERROR_UNSUPPORTED. */
int assist_batch_forces(const struct assist_ephem* ephem, const struct assist_extras* opts, int n, int n_var,
                        const double* t, const double* states, const double* params, double* acc, int* status,
                        int precision);

/* Batched ephemeris queries on the GPU: query q is ASSIST body body[q] (0..26) at t[q] (days
 * from jd_ref), as reference assist_all_ephem / assist_get_particle (src/forces.c:175-263,
 * src/assist.c:530-554) answer one at a time.  out: n x 10 = GM, x, y, z, vx, vy, vz, ax, ay, az
 * (au, au/d, au/d^2; asteroids: barycentric position, NaN velocity and acceleration); status:
 * per query (ASSIST_ERROR_COVERAGE, ASSIST_ERROR_NEPHEM, ASSIST_ERROR_NAST, ASSIST_ERROR_AST_FILE;
 * out is zero then) or NULL.  Host buffers here; device buffers and a stream below. */
int assist_batch_get_particles(const struct assist_ephem* ephem, int n, const int* body, const double* t, double* out,
                               int* status);
int assist_batch_get_particles_device(const struct assist_ephem* ephem, int n, const int* d_body, const double* d_t,
                                      double* d_out, int* d_status, void* stream);

/* Integrates n systems with IAS15, system i from t0[i] to t1[i] (either direction), states
 * updated in place.  Host pointers; copies to and from the GPU are included. */
int assist_batch_integrate(const struct assist_ephem* ephem, const struct assist_extras* opts,
                           const struct assist_batch_config* cfg, int n, int n_var, const double* t0,
                           const double* t1, double* states, const double* params, int* status,
                           long long* steps, struct assist_batch_stats* stats);

/* Same with device pointers (inputs already resident in HBM); `stream` is a hipStream_t or
 * NULL.  Returns after the kernel completes; stats->kernel_ms is the kernel's device time. */
int assist_batch_integrate_device(const struct assist_ephem* ephem, const struct assist_extras* opts,
                                  const struct assist_batch_config* cfg, int n, int n_var, const double* d_t0,
                                  const double* d_t1, double* d_states, const double* d_params, int* d_status,
                                  long long* d_steps, long long* d_evals, void* stream,
                                  struct assist_batch_stats* stats);

/* Optional inputs and outputs of assist_batch_integrate_device_ex (device pointers).  Initialise
 * with assist_batch_extra_init; the entry points reject a struct whose size field differs. */
struct assist_batch_extra {
    size_t size;              /* sizeof(struct assist_batch_extra) */
    const double* dt0;        /* [n] first trial step per system (days; sign taken from t1 - t0), or NULL */
    const double* bw;         /* [n][21] warm start of the real particle's first step, b[k][c] =
                                 bw[3k + c] |a0| dt^(k+1), or NULL */
    double* trace;            /* [capacity][ASSIST_TRACE_WORDS] sampled accepted steps, or NULL */
    int* trace_count;         /* device counter of records requested (may exceed the capacity) */
    int trace_capacity;
    int trace_every;          /* record accepted steps s >= trace_skip with s % trace_every == 0 */
    int trace_skip;
    int reserved_;
    /* Dense output (all NULL for none): system i writes its state (every vector) at the times
     * out_t[out_offset[i] .. out_offset[i+1]-1], sorted in the direction of integration and not
     * before t0, to out_states[j][1 + n_var][6], interpolating the accepted step that contains each
     * time with its IAS15 polynomial (reference src/assist.c:556-597, 642-680).  Steps are then
     * not shortened to land on t1 (d_t1 may be NULL); the system stops after the step that covers
     * its last output time, and `states` receives the state at the end of that step. */
    const int* out_offset;
    const double* out_t;
    double* out_states;
    /* Checkpoint and resume (all NULL for none), ASSIST_SNAPSHOT_WORDS(1 + n_var) doubles per
     * system.  snapshot_in: each system resumes from its snapshot; t0 and dt0 are not read, t1
     * is the end as before, and the lanes per system recorded in the snapshot are used unless
     * the config sets them.  Not with dense output (ASSIST_BATCH_ERROR_UNSUPPORTED).
     * snapshot_out: each system's snapshot when it stops (finished, paused or failed).
     * t_pause: system i stops with status ASSIST_SYSTEM_PAUSED at the first step boundary at or
     * past t_pause[i] in its direction (unless that boundary is t1). */
    const double* snapshot_in;
    double* snapshot_out;
    const double* t_pause;
    /* Step archive (NULL for none; the counterpart of REBOUND's SimulationArchive saved every
     * step, which reference src/assist.c:599-633 reads): system i writes its snapshot at its
     * start and after every accepted step to archive[i][k][ASSIST_SNAPSHOT_WORDS(1 + n_var)],
     * k < archive_capacity, and the number of snapshots it took to archive_count[i] (more than
     * the capacity when it ran out; the later ones are not written). */
    double* archive;
    int* archive_count;
    int archive_capacity;
    int reserved2_;
};

/* Same as assist_batch_integrate_device, with the learned start and step traces of `ex`. */
int assist_batch_integrate_device_ex(const struct assist_ephem* ephem, const struct assist_extras* opts,
                                     const struct assist_batch_config* cfg, int n, int n_var, const double* d_t0,
                                     const double* d_t1, double* d_states, const double* d_params, int* d_status,
                                     long long* d_steps, long long* d_evals, void* stream,
                                     struct assist_batch_stats* stats, const struct assist_batch_extra* ex);

/* Allocation-free variant of assist_batch_integrate_device (no counterpart in the reference):
 * for callers that cannot allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  It calls no hipMalloc, hipFree or other allocating
 * routine: its scratch (the work-queue counter and the b/e records of at most 65 536 resident
 * lanes, 22 MB) is a static device array of the library, which every GPU the library runs on holds
 * until the process ends, and the grid is capped to it.  It runs without the substep cache, the
 * autotuner, the queue order, graph capture and timing (same results bit for bit; use_graph = 1
 * returns ASSIST_BATCH_ERROR_UNSUPPORTED).  The ephemeris must already be resident on the GPU
 * (any other batched call with it makes it so; otherwise ASSIST_BATCH_ERROR_UNSUPPORTED).  Calls
 * that overlap on one GPU share that array and race: the library's lock serialises calls from
 * threads of one process, but not from several processes on one GPU.  Returns after the kernel
 * completes on `stream`. */
int assist_batch_integrate_device_noalloc(const struct assist_ephem* ephem, const struct assist_extras* opts,
                                          const struct assist_batch_config* cfg, int n, int n_var,
                                          const double* d_t0, const double* d_t1, double* d_states,
                                          const double* d_params, int* d_status, long long* d_steps,
                                          long long* d_evals, void* stream);

/* Dense output with host buffers: out_offset[n + 1], out_t[out_offset[n]], out_states
 * [out_offset[n]][1 + n_var][6]; the output times of each system are checked to be sorted in one
 * direction from t0[i] (ASSIST_BATCH_ERROR_ARG otherwise).  states receive the end of each
 * system's last step. */
int assist_batch_integrate_dense(const struct assist_ephem* ephem, const struct assist_extras* opts,
                                 const struct assist_batch_config* cfg, int n, int n_var, const double* t0,
                                 double* states, const double* params, const int* out_offset, const double* out_t,
                                 double* out_states, int* status, long long* steps, struct assist_batch_stats* stats);

/* States at arbitrary times from step archives (struct assist_batch_extra.archive): query q
 * asks for system q_sys[q] at q_t[q]; out_states[q][1 + n_var][6].  The state is the IAS15
 * polynomial of the archived step that contains the time, from the earlier snapshot's state and
 * the later one's a0, b and step (reference assist_interpolate_simulation, src/assist.c:682-752,
 * with the blob lookup of assist_create_interpolated_simulation, :599-633); a time equal to a
 * snapshot's returns its state.  Unlike the reference, the first step interpolates too (every
 * snapshot holds its step's a0).  Per-query status: 0, or ASSIST_ERROR_COVERAGE outside the
 * archived steps (out is zero then).  Device buffers and a stream; host buffers below. */
int assist_batch_archive_interpolate_device(int n, int n_var, const double* d_archive, int capacity,
                                            const int* d_count, int nq, const int* d_q_sys, const double* d_q_t,
                                            double* d_out_states, int* d_status, void* stream);
int assist_batch_archive_interpolate(int n, int n_var, const double* archive, int capacity, const int* count, int nq,
                                     const int* q_sys, const double* q_t, double* out_states, int* status);

/* The reference's dense-output call (reference src/assist.c:642-680): brings the simulation to
 * time t by integrating every real particle (with its variational particles) as its own system
 * and interpolating inside the step that covers t; sim->t becomes t. */
void assist_integrate_or_interpolate(struct assist_extras* ax, double t);

/* Device buffer helpers for callers without their own allocator. */
void* assist_device_malloc(size_t bytes);
void assist_device_free(void* p);
int assist_memcpy_h2d(void* dst, const void* src, size_t bytes);
int assist_memcpy_d2h(void* dst, const void* src, size_t bytes);
int assist_memcpy_d2d(void* dst, const void* src, size_t bytes);
int assist_device_synchronize(void);

/* ------------------------------------------------------------------------------------------
 * Part 3: learned IAS15 start (no counterpart in the reference)
 *
 * A multilayer perceptron (16 -> 128 -> 128 -> 128 -> 32, tanh; BF16 weights on the matrix
 * cores, FP32 accumulation) maps a system's state to the step IAS15 settles on and a warm start
 * for the first step's b coefficients.  It is trained on the GPU by the entry points below from
 * step traces of the plain integrator (assist_batch_integrate_device_ex).
 * ---------------------------------------------------------------------------------------- */
#define ASSIST_MLP_IN 16
#define ASSIST_MLP_HIDDEN 128
#define ASSIST_MLP_OUT 32
/* Trace record: t, x[3], v[3] at the start of the step, the previous step, the accepted step,
 * IAS15's next-step proposal before the growth cap, a0[3], b[7][3] (k-major). */
#define ASSIST_TRACE_WORDS 34

struct assist_mlp;
int assist_mlp_num_params(void);
/* Workspace for batches of up to max_batch rows; Glorot-uniform weights from `seed`. */
struct assist_mlp* assist_mlp_create(int max_batch, unsigned long long seed);
void assist_mlp_free(struct assist_mlp* net);
int assist_mlp_max_batch(const struct assist_mlp* net);
/* FP32 parameters, layer by layer: W[in][out] row-major, then b[out]. */
int assist_mlp_set_params(struct assist_mlp* net, const float* host);
int assist_mlp_get_params(const struct assist_mlp* net, float* host);
/* Per-output weights of the squared-error loss (host array of ASSIST_MLP_OUT). */
int assist_mlp_set_output_weights(struct assist_mlp* net, const float* host_w);
/* d_X[M][16] -> d_Y[M][32] (device pointers, M <= max_batch). */
int assist_mlp_forward(struct assist_mlp* net, const float* d_X, int M, float* d_Y, void* stream);
/* The same outputs, bit for bit, for any M, keeping no activations (one fused kernel: the
 * prediction path of assist_batch_predict_start). */
int assist_mlp_predict(struct assist_mlp* net, const float* d_X, int M, float* d_Y, void* stream);
/* Forward and backward pass of the loss against d_T[M][32]; gradients stay on the device
 * (host_grad, if given, receives a copy) and *loss receives the loss. */
int assist_mlp_loss_grad(struct assist_mlp* net, const float* d_X, const float* d_T, int M, float* loss,
                         float* host_grad, void* stream);
/* One Adam step (bias-corrected; step counts from 1) on the gradients of the last loss_grad. */
int assist_mlp_adam_step(struct assist_mlp* net, float lr, float beta1, float beta2, float eps, int step, void* stream);
/* Features of rows [t, x[3], v[3], dt_last, ...] (stride doubles apart): d_X[n][16] and, when
 * d_omega is given, the heliocentric mean motion sqrt(GM_sun / r^3) per row. */
int assist_step_features(const struct assist_ephem* ephem, int n, const double* d_rows, int stride, float* d_X,
                         double* d_omega);
/* First step (d_dt0[n]) and warm start (d_bw[n][21], may be NULL) of a batch from the network,
 * iterating the last-step input `iterations` times from 1 day; steps are clamped to
 * [min_dt, |t1 - t0|]. */
int assist_batch_predict_start(const struct assist_ephem* ephem, struct assist_mlp* net, int n, int n_var,
                               const double* d_t0, const double* d_t1, const double* d_states, double min_dt,
                               int iterations, double* d_dt0, double* d_bw);

#ifdef __cplusplus
}
#endif
#endif
