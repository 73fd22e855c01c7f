// Reference API surface of libassist (host side).
//
// Lifecycle, discovery, format detection, time bounds and ephemeris queries
// follow reference src/assist.c:45-554 and src/forces.c:175-263; the force hook
// and integration run on the GPU through the batched entry points (batch.cpp).
#include <fcntl.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <vector>

#include "../../include/ascii_ephem.h"
#include "internal.h"

#define ASSIST_STR2(s) #s
#define ASSIST_STR(s) ASSIST_STR2(s)
#ifndef ASSISTGITHASH
#define ASSISTGITHASH notavailable0000000000000000000000000001
#endif

extern "C" {

const char* assist_build_str = __DATE__ " " __TIME__;
const char* assist_version_str = "1.2.0";
const char* assist_githash_str = ASSIST_STR(ASSISTGITHASH);

const char* assist_error_messages[] = {
    "No error has occured.",
    "The JPL planet ephemeris file has not been found.",
    "The JPL asteroid ephemeris file has not been found. Asteroid forces have been disabled.",
    "The requested asteroid ID has not been found.",
    "The requested planet ID has not been found.",
    "The requested time is outside the coverage provided by the ephemeris file.",
};
const int assist_error_messages_N = ASSIST_ERROR_N;

// ---------------------------------------------------------------------------------------------
// format detection and discovery (reference src/assist.c:66-154)
// ---------------------------------------------------------------------------------------------
int assist_detect_ascii_bin_signature(int fd) {
    char names[18];
    if (lseek(fd, 0x00FC, SEEK_SET) != 0x00FC) return 0;
    if (read(fd, names, sizeof names) != (ssize_t)sizeof names) return 0;
    int valid = 0;
    for (int i = 0; i < 3; i++) {
        const char* nm = names + 6 * i;
        const bool letter0 = (nm[0] >= 'A' && nm[0] <= 'Z') || (nm[0] >= 'a' && nm[0] <= 'z');
        if (!letter0) continue;
        int alnum = 0, ok = 1;
        for (int j = 0; j < 6; j++) {
            const unsigned char c = (unsigned char)nm[j];
            if ((c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9')) alnum++;
            else if (c == ' ' || c == 0) {
            } else { ok = 0; break; }
        }
        if (ok && alnum >= 2) valid++;
    }
    return valid >= 2;
}

ephemeris_file_format_t assist_detect_ephemeris_file_format(int fd) {
    char buf[1024];
    lseek(fd, 0, SEEK_SET);
    const ssize_t n = read(fd, buf, sizeof buf);
    if (n <= 0) return FILE_FORMAT_UNKNOWN;
    if (n >= 8 && strncmp(buf, "DAF/SPK ", 8) == 0) return FILE_FORMAT_VALID_BSP;
    if (assist_detect_ascii_bin_signature(fd)) {
        lseek(fd, 0, SEEK_SET);
        return FILE_FORMAT_ASCII_BIN;
    }
    return FILE_FORMAT_UNKNOWN;
}

int assist_discover_planets_path(char* out_path, size_t out_path_size, const char* assist_dir) {
    if (!out_path || out_path_size == 0 || !assist_dir) return 0;
    static const char* candidates[] = {"/data/de441.bsp", "/data/de440.bsp", "/data/linux_m13000p17000.441",
                                       "/data/linux_p1550p2650.440"};
    for (const char* c : candidates) {
        snprintf(out_path, out_path_size, "%s%s", assist_dir, c);
        if (access(out_path, R_OK) == 0) return 1;
    }
    return 0;
}

// ---------------------------------------------------------------------------------------------
// ephemeris lifecycle (reference src/assist.c:188-373)
// ---------------------------------------------------------------------------------------------
static int planets_calc_cb(const struct assist_ephem* e, double jd_ref, double jd_rel, int body, double* const GM,
                           double* const x, double* const y, double* const z, double* const vx, double* const vy,
                           double* const vz, double* const ax, double* const ay, double* const az) {
    return assist_host::planets_calc(e, jd_ref, jd_rel, body, GM, x, y, z, vx, vy, vz, ax, ay, az);
}

static int ascii_planets_calc_cb(const struct assist_ephem* e, double jd_ref, double jd_rel, int body, double* const GM,
                                 double* const x, double* const y, double* const z, double* const vx,
                                 double* const vy, double* const vz, double* const ax, double* const ay,
                                 double* const az) {
    return assist_ascii_calc_from_ephem(e, jd_ref, jd_rel, body, GM, x, y, z, vx, vy, vz, ax, ay, az);
}

int assist_ephem_init(struct assist_ephem* ephem, char* user_planets_path, char* user_asteroids_path) {
    const int FN = 1024;
    char planets_path[FN], asteroids_path[FN];
    ephem->jd_ref = 2451545.0;
    ephem->spk_planets = nullptr;
    ephem->spk_asteroids = nullptr;
    ephem->ascii_planets = nullptr;
    const char* dir = getenv("ASSIST_DIR");
    if (!user_planets_path && !dir) return ASSIST_ERROR_EPHEM_FILE;
    if (!user_planets_path) {
        if (!assist_discover_planets_path(planets_path, FN, dir)) snprintf(planets_path, FN, "%s%s", dir, "/data/de441.bsp");
    } else {
        snprintf(planets_path, FN, "%s", user_planets_path);
    }
    ephemeris_file_format_t fmt = FILE_FORMAT_UNKNOWN;
    {
        int fd = open(planets_path, O_RDONLY);
        if (fd >= 0) {
            fmt = assist_detect_ephemeris_file_format(fd);
            close(fd);
        }
    }
    int ast_missing = 0;
    if (!user_asteroids_path) {
        if (!dir) ast_missing = 1;
        else snprintf(asteroids_path, FN, "%s%s", dir, "/data/sb441-n16.bsp");
    } else {
        snprintf(asteroids_path, FN, "%s", user_asteroids_path);
    }
    if (!ast_missing && !(ephem->spk_asteroids = assist_spk_init(asteroids_path))) ast_missing = 1;
    if (fmt == FILE_FORMAT_ASCII_BIN) {
        // planets from a JPL binary file: its constants table gives the force-model constants and
        // the asteroid masses (reference src/assist.c:291-327)
        ephem->ascii_planets = assist_ascii_init(planets_path);
        if (!ephem->ascii_planets) {
            if (ephem->spk_asteroids) { assist_spk_free(ephem->spk_asteroids); ephem->spk_asteroids = nullptr; }
            return ASSIST_ERROR_EPHEM_FILE;
        }
        const ascii_s* a = ephem->ascii_planets;
        ephem->planets_source = FILE_FORMAT_ASCII_BIN;
        for (int k = 0; k < ASSIST_BODY_NPLANETS; k++) ephem->spk_target_index[k] = -1;
        ephem->spk_emb_index = -1;
        ephem->J2E = a->J2E;
        ephem->J3E = a->J3E;
        ephem->J4E = a->J4E;
        ephem->J2SUN = a->J2SUN;
        ephem->AU = a->AU;
        ephem->RE = a->RE;
        ephem->CLIGHT = a->CLIGHT;
        ephem->ASUN = a->ASUN;
        ephem->EMRAT = a->cem;
        ephem->Re_eq = ephem->RE / ephem->AU;
        ephem->Rs_eq = ephem->ASUN / ephem->AU;
        ephem->c_AU_per_day = (ephem->CLIGHT / ephem->AU) * 86400.0;
        ephem->c_squared = ephem->c_AU_per_day * ephem->c_AU_per_day;
        ephem->over_c_squared = 1.0 / ephem->c_squared;
        ephem->planets_calc = ascii_planets_calc_cb;
        if (ephem->spk_asteroids) {
            for (int i = 0; i < ephem->spk_asteroids->num; i++) {
                spk_target& t = ephem->spk_asteroids->targets[i];
                if (t.code >= 2000000) {
                    char key[16];  // only the first 6 characters take part in the match
                    snprintf(key, sizeof key, "MA%04d", t.code - 2000000);
                    double gm = 0.0;
                    if (assist_ascii_find_constant(a, key, &gm)) t.mass = gm;
                } else if (t.code == 399) {
                    t.mass = a->mass[ASSIST_BODY_EARTH];
                } else if (t.code == 301) {
                    t.mass = a->mass[ASSIST_BODY_MOON];
                } else if (t.code == 10) {
                    t.mass = a->mass[ASSIST_BODY_SUN];
                }
            }
        }
        if (ast_missing) fprintf(stderr, "(ASSIST) %s\n", assist_error_messages[ASSIST_ERROR_AST_FILE]);
        return ASSIST_SUCCESS;
    }
    if (fmt != FILE_FORMAT_VALID_BSP) {
        fprintf(stderr, "(ASSIST) Error: Failed to initialize planets ephemeris from '%s'.\n", planets_path);
        fprintf(stderr, "(ASSIST) Supported planet files: NAIF SPK kernels (.bsp) and JPL binary ephemerides (.440/.441).\n");
        if (ephem->spk_asteroids) { assist_spk_free(ephem->spk_asteroids); ephem->spk_asteroids = nullptr; }
        return ASSIST_ERROR_EPHEM_FILE;
    }
    ephem->spk_planets = assist_spk_init(planets_path);
    if (!ephem->spk_planets) {
        if (ephem->spk_asteroids) { assist_spk_free(ephem->spk_asteroids); ephem->spk_asteroids = nullptr; }
        return ASSIST_ERROR_EPHEM_FILE;
    }
    ephem->planets_source = FILE_FORMAT_VALID_BSP;
    static const int naif[ASSIST_BODY_NPLANETS] = {10, 1, 2, 399, 301, 4, 5, 6, 7, 8, 9};
    for (int k = 0; k < ASSIST_BODY_NPLANETS; k++) ephem->spk_target_index[k] = assist_host::find_target(ephem->spk_planets, naif[k]);
    ephem->spk_emb_index = assist_host::find_target(ephem->spk_planets, 3);
    // constants and masses from the planets file's comment area (reference src/assist.c:330-345)
    struct spk_constants_and_masses c = assist_load_spk_constants_and_masses(planets_path);
    assist_apply_spk_constants(ephem, &c);
    assist_spk_join_masses(ephem->spk_planets, &c.masses, ephem->EMRAT);
    assist_spk_join_masses(ephem->spk_asteroids, &c.masses, ephem->EMRAT);
    assist_free_spk_constants_and_masses(&c);
    ephem->planets_calc = planets_calc_cb;
    if (ast_missing) fprintf(stderr, "(ASSIST) %s\n", assist_error_messages[ASSIST_ERROR_AST_FILE]);
    return ASSIST_SUCCESS;
}

struct assist_ephem* assist_ephem_create(char* planets_file_name, char* asteroids_file_name) {
    auto* e = (struct assist_ephem*)calloc(1, sizeof(struct assist_ephem));
    const int err = assist_ephem_init(e, planets_file_name, asteroids_file_name);
    if (err != ASSIST_SUCCESS) {
        fprintf(stderr, "(ASSIST) An error occured while trying to initialize the ephemeris structure.\n");
        fprintf(stderr, "(ASSIST) %s\n", assist_error_messages[err]);
        assist_ephem_free(e);
        return nullptr;
    }
    return e;
}

void assist_ephem_free_pointers(struct assist_ephem* ephem) {
    if (!ephem) return;
    if (ephem->spk_planets) assist_spk_free(ephem->spk_planets);
    if (ephem->spk_asteroids) assist_spk_free(ephem->spk_asteroids);
    if (ephem->ascii_planets) assist_ascii_free(ephem->ascii_planets);
    ephem->spk_planets = nullptr;
    ephem->spk_asteroids = nullptr;
    ephem->ascii_planets = nullptr;
}

void assist_ephem_free(struct assist_ephem* ephem) {
    assist_ephem_free_pointers(ephem);
    free(ephem);
}

void assist_ephem_time_bounds(const struct assist_ephem* ephem, double* t_beg, double* t_end) {
    if (!ephem) return;
    double beg = -INFINITY, end = INFINITY;
    if (!ephem->spk_planets && ephem->ascii_planets) {
        beg = ephem->ascii_planets->beg;
        end = ephem->ascii_planets->end;
    }
    const spk_s* files[2] = {ephem->spk_planets, ephem->spk_asteroids};
    for (const spk_s* s : files) {
        if (!s) continue;
        for (int i = 0; i < s->num; i++) {
            if (s->targets[i].beg > beg) beg = s->targets[i].beg;
            if (s->targets[i].end < end) end = s->targets[i].end;
        }
    }
    if (t_beg) *t_beg = beg - ephem->jd_ref;
    if (t_end) *t_end = end - ephem->jd_ref;
}

// reference src/forces.c:175-263; the cache (when given) is transparent, as in the reference
int assist_all_ephem(const struct assist_ephem* ephem, struct assist_ephem_cache* cache, const int i, const double t,
                     double* const GM, double* const x, double* const y, double* const z, double* const vx,
                     double* const vy, double* const vz, double* const ax, double* const ay, double* const az) {
    if (cache) {
        double* ct = cache->t + 7 * i;
        for (int s = 0; s < 7; s++) {
            if (t == ct[s]) {
                const assist_cache_item* it = cache->items + i * 7 + s;
                *GM = it->GM; *x = it->x; *y = it->y; *z = it->z;
                *vx = it->vx; *vy = it->vy; *vz = it->vz; *ax = it->ax; *ay = it->ay; *az = it->az;
                return ASSIST_SUCCESS;
            }
        }
    }
    const double jd_ref = ephem->jd_ref;
    if (i < ASSIST_BODY_NPLANETS) {
        const int f = ephem->planets_calc
                          ? ephem->planets_calc(ephem, jd_ref, t, i, GM, x, y, z, vx, vy, vz, ax, ay, az)
                          : assist_host::planets_calc(ephem, jd_ref, t, i, GM, x, y, z, vx, vy, vz, ax, ay, az);
        if (f != ASSIST_SUCCESS) return f;
    } else {
        double p[3];
        int f = assist_host::asteroid_pos(ephem->spk_asteroids, jd_ref, t, i - ASSIST_BODY_NPLANETS, GM, p);
        if (f != ASSIST_SUCCESS) return f;
        double GMs, xs, ys, zs, d1, d2, d3, d4, d5, d6;
        f = assist_all_ephem(ephem, cache, ASSIST_BODY_SUN, t, &GMs, &xs, &ys, &zs, &d1, &d2, &d3, &d4, &d5, &d6);
        if (f != ASSIST_SUCCESS) return f;
        *x = p[0] + xs; *y = p[1] + ys; *z = p[2] + zs;
        *vx = *vy = *vz = *ax = *ay = *az = NAN;
    }
    if (cache) {
        double* ct = cache->t + 7 * i;
        int os = 0;
        double ot = ct[0];
        for (int s = 1; s < 7; s++) {
            if (cache->dt_sign > 0 ? ct[s] < ot : ct[s] > ot) { ot = ct[s]; os = s; }
        }
        ct[os] = t;
        assist_cache_item* it = cache->items + i * 7 + os;
        it->GM = *GM; it->x = *x; it->y = *y; it->z = *z;
        it->vx = *vx; it->vy = *vy; it->vz = *vz; it->ax = *ax; it->ay = *ay; it->az = *az;
    }
    return ASSIST_SUCCESS;
}

struct reb_particle assist_get_particle_with_error(const struct assist_ephem* ephem, const int particle_id,
                                                   const double t, int* error) {
    struct reb_particle p;
    memset(&p, 0, sizeof p);
    double GM = 0;
    *error = assist_all_ephem(ephem, nullptr, particle_id, t, &GM, &p.x, &p.y, &p.z, &p.vx, &p.vy, &p.vz, &p.ax, &p.ay, &p.az);
    p.m = GM;
    return p;
}

struct reb_particle assist_get_particle(const struct assist_ephem* ephem, const int particle_id, const double t) {
    int err = 0;
    struct reb_particle p = assist_get_particle_with_error(ephem, particle_id, t, &err);
    if (err != ASSIST_SUCCESS) {
        fprintf(stderr, "(ASSIST) An error occured while trying to initialize particle from ephemeris data.\n");
        fprintf(stderr, "(ASSIST) %s\n", assist_error_messages[err]);
    }
    return p;
}

// ---------------------------------------------------------------------------------------------
// extras (reference src/assist.c:379-501)
// ---------------------------------------------------------------------------------------------
void assist_init(struct assist_extras* assist, struct reb_simulation* sim, struct assist_ephem* ephem) {
    assist->sim = sim;
    assist->ephem_cache = (assist_ephem_cache*)calloc(1, sizeof(assist_ephem_cache));
    int n_total = ASSIST_BODY_NPLANETS + (ephem->spk_asteroids ? ephem->spk_asteroids->num : 0);
    assist->gr_eih_sources = 1;
    assist->ephem_cache->items = (assist_cache_item*)calloc((size_t)n_total * 7, sizeof(assist_cache_item));
    assist->ephem_cache->t = (double*)malloc((size_t)n_total * 7 * sizeof(double));
    for (int i = 0; i < 7 * n_total; i++) assist->ephem_cache->t[i] = -1e306;
    assist->ephem = ephem;
    assist->particle_params = nullptr;
    assist->forces = ASSIST_FORCE_SUN | ASSIST_FORCE_PLANETS | ASSIST_FORCE_ASTEROIDS | ASSIST_FORCE_NON_GRAVITATIONAL |
                     ASSIST_FORCE_EARTH_HARMONICS | ASSIST_FORCE_SUN_HARMONICS | ASSIST_FORCE_GR_EIH;
    assist->last_state = nullptr;
    assist->current_state = nullptr;
    assist->alpha = 1.0;
    assist->nk = 0.0;
    assist->nm = 2.0;
    assist->nn = 5.093;
    assist->r0 = 1.0;
    if (sim) {
        sim->extras = assist;
        sim->additional_forces = assist_additional_forces;
        sim->ri_ias15.adaptive_mode = 1;
    }
}

struct assist_extras* assist_attach(struct reb_simulation* sim, struct assist_ephem* ephem) {
    if (!sim) {
        fprintf(stderr, "(ASSIST) Error: Simulation pointer passed to assist_attach was NULL.\n");
        return nullptr;
    }
    int should_free = 0;
    if (!ephem) {
        ephem = assist_ephem_create(nullptr, nullptr);
        if (!ephem) {
            fprintf(stderr, "(ASSIST) Error: Ephemeris pointer passed to assist_attach was NULL. Initialization with default path failed.\n");
            return nullptr;
        }
        should_free = 1;
    }
    auto* a = (struct assist_extras*)calloc(1, sizeof(struct assist_extras));
    assist_init(a, sim, ephem);
    a->extras_should_free_ephem = should_free;
    return a;
}

void assist_detach(struct reb_simulation* sim, struct assist_extras* assist) {
    if (assist->sim && sim) {
        sim->extras = nullptr;
        sim->additional_forces = nullptr;
    }
    assist->sim = nullptr;
}

void assist_free_pointers(struct assist_extras* assist) {
    if (assist->sim) {
        assist_detach(assist->sim, assist);
        assist->sim = nullptr;
    }
    free(assist->last_state);
    free(assist->current_state);
    assist->last_state = assist->current_state = nullptr;
    if (assist->ephem_cache) {
        free(assist->ephem_cache->items);
        free(assist->ephem_cache->t);
        free(assist->ephem_cache);
        assist->ephem_cache = nullptr;
    }
    if (assist->extras_should_free_ephem && assist->ephem) {
        assist_ephem_free(assist->ephem);
        assist->ephem = nullptr;
    }
}

void assist_free(struct assist_extras* assist) {
    assist_free_pointers(assist);
    free(assist);
}

void assist_error(struct assist_extras* assist, const char* const msg) {
    if (!assist->sim)
        fprintf(stderr, "(ASSIST) Error: A Simulation is no longer attached to the ASSIST extras instance. Most likely the Simulation has been freed.\n");
    else
        reb_simulation_error(assist->sim, msg);
}

// ---------------------------------------------------------------------------------------------
// minimal simulation
// ---------------------------------------------------------------------------------------------
void reb_simulation_init(struct reb_simulation* s) {
    memset(s, 0, sizeof(*s));
    s->dt = 0.001;
    s->exact_finish_time = 1;
    s->ri_ias15.epsilon = 1e-9;
    s->ri_ias15.min_dt = 0.0;
    s->ri_ias15.adaptive_mode = 1;
}

struct reb_simulation* reb_simulation_create(void) {
    auto* s = (struct reb_simulation*)malloc(sizeof(struct reb_simulation));
    reb_simulation_init(s);
    return s;
}

void reb_simulation_free_pointers(struct reb_simulation* sim) {
    if (!sim) return;
    free(sim->particles);
    free(sim->var_config);
    sim->particles = nullptr;
    sim->var_config = nullptr;
    sim->N = sim->N_var = sim->N_var_config = sim->allocated_N = 0;
}

void reb_simulation_free(struct reb_simulation* sim) {
    if (!sim) return;
    reb_simulation_free_pointers(sim);
    free(sim);
}

void reb_simulation_error(struct reb_simulation* sim, const char* msg) {
    snprintf(sim->error_message, sizeof sim->error_message, "%s", msg);
    fprintf(stderr, "(ASSIST) Error: %s\n", msg);
    sim->status = 1;
}

static void grow(struct reb_simulation* sim) {
    if (sim->N + 1 > sim->allocated_N) {
        sim->allocated_N = sim->allocated_N ? 2 * sim->allocated_N : 16;
        sim->particles = (reb_particle*)realloc(sim->particles, (size_t)sim->allocated_N * sizeof(reb_particle));
    }
}

void reb_simulation_add(struct reb_simulation* sim, struct reb_particle p) {
    grow(sim);
    sim->particles[sim->N++] = p;
}

int reb_simulation_add_variation_1st_order(struct reb_simulation* sim, int testparticle) {
    grow(sim);
    const int idx = sim->N;
    struct reb_particle z;
    memset(&z, 0, sizeof z);
    sim->particles[sim->N++] = z;
    sim->N_var++;
    sim->var_config = (reb_variational_configuration*)realloc(sim->var_config, (size_t)(sim->N_var_config + 1) * sizeof(reb_variational_configuration));
    sim->var_config[sim->N_var_config].index = idx;
    sim->var_config[sim->N_var_config].testparticle = testparticle;
    sim->N_var_config++;
    return idx;
}

}  // extern "C"
