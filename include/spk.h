/*
 * SPK (DAF type-2) ephemeris files: the reference's C interface for them, reference
 * src/spk.h:1-122, with the same struct layouts, function names, argument meanings and return
 * codes.  The reference's own unit tests read these structs directly
 * (unit_tests/spk_init/problem.c:10-62 walks spk_s.targets[] past `num` and expects code 0
 * there, which this library guarantees by zeroing the spare capacity).
 *
 * `struct spk_s` carries one field beyond the reference layout, at its end (`device_cache`,
 * the GPU copy of the file's records).  The library allocates every spk_s, so callers that know
 * only the reference layout read the same fields at the same offsets.
 */
#ifndef ASSIST_AMD_SPK_H
#define ASSIST_AMD_SPK_H

#include <stddef.h>
#include <stdint.h>

#include "assist.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RECORD_LENGTH 1024          /* reference src/spk.h:12 */
#define MAX_COMMENT_LENGTH 1000000  /* reference src/spk.h:13 */

/* Position, velocity and acceleration (km, km/s, km/s^2) -- reference src/spk.h:15-19. */
struct mpos_s {
    double u[3];
    double v[3];
    double w[3];
};

/* GM constants found in a file's comment area -- reference src/spk.h:21-25. */
struct mass_data {
    char** names;
    double* values;
    size_t count;
};

/* reference src/spk.h:28-41 */
struct spk_constants_and_masses {
    double AU;
    double EMRAT;
    double J2E;
    double J3E;
    double J4E;
    double J2SUN;
    double RE;
    double CLIGHT;
    double ASUN;
    struct mass_data masses;
};

/* One target of an SPK file -- reference src/spk.h:44-56. */
struct spk_target {
    int code;           /* NAIF code */
    int cen;            /* centre code */
    double mass;        /* GM (au^3/day^2); 0 until assist_spk_join_masses finds one */
    double beg;         /* JD of the first segment's start */
    double end;         /* JD of the last segment's end */
    double res;         /* length of the first segment (days) */
    int* one;           /* 1-based word address of each segment's first record */
    int* two;           /* 1-based word address of each segment's last word (its directory) */
    int ind;            /* index of the last segment (one/two hold ind + 1 entries) */
    int allocated_ind;  /* capacity of one/two */
};

/* The targets of one SPK file -- reference src/spk.h:59-65, plus device_cache. */
struct spk_s {
    struct spk_target* targets;
    int num;            /* targets in the file */
    int allocated_num;  /* capacity of targets[]; entries num .. allocated_num-1 are zero */
    void* map;          /* the file, memory-mapped read-only */
    size_t len;         /* length of the mapping */
    void* device_cache; /* (not in the reference) GPU copy of the records, built on first use */
};

/* One summary of a summary record -- reference src/spk.h:68-77. */
struct sum_s {
    double beg;
    double end;
    int tar;
    int cen;
    int ref;
    int ver;
    int one;
    int two;
};

/* A DAF record -- reference src/spk.h:80-99. */
union record_t {
    char buf[RECORD_LENGTH];
    struct {
        double next;
        double prev;
        double nsum;
        struct sum_s s[25];
    } summary;
    struct {
        char locidw[8];
        int nd;
        int ni;
        char locifn[60];
        int fward;
        int bward;
        int free;
    } file;
};

/* Entry points -- reference src/spk.h:105-122 (definitions src/spk.c). */

/* Unmaps and frees `pl`; -1 for NULL (reference src/spk.c:27-45). */
int assist_spk_free(struct spk_s* pl);
/* Reads the file's summaries into targets (first-seen order) and maps the file; NULL on any
 * error.  Masses stay 0 until assist_spk_join_masses (reference src/spk.c:267-402). */
struct spk_s* assist_spk_init(const char* path);
/* The DAF file record of `fd` (calloc'd; free() it), or NULL when the file is not a DAF/SPK file
 * with 40-byte summaries (reference src/spk.c:213-263).  Unlike the reference, `fd` stays open on
 * failure: the caller owns it. */
union record_t* assist_load_spk_file_record(int fd);
/* The comment area (records 2 .. first_summary_record - 1), NUL and EOT padding dropped and NULs
 * turned into newlines, as one malloc'd string in *comments (NULL when there is none or on a read
 * error; the reference exits the process there) (reference src/spk.c:49-110). */
void parse_comments(int fd, int first_summary_record, char** comments);
/* Constants (cau/AU, EMRAT, J2E, J3E, J4E, J2SUN, RE, CLIGHT, ASUN) and GM / MA masses from the
 * lines after "Initial conditions and constants used for integration:"; all zero when the file
 * cannot be read (reference src/spk.c:696-766). */
struct spk_constants_and_masses assist_load_spk_constants_and_masses(const char* path);
/* Copies the constants into `ephem` and derives Re_eq, Rs_eq, c_AU_per_day, c_squared and
 * over_c_squared (reference src/spk.c:769-789). */
void assist_apply_spk_constants(struct assist_ephem* ephem, const struct spk_constants_and_masses* data);
void assist_free_spk_constants_and_masses(struct spk_constants_and_masses* data);
/* Sets the mass of every target whose mass is 0 from GMS/GM1..GM9/GMB (the Earth and Moon split
 * GMB by EMRAT) or MAnnnn for asteroid 2000000 + nnnn (reference src/spk.c:137-210). */
void assist_spk_join_masses(struct spk_s* sp, const struct mass_data* masses, double emrat);
/* Asteroid m (0-based target index): heliocentric position in au (km / 149597870.7) and GM
 * (reference src/spk.c:405-481). */
enum ASSIST_STATUS assist_spk_calc(const struct spk_s* pl, double jd_ref, double jd_rel, int m, double* GM, double* x,
                                   double* y, double* z);
/* Planet-file target with NAIF code `code`: barycentric position, velocity and acceleration in
 * au, au/day, au/day^2 (Earth 399 and Moon 301 add the EMB) (reference src/spk.c:550-630). */
enum ASSIST_STATUS assist_spk_calc_planets(const struct assist_ephem* ephem, double jd_ref, double jd_rel, int code,
                                           double* GM, double* x, double* y, double* z, double* vx, double* vy,
                                           double* vz, double* ax, double* ay, double* az);
/* Same by ASSIST body index, with the Earth from the EMB and the Moon when 399 is absent
 * (reference src/spk.c:632-693). */
enum ASSIST_STATUS assist_spk_calc_planets_by_assist(const struct assist_ephem* ephem, double jd_ref, double jd_rel,
                                                     int assist_body, double* GM, double* out_x, double* out_y,
                                                     double* out_z, double* out_vx, double* out_vy, double* out_vz,
                                                     double* out_ax, double* out_ay, double* out_az);
/* The target with NAIF code `code`, or NULL (reference src/spk.c:483-490). */
struct spk_target* assist_spk_find_target(const struct spk_s* pl, int code);
/* Chebyshev position, velocity and acceleration of `target` in km, km/s, km/s^2 (reference
 * src/spk.c:492-547). */
struct mpos_s assist_spk_target_pos(const struct spk_s* pl, const struct spk_target* target, double jd_ref,
                                    double jd_rel);

#ifdef __cplusplus
}
#endif
#endif
