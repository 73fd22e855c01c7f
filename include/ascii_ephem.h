/*
 * JPL binary ephemeris (.440/.441) planets backend of libassist.
 *
 * The reference's interface for this file format, reference src/ascii_ephem.h:1-93: the same
 * function names, argument meanings and return codes, and the same `struct ascii_s` field
 * order (one field appended at the end).  The files are JPL's ASCII ephemeris blocks converted
 * to native little-endian binary: a header block (titles, constant names from byte 0x00FC,
 * header values from 0x0A5C), a block of constant values, then one block per `inc` days holding
 * every column's Chebyshev coefficients for its `niv` sub-intervals.
 *
 * On the GPU the planet columns are re-laid at first use into the device's padded record
 * format, one record per sub-interval, so the force kernels evaluate SPK and JPL binary
 * planets with the same code.
 */
#ifndef ASSIST_AMD_ASCII_EPHEM_H
#define ASSIST_AMD_ASCII_EPHEM_H

#include <stddef.h>
#include <stdint.h>

#include "assist.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Column order of a block, reference src/ascii_ephem.h:42-61. */
enum {
    ASCII_MER,  /* Mercury */
    ASCII_VEN,  /* Venus */
    ASCII_EMB,  /* Earth-Moon barycentre */
    ASCII_MAR,  /* Mars */
    ASCII_JUP,  /* Jupiter */
    ASCII_SAT,  /* Saturn */
    ASCII_URA,  /* Uranus */
    ASCII_NEP,  /* Neptune */
    ASCII_PLU,  /* Pluto */
    ASCII_LUN,  /* Moon, geocentric */
    ASCII_SUN,  /* Sun */
    ASCII_NUT,  /* nutations (2 components) */
    ASCII_LIB,  /* lunar librations */
    ASCII_MAN,  /* lunar mantle */
    ASCII_TDB,  /* TT-TDB (1 component) */
    ASCII_N,
};

/* Same field order as reference src/ascii_ephem.h:63-86. */
struct ascii_s {
    double beg, end;       /* first and last JD covered */
    double inc;            /* days per block */
    double cau;            /* km per au (header value; positions are divided by it) */
    double cem;            /* Earth/Moon mass ratio */
    int32_t num;           /* number of constants */
    int32_t ver;           /* ephemeris number, e.g. 440 */
    int32_t off[ASCII_N];  /* 0-based word offset of each column in a block */
    int32_t ncf[ASCII_N];  /* Chebyshev coefficients per component */
    int32_t niv[ASCII_N];  /* sub-intervals per block */
    int32_t ncm[ASCII_N];  /* components */
    double mass[ASCII_N];  /* GM in ASSIST body order (entries 0..10) */
    double J2E;
    double J3E;
    double J4E;
    double J2SUN;
    double AU;
    double RE;
    double CLIGHT;
    double ASUN;
    size_t len, rec;       /* file size and block size in bytes */
    void* map;             /* the memory-mapped file */
    double* con;           /* constant values */
    char** str;            /* constant names (6 characters, NUL-terminated) */
    void* device_cache;    /* not in the reference: the GPU copy of the planet columns */
};

/* reference src/ascii_ephem.c:105-252; NULL when the file cannot be read or is malformed */
struct ascii_s* assist_ascii_init(char* path);
/* reference src/ascii_ephem.c:254-266 */
void assist_ascii_free(struct ascii_s* ascii);
/* reference src/ascii_ephem.c:27-65: one column's position u (km), velocity v (km/s) and
 * acceleration w (km/s^2) at block fraction t0 of a block t1 days long. */
void assist_ascii_work(double* P, int ncm, int ncf, int niv, double t0, double t1, double* u, double* v, double* w);
/* reference src/ascii_ephem.c:275-384: GM and barycentric state of ASSIST body `body` in au,
 * au/d, au/d^2 at jd_ref + jd_rel; ASSIST_ERROR_COVERAGE outside the file, ASSIST_ERROR_NEPHEM
 * for a body index outside 0..10. */
enum ASSIST_STATUS assist_ascii_calc(struct ascii_s* pl, double jd_ref, double jd_rel, int body, double* const GM,
                                     double* const x, double* const y, double* const z, double* const vx,
                                     double* const vy, double* const vz, double* const ax, double* const ay,
                                     double* const az);
/* reference src/ascii_ephem.c:386-393 (the `planets_calc` of an ephemeris with JPL binary planets) */
enum ASSIST_STATUS assist_ascii_calc_from_ephem(const struct assist_ephem* ephem, double jd_ref, double jd_rel,
                                                int body, double* const GM, double* const x, double* const y,
                                                double* const z, double* const vx, double* const vy,
                                                double* const vz, double* const ax, double* const ay,
                                                double* const az);
/* reference src/ascii_ephem.c:80-102: 1 and the value of the constant named `name` (space
 * padded to 6 characters, exact match), else 0 and *out_value = 0 */
int assist_ascii_find_constant(const struct ascii_s* ascii, const char* name, double* out_value);

#ifdef __cplusplus
}
#endif
#endif
