/*
 * Oracle (test infrastructure): IAS15, restated from Rein & Spiegel (2015),
 * "IAS15: a fast, adaptive, high-order integrator for gravitational dynamics,
 * accurate to machine precision over a billion orbits", MNRAS 446, 1424.
 *
 * 15th-order Gauss-Radau predictor-corrector (Everhart 1985) with
 *  - the acceleration expanded in the Newton basis on the 7 Gauss-Radau
 *    spacings (g values) and in powers of h (b values),
 *  - predictor-corrector iterations until max|delta b6| / max|a| < 1e-16
 *    (at most 12; stop early when the error stops decreasing after 3); as an
 *    option of the product (cfg->pc_tol > 0), until the real particles' own
 *    max|delta b6| / max|a| is below pc_tol, and (cfg->pc_tol1 > 0) after the
 *    first pass of an attempt when that pass changed them by less than pc_tol1
 *    (times max(0.1, pc_aref / max|a|) where max|a| exceeds cfg->pc_aref > 0),
 *  - the global step-size criterion dt' = dt (eps / (max|b6| / max|a|))^(1/7)
 *    over the real particles (ASSIST selects this "legacy global" mode,
 *    reference src/assist.c:446), rejection when dt' < dt/4, growth <= 4x,
 *  - b and e re-expanded for the next step from the last accepted step
 *    (binomial re-centering scaled by the step ratio, plus the last correction),
 *  - Kahan-compensated position and velocity updates.
 * The variational particles ride along in the same predictor-corrector; only
 * real particles enter the step-size criterion.
 *
 * The coefficient tables are computed here in long double from the defining
 * polynomial (P7 + P8)(2h - 1), independently of the product's tables.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double H[8], RR[28], C[21], D[21];
static int tables_ready = 0;

static long double legendre(int n, long double x) {
    long double p0 = 1.0L, p1 = x;
    if (n == 0) return p0;
    for (int k = 2; k <= n; k++) {
        long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

static void build_tables(void) {
    long double h[8];
    h[0] = 0.0L;
    /* roots of P7 + P8 on (-1, 1) other than -1, by bracketing + bisection/Newton */
    int found = 0;
    const int NS = 20000;
    long double xa = -1.0L + 1e-6L, fa = legendre(7, xa) + legendre(8, xa);
    for (int i = 1; i <= NS && found < 7; i++) {
        long double xb = -1.0L + 1e-6L + (2.0L - 2e-6L) * i / NS;
        long double fb = legendre(7, xb) + legendre(8, xb);
        if ((fa < 0) != (fb < 0)) {
            long double lo = xa, hi = xb, flo = fa;
            for (int it = 0; it < 200; it++) {
                long double mid = 0.5L * (lo + hi);
                long double fm = legendre(7, mid) + legendre(8, mid);
                if ((fm < 0) == (flo < 0)) { lo = mid; flo = fm; } else hi = mid;
            }
            h[1 + found++] = 0.5L * (0.5L * (lo + hi) + 1.0L);
        }
        xa = xb;
        fa = fb;
    }
    int k = 0;
    for (int i = 1; i < 8; i++)
        for (int j = 0; j < i; j++) RR[k++] = (double)(h[i] - h[j]);
    /* Newton basis N_k(h) = h prod_{m=1}^{k-1} (h - h_m) in power coefficients */
    long double N[8][9];
    memset(N, 0, sizeof N);
    N[1][1] = 1.0L;
    for (int kk = 2; kk <= 7; kk++) {
        for (int p = 0; p <= 8; p++) {
            long double v = 0.0L;
            if (p >= 1) v += N[kk - 1][p - 1];
            v -= N[kk - 1][p] * h[kk - 1];
            N[kk][p] = v;
        }
    }
    k = 0;
    for (int kk = 2; kk <= 7; kk++)
        for (int j = 0; j < kk - 1; j++) C[k++] = (double)N[kk][j + 1];
    /* d: inverse of the unit upper-triangular M[j][k] = [h^(j+1)] N_{k+1} */
    long double M[7][7], Inv[7][7];
    for (int j = 0; j < 7; j++)
        for (int kk = 0; kk < 7; kk++) M[j][kk] = N[kk + 1][j + 1];
    for (int col = 0; col < 7; col++) {
        for (int row = 6; row >= 0; row--) {
            long double v = (row == col) ? 1.0L : 0.0L;
            for (int m = row + 1; m < 7; m++) v -= M[row][m] * Inv[m][col];
            Inv[row][col] = v / M[row][row];
        }
    }
    k = 0;
    for (int j = 1; j < 7; j++)
        for (int kk = 0; kk < j; kk++) D[k++] = (double)Inv[kk][j];
    for (int i = 0; i < 8; i++) H[i] = (double)h[i];
    tables_ready = 1;
}

void orc_ias15_tables(double h[8], double rr[28], double c[21], double d[21]) {
    if (!tables_ready) build_tables();
    memcpy(h, H, sizeof H);
    memcpy(rr, RR, sizeof RR);
    memcpy(c, C, sizeof C);
    memcpy(d, D, sizeof D);
}

void orc_ias15_cfg_default(orc_ias15_cfg *c) {
    c->epsilon = 1e-9;
    c->min_dt = 0.0;
    c->dt = 0.001;
    c->exact_finish = 1;
    c->max_steps = 100000000L;
    c->pc_tol = 0.0;
    c->pc_tol1 = 0.0;
    c->pc_aref = 0.0;
}

static inline void add_cs(double *p, double *cs, double in) {
    const double y = in - *cs;
    const double t = *p + y;
    *cs = (t - *p) - y;
    *p = t;
}

typedef struct {
    int n3;
    double *x0, *v0, *a0, *csx, *csv, *at;
    double *b[7], *g[7], *e[7], *br[7], *er[7];
} ias_buf;

static void buf_alloc(ias_buf *B, int n3) {
    B->n3 = n3;
    int nb = 6 + 35;
    double *m = calloc((size_t)nb * n3, sizeof(double));
    double **slots[6] = {&B->x0, &B->v0, &B->a0, &B->csx, &B->csv, &B->at};
    for (int i = 0; i < 6; i++) *slots[i] = m + (size_t)i * n3;
    for (int k = 0; k < 7; k++) {
        B->b[k] = m + (size_t)(6 + k) * n3;
        B->g[k] = m + (size_t)(13 + k) * n3;
        B->e[k] = m + (size_t)(20 + k) * n3;
        B->br[k] = m + (size_t)(27 + k) * n3;
        B->er[k] = m + (size_t)(34 + k) * n3;
    }
}

static void buf_free(ias_buf *B) { free(B->x0); }

/* b, e for a step of ratio q = dt/dt_last_success from the last accepted br, er. */
static void predict_from_last(ias_buf *B, double ratio) {
    const int n3 = B->n3;
    if (ratio > 20.) {
        for (int k = 0; k < 7; k++) {
            memset(B->e[k], 0, n3 * sizeof(double));
            memset(B->b[k], 0, n3 * sizeof(double));
        }
        return;
    }
    const double q1 = ratio, q2 = q1 * q1, q3 = q1 * q2, q4 = q2 * q2, q5 = q2 * q3, q6 = q3 * q3, q7 = q3 * q4;
    for (int k = 0; k < n3; k++) {
        const double b0 = B->br[0][k], b1 = B->br[1][k], b2 = B->br[2][k], b3 = B->br[3][k];
        const double b4 = B->br[4][k], b5 = B->br[5][k], b6 = B->br[6][k];
        double be[7];
        for (int m = 0; m < 7; m++) be[m] = B->br[m][k] - B->er[m][k];
        /* Taylor re-centring of the h-polynomial at h = 1 (binomial coefficients) */
        B->e[0][k] = q1 * (b6 * 7.0 + b5 * 6.0 + b4 * 5.0 + b3 * 4.0 + b2 * 3.0 + b1 * 2.0 + b0);
        B->e[1][k] = q2 * (b6 * 21.0 + b5 * 15.0 + b4 * 10.0 + b3 * 6.0 + b2 * 3.0 + b1);
        B->e[2][k] = q3 * (b6 * 35.0 + b5 * 20.0 + b4 * 10.0 + b3 * 4.0 + b2);
        B->e[3][k] = q4 * (b6 * 35.0 + b5 * 15.0 + b4 * 5.0 + b3);
        B->e[4][k] = q5 * (b6 * 21.0 + b5 * 6.0 + b4);
        B->e[5][k] = q6 * (b6 * 7.0 + b5);
        B->e[6][k] = q7 * b6;
        for (int m = 0; m < 7; m++) B->b[m][k] = B->e[m][k] + be[m];
    }
}

static void eval(const orc_ephem *eph, const orc_opts *o, orc_sys *s, orc_stats *st) {
    orc_additional_forces(eph, o, s);
    st->evals++;
}

int orc_integrate(const orc_ephem *eph, const orc_opts *o, orc_ias15_cfg *cfg, orc_sys *s, double t_end,
                  orc_stats *st) {
    if (!tables_ready) build_tables();
    const int N = s->n_real + s->n_var, n3 = 3 * N, n3r = 3 * s->n_real;
    const double safety = 0.25;
    memset(st, 0, sizeof(*st));
    ias_buf B;
    buf_alloc(&B, n3);
    for (int i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) {
            B.x0[3 * i + c] = s->p[9 * i + c];
            B.v0[3 * i + c] = s->p[9 * i + 3 + c];
        }
    double t = s->t;
    const double dir = (t_end >= t) ? 1.0 : -1.0;
    orc_cache_direction(dir > 0 ? 1 : -1);
    double dt = cfg->dt;
    if (dt * dir < 0) dt = -dt;
    double dt_last_success = 0.0;
    s->status = 0;
    while (t != t_end && (t_end - t) * dir > 0) {
        if (st->steps >= cfg->max_steps) { st->status = -2; break; }
        double dts = dt;
        int last = 0;
        if (cfg->exact_finish && (t + dts - t_end) * dir >= 0) { dts = t_end - t; last = 1; }
        if (!(fabs(dts) > 0) || !isfinite(dts)) { st->status = -3; break; }
        /* acceleration at the start of the step */
        for (int i = 0; i < N; i++)
            for (int c = 0; c < 3; c++) {
                s->p[9 * i + c] = B.x0[3 * i + c];
                s->p[9 * i + 3 + c] = B.v0[3 * i + c];
            }
        s->t = t;
        eval(eph, o, s, st);
        if (s->status) { st->status = s->status; break; }
        for (int i = 0; i < N; i++)
            for (int c = 0; c < 3; c++) B.a0[3 * i + c] = s->p[9 * i + 6 + c];
        int accepted = 0, attempts = 0;
        double dt_new = dts;
        while (!accepted) {
            if (++attempts > 64) { st->status = -3; break; }
            if (dt_last_success != 0.0) predict_from_last(&B, dts / dt_last_success);
            else
                for (int k = 0; k < 7; k++) { memset(B.b[k], 0, n3 * sizeof(double)); memset(B.e[k], 0, n3 * sizeof(double)); }
            for (int k = 0; k < n3; k++) {
                const double b0 = B.b[0][k], b1 = B.b[1][k], b2 = B.b[2][k], b3 = B.b[3][k], b4 = B.b[4][k], b5 = B.b[5][k], b6 = B.b[6][k];
                B.g[0][k] = b6 * D[15] + b5 * D[10] + b4 * D[6] + b3 * D[3] + b2 * D[1] + b1 * D[0] + b0;
                B.g[1][k] = b6 * D[16] + b5 * D[11] + b4 * D[7] + b3 * D[4] + b2 * D[2] + b1;
                B.g[2][k] = b6 * D[17] + b5 * D[12] + b4 * D[8] + b3 * D[5] + b2;
                B.g[3][k] = b6 * D[18] + b5 * D[13] + b4 * D[9] + b3;
                B.g[4][k] = b6 * D[19] + b5 * D[14] + b4;
                B.g[5][k] = b6 * D[20] + b5;
                B.g[6][k] = b6;
            }
            double pc_err = 1e300, pc_err_last = 2.0, pc_mak = 0.0;
            int it = 0;
            const double pc_tol = cfg->pc_tol > 0 ? cfg->pc_tol : 1e-16;
            const int pc_n = cfg->pc_tol > 0 ? n3r : n3; /* components in the test */
            for (;;) {
                if (pc_err < pc_tol) break;
                if (it == 1 && cfg->pc_tol1 > 0) {
                    const double tol1 = (cfg->pc_aref > 0 && pc_mak > cfg->pc_aref) ? cfg->pc_tol1 * fmax(0.1, cfg->pc_aref / pc_mak) : cfg->pc_tol1;
                    if (pc_err < tol1) break;
                }
                if (it > 2 && pc_err_last <= pc_err) break;
                if (it >= 12) { st->pc_maxed++; break; }
                pc_err_last = pc_err;
                it++;
                double maxak = 0.0, maxdb6 = 0.0;
                for (int n = 1; n < 8; n++) {
                    const double h = H[n];
                    double sx[9], sv[8];
                    sx[0] = dts * h;
                    sx[1] = sx[0] * sx[0] / 2.;
                    sx[2] = sx[1] * h / 3.;
                    sx[3] = sx[2] * h / 2.;
                    sx[4] = 3. * sx[3] * h / 5.;
                    sx[5] = 2. * sx[4] * h / 3.;
                    sx[6] = 5. * sx[5] * h / 7.;
                    sx[7] = 3. * sx[6] * h / 4.;
                    sx[8] = 7. * sx[7] * h / 9.;
                    sv[0] = dts * h;
                    sv[1] = sv[0] * h / 2.;
                    sv[2] = 2. * sv[1] * h / 3.;
                    sv[3] = 3. * sv[2] * h / 4.;
                    sv[4] = 4. * sv[3] * h / 5.;
                    sv[5] = 5. * sv[4] * h / 6.;
                    sv[6] = 6. * sv[5] * h / 7.;
                    sv[7] = 7. * sv[6] * h / 8.;
                    for (int i = 0; i < N; i++)
                        for (int c = 0; c < 3; c++) {
                            const int k = 3 * i + c;
                            double **b = B.b;
                            s->p[9 * i + c] = -B.csx[k] + ((sx[8] * b[6][k] + sx[7] * b[5][k] + sx[6] * b[4][k] + sx[5] * b[3][k] +
                                                            sx[4] * b[2][k] + sx[3] * b[1][k] + sx[2] * b[0][k] + sx[1] * B.a0[k] +
                                                            sx[0] * B.v0[k]) + B.x0[k]);
                            s->p[9 * i + 3 + c] = -B.csv[k] + ((sv[7] * b[6][k] + sv[6] * b[5][k] + sv[5] * b[4][k] + sv[4] * b[3][k] +
                                                                sv[3] * b[2][k] + sv[2] * b[1][k] + sv[1] * b[0][k] + sv[0] * B.a0[k]) +
                                                               B.v0[k]);
                        }
                    s->t = t + dts * h;
                    eval(eph, o, s, st);
                    if (s->status) break;
                    for (int i = 0; i < N; i++)
                        for (int c = 0; c < 3; c++) B.at[3 * i + c] = s->p[9 * i + 6 + c];
                    double **g = B.g, **b = B.b;
                    for (int k = 0; k < n3; k++) {
                        const double gk = B.at[k] - B.a0[k];
                        double tmp;
                        switch (n) {
                        case 1:
                            tmp = g[0][k];
                            g[0][k] = gk / RR[0];
                            b[0][k] += g[0][k] - tmp;
                            break;
                        case 2:
                            tmp = g[1][k];
                            g[1][k] = (gk / RR[1] - g[0][k]) / RR[2];
                            tmp = g[1][k] - tmp;
                            b[0][k] += tmp * C[0];
                            b[1][k] += tmp;
                            break;
                        case 3:
                            tmp = g[2][k];
                            g[2][k] = ((gk / RR[3] - g[0][k]) / RR[4] - g[1][k]) / RR[5];
                            tmp = g[2][k] - tmp;
                            b[0][k] += tmp * C[1];
                            b[1][k] += tmp * C[2];
                            b[2][k] += tmp;
                            break;
                        case 4:
                            tmp = g[3][k];
                            g[3][k] = (((gk / RR[6] - g[0][k]) / RR[7] - g[1][k]) / RR[8] - g[2][k]) / RR[9];
                            tmp = g[3][k] - tmp;
                            b[0][k] += tmp * C[3];
                            b[1][k] += tmp * C[4];
                            b[2][k] += tmp * C[5];
                            b[3][k] += tmp;
                            break;
                        case 5:
                            tmp = g[4][k];
                            g[4][k] = ((((gk / RR[10] - g[0][k]) / RR[11] - g[1][k]) / RR[12] - g[2][k]) / RR[13] - g[3][k]) / RR[14];
                            tmp = g[4][k] - tmp;
                            b[0][k] += tmp * C[6];
                            b[1][k] += tmp * C[7];
                            b[2][k] += tmp * C[8];
                            b[3][k] += tmp * C[9];
                            b[4][k] += tmp;
                            break;
                        case 6:
                            tmp = g[5][k];
                            g[5][k] = (((((gk / RR[15] - g[0][k]) / RR[16] - g[1][k]) / RR[17] - g[2][k]) / RR[18] - g[3][k]) / RR[19] -
                                       g[4][k]) / RR[20];
                            tmp = g[5][k] - tmp;
                            b[0][k] += tmp * C[10];
                            b[1][k] += tmp * C[11];
                            b[2][k] += tmp * C[12];
                            b[3][k] += tmp * C[13];
                            b[4][k] += tmp * C[14];
                            b[5][k] += tmp;
                            break;
                        default: {
                            tmp = g[6][k];
                            g[6][k] = ((((((gk / RR[21] - g[0][k]) / RR[22] - g[1][k]) / RR[23] - g[2][k]) / RR[24] - g[3][k]) / RR[25] -
                                        g[4][k]) / RR[26] - g[5][k]) / RR[27];
                            tmp = g[6][k] - tmp;
                            b[0][k] += tmp * C[15];
                            b[1][k] += tmp * C[16];
                            b[2][k] += tmp * C[17];
                            b[3][k] += tmp * C[18];
                            b[4][k] += tmp * C[19];
                            b[5][k] += tmp * C[20];
                            b[6][k] += tmp;
                            if (k < pc_n) {
                                const double ak = fabs(B.at[k]);
                                if (isnormal(ak) && ak > maxak) maxak = ak;
                                const double db6 = fabs(tmp);
                                if (isnormal(db6) && db6 > maxdb6) maxdb6 = db6;
                            }
                        }
                        }
                    }
                }
                if (s->status) break;
                st->iterations++;
                pc_err = maxdb6 / maxak;
                pc_mak = maxak;
            }
            if (s->status) break;
            /* step-size control (global criterion over real particles) */
            dt_new = dts;
            if (cfg->epsilon > 0) {
                double maxak = 0.0, maxb6 = 0.0;
                for (int k = 0; k < n3r; k++) {
                    const double ak = fabs(B.at[k]);
                    if (isnormal(ak) && ak > maxak) maxak = ak;
                    const double b6k = fabs(B.b[6][k]);
                    if (isnormal(b6k) && b6k > maxb6) maxb6 = b6k;
                }
                const double err = maxb6 / maxak;
                if (isnormal(err)) dt_new = pow(cfg->epsilon / err, 1.0 / 7.0) * dts;
                else dt_new = dts / safety;
                if (fabs(dt_new) < cfg->min_dt) dt_new = copysign(cfg->min_dt, dt_new);
                if (fabs(dt_new / dts) < safety) {
                    st->rejected++;
                    dts = dt_new;
                    last = 0;
                    continue;
                }
                if (dt_new / dts > 1. / safety) dt_new = dts / safety;
            }
            accepted = 1;
        }
        if (!accepted) { if (!st->status) st->status = s->status; break; }
        /* advance x0, v0 to the end of the step (Kahan compensated) */
        const double dt2 = dts * dts;
        for (int k = 0; k < n3; k++) {
            double **b = B.b;
            add_cs(&B.x0[k], &B.csx[k], b[6][k] / 72. * dt2);
            add_cs(&B.x0[k], &B.csx[k], b[5][k] / 56. * dt2);
            add_cs(&B.x0[k], &B.csx[k], b[4][k] / 42. * dt2);
            add_cs(&B.x0[k], &B.csx[k], b[3][k] / 30. * dt2);
            add_cs(&B.x0[k], &B.csx[k], b[2][k] / 20. * dt2);
            add_cs(&B.x0[k], &B.csx[k], b[1][k] / 12. * dt2);
            add_cs(&B.x0[k], &B.csx[k], b[0][k] / 6. * dt2);
            add_cs(&B.x0[k], &B.csx[k], B.a0[k] / 2. * dt2);
            add_cs(&B.x0[k], &B.csx[k], B.v0[k] * dts);
            add_cs(&B.v0[k], &B.csv[k], b[6][k] / 8. * dts);
            add_cs(&B.v0[k], &B.csv[k], b[5][k] / 7. * dts);
            add_cs(&B.v0[k], &B.csv[k], b[4][k] / 6. * dts);
            add_cs(&B.v0[k], &B.csv[k], b[3][k] / 5. * dts);
            add_cs(&B.v0[k], &B.csv[k], b[2][k] / 4. * dts);
            add_cs(&B.v0[k], &B.csv[k], b[1][k] / 3. * dts);
            add_cs(&B.v0[k], &B.csv[k], b[0][k] / 2. * dts);
            add_cs(&B.v0[k], &B.csv[k], B.a0[k] * dts);
        }
        t = last ? t_end : t + dts;
        st->steps++;
        st->dt_last_done = dts;
        dt_last_success = dts;
        for (int k = 0; k < 7; k++) {
            memcpy(B.er[k], B.e[k], n3 * sizeof(double));
            memcpy(B.br[k], B.b[k], n3 * sizeof(double));
        }
        if (!(last && cfg->epsilon <= 0)) dt = dt_new;
        int bad = 0;
        for (int k = 0; k < n3; k++) if (!isfinite(B.x0[k]) || !isfinite(B.v0[k])) bad = 1;
        if (bad) { st->status = -1; break; }
    }
    for (int i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) {
            s->p[9 * i + c] = B.x0[3 * i + c];
            s->p[9 * i + 3 + c] = B.v0[3 * i + c];
        }
    s->t = t;
    st->dt_next = dt;
    cfg->dt = dt;
    buf_free(&B);
    return st->status;
}
