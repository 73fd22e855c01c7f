/*
 * Oracle (test infrastructure): the ASSIST force model restated term by term.
 * Each function follows the cited reference lines formula by formula and keeps
 * the reference's accumulation order, so it is the parity target for the GPU.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

#define PX(s, i) ((s)->p + 9 * (i))
#define PI_ 3.14159265358979323846

void orc_opts_default(orc_opts *o) {
    /* reference src/assist.c:415,424-438 */
    o->forces = 0x7F;
    o->gr_eih_sources = 1;
    o->geocentric = 0;
    o->alpha = 1.0; o->nk = 0.0; o->nm = 2.0; o->nn = 5.093; o->r0 = 1.0;
}

static int get(const orc_ephem *e, orc_sys *s, int i, double t, double o[10]) {
    int f = orc_all_ephem(e, i, t, o);
    if (f != ORC_SUCCESS && s->status == 0) s->status = f;
    return f;
}

/* reference src/forces.c:266-433 */
static void force_direct(const orc_ephem *e, const orc_opts *op, orc_sys *s, double xo, double yo, double zo) {
    static const int order[ORC_NPLANETS] = {10, 4, 5, 1, 9, 8, 3, 2, 7, 6, 0};
    const int N_real = s->n_real;
    const double t = s->t;
    int ast_num = orc_ephem_num_asteroids(e);
    double b[10];
    for (int k = 0; k < ORC_NPLANETS + ast_num; k++) {
        int i = (k >= ast_num) ? order[k - ast_num] : k + ORC_NPLANETS;
        if (i == 0 && !(op->forces & ORC_FORCE_SUN)) continue;
        if (i > 0 && i < ORC_NPLANETS && !(op->forces & ORC_FORCE_PLANETS)) continue;
        if (i >= ORC_NPLANETS && !(op->forces & ORC_FORCE_ASTEROIDS)) continue;
        if (get(e, s, i, t, b) != ORC_SUCCESS) return;
        const double GM = b[0], x = b[1], y = b[2], z = b[3];
        for (int j = 0; j < N_real; j++) {
            double *p = PX(s, j);
            const double dx = p[0] + (xo - x);
            const double dy = p[1] + (yo - y);
            const double dz = p[2] + (zo - z);
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double _r = sqrt(r2);
            const double prefac = GM / (_r * _r * _r);
            p[6] -= prefac * dx;
            p[7] -= prefac * dy;
            p[8] -= prefac * dz;
        }
    }
    /* variational part: ignores the force flags (reference src/forces.c:359) */
    for (int k = 0; k < ORC_NPLANETS + ast_num; k++) {
        int i = (k >= ast_num) ? order[k - ast_num] : k + ORC_NPLANETS;
        if (get(e, s, i, t, b) != ORC_SUCCESS) return;
        if (s->n_var == 0) return;
        const double GM = b[0], x = b[1], y = b[2], z = b[3];
        for (int j = 0; j < N_real; j++) {
            const double *p = PX(s, j);
            const double dx = p[0] + (xo - x);
            const double dy = p[1] + (yo - y);
            const double dz = p[2] + (zo - z);
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double _r = sqrt(r2);
            const double r3inv = 1. / (r2 * _r);
            const double r5inv = 3. * r3inv / r2;
            const double dxdx = dx * dx * r5inv - r3inv;
            const double dydy = dy * dy * r5inv - r3inv;
            const double dzdz = dz * dz * r5inv - r3inv;
            const double dxdy = dx * dy * r5inv;
            const double dxdz = dx * dz * r5inv;
            const double dydz = dy * dz * r5inv;
            for (int v = 0; v < s->n_var; v++) {
                if (s->var_tp[v] != j) continue;
                double *q = PX(s, N_real + v);
                const double ddx = q[0], ddy = q[1], ddz = q[2];
                const double dax = ddx * dxdx + ddy * dxdy + ddz * dxdz;
                const double day = ddx * dxdy + ddy * dydy + ddz * dydz;
                const double daz = ddx * dxdz + ddy * dydz + ddz * dzdz;
                q[6] += GM * dax;
                q[7] += GM * day;
                q[8] += GM * daz;
            }
        }
    }
}

/* reference src/forces.c:435-642 */
static void force_earth_J2J4(const orc_ephem *e, orc_sys *s, double xo, double yo, double zo) {
    const int N_real = s->n_real;
    double b[10];
    get(e, s, 3, s->t, b);
    const double GMearth = b[0];
    const double xr = b[1], yr = b[2], zr = b[3];
    const double J2e = e->J2E, J3e = e->J3E, J4e = e->J4E, Re_eq = e->Re_eq;
    const double RAe = 0.0 * PI_ / 180., Dece = 90.0 * PI_ / 180.;
    const double cosa = cos(RAe), sina = sin(RAe), cosd = cos(Dece), sind = sin(Dece);
    for (int j = 0; j < N_real; j++) {
        double *p = PX(s, j);
        double dx = p[0] + (xo - xr);
        double dy = p[1] + (yo - yr);
        double dz = p[2] + (zo - zr);
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r = sqrt(r2);
        double dxp = -dx * sina + dy * cosa;
        double dyp = -dx * cosa * sind - dy * sina * sind + dz * cosd;
        double dzp = dx * cosa * cosd + dy * sina * cosd + dz * sind;
        dx = dxp; dy = dyp; dz = dzp;
        const double costheta2 = dz * dz / r2;
        const double J2e_prefac = 3. * J2e * Re_eq * Re_eq / r2 / r2 / r / 2.;
        const double J2e_fac = 5. * costheta2 - 1.;
        double resx = GMearth * J2e_prefac * J2e_fac * dx;
        double resy = GMearth * J2e_prefac * J2e_fac * dy;
        double resz = GMearth * J2e_prefac * (J2e_fac - 2.) * dz;
        const double J3e_prefac = 5. * J3e * Re_eq * Re_eq * Re_eq / r2 / r2 / r / 2.;
        const double J3e_fac = 3. - 7. * costheta2;
        resx += -GMearth * J3e_prefac * (1. / r2) * J3e_fac * dx * dz;
        resy += -GMearth * J3e_prefac * (1. / r2) * J3e_fac * dy * dz;
        resz += -GMearth * J3e_prefac * (6. * costheta2 - 7. * costheta2 * costheta2 - 0.6);
        const double J4e_prefac = 5. * J4e * Re_eq * Re_eq * Re_eq * Re_eq / r2 / r2 / r2 / r / 8.;
        const double J4e_fac = 63. * costheta2 * costheta2 - 42. * costheta2 + 3.;
        resx += GMearth * J4e_prefac * J4e_fac * dx;
        resy += GMearth * J4e_prefac * J4e_fac * dy;
        resz += GMearth * J4e_prefac * (J4e_fac + 12. - 28. * costheta2) * dz;
        double resxp = -resx * sina - resy * cosa * sind + resz * cosa * cosd;
        double resyp = resx * cosa - resy * sina * sind + resz * sina * cosd;
        double reszp = +resy * cosd + resz * sind;
        p[6] += resxp;
        p[7] += resyp;
        p[8] += reszp;

        const double J2e_fac2 = 7. * costheta2 - 1.;
        const double J2e_fac3 = 35. * costheta2 * costheta2 - 30. * costheta2 + 3.;
        const double dxdx = GMearth * J2e_prefac * (J2e_fac - 5. * J2e_fac2 * dx * dx / r2);
        const double dydy = GMearth * J2e_prefac * (J2e_fac - 5. * J2e_fac2 * dy * dy / r2);
        const double dzdz = GMearth * J2e_prefac * (-1.) * J2e_fac3;
        const double dxdy = GMearth * J2e_prefac * (-5.) * J2e_fac2 * dx * dy / r2;
        const double dydz = GMearth * J2e_prefac * (-5.) * (J2e_fac2 - 2.) * dy * dz / r2;
        const double dxdz = GMearth * J2e_prefac * (-5.) * (J2e_fac2 - 2.) * dx * dz / r2;
        const double costheta = dz / r;
        const double J3e_fac2 = 21 * (-3. * costheta2 + 1.) / r2;
        const double J3e_fac3 = 3 * (-21. * costheta2 * costheta2 + 14. * costheta2 - 1.) / r2;
        const double J3e_fac4 = (-63. * costheta2 * costheta2 + 70. * costheta2 - 15.) * costheta / r;
        const double dxdxJ3 = GMearth * J3e_prefac * costheta * (J3e_fac2 * dx * dx - J3e_fac) / r;
        const double dydyJ3 = GMearth * J3e_prefac * costheta * (J3e_fac2 * dy * dy - J3e_fac) / r;
        const double dzdzJ3 = GMearth * J3e_prefac * J3e_fac4;
        const double dxdyJ3 = GMearth * J3e_prefac * J3e_fac2 * costheta * dx * dy / r;
        const double dydzJ3 = GMearth * J3e_prefac * J3e_fac3 * dy;
        const double dxdzJ3 = GMearth * J3e_prefac * J3e_fac3 * dx;
        const double J4e_fac2 = 33. * costheta2 * costheta2 - 18. * costheta2 + 1.;
        const double J4e_fac3 = 33. * costheta2 * costheta2 - 30. * costheta2 + 5.;
        const double J4e_fac4 = 231. * costheta2 * costheta2 * costheta2 - 315. * costheta2 * costheta2 + 105. * costheta2 - 5.;
        const double dxdxJ4 = GMearth * J4e_prefac * (J4e_fac - 21. * J4e_fac2 * dx * dx / r2);
        const double dydyJ4 = GMearth * J4e_prefac * (J4e_fac - 21. * J4e_fac2 * dy * dy / r2);
        const double dzdzJ4 = GMearth * J4e_prefac * (-3.) * J4e_fac4;
        const double dxdyJ4 = GMearth * J4e_prefac * (-21.) * J4e_fac2 * dx * dy / r2;
        const double dydzJ4 = GMearth * J4e_prefac * (-21.) * J4e_fac3 * dy * dz / r2;
        const double dxdzJ4 = GMearth * J4e_prefac * (-21.) * J4e_fac3 * dx * dz / r2;
        for (int v = 0; v < s->n_var; v++) {
            if (s->var_tp[v] != j) continue;
            double *q = PX(s, N_real + v);
            const double ddx = q[0], ddy = q[1], ddz = q[2];
            double ddxp = -ddx * sina + ddy * cosa;
            double ddyp = -ddx * cosa * sind - ddy * sina * sind + ddz * cosd;
            double ddzp = ddx * cosa * cosd + ddy * sina * cosd + ddz * sind;
            double dax = ddxp * dxdx + ddyp * dxdy + ddzp * dxdz;
            double day = ddxp * dxdy + ddyp * dydy + ddzp * dydz;
            double daz = ddxp * dxdz + ddyp * dydz + ddzp * dzdz;
            dax += ddxp * dxdxJ3 + ddyp * dxdyJ3 + ddzp * dxdzJ3;
            day += ddxp * dxdyJ3 + ddyp * dydyJ3 + ddzp * dydzJ3;
            daz += ddxp * dxdzJ3 + ddyp * dydzJ3 + ddzp * dzdzJ3;
            dax += ddxp * dxdxJ4 + ddyp * dxdyJ4 + ddzp * dxdzJ4;
            day += ddxp * dxdyJ4 + ddyp * dydyJ4 + ddzp * dydzJ4;
            daz += ddxp * dxdzJ4 + ddyp * dydzJ4 + ddzp * dzdzJ4;
            double daxp = -dax * sina - day * cosa * sind + daz * cosa * cosd;
            double dayp = dax * cosa - day * sina * sind + daz * sina * cosd;
            double dazp = +day * cosd + daz * sind;
            q[6] += daxp;
            q[7] += dayp;
            q[8] += dazp;
        }
    }
}

/* reference src/forces.c:644-772 */
static void force_solar_J2(const orc_ephem *e, orc_sys *s, double xo, double yo, double zo) {
    const int N_real = s->n_real;
    double b[10];
    get(e, s, 0, s->t, b);
    const double GMsun = b[0], xr = b[1], yr = b[2], zr = b[3];
    const double Rs_eq = e->Rs_eq, J2s = e->J2SUN;
    const double RAs = 286.13 * PI_ / 180., Decs = 63.87 * PI_ / 180.;
    const double cosa = cos(RAs), sina = sin(RAs), cosd = cos(Decs), sind = sin(Decs);
    for (int j = 0; j < N_real; j++) {
        double *p = PX(s, j);
        double dx = p[0] + (xo - xr);
        double dy = p[1] + (yo - yr);
        double dz = p[2] + (zo - zr);
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r = sqrt(r2);
        double dxp = -dx * sina + dy * cosa;
        double dyp = -dx * cosa * sind - dy * sina * sind + dz * cosd;
        double dzp = dx * cosa * cosd + dy * sina * cosd + dz * sind;
        dx = dxp; dy = dyp; dz = dzp;
        const double costheta2 = dz * dz / r2;
        const double J2s_prefac = 3. * J2s * Rs_eq * Rs_eq / r2 / r2 / r / 2.;
        const double J2s_fac = 5. * costheta2 - 1.;
        const double J2s_fac2 = 7. * costheta2 - 1.;
        const double J2s_fac3 = 35. * costheta2 * costheta2 - 30. * costheta2 + 3.;
        double resx = GMsun * J2s_prefac * J2s_fac * dx;
        double resy = GMsun * J2s_prefac * J2s_fac * dy;
        double resz = GMsun * J2s_prefac * (J2s_fac - 2.) * dz;
        double resxp = -resx * sina - resy * cosa * sind + resz * cosa * cosd;
        double resyp = resx * cosa - resy * sina * sind + resz * sina * cosd;
        double reszp = +resy * cosd + resz * sind;
        p[6] += resxp;
        p[7] += resyp;
        p[8] += reszp;
        const double dxdx = GMsun * J2s_prefac * (J2s_fac - 5. * J2s_fac2 * dx * dx / r2);
        const double dydy = GMsun * J2s_prefac * (J2s_fac - 5. * J2s_fac2 * dy * dy / r2);
        const double dzdz = GMsun * J2s_prefac * (-1.) * J2s_fac3;
        const double dxdy = GMsun * J2s_prefac * (-5.) * J2s_fac2 * dx * dy / r2;
        const double dydz = GMsun * J2s_prefac * (-5.) * (J2s_fac2 - 2.) * dy * dz / r2;
        const double dxdz = GMsun * J2s_prefac * (-5.) * (J2s_fac2 - 2.) * dx * dz / r2;
        for (int v = 0; v < s->n_var; v++) {
            if (s->var_tp[v] != j) continue;
            double *q = PX(s, N_real + v);
            double ddx = q[0], ddy = q[1], ddz = q[2];
            double ddxp = -ddx * sina + ddy * cosa;
            double ddyp = -ddx * cosa * sind - ddy * sina * sind + ddz * cosd;
            double ddzp = ddx * cosa * cosd + ddy * sina * cosd + ddz * sind;
            ddx = ddxp; ddy = ddyp; ddz = ddzp;
            double daxp = ddx * dxdx + ddy * dxdy + ddz * dxdz;
            double dayp = ddx * dxdy + ddy * dydy + ddz * dydz;
            double dazp = ddx * dxdz + ddy * dydz + ddz * dzdz;
            double dax = -daxp * sina - dayp * cosa * sind + dazp * cosa * cosd;
            double day = daxp * cosa - dayp * sina * sind + dazp * sina * cosd;
            double daz = +dayp * cosd + dazp * sind;
            q[6] += dax;
            q[7] += day;
            q[8] += daz;
        }
    }
}

/* reference src/forces.c:774-1057 */
static void force_nongrav(const orc_ephem *e, const orc_opts *op, orc_sys *s, double xo, double yo,
                          double zo, double vxo, double vyo, double vzo) {
    const int N_real = s->n_real;
    const double *pp = s->params;
    if (pp == NULL) return;
    double b[10];
    get(e, s, 0, s->t, b);
    const double xr = b[1], yr = b[2], zr = b[3], vxr = b[4], vyr = b[5], vzr = b[6];
    for (int j = 0; j < N_real; j++) {
        double A1 = pp[3 * j + 0], A2 = pp[3 * j + 1], A3 = pp[3 * j + 2];
        double alpha = op->alpha, nk = op->nk, nm = op->nm, nn = op->nn, r0 = op->r0;
        if (A1 == 0. && A2 == 0. && A3 == 0.) continue;
        double *p = PX(s, j);
        double dx = p[0] + (xo - xr);
        double dy = p[1] + (yo - yr);
        double dz = p[2] + (zo - zr);
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r = sqrt(r2);
        const double g = alpha * pow(r / r0, -nm) * pow(1.0 + pow(r / r0, nn), -nk);
        double dvx = p[3] + (vxo - vxr);
        double dvy = p[4] + (vyo - vyr);
        double dvz = p[5] + (vzo - vzr);
        double hx = dy * dvz - dz * dvy;
        double hy = dz * dvx - dx * dvz;
        double hz = dx * dvy - dy * dvx;
        double h2 = hx * hx + hy * hy + hz * hz;
        double h = sqrt(h2);
        double tx = hy * dz - hz * dy;
        double ty = hz * dx - hx * dz;
        double tz = hx * dy - hy * dx;
        const double t2 = tx * tx + ty * ty + tz * tz;
        const double _t = sqrt(t2);
        p[6] += A1 * g * dx / r + A2 * g * tx / _t + A3 * g * hx / h;
        p[7] += A1 * g * dy / r + A2 * g * ty / _t + A3 * g * hy / h;
        p[8] += A1 * g * dz / r + A2 * g * tz / _t + A3 * g * hz / h;

        const double r3 = r * r * r;
        const double v2 = dvx * dvx + dvy * dvy + dvz * dvz;
        const double rdotv = dx * dvx + dy * dvy + dz * dvz;
        const double vdott = dvx * tx + dvy * ty + dvz * tz;
        const double dgdr = (alpha / r0) * (-nm * pow(r / r0, -nm - 1) * pow(1.0 + pow(r / r0, nn), -nk)
                                            + pow(r / r0, -nm) * (-nk * nn) * pow(r / r0, nn - 1) * pow(1.0 + pow(r / r0, nn), -nk - 1));
        const double dgx = dgdr * dx / r, dgy = dgdr * dy / r, dgz = dgdr * dz / r;
        const double hxh3 = hx / (h * h * h), hyh3 = hy / (h * h * h), hzh3 = hz / (h * h * h);
        const double txt3 = tx / (_t * _t * _t), tyt3 = ty / (_t * _t * _t), tzt3 = tz / (_t * _t * _t);
        const double dxdA1 = g * dx / r, dydA1 = g * dy / r, dzdA1 = g * dz / r;
        const double dxdA2 = g * tx / _t, dydA2 = g * ty / _t, dzdA2 = g * tz / _t;
        const double dxdA3 = g * hx / h, dydA3 = g * hy / h, dzdA3 = g * hz / h;
        const double dxdx = A1 * (dgx * dx / r + g * (1. / r - dx * dx / r3))
            + A2 * (dgx * tx / _t + g * ((dx * dvx - rdotv) / _t - txt3 * (2. * dx * vdott - rdotv * tx)))
            + A3 * (dgx * hx / h + g * (-hxh3) * (v2 * dx - rdotv * dvx));
        const double dydy = A1 * (dgy * dy / r + g * (1. / r - dy * dy / r3))
            + A2 * (dgy * ty / _t + g * ((dy * dvy - rdotv) / _t - tyt3 * (2. * dy * vdott - rdotv * ty)))
            + A3 * (dgy * hy / h + g * (-hyh3) * (v2 * dy - rdotv * dvy));
        const double dzdz = A1 * (dgz * dz / r + g * (1. / r - dz * dz / r3))
            + A2 * (dgz * tz / _t + g * ((dz * dvz - rdotv) / _t - tzt3 * (2. * dz * vdott - rdotv * tz)))
            + A3 * (dgz * hz / h + g * (-hzh3) * (v2 * dz - rdotv * dvz));
        const double dxdy = A1 * (dgy * dx / r + g * (-dx * dy / r3))
            + A2 * (dgy * tx / _t + g * ((2 * dy * dvx - dx * dvy) / _t - txt3 * (2 * dy * vdott - rdotv * ty)))
            + A3 * (dgy * hx / h + g * (dvz / h - hxh3 * (v2 * dy - rdotv * dvy)));
        const double dydx = A1 * (dgx * dy / r + g * (-dy * dx / r3))
            + A2 * (dgx * ty / _t + g * ((2 * dx * dvy - dy * dvx) / _t - tyt3 * (2 * dx * vdott - rdotv * tx)))
            + A3 * (dgx * hy / h + g * (-dvz / h - hyh3 * (v2 * dx - rdotv * dvx)));
        const double dxdz = A1 * (dgz * dx / r + g * (-dx * dz / r3))
            + A2 * (dgz * tx / _t + g * ((2 * dz * dvx - dx * dvz) / _t - txt3 * (2 * dz * vdott - rdotv * tz)))
            + A3 * (dgz * hx / h + g * (-dvy / h - hxh3 * (v2 * dz - rdotv * dvz)));
        const double dzdx = A1 * (dgx * dz / r + g * (-dz * dx / r3))
            + A2 * (dgx * tz / _t + g * ((2 * dx * dvz - dz * dvx) / _t - tzt3 * (2 * dx * vdott - rdotv * tx)))
            + A3 * (dgx * hz / h + g * (dvy / h - hzh3 * (v2 * dx - rdotv * dvx)));
        const double dydz = A1 * (dgz * dy / r + g * (-dy * dz / r3))
            + A2 * (dgz * ty / _t + g * ((2 * dz * dvy - dy * dvz) / _t - tyt3 * (2 * dz * vdott - rdotv * tz)))
            + A3 * (dgz * hy / h + g * (dvx / h - hyh3 * (v2 * dz - rdotv * dvz)));
        const double dzdy = A1 * (dgy * dz / r + g * (-dz * dy / r3))
            + A2 * (dgy * tz / _t + g * ((2 * dy * dvz - dz * dvy) / _t - tzt3 * (2 * dy * vdott - rdotv * ty)))
            + A3 * (dgy * hz / h + g * (-dvx / h - hzh3 * (v2 * dy - rdotv * dvy)));
        const double dxdvx = A1 * (0.) + A2 * g * ((dy * dy + dz * dz) / _t - txt3 * r2 * tx) + A3 * g * (-hxh3 * (r2 * dvx - dx * rdotv));
        const double dydvy = A1 * (0.) + A2 * g * ((dx * dx + dz * dz) / _t - tyt3 * r2 * ty) + A3 * g * (-hyh3 * (r2 * dvy - dy * rdotv));
        const double dzdvz = A1 * (0.) + A2 * g * ((dx * dx + dy * dy) / _t - tzt3 * r2 * tz) + A3 * g * (-hzh3 * (r2 * dvz - dz * rdotv));
        const double dxdvy = A1 * (0.) + A2 * g * (-dy * dx / _t - tyt3 * r2 * tx) + A3 * g * (-dz / h - hxh3 * (r2 * dvy - dy * rdotv));
        const double dydvx = A1 * (0.) + A2 * g * (-dx * dy / _t - txt3 * r2 * ty) + A3 * g * (dz / h - hyh3 * (r2 * dvx - dx * rdotv));
        const double dxdvz = A1 * (0.) + A2 * g * (-dz * dx / _t - tzt3 * r2 * tx) + A3 * g * (dy / h - hxh3 * (r2 * dvz - dz * rdotv));
        const double dzdvx = A1 * (0.) + A2 * g * (-dx * dz / _t - txt3 * r2 * tz) + A3 * g * (-dy / h - hzh3 * (r2 * dvx - dx * rdotv));
        const double dydvz = A1 * (0.) + A2 * g * (-dz * dy / _t - tzt3 * r2 * ty) + A3 * g * (-dx / h - hyh3 * (r2 * dvz - dz * rdotv));
        const double dzdvy = A1 * (0.) + A2 * g * (-dy * dz / _t - tyt3 * r2 * tz) + A3 * g * (dx / h - hzh3 * (r2 * dvy - dy * rdotv));
        for (int v = 0; v < s->n_var; v++) {
            if (s->var_tp[v] != j) continue;
            double *q = PX(s, N_real + v);
            double ddx = q[0], ddy = q[1], ddz = q[2], ddvx = q[3], ddvy = q[4], ddvz = q[5];
            double dA1 = pp[3 * (N_real + v) + 0], dA2 = pp[3 * (N_real + v) + 1], dA3 = pp[3 * (N_real + v) + 2];
            const double dax = ddx * dxdx + ddy * dxdy + ddz * dxdz + ddvx * dxdvx + ddvy * dxdvy + ddvz * dxdvz + dA1 * dxdA1 + dA2 * dxdA2 + dA3 * dxdA3;
            const double day = ddx * dydx + ddy * dydy + ddz * dydz + ddvx * dydvx + ddvy * dydvy + ddvz * dydvz + dA1 * dydA1 + dA2 * dydA2 + dA3 * dydA3;
            const double daz = ddx * dzdx + ddy * dzdy + ddz * dzdz + ddvx * dzdvx + ddvy * dzdvy + ddvz * dzdvz + dA1 * dzdA1 + dA2 * dzdA2 + dA3 * dzdA3;
            q[6] += dax;
            q[7] += day;
            q[8] += daz;
        }
    }
}

/* reference src/forces.c:1059-1161 */
static void force_gr_potential(const orc_ephem *e, orc_sys *s, double xo, double yo, double zo) {
    const double C2 = e->c_squared;
    double b[10];
    get(e, s, 0, s->t, b);
    const double GMsun = b[0], xr = b[1], yr = b[2], zr = b[3];
    for (int j = 0; j < s->n_real; j++) {
        double *pj = PX(s, j);
        double px = pj[0] + (xo - xr), py = pj[1] + (yo - yr), pz = pj[2] + (zo - zr);
        const double r2 = px * px + py * py + pz * pz;
        const double r = sqrt(r2);
        const double prefac = -6.0 * GMsun * GMsun / (C2 * r2 * r2);
        pj[6] += prefac * px;
        pj[7] += prefac * py;
        pj[8] += prefac * pz;
        const double dxdx = prefac + -4.0 * prefac * (px / r) * (px / r);
        const double dxdy = -4.0 * prefac * (px / r) * (py / r);
        const double dxdz = -4.0 * prefac * (px / r) * (pz / r);
        const double dydx = -4.0 * prefac * (py / r) * (px / r);
        const double dydy = prefac + -4.0 * prefac * (py / r) * (py / r);
        const double dydz = -4.0 * prefac * (py / r) * (pz / r);
        const double dzdx = -4.0 * prefac * (pz / r) * (px / r);
        const double dzdy = -4.0 * prefac * (pz / r) * (py / r);
        const double dzdz = prefac + -4.0 * prefac * (pz / r) * (pz / r);
        for (int v = 0; v < s->n_var; v++) {
            if (s->var_tp[v] != j) continue;
            double *q = PX(s, s->n_real + v);
            const double ddx = q[0], ddy = q[1], ddz = q[2];
            q[6] += ddx * dxdx + ddy * dxdy + ddz * dxdz;
            q[7] += ddx * dydx + ddy * dydy + ddz * dydz;
            q[8] += ddx * dzdx + ddy * dzdy + ddz * dzdz;
        }
    }
}

/* reference src/forces.c:1163-1286 */
static void force_gr_simple(const orc_ephem *e, orc_sys *s, double xo, double yo, double zo, double vxo,
                            double vyo, double vzo) {
    const double C2 = e->c_squared;
    double b[10];
    get(e, s, 0, s->t, b);
    const double GMsun = b[0], xr = b[1], yr = b[2], zr = b[3], vxr = b[4], vyr = b[5], vzr = b[6];
    for (int j = 0; j < s->n_real; j++) {
        double *pj = PX(s, j);
        double px = pj[0] + (xo - xr), py = pj[1] + (yo - yr), pz = pj[2] + (zo - zr);
        double pvx = pj[3] + (vxo - vxr), pvy = pj[4] + (vyo - vyr), pvz = pj[5] + (vzo - vzr);
        const double v2 = pvx * pvx + pvy * pvy + pvz * pvz;
        const double r = sqrt(px * px + py * py + pz * pz);
        const double A = 4.0 * GMsun / r - v2;
        const double B = 4.0 * (px * pvx + py * pvy + pz * pvz);
        const double prefac = GMsun / (r * r * r * C2);
        pj[6] += prefac * (A * px + B * pvx);
        pj[7] += prefac * (A * py + B * pvy);
        pj[8] += prefac * (A * pz + B * pvz);
        const double dpdr = -3.0 * prefac / r;
        const double dxdx = dpdr * px / r * (A * px + B * pvx) + prefac * (A - px * (px / r) * 4.0 * GMsun / (r * r) + 4.0 * pvx * pvx);
        const double dxdy = dpdr * py / r * (A * px + B * pvx) + prefac * (-px * (py / r) * 4.0 * GMsun / (r * r) + 4.0 * pvy * pvx);
        const double dxdz = dpdr * pz / r * (A * px + B * pvx) + prefac * (-px * (pz / r) * 4.0 * GMsun / (r * r) + 4.0 * pvz * pvx);
        const double dxdvx = prefac * (-2.0 * pvx * px + 4.0 * px * pvx + B);
        const double dxdvy = prefac * (-2.0 * pvy * px + 4.0 * py * pvx);
        const double dxdvz = prefac * (-2.0 * pvz * px + 4.0 * pz * pvx);
        const double dydx = dpdr * px / r * (A * py + B * pvy) + prefac * (-py * (px / r) * 4.0 * GMsun / (r * r) + 4.0 * pvx * pvy);
        const double dydy = dpdr * py / r * (A * py + B * pvy) + prefac * (A - py * (py / r) * 4.0 * GMsun / (r * r) + 4.0 * pvy * pvy);
        const double dydz = dpdr * pz / r * (A * py + B * pvy) + prefac * (-py * (pz / r) * 4.0 * GMsun / (r * r) + 4.0 * pvz * pvy);
        const double dydvx = prefac * (-2.0 * pvx * py + 4.0 * px * pvy);
        const double dydvy = prefac * (-2.0 * pvy * py + 4.0 * py * pvy + B);
        const double dydvz = prefac * (-2.0 * pvz * py + 4.0 * pz * pvy);
        const double dzdx = dpdr * px / r * (A * pz + B * pvz) + prefac * (-pz * (px / r) * 4.0 * GMsun / (r * r) + 4.0 * pvx * pvz);
        const double dzdy = dpdr * py / r * (A * pz + B * pvz) + prefac * (-pz * (py / r) * 4.0 * GMsun / (r * r) + 4.0 * pvy * pvz);
        const double dzdz = dpdr * pz / r * (A * pz + B * pvz) + prefac * (A - pz * (pz / r) * 4.0 * GMsun / (r * r) + 4.0 * pvz * pvz);
        const double dzdvx = prefac * (-2.0 * pvx * pz + 4.0 * px * pvz);
        const double dzdvy = prefac * (-2.0 * pvy * pz + 4.0 * py * pvz);
        const double dzdvz = prefac * (-2.0 * pvz * pz + 4.0 * pz * pvz + B);
        for (int v = 0; v < s->n_var; v++) {
            if (s->var_tp[v] != j) continue;
            double *q = PX(s, s->n_real + v);
            const double ddx = q[0], ddy = q[1], ddz = q[2], ddvx = q[3], ddvy = q[4], ddvz = q[5];
            q[6] += ddx * dxdx + ddy * dxdy + ddz * dxdz + ddvx * dxdvx + ddvy * dxdvy + ddvz * dxdvz;
            q[7] += ddx * dydx + ddy * dydy + ddz * dydz + ddvx * dydvx + ddvy * dydvy + ddvz * dydvz;
            q[8] += ddx * dzdx + ddy * dzdy + ddz * dzdz + ddvx * dzdvx + ddvy * dzdvy + ddvz * dzdvz;
        }
    }
}

/* reference src/forces.c:1288-1983 (beta = gamma = 1) */
static void force_eih(const orc_ephem *e, const orc_opts *op, orc_sys *s, double xo, double yo, double zo,
                      double vxo, double vyo, double vzo, double axo, double ayo, double azo) {
    const double over_C2 = e->over_c_squared;
    const double beta = 1.0, gamma = 1.0;
    const int N_real = s->n_real, S = op->gr_eih_sources;
    const double t = s->t;
    double bj[10], bk[10];
    for (int i = 0; i < N_real; i++) {
        double *pi = PX(s, i);
        double term7x_sum = 0.0, term7y_sum = 0.0, term7z_sum = 0.0;
        double term8x_sum = 0.0, term8y_sum = 0.0, term8z_sum = 0.0;
        for (int j = 0; j < S; j++) {
            get(e, s, j, t, bj);
            const double GMj = bj[0], xj = bj[1], yj = bj[2], zj = bj[3], vxj = bj[4], vyj = bj[5], vzj = bj[6];
            double axj, ayj, azj;
            const double dxij = pi[0] + (xo - xj);
            const double dyij = pi[1] + (yo - yj);
            const double dzij = pi[2] + (zo - zj);
            const double rij2 = dxij * dxij + dyij * dyij + dzij * dzij;
            const double _rij = sqrt(rij2);
            const double prefacij = GMj / (rij2 * _rij);
            const double vi2 = pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5];
            const double term2 = gamma * over_C2 * vi2;
            const double vj2 = (vxj - vxo) * (vxj - vxo) + (vyj - vyo) * (vyj - vyo) + (vzj - vzo) * (vzj - vzo);
            const double term3 = (1 + gamma) * over_C2 * vj2;
            const double vidotvj = pi[3] * (vxj - vxo) + pi[4] * (vyj - vyo) + pi[5] * (vzj - vzo);
            const double term4 = -2 * (1 + gamma) * over_C2 * vidotvj;
            const double rijdotvj = dxij * (vxj - vxo) + dyij * (vyj - vyo) + dzij * (vzj - vzo);
            const double term5 = -1.5 * over_C2 * (rijdotvj * rijdotvj) / (_rij * _rij);
            const double fx = (2 + 2 * gamma) * pi[3] - (1 + 2 * gamma) * (vxj - vxo);
            const double fy = (2 + 2 * gamma) * pi[4] - (1 + 2 * gamma) * (vyj - vyo);
            const double fz = (2 + 2 * gamma) * pi[5] - (1 + 2 * gamma) * (vzj - vzo);
            const double f = dxij * fx + dyij * fy + dzij * fz;
            const double prefacij_f = prefacij * f;
            term7x_sum += prefacij_f * (pi[3] - (vxj - vxo));
            term7y_sum += prefacij_f * (pi[4] - (vyj - vyo));
            term7z_sum += prefacij_f * (pi[5] - (vzj - vzo));
            double term0 = 0.0, term1 = 0.0;
            axj = 0.0; ayj = 0.0; azj = 0.0;
            for (int k = 0; k < ORC_NPLANETS; k++) {
                get(e, s, k, t, bk);
                const double GMk = bk[0], xk = bk[1], yk = bk[2], zk = bk[3];
                const double dxik = pi[0] + (xo - xk);
                const double dyik = pi[1] + (yo - yk);
                const double dzik = pi[2] + (zo - zk);
                const double rik2 = dxik * dxik + dyik * dyik + dzik * dzik;
                const double _rik = sqrt(rik2);
                term0 += GMk / _rik;
                if (k != j) {
                    const double dxjk = xj - xk, dyjk = yj - yk, dzjk = zj - zk;
                    const double rjk2 = dxjk * dxjk + dyjk * dyjk + dzjk * dzjk;
                    const double _rjk = sqrt(rjk2);
                    term1 += GMk / _rjk;
                    const double fac = GMk / (rjk2 * _rjk);
                    axj -= fac * dxjk;
                    ayj -= fac * dyjk;
                    azj -= fac * dzjk;
                }
            }
            term0 *= -2 * (beta + gamma) * over_C2;
            term1 *= -(2 * beta - 1) * over_C2;
            const double rijdotaj = dxij * (axj - axo) + dyij * (ayj - ayo) + dzij * (azj - azo);
            const double term6 = -0.5 * over_C2 * rijdotaj;
            const double term8_fac = GMj / _rij * (3 + 4 * gamma) / 2;
            term8x_sum += term8_fac * axj;
            term8y_sum += term8_fac * ayj;
            term8z_sum += term8_fac * azj;
            const double factor = term0 + term1 + term2 + term3 + term4 + term5 + term6;
            pi[6] += -prefacij * dxij * factor;
            pi[7] += -prefacij * dyij * factor;
            pi[8] += -prefacij * dzij * factor;
        }
        pi[6] += term7x_sum * over_C2 + term8x_sum * over_C2;
        pi[7] += term7y_sum * over_C2 + term8y_sum * over_C2;
        pi[8] += term7z_sum * over_C2 + term8z_sum * over_C2;
    }
    if (s->n_var == 0) return;
    for (int i = 0; i < N_real; i++) {
        const double *pi = PX(s, i);
        double J[3][6] = {{0}};
        double d7[3][6] = {{0}}; /* d term7_sum / d(x,y,z,vx,vy,vz) */
        double d8[3][3] = {{0}};
        for (int j = 0; j < S; j++) {
            get(e, s, j, t, bj);
            const double GMj = bj[0], xj = bj[1], yj = bj[2], zj = bj[3], vxj = bj[4], vyj = bj[5], vzj = bj[6];
            double axj, ayj, azj;
            const double dxij = pi[0] + (xo - xj);
            const double dyij = pi[1] + (yo - yj);
            const double dzij = pi[2] + (zo - zj);
            const double rij2 = dxij * dxij + dyij * dyij + dzij * dzij;
            const double _rij = sqrt(rij2);
            const double prefacij = GMj / (_rij * _rij * _rij);
            const double dprefacijdx = -3.0 * GMj / (_rij * _rij * _rij * _rij * _rij) * dxij;
            const double dprefacijdy = -3.0 * GMj / (_rij * _rij * _rij * _rij * _rij) * dyij;
            const double dprefacijdz = -3.0 * GMj / (_rij * _rij * _rij * _rij * _rij) * dzij;
            const double vi2 = pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5];
            const double term2 = gamma * over_C2 * vi2;
            const double dterm2dvx = 2.0 * gamma * over_C2 * pi[3];
            const double dterm2dvy = 2.0 * gamma * over_C2 * pi[4];
            const double dterm2dvz = 2.0 * gamma * over_C2 * pi[5];
            const double vj2 = (vxj - vxo) * (vxj - vxo) + (vyj - vyo) * (vyj - vyo) + (vzj - vzo) * (vzj - vzo);
            const double term3 = (1 + gamma) * over_C2 * vj2;
            const double vidotvj = pi[3] * (vxj - vxo) + pi[4] * (vyj - vyo) + pi[5] * (vzj - vzo);
            const double term4 = -2 * (1 + gamma) * over_C2 * vidotvj;
            const double dterm4dvx = -2 * (1 + gamma) * over_C2 * (vxj - vxo);
            const double dterm4dvy = -2 * (1 + gamma) * over_C2 * (vyj - vyo);
            const double dterm4dvz = -2 * (1 + gamma) * over_C2 * (vzj - vzo);
            const double rijdotvj = dxij * (vxj - vxo) + dyij * (vyj - vyo) + dzij * (vzj - vzo);
            const double term5 = -1.5 * over_C2 * (rijdotvj * rijdotvj) / (_rij * _rij);
            const double term5_fac = 3.0 * over_C2 * rijdotvj / _rij;
            const double dterm5dx = -term5_fac * ((vxj - vxo) / _rij - rijdotvj * dxij / (_rij * _rij * _rij));
            const double dterm5dy = -term5_fac * ((vyj - vyo) / _rij - rijdotvj * dyij / (_rij * _rij * _rij));
            const double dterm5dz = -term5_fac * ((vzj - vzo) / _rij - rijdotvj * dzij / (_rij * _rij * _rij));
            const double fx = (2 + 2 * gamma) * pi[3] - (1 + 2 * gamma) * (vxj - vxo);
            const double fy = (2 + 2 * gamma) * pi[4] - (1 + 2 * gamma) * (vyj - vyo);
            const double fz = (2 + 2 * gamma) * pi[5] - (1 + 2 * gamma) * (vzj - vzo);
            const double f = dxij * fx + dyij * fy + dzij * fz;
            const double dfdx[3] = {fx, fy, fz};
            const double dfdv[3] = {dxij * (2 + 2 * gamma), dyij * (2 + 2 * gamma), dzij * (2 + 2 * gamma)};
            const double dpref[3] = {dprefacijdx, dprefacijdy, dprefacijdz};
            const double rel[3] = {pi[3] - (vxj - vxo), pi[4] - (vyj - vyo), pi[5] - (vzj - vzo)};
            for (int c = 0; c < 3; c++) {
                for (int q = 0; q < 3; q++) {
                    d7[c][q] += dpref[q] * f * rel[c] + prefacij * dfdx[q] * rel[c];
                    double dv = prefacij * dfdv[q] * rel[c];
                    if (q == c) dv = prefacij * dfdv[q] * rel[c] + prefacij * f;
                    d7[c][3 + q] += dv;
                }
            }
            double term0 = 0.0, dterm0dx = 0.0, dterm0dy = 0.0, dterm0dz = 0.0;
            double term1 = 0.0;
            const double dterm1dvx = 0.0, dterm1dvy = 0.0, dterm1dvz = 0.0;
            const double dterm1dx = 0.0, dterm1dy = 0.0, dterm1dz = 0.0;
            axj = 0.0; ayj = 0.0; azj = 0.0;
            for (int k = 0; k < ORC_NPLANETS; k++) {
                get(e, s, k, t, bk);
                const double GMk = bk[0], xk = bk[1], yk = bk[2], zk = bk[3];
                const double dxik = pi[0] + (xo - xk);
                const double dyik = pi[1] + (yo - yk);
                const double dzik = pi[2] + (zo - zk);
                const double rik2 = dxik * dxik + dyik * dyik + dzik * dzik;
                const double _rik = sqrt(rik2);
                term0 += GMk / _rik;
                dterm0dx -= GMk / (_rik * _rik * _rik) * dxik;
                dterm0dy -= GMk / (_rik * _rik * _rik) * dyik;
                dterm0dz -= GMk / (_rik * _rik * _rik) * dzik;
                if (k != j) {
                    const double dxjk = xj - xk, dyjk = yj - yk, dzjk = zj - zk;
                    const double rjk2 = dxjk * dxjk + dyjk * dyjk + dzjk * dzjk;
                    const double _rjk = sqrt(rjk2);
                    term1 += GMk / _rjk;
                    axj -= GMk * dxjk / (_rjk * _rjk * _rjk);
                    ayj -= GMk * dyjk / (_rjk * _rjk * _rjk);
                    azj -= GMk * dzjk / (_rjk * _rjk * _rjk);
                }
            }
            term0 *= -2 * (beta + gamma) * over_C2;
            dterm0dx *= -2 * (beta + gamma) * over_C2;
            dterm0dy *= -2 * (beta + gamma) * over_C2;
            dterm0dz *= -2 * (beta + gamma) * over_C2;
            term1 *= -(2 * beta - 1) * over_C2;
            const double rijdotaj = dxij * (axj - axo) + dyij * (ayj - ayo) + dzij * (azj - azo);
            const double term6 = -0.5 * over_C2 * rijdotaj;
            const double dterm6dx = -0.5 * over_C2 * (axj - axo);
            const double dterm6dy = -0.5 * over_C2 * (ayj - ayo);
            const double dterm6dz = -0.5 * over_C2 * (azj - azo);
            const double aj[3] = {axj, ayj, azj};
            const double dij[3] = {dxij, dyij, dzij};
            for (int c = 0; c < 3; c++)
                for (int q = 0; q < 3; q++)
                    d8[c][q] += -GMj * aj[c] / (_rij * _rij * _rij) * dij[q] * (3 + 4 * gamma) / 2;
            const double factor = term0 + term1 + term2 + term3 + term4 + term5 + term6;
            const double dfac_x[3] = {dterm0dx + dterm1dx + dterm5dx + dterm6dx, dterm0dy + dterm1dy + dterm5dy + dterm6dy,
                                      dterm0dz + dterm1dz + dterm5dz + dterm6dz};
            const double dfac_v[3] = {dterm1dvx + dterm2dvx + dterm4dvx, dterm1dvy + dterm2dvy + dterm4dvy,
                                      dterm1dvz + dterm2dvz + dterm4dvz};
            for (int c = 0; c < 3; c++) {
                for (int q = 0; q < 3; q++) {
                    double v = -dpref[q] * dij[c] * factor;
                    if (q == c) v = v - prefacij * factor;
                    v = v - prefacij * dij[c] * dfac_x[q];
                    J[c][q] += v;
                    J[c][3 + q] += -prefacij * dij[c] * dfac_v[q];
                }
            }
        }
        for (int c = 0; c < 3; c++) {
            for (int q = 0; q < 3; q++) {
                J[c][q] += d7[c][q] * over_C2 + d8[c][q] * over_C2;
                J[c][3 + q] += d7[c][3 + q] * over_C2;
            }
        }
        for (int v = 0; v < s->n_var; v++) {
            if (s->var_tp[v] != i) continue;
            double *q = PX(s, N_real + v);
            for (int c = 0; c < 3; c++)
                q[6 + c] += q[0] * J[c][0] + q[1] * J[c][1] + q[2] * J[c][2] + q[3] * J[c][3] + q[4] * J[c][4] + q[5] * J[c][5];
        }
    }
}

/* reference src/forces.c:49-173 */
void orc_additional_forces(const orc_ephem *e, const orc_opts *o, orc_sys *s) {
    const int N = s->n_real + s->n_var;
    for (int i = 0; i < N; i++) PX(s, i)[6] = PX(s, i)[7] = PX(s, i)[8] = 0.0;
    double xo = 0, yo = 0, zo = 0, vxo = 0, vyo = 0, vzo = 0, axo = 0, ayo = 0, azo = 0;
    double b[10];
    if (o->geocentric == 1) {
        if (get(e, s, 3, s->t, b) != ORC_SUCCESS) return;
        xo = b[1]; yo = b[2]; zo = b[3]; vxo = b[4]; vyo = b[5]; vzo = b[6]; axo = b[7]; ayo = b[8]; azo = b[9];
    }
    if (o->forces & ORC_FORCE_NON_GRAVITATIONAL) force_nongrav(e, o, s, xo, yo, zo, vxo, vyo, vzo);
    if (o->forces & ORC_FORCE_EARTH_HARMONICS) force_earth_J2J4(e, s, xo, yo, zo);
    if (o->forces & ORC_FORCE_SUN_HARMONICS) force_solar_J2(e, s, xo, yo, zo);
    if (o->forces & ORC_FORCE_GR_EIH) force_eih(e, o, s, xo, yo, zo, vxo, vyo, vzo, axo, ayo, azo);
    if (o->forces & ORC_FORCE_GR_POTENTIAL) force_gr_potential(e, s, xo, yo, zo);
    if (o->forces & ORC_FORCE_GR_SIMPLE) force_gr_simple(e, s, xo, yo, zo, vxo, vyo, vzo);
    if (o->forces & (ORC_FORCE_SUN | ORC_FORCE_PLANETS | ORC_FORCE_ASTEROIDS)) force_direct(e, o, s, xo, yo, zo);
    if (o->geocentric == 1) {
        get(e, s, 3, s->t, b);
        for (int j = 0; j < s->n_real; j++) {
            PX(s, j)[6] -= b[7];
            PX(s, j)[7] -= b[8];
            PX(s, j)[8] -= b[9];
        }
    }
}
