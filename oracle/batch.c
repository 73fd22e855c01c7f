/*
 * Oracle (test infrastructure): batched drivers.  Every system (one real
 * particle plus n_var variational particles) is an independent reference-style
 * simulation with its own IAS15 state, so the CPU baseline parallelises over
 * systems with OpenMP (dynamic schedule: spans and step counts differ).
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

int orc_batch_forces(const orc_ephem *e, const orc_opts *o, int nsys, int n_var, const double *t,
                     const double *states, const double *params, double *acc, int *status) {
    const int nv = 1 + n_var;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int i = 0; i < nsys; i++) {
        double p[9 * 16];
        int tp[16] = {0};
        orc_sys s;
        s.n_real = 1;
        s.n_var = n_var;
        s.var_tp = tp;
        s.t = t[i];
        s.p = p;
        s.params = params ? params + (size_t)3 * nv * i : NULL;
        s.status = 0;
        for (int v = 0; v < nv; v++)
            for (int c = 0; c < 6; c++) p[9 * v + c] = states[((size_t)i * nv + v) * 6 + c];
        orc_additional_forces(e, o, &s);
        for (int v = 0; v < nv; v++)
            for (int c = 0; c < 3; c++) acc[((size_t)i * nv + v) * 3 + c] = p[9 * v + 6 + c];
        if (status) status[i] = s.status;
        bad += s.status != 0;
    }
    return bad;
}

int orc_batch_integrate(const orc_ephem *e, const orc_opts *o, const orc_ias15_cfg *cfg, int nsys, int n_var,
                        const double *t0, const double *t1, double *states, const double *params, int nthreads,
                        int *status, long *steps, long *evals, long *rejected, double *t_out) {
    const int nv = 1 + n_var;
    int bad = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
    for (int i = 0; i < nsys; i++) {
        double p[9 * 16];
        int tp[16] = {0};
        orc_sys s;
        orc_stats st;
        orc_ias15_cfg c = *cfg;
        s.n_real = 1;
        s.n_var = n_var;
        s.var_tp = tp;
        s.t = t0[i];
        s.p = p;
        s.params = params ? params + (size_t)3 * nv * i : NULL;
        s.status = 0;
        for (int v = 0; v < nv; v++)
            for (int k = 0; k < 6; k++) p[9 * v + k] = states[((size_t)i * nv + v) * 6 + k];
        orc_integrate(e, o, &c, &s, t1[i], &st);
        for (int v = 0; v < nv; v++)
            for (int k = 0; k < 6; k++) states[((size_t)i * nv + v) * 6 + k] = p[9 * v + k];
        if (status) status[i] = st.status;
        if (steps) steps[i] = st.steps;
        if (evals) evals[i] = st.evals;
        if (rejected) rejected[i] = st.rejected;
        if (t_out) t_out[i] = s.t;
        bad += st.status != 0;
    }
    return bad;
}
