/*
 * Host timing contract of the RBF-FD SWE driver (swe_code/main/profiling.c):
 * wall-clock accumulators per attempt, reported in seconds per timestep
 * with average / min / max / "stddev" per phase, where the stddev is the
 * reference's sqrt(sum((avg - x)^2)) / nattempts over every attempt
 * (profiling.c:129-165) and Eval_Rhs averages, minimises and maximises only
 * attempts >= 2 (profiling.c:119-125): with nattempts <= 2 its Average, Min
 * and Max are 0 and its STDDEV is taken against that 0.  nsteps <= 0 is
 * reported per single step (the reference would divide by zero).
 */
#include "swe_host.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

double swe_get_time(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int swe_timers_init(swe_timers *t, int nattempts)
{
    memset(t, 0, sizeof(*t));
    if (nattempts <= 0)
        return SWE_HOST_EARG;
    t->nattempts = nattempts;
    double **arr[5] = {&t->t_main, &t->t_eval_rhs, &t->t_eval_K, &t->t_update_D, &t->t_update_H};
    for (int k = 0; k < 5; k++)
        if (!(*arr[k] = (double *)calloc(nattempts, sizeof(double)))) {
            swe_timers_free(t);
            return SWE_HOST_ENOMEM;
        }
    return 0;
}

void swe_timers_free(swe_timers *t)
{
    free(t->t_main); free(t->t_eval_rhs); free(t->t_eval_K);
    free(t->t_update_D); free(t->t_update_H);
    memset(t, 0, sizeof(*t));
}

static void phase_stats(const double *x, int n, int first, int nsteps, double out[4])
{
    double avg = 0.0, mn = 0.0, mx = 0.0, ss = 0.0;
    const int used = n - first;
    for (int i = first; i < n; i++) {
        const double v = x[i] / nsteps;
        avg += v / used;
        if (i == first || v < mn) mn = v;
        if (i == first || v > mx) mx = v;
    }
    for (int i = 0; i < n; i++) {
        const double v = x[i] / nsteps;
        ss += pow(avg - v, 2);
    }
    out[0] = avg;
    out[1] = mn;
    out[2] = mx;
    out[3] = sqrt(ss) / n;
}

void swe_process_profiling(const swe_timers *t, int nsteps, int print, double out[20])
{
    const double *ph[5] = {t->t_main, t->t_eval_rhs, t->t_eval_K, t->t_update_D, t->t_update_H};
    static const char *names[5] = {"Main RK4 Loop", "Eval_Rhs     ", "Eval_K       ",
                                   "Update_D     ", "Update_H     "};
    double local[20];
    double *o = out ? out : local;
    for (int k = 0; k < 5; k++) {
        const int first = k == 1 ? 2 : 0;   /* warm-up attempts out of Eval_Rhs */
        phase_stats(ph[k], t->nattempts, first, nsteps > 0 ? nsteps : 1, &o[4 * k]);
    }
    if (!print)
        return;
    printf("\n============================================================= Profiling Results "
           "==============================================================\n\n");
    printf("Total Initialization Time (seconds): \t%e\n", t->t_init);
    printf("Device Setup Time (seconds): \t%e\n", t->t_dev);
    for (int k = 0; k < 5; k++)
        printf("%s (seconds/timestep) -> \tAverage: \t%e \tMin: \t%e \tMax: \t%e \tSTDDEV: \t%e\n",
               names[k], o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
    printf("\n==================================================================================="
           "===========================================================\n");
    fflush(stdout);
}
