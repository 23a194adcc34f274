/*
 * CPU oracle for the RBF-FD shallow-water RK4 hot path -- TEST
 * INFRASTRUCTURE ONLY (see swe_oracle.h for who may use it and how each
 * function is pinned).
 *
 * The arithmetic (operation order, mul-then-add accumulation, sequential
 * stencil loop, constants dt/2.0, 2.0, 1.0, dt/6.0) follows the reference
 * exactly, because parity is checked bitwise against the reference build:
 *   eval_RHS_swe   swe_code/main/rk4_rbffd_swe.c:189-362
 *   RK_substep     swe_code/main/rk4_rbffd_swe.c:30-72
 *   axpy kernels   swe_code/main/rk4_rbffd_swe.c:74-186
 *   time loop      swe_code/main/main.c:336-358
 *   verify_output  swe_code/main/main.c:439-477
 *   index macros   swe_code/main/include/swe_config.h:21-54
 *
 * Three flavours are built from this file:
 *   strict  (-O2 -ffp-contract=off): sequential j, no FMA -> the checker;
 *   fast    (-O3 -march=... -DORACLE_INNER_SIMD): the reference's best CPU
 *           configuration (omp simd reduction over j, rk4_rbffd_swe.c:275-278),
 *           used only as the timed CPU baseline;
 *   strict32 (-DORACLE_FP32): the model type is float, as in the reference's
 *           shipped RUN_32BIT build (swe_config.h:75-80).  Every expression
 *           keeps the reference's C types, so its promotions carry over: the
 *           coefficients dt/2.0 (eval_K) are computed in double and rounded
 *           to float as fType arguments, while update_H's (dt/6.0)*D and the
 *           sum into H are double, rounded to float on the store
 *           (rk4_rbffd_swe.c:86-121).
 */
#include "swe_oracle.h"

#include <math.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* swe_config.h:31-37 (DM_LIN_ID / VECT_LIN_ID) and :50-54 (SVM_LIN_ID). */
static inline long dm_lin(int i, int j, int nnbr, int sfdl, int S)
{
    if (sfdl)
        return (long)(i / S) * S * nnbr + (long)j * S + (i % S);
    return (long)i * nnbr + j;
}

static inline long vect_lin(int i, int c, int sfdl, int S)
{
    if (sfdl)
        return (long)(i / S) * S * 3 + (long)c * S + (i % S);
    return (long)i * 3 + c;
}

static inline long svm_lin(int pid, int sv, int cfdl, int S)
{
    if (cfdl)
        return 4L * pid + sv;
    return (long)(pid / S) * S * 4 + (long)sv * S + (pid % S);
}

/* One body, instantiated per (CFDL, SFDL) pair so that the layout branches
 * fold away in the node loop. */
#ifdef ORACLE_INNER_SIMD
#define ORACLE_J_PRAGMA _Pragma("omp simd reduction(+:dudx,dvdx,dwdx,dhdx,dudy,dvdy,dwdy,dhdy,dudz,dvdz,dwdz,dhdz,lu,lv,lw,lh)")
#else
#define ORACLE_J_PRAGMA
#endif

#define ORACLE_RHS_BODY(CF, SF)                                                 \
    const int S = m->simd, nnbr = m->Nnbr, pnnbr = m->padded_Nnbr;              \
    const int cps = m->compute_pid_s, n = m->compute_size;                      \
    const oracle_real gh0 = m->gh0;                                             \
    _Pragma("omp parallel for schedule(static)")                               \
    for (int i = 0; i < n; i++) {                                               \
        oracle_real dudx = 0.0, dvdx = 0.0, dwdx = 0.0, dhdx = 0.0;                  \
        oracle_real dudy = 0.0, dvdy = 0.0, dwdy = 0.0, dhdy = 0.0;                  \
        oracle_real dudz = 0.0, dvdz = 0.0, dwdz = 0.0, dhdz = 0.0;                  \
        oracle_real lu = 0.0, lv = 0.0, lw = 0.0, lh = 0.0;                          \
        ORACLE_J_PRAGMA                                                         \
        for (int j = 0; j < nnbr; j++) {                                        \
            const long d = dm_lin(i, j, pnnbr, SF, S);                          \
            const int q = m->idx[d];                                            \
            const oracle_real su =K[svm_lin(q, 0, CF, S)];                          \
            const oracle_real sv =K[svm_lin(q, 1, CF, S)];                          \
            const oracle_real sw =K[svm_lin(q, 2, CF, S)];                          \
            const oracle_real sh =K[svm_lin(q, 3, CF, S)];                          \
            const oracle_real wx = m->Dx[d], wy = m->Dy[d], wz = m->Dz[d]; \
            const oracle_real wl = m->L[d];                                     \
            dudx += wx * su; dvdx += wx * sv; dwdx += wx * sw; dhdx += wx * sh; \
            dudy += wy * su; dvdy += wy * sv; dwdy += wy * sw; dhdy += wy * sh; \
            dudz += wz * su; dvdz += wz * sv; dwdz += wz * sw; dhdz += wz * sh; \
            lu += wl * su;   lv += wl * sv;   lw += wl * sw;   lh += wl * sh;   \
        }                                                                       \
        const oracle_real u =K[svm_lin(i + cps, 0, CF, S)];                         \
        const oracle_real v =K[svm_lin(i + cps, 1, CF, S)];                         \
        const oracle_real w =K[svm_lin(i + cps, 2, CF, S)];                         \
        const oracle_real h =K[svm_lin(i + cps, 3, CF, S)];                         \
        const oracle_real fi =m->f[i], xi = m->x[i], yi = m->y[i], zi = m->z[i];    \
        const oracle_real ru =-((u * dudx) + (v * dudy) + (w * dudz)                \
                            + (fi * ((yi * w) - (zi * v))) + dhdx);             \
        const oracle_real rv =-((u * dvdx) + (v * dvdy) + (w * dvdz)                \
                            + (fi * ((zi * u) - (xi * w))) + dhdy);             \
        const oracle_real rw =-((u * dwdx) + (v * dwdy) + (w * dwdz)                \
                            + (fi * ((xi * v) - (yi * u))) + dhdz);             \
        const long cx = vect_lin(i, 0, SF, S), cy = vect_lin(i, 1, SF, S);      \
        const long cz = vect_lin(i, 2, SF, S);                                  \
        F[svm_lin(i, 0, CF, S)] = (m->p_u[cx] * ru) + (m->p_u[cy] * rv)         \
                                  + (m->p_u[cz] * rw) + lu;                     \
        F[svm_lin(i, 1, CF, S)] = (m->p_v[cx] * ru) + (m->p_v[cy] * rv)         \
                                  + (m->p_v[cz] * rw) + lv;                     \
        F[svm_lin(i, 2, CF, S)] = (m->p_w[cx] * ru) + (m->p_w[cy] * rv)         \
                                  + (m->p_w[cz] * rw) + lw;                     \
        F[svm_lin(i, 3, CF, S)] =                                               \
            -((u * (dhdx - m->gradghm[cx])) + (v * (dhdy - m->gradghm[cy]))     \
              + (w * (dhdz - m->gradghm[cz]))                                   \
              + ((h + gh0 - m->ghm[i]) * (dudx + dvdy + dwdz)))                 \
            + lh;                                                               \
    }

static void rhs_00(const oracle_psmd *m, const oracle_real *K, oracle_real *F) { ORACLE_RHS_BODY(0, 0) }
static void rhs_10(const oracle_psmd *m, const oracle_real *K, oracle_real *F) { ORACLE_RHS_BODY(1, 0) }
static void rhs_01(const oracle_psmd *m, const oracle_real *K, oracle_real *F) { ORACLE_RHS_BODY(0, 1) }
static void rhs_11(const oracle_psmd *m, const oracle_real *K, oracle_real *F) { ORACLE_RHS_BODY(1, 1) }

void oracle_eval_rhs(const oracle_psmd *m, const oracle_real *K, oracle_real *F,
                     oracle_timers *t)
{
    const double t0 = now_s();
    if (m->cfdl && m->sfdl)
        rhs_11(m, K, F);
    else if (m->cfdl)
        rhs_10(m, K, F);
    else if (m->sfdl)
        rhs_01(m, K, F);
    else
        rhs_00(m, K, F);
    if (t)
        t->t_eval_rhs += now_s() - t0;
}

/* eval_K, rk4_rbffd_swe.c:123-158: K = H + c*F over the compute range. */
static void eval_k(const oracle_psmd *m, const oracle_real *H, oracle_real *K,
                   const oracle_real *F, oracle_real c, oracle_timers *t)
{
    const double t0 = now_s();
    const long off = 4L * m->compute_pid_s, n4 = 4L * m->compute_size;
#pragma omp parallel for simd schedule(static)
    for (long i = 0; i < n4; i++)
        K[off + i] = H[off + i] + (c * F[i]);
    if (t)
        t->t_eval_K += now_s() - t0;
}

/* update_D, rk4_rbffd_swe.c:160-186: D += coeff*F. */
static void update_d(const oracle_psmd *m, oracle_real *D, const oracle_real *F,
                     oracle_real coeff, oracle_timers *t)
{
    const double t0 = now_s();
    const long n4 = 4L * m->compute_size;
#pragma omp parallel for simd schedule(static)
    for (long i = 0; i < n4; i++)
        D[i] += (coeff * F[i]);
    if (t)
        t->t_update_D += now_s() - t0;
}

/* update_H, rk4_rbffd_swe.c:86-121: H += (dt/6)*D. */
static void update_h(const oracle_psmd *m, oracle_real *H, const oracle_real *D,
                     oracle_real dt, oracle_timers *t)
{
    const double t0 = now_s();
    const long off = 4L * m->compute_pid_s, n4 = 4L * m->compute_size;
#pragma omp parallel for simd schedule(static)
    for (long i = 0; i < n4; i++)
        H[off + i] += ((dt / 6.0) * D[i]);
    if (t)
        t->t_update_H += now_s() - t0;
}

void oracle_rk_substep(const oracle_psmd *m, oracle_real *H, oracle_real *K, oracle_real *F,
                       oracle_real *D, int substep_id, oracle_timers *t)
{
    const oracle_real dt = m->dt;
    switch (substep_id) {
    case 0:
        oracle_eval_rhs(m, H, F, t);
        memcpy(D, F, sizeof(oracle_real) * 4L * m->compute_size);
        eval_k(m, H, K, F, dt / 2.0, t);
        break;
    case 1:
        oracle_eval_rhs(m, K, F, t);
        update_d(m, D, F, 2.0, t);
        eval_k(m, H, K, F, dt / 2.0, t);
        break;
    case 2:
        oracle_eval_rhs(m, K, F, t);
        update_d(m, D, F, 2.0, t);
        eval_k(m, H, K, F, dt, t);
        break;
    case 3:
        oracle_eval_rhs(m, K, F, t);
        update_d(m, D, F, 1.0, t);
        update_h(m, H, D, dt, t);
        break;
    default:
        break;
    }
}

void oracle_run(const oracle_psmd *m, oracle_real *H, oracle_real *K, oracle_real *F,
                oracle_real *D, int nsteps, oracle_timers *t)
{
    for (int step = 1; step <= nsteps; step++)
        for (int s = 0; s < 4; s++)
            oracle_rk_substep(m, H, K, F, D, s, t);
}

int oracle_verify(const oracle_psmd *m, const oracle_real *H1, const oracle_real *H2,
                  double tol)
{
    int count = 0;
    const int lo = m->compute_pid_s, hi = m->compute_pid_s + m->compute_size;
    for (int i = lo; i < hi; i++) {
        for (int sv = 0; sv < 4; sv++) {
            const long k = svm_lin(i, sv, m->cfdl, m->simd);
            oracle_real diff = H1[k] - H2[k];
            diff = diff < 0 ? -diff : diff;
            if (diff > tol || diff != diff)
                count++;
        }
    }
    return count;
}

int oracle_flavour_inner_simd(void)
{
#ifdef ORACLE_INNER_SIMD
    return 1;
#else
    return 0;
#endif
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int oracle_real_size(void) { return (int)sizeof(oracle_real); }
