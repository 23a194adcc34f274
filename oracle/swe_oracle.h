/*
 * CPU oracle for the RBF-FD shallow-water RK4 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU path (swe_code/main/rk4_rbffd_swe.c, swe_code/main/main.c:439-477)
 * used as the parity checker for the HIP kernels and as the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (libswe_hip.so and the C host)
 * never links or calls anything here.
 *
 * Pinning of each function (what proves the restatement right; the tests
 * are tests/test_reference_parity.py, tests/test_golden.py and
 * tests/test_oracle_pinning.py):
 *   oracle_eval_rhs     - bitwise equal to the reference build (oracle/_ref)
 *                         on the shipped 2562 input and generated meshes;
 *                         against dense stencil operators on a 162-node mesh
 *                         (zero-velocity and zero-depth states), the
 *                         constant-state closed form and tangency.
 *   oracle_rk_substep   - bitwise equal to the reference build after 1/10/100
 *                         steps; one step of the linear system dH/dt = L H
 *                         equals the textbook RK4 polynomial.
 *   oracle_verify       - known-answer H_100 of the shipped 2562 input
 *                         (count == 0 at 1e-9) and planted-error counts.
 */
#ifndef SWE_ORACLE_H
#define SWE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Model precision (the reference's fType, swe_config.h:70-84): double, or
 * float in the FP32 build of this file (ORACLE_FP32, liboracle32.so), which
 * follows the reference's RUN_32BIT build, C promotions included. */
#ifdef ORACLE_FP32
typedef float oracle_real;
#else
typedef double oracle_real;
#endif

/* Local patch static model data in the reference host layout
 * (swe_code/main/include/swe_config.h:187-233, serial single patch).
 * Index macros (swe_config.h:21-54) are evaluated with `cfdl`, `sfdl`
 * and `simd` (= SIMD_LENGTH) given at run time. */
typedef struct oracle_psmd {
    int compute_size;   /* padded node count (serial: == patch_size)       */
    int patch_size;
    int compute_pid_s;  /* 0 in the serial case (init_patches.c:287-303)   */
    int Nnbr;           /* real stencil width (loop bound, rk4:285)        */
    int padded_Nnbr;    /* row stride of idx/Dx/Dy/Dz/L                    */
    int cfdl, sfdl, simd;
    oracle_real gh0, dt;
    const oracle_real *x, *y, *z, *f, *ghm;
    const oracle_real *p_u, *p_v, *p_w, *gradghm;
    const int *idx;
    const oracle_real *Dx, *Dy, *Dz, *L;
} oracle_psmd;

/* Per-phase wall-clock accumulators, seconds (timing_struct,
 * swe_config.h:262-279, one attempt). */
typedef struct oracle_timers {
    double t_eval_rhs, t_eval_K, t_update_D, t_update_H;
} oracle_timers;

/* eval_RHS_swe (rk4_rbffd_swe.c:189-362): F = RHS(K). */
void oracle_eval_rhs(const oracle_psmd *m, const oracle_real *K, oracle_real *F,
                     oracle_timers *t);

/* RK_substep (rk4_rbffd_swe.c:30-72) with eval_K/update_D/update_H/
 * copy_fp_arr (:74-186). */
void oracle_rk_substep(const oracle_psmd *m, oracle_real *H, oracle_real *K, oracle_real *F,
                       oracle_real *D, int substep_id, oracle_timers *t);

/* The time loop of main.c:336-358 for one attempt: H advanced nsteps.
 * Scratch K (patch_size*4), F and D (compute_size*4) are caller-owned. */
void oracle_run(const oracle_psmd *m, oracle_real *H, oracle_real *K, oracle_real *F,
                oracle_real *D, int nsteps, oracle_timers *t);

/* verify_output (main.c:439-477): number of entries over the compute range
 * whose |H1-H2| exceeds tol, or that are NaN. */
int oracle_verify(const oracle_psmd *m, const oracle_real *H1, const oracle_real *H2,
                  double tol);

/* 1 if built with the reference's inner-SIMD reduction (RHS_INNER_SIMD),
 * i.e. the fast timed flavour; 0 for the strict sequential flavour. */
int oracle_flavour_inner_simd(void);
int oracle_num_threads(void);
/* sizeof(oracle_real): 8, or 4 in the FP32 build. */
int oracle_real_size(void);

#ifdef __cplusplus
}
#endif
#endif
