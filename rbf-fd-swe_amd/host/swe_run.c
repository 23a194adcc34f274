/*
 * swe_run: the reference driver (swe_code/main/main.c) on the HIP path.
 *
 * Same run-time parameters (swe_code/main/runtime_params.c:12-40):
 *   SWE_INPUT_FILE, SWE_USE_NETCDF, SWE_NUM_COMPUTE_UNITS, SWE_NSTEPS,
 *   SWE_NATTEMPTS, SWE_USE_RCM, SWE_USE_OCL (accepted; the compute back-end
 *   is always libswe_hip)
 * and run-time replacements of the reference's compile-time switches:
 *   SWE_LAYOUT       default | cfdl | sfdl | cfdl_sfdl   (USE_CFDL/USE_SFDL)
 *   SWE_SIMD_LENGTH  SIMD_LENGTH (default 4)
 *   SWE_HIP_MODE     substep (RK_substep drop-in, 4 calls per step, as
 *                    main.c:336-358) | steps (fused, graph-captured)
 *   SWE_HIP_EXACT    1 = bitwise reference arithmetic (default 0)
 *   SWE_MEMBERS      ensemble members advanced together (default 1)
 *   SWE_OUTPUT_FILE, SWE_HIST_INT   history output, the reference's USE_HIST
 *                    parameters (runtime_params.c:24-27, a stub there): the
 *                    state of member 0 every SWE_HIST_INT steps of the first
 *                    attempt, in the file node order (RCM and padding undone):
 *                    int32 N, then per snapshot int32 step, f64 time,
 *                    f64 H[N][4] (u, v, w, h)
 *   SWE_RESTART_FILE, SWE_RESTART_STEP   restart: every attempt starts from
 *                    the snapshot of that step (default: the file's last) of
 *                    a history file, the run ends at SWE_NSTEPS as if it had
 *                    never stopped (padding rows repeat the last node, as the
 *                    reference pads H_init, matrix_transformations.c:20)
 * Init pipeline main.c:94-190, attempt loop :294-392, verify_output
 * :439-477 (every member against H_100), profiling report profiling.c.
 * SWE_USE_NETCDF=1 reads a NetCDF classic input (input.c:14-186, the
 * reference's USE_NCIO build) with the C host's own reader.
 * Errors: a missing parameter prints and exits 1, an unreadable input
 * exits 2 (input.c CHECK_OPEN_ERR, CHECK_ERR), a HIP error prints the
 * library's message and exits -1 (device_setup.c checkError).
 */
#include "swe_hip.h"
#include "swe_host.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const char *need_env(const char *name)
{
    const char *v = getenv(name);
    if (!v) {
        fprintf(stderr, "swe_run: environment variable %s is not set\n", name);
        exit(1);
    }
    return v;
}

static int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    return v ? atoi(v) : dflt;
}

/* One history snapshot: host layout -> file node order (inverse RCM). */
static int write_snapshot(FILE *fp, const swe_problem *prob, const double *H, int step)
{
    const int N = prob->g.Nnodes;
    double *row = (double *)malloc(sizeof(double) * 4 * (size_t)N);
    if (!row)
        return -1;
    for (int i = 0; i < N; i++) {
        const int file_node = prob->mapping ? prob->mapping[i] : i;
        for (int v = 0; v < 4; v++)
            row[4 * (size_t)file_node + v] = H[swe_svm_lin_id(&prob->layout, i, v)];
    }
    const int32_t st = step;
    const double t = step * prob->g.dt;
    int ok = fwrite(&st, sizeof(st), 1, fp) == 1 && fwrite(&t, sizeof(t), 1, fp) == 1
             && fwrite(row, sizeof(double), 4 * (size_t)N, fp) == 4 * (size_t)N;
    free(row);
    return ok ? 0 : -1;
}

/* Snapshot `step` (< 0: the last) of a history file -> host layout H (the
 * inverse of write_snapshot; padding rows copy the last node).  Returns 0,
 * or -1 when the file does not match the problem or has no such step. */
static int read_snapshot(const char *path, const swe_problem *prob, int step, double *H,
                         int *got_step)
{
    const int N = prob->g.Nnodes;
    FILE *fp = fopen(path, "rb");
    if (!fp)
        return -1;
    int32_t n32 = 0;
    const size_t rec = sizeof(int32_t) + sizeof(double) + sizeof(double) * 4 * (size_t)N;
    double *row = (double *)malloc(sizeof(double) * 4 * (size_t)N);
    int found = 0;
    if (row && fread(&n32, sizeof(n32), 1, fp) == 1 && n32 == N) {
        int32_t st;
        double t;
        while (fread(&st, sizeof(st), 1, fp) == 1 && fread(&t, sizeof(t), 1, fp) == 1) {
            if (step >= 0 && st != step) {
                if (fseek(fp, (long)(rec - sizeof(st) - sizeof(t)), SEEK_CUR) != 0)
                    break;
                continue;
            }
            if (fread(row, sizeof(double), 4 * (size_t)N, fp) != 4 * (size_t)N)
                break;
            found = 1;
            *got_step = st;
            if (step >= 0)
                break;
        }
    }
    fclose(fp);
    if (found) {
        for (int i = 0; i < prob->patch.patch_size; i++) {
            const int node = i < N ? i : N - 1;
            const int file_node = prob->mapping ? prob->mapping[node] : node;
            for (int v = 0; v < 4; v++)
                H[swe_svm_lin_id(&prob->layout, i, v)] = row[4 * (size_t)file_node + v];
        }
    }
    free(row);
    return found ? 0 : -1;
}

static void hip_check(int rc, const char *what)
{
    if (rc != SWE_HIP_OK) {
        printf("Error in %s: %s: %s\n", what, swe_hip_strerror(rc), swe_hip_last_error());
        exit(-1);
    }
}

int main(void)
{
    const double t_start = swe_get_time();
    const char *input = need_env("SWE_INPUT_FILE");
    const int use_netcdf = atoi(need_env("SWE_USE_NETCDF"));
    (void)need_env("SWE_NUM_COMPUTE_UNITS");   /* accepted, one GPU */
    const int nsteps = atoi(need_env("SWE_NSTEPS"));
    const int nattempts = atoi(need_env("SWE_NATTEMPTS"));
    const int use_rcm = atoi(need_env("SWE_USE_RCM"));
    const char *layout_name = getenv("SWE_LAYOUT") ? getenv("SWE_LAYOUT") : "default";
    const char *mode = getenv("SWE_HIP_MODE") ? getenv("SWE_HIP_MODE") : "substep";
    const int exact = env_int("SWE_HIP_EXACT", 0);
    const int members = env_int("SWE_MEMBERS", 1);
    const char *hist_path = getenv("SWE_OUTPUT_FILE");
    const int hist_int = env_int("SWE_HIST_INT", 0);
    const char *restart_path = getenv("SWE_RESTART_FILE");
    const int restart_want = env_int("SWE_RESTART_STEP", -1);
    swe_layout lay = {0, 0, env_int("SWE_SIMD_LENGTH", 4)};
    if (!strcmp(layout_name, "cfdl")) lay.cfdl = 1;
    else if (!strcmp(layout_name, "sfdl")) lay.sfdl = 1;
    else if (!strcmp(layout_name, "cfdl_sfdl")) lay.cfdl = lay.sfdl = 1;
    else if (strcmp(layout_name, "default")) {
        fprintf(stderr, "swe_run: unknown SWE_LAYOUT %s\n", layout_name);
        return 1;
    }
    if (nsteps < 0 || nattempts < 1 || members < 1) {
        fprintf(stderr, "swe_run: bad SWE_NSTEPS/SWE_NATTEMPTS/SWE_MEMBERS\n");
        return 1;
    }
    printf("\n\n============================= Runtime Configurations =============================\n\n");
    printf("Input from file: yes\n\tInput File Path: %s\n\n", input);
    printf("I/O Method: %s\n", use_netcdf == 1 ? "NETCDF" : "BINARY");
    printf("Simulation length (timesteps): %d\n", nsteps);
    printf("Number of attempts: %d\n", nattempts);
    printf("USE RCM Node Ordering: %s\n", use_rcm == 1 ? "yes" : "no");
    printf("Layout: %s (SIMD_LENGTH %d), HIP mode: %s%s, members: %d\n", layout_name, lay.simd,
           mode, exact ? " (exact)" : "", members);

    swe_timers timer;
    if (swe_timers_init(&timer, nattempts)) {
        fprintf(stderr, "swe_run: timer allocation failed\n");
        return 1;
    }
    swe_problem prob;
    int err = swe_problem_load_fmt(input, use_rcm == 1, &lay, 0, use_netcdf == 1, &prob);
    if (err == SWE_HOST_EOPEN) {
        printf("Error opening file %s\n", input);
        return 2;
    }
    if (err && use_netcdf == 1 && err != SWE_HOST_ENOMEM && err != SWE_HOST_EGRAPH) {
        /* the reference's NetCDF reader exits 2 on any library error (CHECK_ERR) */
        printf("Error in input.c: %s: host error %d\n", input, err);
        return 2;
    }
    if (err) {
        fprintf(stderr, "swe_run: cannot prepare %s (host error %d)\n", input, err);
        return 1;
    }
    timer.t_init = swe_get_time() - t_start;

    const swe_patch *p = &prob.patch;
    swe_hip_static st = {p->compute_size, p->patch_size, p->compute_pid_s, p->Nnbr,
                         p->padded_Nnbr, lay.cfdl, lay.sfdl, lay.simd, p->gh0, p->dt,
                         p->x, p->y, p->z, p->f, p->ghm, p->p_u, p->p_v, p->p_w,
                         p->gradghm, p->idx, p->Dx, p->Dy, p->Dz, p->L};
    const double t_dev = swe_get_time();
    swe_hip_ctx *ctx = NULL;
    hip_check(swe_hip_create(&st, members, exact ? SWE_HIP_EXACT : 0u, 0, &ctx), "swe_hip_create");
    timer.t_dev = swe_get_time() - t_dev;
    const int substep_mode = strcmp(mode, "steps") != 0;
    double *H = (double *)malloc(sizeof(double) * 4 * (size_t)p->patch_size);
    if (!H)
        return 1;
    /* the initial state of every attempt: H_init, or a restart snapshot */
    const double *H_start = prob.H_init;
    double *H_restart = NULL;
    int start_step = 0;
    if (restart_path) {
        H_restart = (double *)malloc(sizeof(double) * 4 * (size_t)p->patch_size);
        if (!H_restart || read_snapshot(restart_path, &prob, restart_want, H_restart, &start_step)) {
            printf("Error reading restart snapshot %d of %s\n", restart_want, restart_path);
            return 2;
        }
        if (start_step > nsteps) {
            fprintf(stderr, "swe_run: restart step %d beyond SWE_NSTEPS %d\n", start_step, nsteps);
            return 1;
        }
        H_start = H_restart;
        printf("Restart from step %d of %s\n", start_step, restart_path);
    }
    FILE *hist = NULL;
    if (hist_path && hist_int > 0) {
        hist = fopen(hist_path, "wb");
        if (!hist) {
            printf("Error opening file %s\n", hist_path);
            return 2;
        }
        const int32_t n32 = prob.g.Nnodes;
        fwrite(&n32, sizeof(n32), 1, hist);
        printf("Write history to output file: yes\n\tOutput File Path: %s\n\tHistory Interval "
               "(timesteps): %d\n", hist_path, hist_int);
    }
    for (int attempt = 0; attempt < nattempts; attempt++) {
        for (int m = 0; m < members; m++)
            hip_check(swe_hip_set_state(ctx, m, H_start), "swe_hip_set_state");
        hip_check(swe_hip_set_timing(ctx, substep_mode), "swe_hip_set_timing");
        timer.attempt = attempt;
        const double t0 = swe_get_time();
        const int chunk = (hist && attempt == 0) ? hist_int : nsteps;
        for (int done = start_step; done < nsteps;) {
            /* history snapshots stay on multiples of SWE_HIST_INT */
            const int to_mark = chunk - done % chunk;
            const int k = nsteps - done < to_mark ? nsteps - done : to_mark;
            if (substep_mode) {
                for (int i = 0; i < k; i++)
                    for (int rk = 0; rk < 4; rk++)
                        hip_check(swe_hip_rk_substep(ctx, rk), "swe_hip_rk_substep");
            } else {
                hip_check(swe_hip_rk_steps(ctx, k), "swe_hip_rk_steps");
            }
            done += k;
            if (hist && attempt == 0 && done % hist_int == 0) {
                hip_check(swe_hip_get_state(ctx, 0, H), "swe_hip_get_state");
                if (write_snapshot(hist, &prob, H, done)) {
                    fprintf(stderr, "swe_run: history write failed\n");
                    return 1;
                }
            }
        }
        hip_check(swe_hip_synchronize(ctx), "swe_hip_synchronize");
        timer.t_main[attempt] = swe_get_time() - t0;
        swe_hip_timers ht;
        hip_check(swe_hip_get_timers(ctx, &ht, 1), "swe_hip_get_timers");
        timer.t_eval_rhs[attempt] = ht.t_eval_rhs;
        timer.t_eval_K[attempt] = ht.t_eval_K;
        timer.t_update_D[attempt] = ht.t_update_D;
        timer.t_update_H[attempt] = ht.t_update_H;
        int worst = 0;
        for (int m = 0; m < members; m++) {
            hip_check(swe_hip_get_state(ctx, m, H), "swe_hip_get_state");
            const int c = swe_verify_output(p, &lay, H, prob.H_100, 1.0e-9);
            if (c > worst)
                worst = c;
        }
        printf("Attempt %d:\t", attempt + 1);
        if (worst != 0)
            printf("\n\nVERIFICATION FAILED: There were a total of %d values that exceeded an error of %.1e\n\n",
                   worst, 1.0e-9);
        else
            printf("VERIFICATION SUCCEEDED\n");
        fflush(stdout);
    }
    if (hist)
        fclose(hist);
    free(H_restart);
    swe_process_profiling(&timer, nsteps - start_step, 1, NULL);
    free(H);
    swe_hip_destroy(ctx);
    swe_problem_free(&prob);
    swe_timers_free(&timer);
    return 0;
}
