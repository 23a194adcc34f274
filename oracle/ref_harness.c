/*
 * Harness around the reference's own C sources -- TEST INFRASTRUCTURE ONLY.
 *
 * Built by oracle/build_ref.sh together with the reference's source files,
 * compiled where they lie under the reference tree, into oracle/_ref/.  It
 * replaces only the reference's main() (swe_code/main/main.c:45), for either
 * model precision (the header's fType), and it owns the
 * globals main.c defines (main.c:34-41) and runs the same serial
 * initialisation sequence (main.c:94-190), then exposes the reference's
 * RK_substep / eval_RHS_swe and the prepared arrays to the tests through a
 * flat C ABI, so the oracle and the product host can be compared with the
 * reference bit for bit.
 */
#include <swe_config.h>
#include <input.h>
#include <rcm.h>
#include <reorder_nodes.h>
#include <layout.h>
#include <matrix_transformations.h>
#include <init_patches.h>
#include <rk4_rbffd_swe.h>
#include <profiling.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

sim_params_struct sim_params;
PSMD_struct *LPSMD;
timing_struct local_timer;
int mpi_rank = 0;
int mpi_size = 1;

void eval_RHS_swe(fType *K, fType *F); /* defined in rk4_rbffd_swe.c:189 */

static fType *g_H_init, *g_H_100;
static int *g_mapping;
static int g_Nnodes, g_padded_Nnodes;

/* sizeof(fType): 8 for the FP64 header variant, 4 for the reference's
 * shipped RUN_32BIT header; every array below is of fType. */
int ref_ftype_size(void) { return (int)sizeof(fType); }

/* main.c:94-190 for mpi_size == 1 and binary input. Returns 0 on success. */
int ref_setup(const char *path, int use_rcm, int nattempts)
{
    memset(&sim_params, 0, sizeof(sim_params));
    strncpy(sim_params.inputFile, path, MAX_FILE_PATH_SIZE - 1);
    sim_params.nattempts = nattempts;
    sim_params.USE_RCM = use_rcm;
    local_timer = init_timer(nattempts);
    LPSMD = (PSMD_struct *)malloc(sizeof(PSMD_struct));

    GSMD_struct G = get_GSMD_bin(sim_params.inputFile);
    fType *Hg = get_ICs_bin(sim_params.inputFile);
    fType *H100g = get_FCs_bin(sim_params.inputFile);
    g_Nnodes = G.Nnodes;
    free(g_mapping);
    g_mapping = NULL;
    if (use_rcm) {
        g_mapping = rcm_mapping(G);
        G = reorder_GSMD_struct(g_mapping, G);
        Hg = reorder_2D_fp_arr(Hg, g_mapping, G.Nnodes, 4);
        H100g = reorder_2D_fp_arr(H100g, g_mapping, G.Nnodes, 4);
    }
    G = pad_GSMD_data(G);
    Hg = pad_2D_fp_arr(Hg, G.Nnodes, G.padded_Nnodes, 4, 4);
    H100g = pad_2D_fp_arr(H100g, G.Nnodes, G.padded_Nnodes, 4, 4);
#ifdef USE_SFDL
    G = convert2_SFDL(G);
#endif
#ifndef USE_CFDL
    Hg = transpose_2D_fp_arr(Hg, G.padded_Nnodes, 4);
    Hg = tile_2D_fp_arr(Hg, 4, G.padded_Nnodes, SIMD_LENGTH);
    H100g = transpose_2D_fp_arr(H100g, G.padded_Nnodes, 4);
    H100g = tile_2D_fp_arr(H100g, 4, G.padded_Nnodes, SIMD_LENGTH);
#endif
    g_padded_Nnodes = G.padded_Nnodes;
    patch_struct *GPD = get_GPD(G);
    init_LPSMD(GPD, G);
    g_H_init = get_patch_ICs(Hg, GPD);
    g_H_100 = get_patch_ICs(H100g, GPD);
    return 0;
}

/* Sizes and layout of the prepared patch:
 * [Nnodes, padded_Nnodes, compute_size, patch_size, compute_pid_s, Nnbr,
 *  padded_Nnbr, SIMD_LENGTH, cfdl, sfdl] */
void ref_info(int *out)
{
    out[0] = g_Nnodes;
    out[1] = g_padded_Nnodes;
    out[2] = LPSMD->compute_size;
    out[3] = LPSMD->patch_size;
    out[4] = LPSMD->compute_pid_s;
    out[5] = LPSMD->Nnbr;
    out[6] = LPSMD->padded_Nnbr;
    out[7] = SIMD_LENGTH;
#ifdef USE_CFDL
    out[8] = 1;
#else
    out[8] = 0;
#endif
#ifdef USE_SFDL
    out[9] = 1;
#else
    out[9] = 0;
#endif
}

void ref_scalars(double *out)
{
    out[0] = (double)LPSMD->gh0;
    out[1] = (double)LPSMD->dt;
}

/* Pointers to the 13 prepared arrays, in the order
 * x y z f ghm p_u p_v p_w gradghm Dx Dy Dz L, then idx separately. */
void ref_arrays(const fType **fp, const int **idx)
{
    fp[0] = LPSMD->x;   fp[1] = LPSMD->y;   fp[2] = LPSMD->z;
    fp[3] = LPSMD->f;   fp[4] = LPSMD->ghm;
    fp[5] = LPSMD->p_u; fp[6] = LPSMD->p_v; fp[7] = LPSMD->p_w;
    fp[8] = LPSMD->gradghm;
    fp[9] = LPSMD->Dx;  fp[10] = LPSMD->Dy; fp[11] = LPSMD->Dz;
    fp[12] = LPSMD->L;
    *idx = LPSMD->idx;
}

const fType *ref_H_init(void) { return g_H_init; }
const fType *ref_H_100(void) { return g_H_100; }
const int *ref_mapping(void) { return g_mapping; }

void ref_rk_substep(fType *H, fType *K, fType *F, fType *D, int s)
{
    local_timer.attempt = 0;
    RK_substep(H, K, F, D, s);
}

void ref_eval_rhs(fType *K, fType *F)
{
    local_timer.attempt = 0;
    eval_RHS_swe(K, F);
}

/* rcm_mapping (rcm.c:9) of an arbitrary stencil graph [N][W] (0-based). */
void ref_rcm_of(int N, int W, const int *idx, int *out)
{
    GSMD_struct G;
    memset(&G, 0, sizeof(G));
    G.Nnodes = N;
    G.Nnbr = W;
    G.idx = (int *)idx;
    int *m = rcm_mapping(G);
    memcpy(out, m, sizeof(int) * N);
    free(m);
}

/* process_profiling_results (profiling.c:54-172) on caller-given per-attempt
 * phase times (seconds): its report, printed to stdout, is captured into
 * buf (at most cap - 1 bytes, NUL-terminated).  Returns the report length,
 * or -1 when stdout cannot be redirected. */
int ref_profiling(int nattempts, int nsteps, const double *t_main, const double *t_eval_rhs,
                  const double *t_eval_K, const double *t_update_D, const double *t_update_H,
                  char *buf, int cap)
{
    sim_params.nattempts = nattempts;
    sim_params.nsteps = nsteps;
    sim_params.OCL = 0;
    local_timer = init_timer(nattempts);
    memcpy(local_timer.t_main, t_main, sizeof(double) * nattempts);
    memcpy(local_timer.t_eval_rhs, t_eval_rhs, sizeof(double) * nattempts);
    memcpy(local_timer.t_eval_K, t_eval_K, sizeof(double) * nattempts);
    memcpy(local_timer.t_update_D, t_update_D, sizeof(double) * nattempts);
    memcpy(local_timer.t_update_H, t_update_H, sizeof(double) * nattempts);
    FILE *tmp = tmpfile();
    if (!tmp)
        return -1;
    fflush(stdout);
    const int saved = dup(1);
    if (saved < 0 || dup2(fileno(tmp), 1) < 0) {
        fclose(tmp);
        return -1;
    }
    process_profiling_results();
    fflush(stdout);
    dup2(saved, 1);
    close(saved);
    rewind(tmp);
    const size_t got = fread(buf, 1, (size_t)(cap - 1), tmp);
    buf[got] = '\0';
    fclose(tmp);
    return (int)got;
}
