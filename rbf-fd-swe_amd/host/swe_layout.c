/*
 * Host data layouts: index functions, padding, the SFDL transpose+tile, the
 * serial single-patch set-up and the init pipeline of the reference driver.
 *
 *   index functions  swe_code/main/include/swe_config.h:21-54
 *   padding          swe_code/layout/layout.c:10-45,
 *                    swe_code/layout/matrix_transformations.c:9-59
 *                    (new rows/columns replicate the last real one)
 *   SFDL             swe_code/layout/layout.c:48-78 (transpose, then tiles
 *                    of SIMD_LENGTH columns, matrix_transformations.c:61-136)
 *   serial patch     swe_code/mpi/init_patches.c:21-92 with mpi_size == 1,
 *                    :172-195, :287-303, :308-341
 *   init pipeline    swe_code/main/main.c:94-190
 */
#include "swe_host.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int swe_roundup(const swe_layout *l, int n) { return (((n - 1) / l->simd) + 1) * l->simd; }

long swe_dm_lin_id(const swe_layout *l, int node, int nbr, int nnbr)
{
    const int S = l->simd;
    if (l->sfdl)
        return (long)(node / S) * S * nnbr + (long)nbr * S + node % S;
    return (long)node * nnbr + nbr;
}

long swe_vect_lin_id(const swe_layout *l, int node, int comp)
{
    const int S = l->simd;
    if (l->sfdl)
        return (long)(node / S) * S * 3 + (long)comp * S + node % S;
    return (long)node * 3 + comp;
}

long swe_svm_lin_id(const swe_layout *l, int pid, int sv)
{
    const int S = l->simd;
    if (l->cfdl)
        return 4L * pid + sv;
    return (long)(pid / S) * S * 4 + (long)sv * S + pid % S;
}

/* pad_2D_fp_arr: rows >= dim1 copy row dim1-1, columns >= dim2 copy column
 * dim2-1.  Returns var itself when nothing changes, else a new array (the
 * old one is freed). */
double *swe_pad_2d(double *var, int dim1, int pdim1, int dim2, int pdim2)
{
    if (dim1 == pdim1 && dim2 == pdim2)
        return var;
    double *out = (double *)malloc(sizeof(double) * (size_t)pdim1 * pdim2);
    if (!out)
        return NULL;
    for (int i = 0; i < pdim1; i++) {
        const int si = i < dim1 ? i : dim1 - 1;
        for (int j = 0; j < pdim2; j++) {
            const int sj = j < dim2 ? j : dim2 - 1;
            out[(size_t)i * pdim2 + j] = var[(size_t)si * dim2 + sj];
        }
    }
    free(var);
    return out;
}

static int *pad_2d_int(int *var, int dim1, int pdim1, int dim2, int pdim2)
{
    if (dim1 == pdim1 && dim2 == pdim2)
        return var;
    int *out = (int *)malloc(sizeof(int) * (size_t)pdim1 * pdim2);
    if (!out)
        return NULL;
    for (int i = 0; i < pdim1; i++) {
        const int si = i < dim1 ? i : dim1 - 1;
        for (int j = 0; j < pdim2; j++) {
            const int sj = j < dim2 ? j : dim2 - 1;
            out[(size_t)i * pdim2 + j] = var[(size_t)si * dim2 + sj];
        }
    }
    free(var);
    return out;
}

/* [dim1][dim2] -> [dim2][dim1] */
double *swe_transpose_2d(double *a, int dim1, int dim2)
{
    double *out = (double *)malloc(sizeof(double) * (size_t)dim1 * dim2);
    if (!out)
        return NULL;
    for (int r = 0; r < dim1; r++)
        for (int c = 0; c < dim2; c++)
            out[(size_t)c * dim1 + r] = a[(size_t)r * dim2 + c];
    free(a);
    return out;
}

static int *transpose_2d_int(int *a, int dim1, int dim2)
{
    int *out = (int *)malloc(sizeof(int) * (size_t)dim1 * dim2);
    if (!out)
        return NULL;
    for (int r = 0; r < dim1; r++)
        for (int c = 0; c < dim2; c++)
            out[(size_t)c * dim1 + r] = a[(size_t)r * dim2 + c];
    free(a);
    return out;
}

/* [dim1][dim2] -> [dim2/T][dim1][T]: column blocks of width T made
 * contiguous (dim2 must be a multiple of T). */
double *swe_tile_2d(double *a, int dim1, int dim2, int T)
{
    double *out = (double *)malloc(sizeof(double) * (size_t)dim1 * dim2);
    if (!out)
        return NULL;
    const int ntiles = dim2 / T;
    for (int t = 0; t < ntiles; t++)
        for (int r = 0; r < dim1; r++)
            memcpy(&out[((size_t)t * dim1 + r) * T], &a[(size_t)r * dim2 + (size_t)t * T],
                   sizeof(double) * T);
    free(a);
    return out;
}

static int *tile_2d_int(int *a, int dim1, int dim2, int T)
{
    int *out = (int *)malloc(sizeof(int) * (size_t)dim1 * dim2);
    if (!out)
        return NULL;
    const int ntiles = dim2 / T;
    for (int t = 0; t < ntiles; t++)
        for (int r = 0; r < dim1; r++)
            memcpy(&out[((size_t)t * dim1 + r) * T], &a[(size_t)r * dim2 + (size_t)t * T],
                   sizeof(int) * T);
    free(a);
    return out;
}

int swe_pad_gsmd(swe_gsmd *g, const swe_layout *l)
{
    if (l->simd <= 0)
        return SWE_HOST_EARG;
    const int N = g->Nnodes, Np = swe_roundup(l, g->Nnodes);
    g->padded_Nnodes = Np;
    double **one[5] = {&g->x, &g->y, &g->z, &g->f, &g->ghm};
    for (int k = 0; k < 5; k++)
        if (!(*one[k] = swe_pad_2d(*one[k], N, Np, 1, 1)))
            return SWE_HOST_ENOMEM;
    double **three[4] = {&g->p_u, &g->p_v, &g->p_w, &g->gradghm};
    for (int k = 0; k < 4; k++)
        if (!(*three[k] = swe_pad_2d(*three[k], N, Np, 3, 3)))
            return SWE_HOST_ENOMEM;
    g->padded_Nnbr = l->sfdl ? g->Nnbr : swe_roundup(l, g->Nnbr);
    double **dm[4] = {&g->Dx, &g->Dy, &g->Dz, &g->L};
    for (int k = 0; k < 4; k++)
        if (!(*dm[k] = swe_pad_2d(*dm[k], N, Np, g->Nnbr, g->padded_Nnbr)))
            return SWE_HOST_ENOMEM;
    if (!(g->idx = pad_2d_int(g->idx, N, Np, g->Nnbr, g->padded_Nnbr)))
        return SWE_HOST_ENOMEM;
    return 0;
}

int swe_convert_sfdl(swe_gsmd *g, const swe_layout *l)
{
    const int Np = g->padded_Nnodes, W = g->padded_Nnbr, S = l->simd;
    if (Np % S)
        return SWE_HOST_EARG;
    double **three[4] = {&g->p_u, &g->p_v, &g->p_w, &g->gradghm};
    for (int k = 0; k < 4; k++) {
        if (!(*three[k] = swe_transpose_2d(*three[k], Np, 3)))
            return SWE_HOST_ENOMEM;
        if (!(*three[k] = swe_tile_2d(*three[k], 3, Np, S)))
            return SWE_HOST_ENOMEM;
    }
    double **dm[4] = {&g->Dx, &g->Dy, &g->Dz, &g->L};
    for (int k = 0; k < 4; k++) {
        if (!(*dm[k] = swe_transpose_2d(*dm[k], Np, W)))
            return SWE_HOST_ENOMEM;
        if (!(*dm[k] = swe_tile_2d(*dm[k], W, Np, S)))
            return SWE_HOST_ENOMEM;
    }
    if (!(g->idx = transpose_2d_int(g->idx, Np, W)))
        return SWE_HOST_ENOMEM;
    if (!(g->idx = tile_2d_int(g->idx, W, Np, S)))
        return SWE_HOST_ENOMEM;
    return 0;
}

int swe_state_to_layout(double **H, int Np, const swe_layout *l)
{
    if (l->cfdl)
        return 0;
    if (!(*H = swe_transpose_2d(*H, Np, 4)))
        return SWE_HOST_ENOMEM;
    if (!(*H = swe_tile_2d(*H, 4, Np, l->simd)))
        return SWE_HOST_ENOMEM;
    return 0;
}

/* get_GPD + init_LPSMD for one patch: the compute range is every padded node
 * and the halo covers [ROUNDDOWN(min nbr), ROUNDUP(max nbr + 1)), which for
 * a single patch is the whole padded range. */
int swe_init_patch(const swe_gsmd *g, const swe_layout *l, swe_patch *p)
{
    memset(p, 0, sizeof(*p));
    const int Np = g->padded_Nnodes;
    int lo = 0, hi = Np - 1;
    for (int i = 0; i < Np; i++)
        for (int j = 0; j < g->Nnbr; j++) {
            const int id = g->idx[swe_dm_lin_id(l, i, j, g->padded_Nnbr)];
            if (id < lo) lo = id;
            if (id > hi) hi = id;
        }
    const int halo_s = (lo / l->simd) * l->simd;
    const int halo_e = swe_roundup(l, hi + 1);
    p->compute_size = Np;
    p->patch_size = halo_e - halo_s;
    p->halo_pid_s = 0;
    p->compute_pid_s = 0 - halo_s;
    p->compute_pid_e = p->compute_pid_s + Np;
    p->halo_pid_e = p->compute_pid_e + (halo_e - Np);
    p->Nnbr = g->Nnbr;
    p->padded_Nnbr = g->padded_Nnbr;
    p->gh0 = g->gh0;
    p->dt = g->dt;
    p->x = g->x; p->y = g->y; p->z = g->z; p->f = g->f; p->ghm = g->ghm;
    p->p_u = g->p_u; p->p_v = g->p_v; p->p_w = g->p_w; p->gradghm = g->gradghm;
    p->idx = g->idx;
    p->Dx = g->Dx; p->Dy = g->Dy; p->Dz = g->Dz; p->L = g->L;
    return (p->patch_size == Np && p->compute_pid_s == 0) ? 0 : SWE_HOST_EARG;
}

void swe_problem_free(swe_problem *p)
{
    swe_gsmd_free(&p->g);
    free(p->H_init);
    free(p->H_100);
    free(p->mapping);
    memset(p, 0, sizeof(*p));
}

int swe_problem_load(const char *path, int use_rcm, const swe_layout *l,
                     swe_problem *p)
{
    return swe_problem_load_ex(path, use_rcm, l, 0, p);
}

int swe_problem_load_ex(const char *path, int use_rcm, const swe_layout *l, int fp32,
                        swe_problem *p)
{
    return swe_problem_load_fmt(path, use_rcm, l, fp32, 0, p);
}

int swe_problem_load_fmt(const char *path, int use_rcm, const swe_layout *l, int fp32,
                         int netcdf, swe_problem *p)
{
    memset(p, 0, sizeof(*p));
    if (!l || l->simd <= 0)
        return SWE_HOST_EARG;
    p->layout = *l;
    int err = netcdf ? swe_get_gsmd_nc_ex(path, fp32, &p->g) : swe_get_gsmd_bin_ex(path, fp32, &p->g);
    if (!err)
        err = netcdf ? swe_get_state_nc_ex(path, 0, fp32, &p->H_init)
                     : swe_get_state_bin_ex(path, 0, fp32, &p->H_init);
    if (!err)
        err = netcdf ? swe_get_state_nc_ex(path, 1, fp32, &p->H_100)
                     : swe_get_state_bin_ex(path, 1, fp32, &p->H_100);
    if (!err && use_rcm) {
        err = swe_rcm_mapping(&p->g, &p->mapping);
        if (!err)
            err = swe_reorder_gsmd(&p->g, p->mapping);
        if (!err && !(p->H_init = swe_reorder_2d(p->H_init, p->mapping, p->g.Nnodes, 4)))
            err = SWE_HOST_ENOMEM;
        if (!err && !(p->H_100 = swe_reorder_2d(p->H_100, p->mapping, p->g.Nnodes, 4)))
            err = SWE_HOST_ENOMEM;
    }
    const int N = p->g.Nnodes;
    if (!err)
        err = swe_pad_gsmd(&p->g, l);
    if (!err && !(p->H_init = swe_pad_2d(p->H_init, N, p->g.padded_Nnodes, 4, 4)))
        err = SWE_HOST_ENOMEM;
    if (!err && !(p->H_100 = swe_pad_2d(p->H_100, N, p->g.padded_Nnodes, 4, 4)))
        err = SWE_HOST_ENOMEM;
    if (!err && l->sfdl)
        err = swe_convert_sfdl(&p->g, l);
    if (!err)
        err = swe_state_to_layout(&p->H_init, p->g.padded_Nnodes, l);
    if (!err)
        err = swe_state_to_layout(&p->H_100, p->g.padded_Nnodes, l);
    if (!err)
        err = swe_init_patch(&p->g, l, &p->patch);
    if (err)
        swe_problem_free(p);
    return err;
}

int swe_verify_output(const swe_patch *p, const swe_layout *l,
                      const double *H1, const double *H2, double tol)
{
    int count = 0;
    for (int i = p->compute_pid_s; i < p->compute_pid_e; i++)
        for (int sv = 0; sv < 4; sv++) {
            const long k = swe_svm_lin_id(l, i, sv);
            double d = H1[k] - H2[k];
            d = d < 0 ? -d : d;
            if (d > tol || d != d)
                count++;
        }
    return count;
}
