/*
 * Binary input reader of the RBF-FD SWE host.
 *
 * File format (little-endian), as read by swe_code/io/input.c:189-414 and
 * written by swe_code/io/nc2bin.c:32-80:
 *   int32 Nnodes, Nnbr
 *   f64   a, gh0, dt, gamma
 *   f64   x[N], y[N], z[N], f[N], ghm[N]
 *   f64   p_u[N][3], p_v[N][3], p_w[N][3], gradghm[N][3]
 *   f64   Dx[N][Nnbr], Dy[N][Nnbr], Dz[N][Nnbr], L[N][Nnbr]  (unit sphere)
 *   int32 idx[N][Nnbr]                                       (1-based)
 *   f64   H_init[N][4], H_100[N][4]
 */
#include "swe_host.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int read_exact(FILE *fp, void *dst, size_t elem, size_t count)
{
    return fread(dst, elem, count, fp) == count ? 0 : SWE_HOST_EREAD;
}

static double *read_f64(FILE *fp, size_t count, int *err)
{
    double *a = (double *)malloc(sizeof(double) * (count ? count : 1));
    if (!a) {
        *err = SWE_HOST_ENOMEM;
        return NULL;
    }
    if (read_exact(fp, a, sizeof(double), count)) {
        free(a);
        *err = SWE_HOST_EREAD;
        return NULL;
    }
    return a;
}

/* The header, checked against the file's size: the format's sections have
 * fixed sizes once N and Nnbr are known, so a shorter file (or a corrupted
 * count) is an error before anything is allocated for it. */
static int read_header(FILE *fp, int *N, int *Nnbr)
{
    int32_t hdr[2];
    if (read_exact(fp, hdr, sizeof(int32_t), 2))
        return SWE_HOST_EREAD;
    if (hdr[0] <= 0 || hdr[1] <= 0)
        return SWE_HOST_EREAD;
    const long long n = hdr[0], w = hdr[1];
    const long long need = 4LL * (2 + n * w) + 8LL * (4 + 17 * n + 4 * n * w + 8 * n);
    const long pos = ftell(fp);
    if (pos < 0 || fseek(fp, 0, SEEK_END) != 0)
        return SWE_HOST_EREAD;
    const long size = ftell(fp);
    if (size < 0 || (long long)size < need || fseek(fp, pos, SEEK_SET) != 0)
        return SWE_HOST_EREAD;
    *N = hdr[0];
    *Nnbr = hdr[1];
    return 0;
}

void swe_gsmd_free(swe_gsmd *g)
{
    if (!g)
        return;
    free(g->x); free(g->y); free(g->z); free(g->f); free(g->ghm);
    free(g->p_u); free(g->p_v); free(g->p_w); free(g->gradghm);
    free(g->idx); free(g->Dx); free(g->Dy); free(g->Dz); free(g->L);
    memset(g, 0, sizeof(*g));
}

/* The reference's FP32 model type (RUN_32BIT, swe_config.h:75-80) rounds
 * every value it reads to float ((fType)tmp, input.c:211-306); values are
 * kept here as the doubles of those floats. */
static void round_f32(double *a, size_t count)
{
    for (size_t k = 0; k < count; k++)
        a[k] = (double)(float)a[k];
}

/* get_GSMD_bin, input.c:189-332. */
int swe_get_gsmd_bin(const char *path, swe_gsmd *g) { return swe_get_gsmd_bin_ex(path, 0, g); }

int swe_get_gsmd_bin_ex(const char *path, int fp32, swe_gsmd *g)
{
    memset(g, 0, sizeof(*g));
    FILE *fp = fopen(path, "rb");
    if (!fp)
        return SWE_HOST_EOPEN;
    int err = read_header(fp, &g->Nnodes, &g->Nnbr);
    double c[4];
    if (!err)
        err = read_exact(fp, c, sizeof(double), 4);
    if (err) {
        fclose(fp);
        return err;
    }
    g->a = c[0];
    g->gh0 = c[1];
    g->dt = c[2];
    g->gamma = c[3];
    const size_t N = (size_t)g->Nnodes, NW = N * (size_t)g->Nnbr;
    double **scal[5] = {&g->x, &g->y, &g->z, &g->f, &g->ghm};
    for (int k = 0; k < 5 && !err; k++)
        *scal[k] = read_f64(fp, N, &err);
    double **vec[4] = {&g->p_u, &g->p_v, &g->p_w, &g->gradghm};
    for (int k = 0; k < 4 && !err; k++)
        *vec[k] = read_f64(fp, 3 * N, &err);
    double **dm[4] = {&g->Dx, &g->Dy, &g->Dz, &g->L};
    for (int k = 0; k < 4 && !err; k++)
        *dm[k] = read_f64(fp, NW, &err);
    if (!err) {
        g->idx = (int *)malloc(sizeof(int) * NW);
        if (!g->idx)
            err = SWE_HOST_ENOMEM;
        else
            err = read_exact(fp, g->idx, sizeof(int), NW);
    }
    fclose(fp);
    if (err) {
        swe_gsmd_free(g);
        return err;
    }
    if (fp32) {
        for (int k = 0; k < 5; k++)
            round_f32(*scal[k], N);
        for (int k = 0; k < 4; k++)
            round_f32(*vec[k], 3 * N);
        for (int k = 0; k < 4; k++)
            round_f32(*dm[k], NW);
    }
    /* Unit-sphere weights -> Earth radius and hyperviscosity coefficient,
     * done in double with the file's a and gamma as the reference does
     * (input.c:314-319); FP32 rounds the result on the store. */
    for (size_t k = 0; k < NW; k++) {
        g->Dx[k] = g->Dx[k] / g->a;
        g->Dy[k] = g->Dy[k] / g->a;
        g->Dz[k] = g->Dz[k] / g->a;
        g->L[k] = g->L[k] * g->gamma;
    }
    if (fp32) {
        for (int k = 0; k < 4; k++)
            round_f32(*dm[k], NW);
        double *cst[4] = {&g->a, &g->gh0, &g->dt, &g->gamma};   /* GSMD.a = (fType)a ... */
        for (int k = 0; k < 4; k++)
            round_f32(cst[k], 1);
    }
    /* 1-based -> 0-based (input.c:326-328); reject ids outside the mesh. */
    for (size_t k = 0; k < NW; k++) {
        g->idx[k] -= 1;
        if (g->idx[k] < 0 || g->idx[k] >= g->Nnodes) {
            swe_gsmd_free(g);
            return SWE_HOST_EREAD;
        }
    }
    g->padded_Nnodes = g->Nnodes;
    g->padded_Nnbr = g->Nnbr;
    return 0;
}

/* get_ICs_bin / get_FCs_bin (input.c:336-414): state block `which` (0 =
 * H_init, 1 = H_100) at the offset implied by the header. */
static int read_state(const char *path, int which, double **out)
{
    *out = NULL;
    FILE *fp = fopen(path, "rb");
    if (!fp)
        return SWE_HOST_EOPEN;
    int N, W;
    int err = read_header(fp, &N, &W);
    if (err) {
        fclose(fp);
        return err;
    }
    const long n = N, w = W;
    const long off = (long)sizeof(int) * (2 + n * w)
                     + (long)sizeof(double) * (4 + 5 * n + 12 * n + 4 * n * w + which * 4 * n);
    if (fseek(fp, off, SEEK_SET) != 0) {
        fclose(fp);
        return SWE_HOST_EREAD;
    }
    *out = read_f64(fp, 4 * (size_t)n, &err);
    fclose(fp);
    return err;
}

int swe_get_ics_bin(const char *path, double **H_init) { return read_state(path, 0, H_init); }
int swe_get_fcs_bin(const char *path, double **H_100) { return read_state(path, 1, H_100); }

/* The same with the FP32 model type's rounding (input.c:369-371, :406-408). */
int swe_get_state_bin_ex(const char *path, int which, int fp32, double **H)
{
    int err = read_state(path, which, H);
    if (!err && fp32) {
        int N, W;
        FILE *fp = fopen(path, "rb");
        if (!fp)
            return SWE_HOST_EOPEN;
        err = read_header(fp, &N, &W);
        fclose(fp);
        if (!err)
            round_f32(*H, 4 * (size_t)N);
    }
    return err;
}
