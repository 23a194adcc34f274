/*
 * Reverse Cuthill-McKee node ordering of the stencil graph.
 *
 * Same permutation as swe_code/rcm/rcm.c:9-173:
 *  - breadth-first from node 0; the visit order is the mapping itself;
 *  - a visited node's stencil (position 0, the node itself, skipped) is
 *    ranked by (number of not-yet-mapped entries j>=1 of the neighbour's own
 *    stencil, counted before any of this round's neighbours are mapped;
 *    then node id), and unmapped neighbours are appended in that order.
 *    The reference ranks with a selection pass by degree followed by an id
 *    sort inside equal-degree runs (rcm.c:133-173); with distinct ids that
 *    is exactly the (degree, id) order computed here;
 *  - the order is reversed by swapping i <-> N-1-i for i < (N-1)/2
 *    (rcm.c:51-55), so for even N the middle pair keeps its order.
 * The reference reads uninitialised memory when the graph is disconnected;
 * here that case returns SWE_HOST_EGRAPH.
 */
#include "swe_host.h"

#include <stdlib.h>
#include <string.h>

typedef struct rank_key {
    int degree;
    int id;
} rank_key;

static int key_less(const rank_key *a, const rank_key *b)
{
    return a->degree < b->degree || (a->degree == b->degree && a->id < b->id);
}

int swe_rcm_mapping(const swe_gsmd *g, int **mapping_out)
{
    const int N = g->Nnodes, W = g->Nnbr;
    *mapping_out = NULL;
    if (N <= 0 || W <= 0)
        return SWE_HOST_EARG;
    int *mapping = (int *)malloc(sizeof(int) * N);
    unsigned char *mapped = (unsigned char *)calloc(N, 1);
    rank_key *keys = (rank_key *)malloc(sizeof(rank_key) * W);
    if (!mapping || !mapped || !keys) {
        free(mapping); free(mapped); free(keys);
        return SWE_HOST_ENOMEM;
    }
    mapping[0] = 0;
    mapped[0] = 1;
    int filled = 1;
    for (int head = 0; head < filled && filled < N; head++) {
        const int *row = &g->idx[(long)W * mapping[head]];
        int nk = 0;
        for (int j = 1; j < W; j++) {
            const int *nrow = &g->idx[(long)W * row[j]];
            int deg = 0;
            for (int k = 1; k < W; k++)
                deg += mapped[nrow[k]] == 0;
            rank_key key = {deg, row[j]};
            int pos = nk++;   /* insertion into the (degree, id) order */
            while (pos > 0 && key_less(&key, &keys[pos - 1])) {
                keys[pos] = keys[pos - 1];
                pos--;
            }
            keys[pos] = key;
        }
        for (int k = 0; k < nk && filled < N; k++) {
            const int id = keys[k].id;
            if (!mapped[id]) {
                mapped[id] = 1;
                mapping[filled++] = id;
            }
        }
    }
    free(keys);
    free(mapped);
    if (filled != N) {
        free(mapping);
        return SWE_HOST_EGRAPH;
    }
    for (int i = 0; i < (N - 1) / 2; i++) {
        const int t = mapping[i];
        mapping[i] = mapping[N - 1 - i];
        mapping[N - 1 - i] = t;
    }
    *mapping_out = mapping;
    return 0;
}

/* reorder_2D_fp_arr (reorder_nodes.c:8-25): row i of the result is row
 * mapping[i] of var. */
double *swe_reorder_2d(double *var, const int *mapping, int dim1, int dim2)
{
    double *out = (double *)malloc(sizeof(double) * (size_t)dim1 * dim2);
    if (!out)
        return NULL;
    for (int i = 0; i < dim1; i++)
        memcpy(&out[(size_t)i * dim2], &var[(size_t)mapping[i] * dim2],
               sizeof(double) * dim2);
    free(var);
    return out;
}

static int *reorder_2d_int(int *var, const int *mapping, int dim1, int dim2)
{
    int *out = (int *)malloc(sizeof(int) * (size_t)dim1 * dim2);
    if (!out)
        return NULL;
    for (int i = 0; i < dim1; i++)
        memcpy(&out[(size_t)i * dim2], &var[(size_t)mapping[i] * dim2],
               sizeof(int) * dim2);
    free(var);
    return out;
}

/* reorder_GSMD_struct (reorder_nodes.c:46-87). */
int swe_reorder_gsmd(swe_gsmd *g, const int *mapping)
{
    const int N = g->Nnodes, W = g->Nnbr;
    double **one[5] = {&g->x, &g->y, &g->z, &g->f, &g->ghm};
    for (int k = 0; k < 5; k++)
        if (!(*one[k] = swe_reorder_2d(*one[k], mapping, N, 1)))
            return SWE_HOST_ENOMEM;
    double **three[4] = {&g->p_u, &g->p_v, &g->p_w, &g->gradghm};
    for (int k = 0; k < 4; k++)
        if (!(*three[k] = swe_reorder_2d(*three[k], mapping, N, 3)))
            return SWE_HOST_ENOMEM;
    double **dm[4] = {&g->Dx, &g->Dy, &g->Dz, &g->L};
    for (int k = 0; k < 4; k++)
        if (!(*dm[k] = swe_reorder_2d(*dm[k], mapping, N, W)))
            return SWE_HOST_ENOMEM;
    if (!(g->idx = reorder_2d_int(g->idx, mapping, N, W)))
        return SWE_HOST_ENOMEM;
    int *inv = (int *)malloc(sizeof(int) * N);
    if (!inv)
        return SWE_HOST_ENOMEM;
    for (int i = 0; i < N; i++)
        inv[mapping[i]] = i;
    for (long k = 0; k < (long)N * W; k++)
        g->idx[k] = inv[g->idx[k]];
    free(inv);
    return 0;
}

void swe_stencil_bandwidth(const swe_gsmd *g, int *max_bw, int *avg_bw)
{
    long sum = 0;
    int mx = 0;
    for (int i = 0; i < g->Nnodes; i++) {
        int lo = g->Nnodes, hi = 0;
        for (int j = 0; j < g->Nnbr; j++) {
            const int id = g->idx[(long)i * g->Nnbr + j];
            if (id > hi) hi = id;
            if (id < lo) lo = id;
        }
        sum += hi - lo;
        if (hi - lo > mx)
            mx = hi - lo;
    }
    *max_bw = mx;
    *avg_bw = g->Nnodes ? (int)(sum / g->Nnodes) : 0;
}
