// Launchers of the FP64 single-state stage kernels (pair, tile and quadw mappings).
#include "swe_ctx.h"

namespace swe {
namespace {

template <int STAGE, bool EX, int NN>
void launch_quadw(const swe_hip_ctx *c, const DevMesh &m, DevStage st)
{
    st.nblk = blocks_x(c, 4L * c->n, kSingleBlock);
    const dim3 g(st.nblk), b(kSingleBlock);
    if (NN > 0 && c->prefetch == 8)
        k_stage_quadw<STAGE, EX, NN, 8><<<g, b, 0, c->stream>>>(m, st);
    else if (NN > 0 && c->prefetch == 4)
        k_stage_quadw<STAGE, EX, NN, 4><<<g, b, 0, c->stream>>>(m, st);
    else
        k_stage_quadw<STAGE, EX, NN, 0><<<g, b, 0, c->stream>>>(m, st);
}

template <int STAGE, bool EX, int NN>
void launch_pair(const swe_hip_ctx *c, const DevMesh &m, DevStage st)
{
    if constexpr (NN > 0) {
        st.nblk = blocks_x(c, 2L * c->n, kSingleBlock);
        k_stage_pair<STAGE, EX, NN><<<dim3(st.nblk), dim3(kSingleBlock), 0, c->stream>>>(m, st);
    } else {
        launch_quadw<STAGE, EX, NN>(c, m, st);   // needs a compile-time stencil width
    }
}

// Pair mapping over LDS-staged tiles: one workgroup of 2 x tile_nodes lanes
// per tile; prefetch selects the row staging (0 registers, 4 direct-to-LDS,
// 8 direct-to-LDS + early weights).
template <int STAGE, bool EX, int NN, int TN>
void launch_ptile_tn(const swe_hip_ctx *c, const DevMesh &m, DevStage st)
{
    const DevTiles tl = tiles_of(c, 1);
    st.nblk = (tl.ntiles + 7) / 8 * 8;
    const dim3 g(st.nblk), b(2 * TN);
    if (c->prefetch == 8)
        k_stage_ptile<STAGE, EX, NN, TN, 2><<<g, b, 0, c->stream>>>(m, tl, st);
    else if (c->prefetch == 4)
        k_stage_ptile<STAGE, EX, NN, TN, 1><<<g, b, 0, c->stream>>>(m, tl, st);
    else
        k_stage_ptile<STAGE, EX, NN, TN, 0><<<g, b, 0, c->stream>>>(m, tl, st);
}

template <int STAGE, bool EX, int NN>
void launch_ptile(const swe_hip_ctx *c, const DevMesh &m, DevStage st)
{
    if constexpr (NN > 0) {
        if (c->tile_nodes == 64)
            launch_ptile_tn<STAGE, EX, NN, 64>(c, m, st);
        else if (c->tile_nodes == 256)
            launch_ptile_tn<STAGE, EX, NN, 256>(c, m, st);
        else
            launch_ptile_tn<STAGE, EX, NN, 128>(c, m, st);
    } else {
        launch_quadw<STAGE, EX, NN>(c, m, st);   // needs a compile-time stencil width
    }
}

}  // namespace

template <int STAGE, bool EX, int NN>
void family_single(const swe_hip_ctx *c, const DevStage &st)
{
    const DevMesh m = mesh_of(c);
    // the resident context's single stages (drop-in substeps, RHS) are the
    // pair mapping's
    if (c->kernel == SWE_HIP_KERNEL_PAIR || c->kernel == SWE_HIP_KERNEL_RESIDENT)
        launch_pair<STAGE, EX, NN>(c, m, st);
    else if (c->kernel == SWE_HIP_KERNEL_TILE && c->tiles.ntiles > 0)
        launch_ptile<STAGE, EX, NN>(c, m, st);
    else
        launch_quadw<STAGE, EX, NN>(c, m, st);
}

namespace {

template <int STAGE>
void by_flags(const swe_hip_ctx *c, const DevStage &st)
{
    const bool nn31 = c->nnbr == 31;
    if (exact(c)) {
        if (nn31) family_single<STAGE, true, 31>(c, st);
        else family_single<STAGE, true, 0>(c, st);
    } else {
        if (nn31) family_single<STAGE, false, 31>(c, st);
        else family_single<STAGE, false, 0>(c, st);
    }
}

}  // namespace

void launch_single_stage(const swe_hip_ctx *c, int stage, const DevStage &st)
{
    switch (stage) {
    case ST_0: by_flags<ST_0>(c, st); break;
    case ST_1: by_flags<ST_1>(c, st); break;
    case ST_2: by_flags<ST_2>(c, st); break;
    case ST_3: by_flags<ST_3>(c, st); break;
    default: by_flags<ST_RHS>(c, st); break;
    }
}

}  // namespace swe
