// Launcher of the LDS-staged (tiled) ensemble stage kernel: member lanes over
// LDS tiles.
#include "swe_ctx.h"

namespace swe {
namespace {

// Member lanes over LDS tiles (member block 32, 31-point stencils): loader
// and compute waves over two LDS images, one persistent workgroup of 8 waves
// per CU.
template <int STAGE, bool EX>
void launch_tspec(const swe_hip_ctx *c, const DevMesh &m, DevStage st)
{
    const DevTiles tl = tiles_of(c, 32);
    const long items = (long)tl.ntiles * tl.nmb;
    const long wgs = c->ncu;
    st.nblk = (int)(((items < wgs ? items : wgs) + 7) / 8 * 8);
    const size_t lds = sizeof(double) * (size_t)kTmemLdsDoubles;
    if (!EX && c->prefetch == 8)   // stencil sums on the FP64 matrix cores
        k_stage_tspec<STAGE, false, 31, true><<<dim3(st.nblk), dim3(512), lds, c->stream>>>(
            m, tl, st);
    else
        k_stage_tspec<STAGE, EX, 31><<<dim3(st.nblk), dim3(512), lds, c->stream>>>(m, tl, st);
}

template <int STAGE>
void by_flags(const swe_hip_ctx *c, const DevStage &st)
{
    // tiled contexts have member block 32 and a 31-point stencil (checked at
    // creation)
    const DevMesh m = mesh_of(c);
    if (exact(c))
        launch_tspec<STAGE, true>(c, m, st);
    else
        launch_tspec<STAGE, false>(c, m, st);
}

}  // namespace

void launch_tiled_stage(const swe_hip_ctx *c, int stage, const DevStage &st)
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
