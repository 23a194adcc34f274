// Launchers of the FP32 single-state stage kernels.
#include "swe_ctx.h"

namespace swe {
namespace {

template <int STAGE, int S>
void launch_split32(const swe_hip_ctx *c, const DevMesh32 &m, DevStage32 st)
{
    st.nblk = blocks_x(c, (long)c->n * S, kSingleBlock);
    const dim3 g(st.nblk), b(kSingleBlock);
    if (c->nnbr == 31)
        k_stage_split32<STAGE, 31, S><<<g, b, 0, c->stream>>>(m, st);
    else
        k_stage_split32<STAGE, 0, S><<<g, b, 0, c->stream>>>(m, st);
}

template <int STAGE, bool EX, int TN>
void launch_tile32_tn(const swe_hip_ctx *c, const DevMesh32 &m, DevStage32 st)
{
    const DevTiles tl = tiles_of(c, 1);
    st.nblk = (tl.ntiles + 7) / 8 * 8;
    k_stage_tile32<STAGE, EX, 31, TN><<<dim3(st.nblk), dim3(TN), 0, c->stream>>>(m, tl, st);
}

template <int STAGE, bool EX>
void launch32(const swe_hip_ctx *c, DevStage32 st)
{
    const DevMesh32 m = mesh32_of(c);
    if (c->kernel == SWE_HIP_KERNEL_TILE && c->tiles.ntiles > 0 && c->nnbr == 31) {
        if (c->tile_nodes == 64)
            launch_tile32_tn<STAGE, EX, 64>(c, m, st);
        else if (c->tile_nodes == 128)
            launch_tile32_tn<STAGE, EX, 128>(c, m, st);
        else
            launch_tile32_tn<STAGE, EX, 256>(c, m, st);
        return;
    }
    if constexpr (!EX) {   // fast mode: S lanes per node split the slots
        switch (c->split32) {
        case 2: launch_split32<STAGE, 2>(c, m, st); return;
        case 4: launch_split32<STAGE, 4>(c, m, st); return;
        case 8: launch_split32<STAGE, 8>(c, m, st); return;
        default: break;
        }
    }
    // one-wave workgroups: at N = 40962 one lane per node is 641 waves, which
    // 256-thread workgroups would put on 161 of the 256 CUs
    st.nblk = blocks_x(c, (long)c->n, kSingleBlock);
    const dim3 g(st.nblk), b(kSingleBlock);
    if (c->nnbr == 31)
        k_stage_lane32<STAGE, EX, 31><<<g, b, 0, c->stream>>>(m, st);
    else
        k_stage_lane32<STAGE, EX, 0><<<g, b, 0, c->stream>>>(m, st);
}

template <int STAGE>
void by_flags(const swe_hip_ctx *c, const DevStage32 &st)
{
    if (exact(c))
        launch32<STAGE, true>(c, st);
    else
        launch32<STAGE, false>(c, st);
}

}  // namespace

void launch_fp32_stage(const swe_hip_ctx *c, int stage, const DevStage32 &st)
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
