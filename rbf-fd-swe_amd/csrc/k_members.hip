// Launchers of the ensemble stage kernel (member blocks of 2..64 lanes).
#include "swe_ctx.h"

namespace swe {
namespace {

template <int STAGE, bool EX, int MBv, int NN>
void launch_members(const swe_hip_ctx *c, const DevMesh &m, DevStage st)
{
    if constexpr (MBv == 64 && NN > 0) {
        // Member pairs: 32 lanes per node.
        st.nblk = blocks_x(c, (long)c->n * 32);
        const dim3 g(st.nblk, c->B / 64), b(kBlock);
        if (c->prefetch == 8)
            k_stage_mpair<STAGE, EX, NN, 4><<<g, b, 0, c->stream>>>(m, st);
        else if (c->prefetch == 4)
            k_stage_mpair<STAGE, EX, NN, 2><<<g, b, 0, c->stream>>>(m, st);
        else
            k_stage_mpair<STAGE, EX, NN, 3><<<g, b, 0, c->stream>>>(m, st);
        return;
    }
    st.nblk = blocks_x(c, (long)c->n * MBv);
    const dim3 g(st.nblk, c->B / MBv), b(kBlock);
    // Exact sums need the pipelined loop: left to the compiler, all the
    // independent products are hoisted ahead of the serial sums.
    if (NN > 0 && (EX || c->prefetch == 4))
        k_stage_members<STAGE, EX, MBv, NN, 4><<<g, b, 0, c->stream>>>(m, st);
    else if (NN > 0 && c->prefetch == 8)
        k_stage_members<STAGE, EX, MBv, NN, 8><<<g, b, 0, c->stream>>>(m, st);
    else
        k_stage_members<STAGE, EX, MBv, NN, 0><<<g, b, 0, c->stream>>>(m, st);
}

}  // namespace

template <int STAGE, bool EX, int NN>
void family_members(const swe_hip_ctx *c, const DevStage &st)
{
    const DevMesh m = mesh_of(c);
    switch (c->MB) {
    case 2: launch_members<STAGE, EX, 2, NN>(c, m, st); break;
    case 4: launch_members<STAGE, EX, 4, NN>(c, m, st); break;
    case 8: launch_members<STAGE, EX, 8, NN>(c, m, st); break;
    case 16: launch_members<STAGE, EX, 16, NN>(c, m, st); break;
    case 32: launch_members<STAGE, EX, 32, NN>(c, m, st); break;
    default: launch_members<STAGE, EX, 64, NN>(c, m, st); break;
    }
}

namespace {

template <int STAGE>
void by_flags(const swe_hip_ctx *c, const DevStage &st)
{
    const bool nn31 = c->nnbr == 31;
    if (exact(c)) {
        if (nn31) family_members<STAGE, true, 31>(c, st);
        else family_members<STAGE, true, 0>(c, st);
    } else {
        if (nn31) family_members<STAGE, false, 31>(c, st);
        else family_members<STAGE, false, 0>(c, st);
    }
}

}  // namespace

void launch_members_stage(const swe_hip_ctx *c, int stage, const DevStage &st)
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
