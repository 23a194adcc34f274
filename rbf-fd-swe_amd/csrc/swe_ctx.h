// Internal definitions of libswe_hip shared by its translation units
// (context and mesh structures, launch helpers).  Not part of the C ABI.
#pragma once

#include "swe_hip.h"
#include "swe_kernels.h"
#include "swe_kernels32.h"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace swe {

constexpr int kBlock = 256;            // threads per workgroup, stage kernels
// Threads per workgroup of the single-state stage kernels: one wave, so a
// launch's workgroups spread over the 256 CUs within one wave of each other
// (at N = 40962 the pair kernel has 1281 waves: 256-thread workgroups put
// two on 65 CUs and one on the rest).
constexpr int kSingleBlock = 64;
constexpr int kGraphCount = 4;         // captured step graphs per context
// Resident whole-step kernel (k_resident.hip): one tile of <= 32 x
// kResMaxWaves nodes per compute unit, whose stencils touch <= kResMaxRows
// distinct nodes (100 KB of LDS; a band of 161 nodes in RCM order touches
// ~1100 at N = 40962, a Hilbert-ordered tile ~400).
constexpr int kResMaxWaves = 8;
constexpr int kResMaxRows = 2560;
constexpr int kResAutoMinNodes = 1;       // SWE_HIP_KERNEL_AUTO picks it from here
constexpr int kResMaxNbrs = 64;           // tiles one tile exchanges rows with (one wave)
constexpr int kResAutoMinSteps = 4;       // AUTO contexts: shorter calls replay stage graphs

// Host index macros of the reference (swe_config.h:21-54) at run time.
struct HostLayout {
    int cfdl, sfdl, S;
    long dm(int i, int j, int nnbr) const
    {
        return sfdl ? (long)(i / S) * S * nnbr + (long)j * S + i % S : (long)i * nnbr + j;
    }
    long vect(int i, int c) const
    {
        return sfdl ? (long)(i / S) * S * 3 + (long)c * S + i % S : (long)i * 3 + c;
    }
    long svm(int pid, int sv) const
    {
        return cfdl ? 4L * pid + sv : (long)(pid / S) * S * 4 + (long)sv * S + pid % S;
    }
};

}  // namespace swe

// Device copy of one patch's static data, shared by every context created
// with `share` (reference-counted; owned by the contexts).
struct SweMesh {
    int refs = 1;
    int dev = 0;
    int n = 0, ld = 0, nnbr = 0;
    double gh0 = 0.0, dt = 0.0;
    swe::HostLayout lay{1, 0, 1};
    int *d_nbr = nullptr;
    double *d_wts = nullptr, *d_geo = nullptr;
    float *d_wts32 = nullptr, *d_geo32 = nullptr;   // FP32 meshes (d_wts, d_geo unused)
    bool fp32 = false;
    long long bytes = 0;
    int order = SWE_HIP_ORDER_HOST;
    int sblk = 0;                 // stencil layout: 0 slot-major, 16 node blocks
    std::vector<int> perm;        // device node -> host node (empty: identity)
    std::vector<int> h_nbr;       // host copy of nbr, device order, [nnbr][n]
};

// Tiles of the LDS-staged ensemble kernel (per context: they depend on the
// member block through the LDS budget).
struct SweTiles {
    int ntiles = 0, umax = 0, tmax = 0;
    int *d_tile_start = nullptr, *d_ustart = nullptr, *d_ulist = nullptr;
    unsigned short *d_lidx = nullptr;
    unsigned short *d_lidx_node = nullptr;   // [n][32] node-major copy of d_lidx
    int umax_used = 0;                       // most distinct rows of one tile
    int *d_ulist_tile = nullptr;             // [ntiles][kTmemStride] (member-lane kernel)
    double *d_wts_node = nullptr;            // [n][32][4] (member-lane kernel)
    double *d_geo_node = nullptr;            // [n][32] (member-lane kernel)
};

struct swe_hip_ctx {
    SweMesh *mesh = nullptr;
    int dev = 0;
    hipStream_t stream = nullptr;
    // copies of the mesh fields used on every launch
    int n = 0, ld = 0, nnbr = 0;
    double gh0 = 0.0, dt = 0.0;
    swe::HostLayout lay{1, 0, 1};
    int *d_nbr = nullptr;
    double *d_wts = nullptr, *d_geo = nullptr;
    int sblk = 0;
    int B = 1, MB = 1;
    int kernel = SWE_HIP_KERNEL_QUAD;   // single-state mapping
    int split32 = 1;                    // FP32: lanes per node (1, 2, 4, 8)
    int prefetch = 0;                   // stencil slots per load batch (0, 4, 8)
    int tile_nodes = 0, tile_rows = 0;  // tile shape of the tiled kernels (0: default)
    int ncu = 0;                        // compute units of the device
    unsigned flags = 0;
    double *d_H = nullptr, *d_Ka = nullptr, *d_Kb = nullptr, *d_D = nullptr, *d_F = nullptr;
    // FP32 contexts (SWE_HIP_FP32): the same buffers in float, d_* unused
    bool fp32 = false;
    float *f_H = nullptr, *f_Ka = nullptr, *f_Kb = nullptr, *f_D = nullptr, *f_F = nullptr;
    double *d_stage = nullptr;   // [n*4] staging for member upload/read-back
    int *d_active = nullptr;     // member slots advanced by the ensemble kernel
    SweTiles tiles;              // LDS-staged ensemble kernel (ntiles == 0: unused)
    hipGraphExec_t graphs[swe::kGraphCount] = {};   // 64, 16, 4 and 1 RK steps
    // resident whole-step kernel (kernel == SWE_HIP_KERNEL_RESIDENT)
    int res_block = 0, res_lds = 0;   // threads per workgroup, dynamic LDS bytes
    void *d_res_sync = nullptr;       // abort word + per-tile stage counts
    unsigned *h_res_abort = nullptr;  // pinned: wait time-out flag of the resident launches
    unsigned *d_res_abort = nullptr;  // its device address
    bool res_auto = false;            // chosen by AUTO: short calls use the stage graphs
    long long res_stages = 0;         // stages counted since the counts were cleared
    int *d_res_nbrs = nullptr;        // [ntiles][kResMaxNbrs] tiles each tile waits for
    bool res_pending = false;         // a launch's flag is not checked yet
    unsigned res_spin = 1u << 22;     // polls a wait may take before it gives up
    bool timing = false;
    swe_hip_timers acc{};
    long long bytes = 0;         // state buffers
    std::vector<double> host_tmp;
};

namespace swe {

inline bool exact(const swe_hip_ctx *c) { return (c->flags & SWE_HIP_EXACT) != 0; }

inline long state_count(const swe_hip_ctx *c) { return 4L * c->n * c->B; }

inline int blocks_x(const swe_hip_ctx *c, long threads, int block = kBlock)
{
    long nb = (threads + block - 1) / block;
    nb = (nb + 7) / 8 * 8;   // multiple of 8 for the XCD remap
    return (int)nb;
}

inline DevMesh mesh_of(const swe_hip_ctx *c)
{
    DevMesh m;
    m.nbr = c->d_nbr;
    m.wts = c->d_wts;
    m.geo = c->d_geo;
    m.n = c->n;
    m.ld = c->ld;
    m.nnbr = c->nnbr;
    m.sblk = c->sblk;
    m.gh0 = c->gh0;
    return m;
}

// Records msg as the calling thread's last error (swe_hip_last_error) and
// returns code.
int set_error(int code, const std::string &msg);

// Resident whole-step kernel (k_resident.hip): bytes of its sync state,
// LDS/co-residency set-up once the tiles are built, n steps (asynchronous),
// and the check of the last launch's time-out flag, to be called
// once the context's stream has drained.
long resident_sync_bytes(int ntiles);
int resident_prepare(swe_hip_ctx *c, std::string *err);
int resident_steps(swe_hip_ctx *c, int nsteps, std::string *err);
int resident_check(swe_hip_ctx *c, std::string *err);

// Looks up this device's no-allocation scratch address (swe_noalloc.hip);
// the device must be current.  On failure *err says why.
int noalloc_setup(int dev, std::string *err);

inline DevMesh32 mesh32_of(const swe_hip_ctx *c)
{
    DevMesh32 m;
    m.nbr = c->d_nbr;
    m.wts = c->mesh->d_wts32;
    m.geo = c->mesh->d_geo32;
    m.n = c->n;
    m.ld = c->ld;
    m.nnbr = c->nnbr;
    m.sblk = c->sblk;
    m.gh0 = (float)c->gh0;
    return m;
}

// One FP32 stage (ST_0..ST_3) or RHS (ST_RHS) on the context's stream.
void launch_fp32_stage(const swe_hip_ctx *c, int stage, const DevStage32 &st);

// One fused stage (ST_0..ST_3) or a plain RHS (ST_RHS) on the context's
// stream, by kernel family (each in its own translation unit).
void launch_single_stage(const swe_hip_ctx *c, int stage, const DevStage &st);
void launch_members_stage(const swe_hip_ctx *c, int stage, const DevStage &st);
void launch_tiled_stage(const swe_hip_ctx *c, int stage, const DevStage &st);

inline void launch_stage(const swe_hip_ctx *c, int stage, const DevStage &st)
{
    if (c->B == 1)
        launch_single_stage(c, stage, st);
    else if (c->tiles.ntiles > 0)
        launch_tiled_stage(c, stage, st);
    else
        launch_members_stage(c, stage, st);
}

// Device view of the context's tiles for member block mb.
inline DevTiles tiles_of(const swe_hip_ctx *c, int mb)
{
    DevTiles tl;
    tl.tile_start = c->tiles.d_tile_start;
    tl.ustart = c->tiles.d_ustart;
    tl.ulist = c->tiles.d_ulist;
    tl.lidx = c->tiles.d_lidx;
    tl.lidx_node = c->tiles.d_lidx_node;
    tl.ulist_tile = c->tiles.d_ulist_tile;
    tl.wts_node = c->tiles.d_wts_node;
    tl.geo_node = c->tiles.d_geo_node;
    tl.ntiles = c->tiles.ntiles;
    tl.nmb = c->B / mb;
    return tl;
}

}  // namespace swe
