// FP32 stage kernels: the reference's shipped model precision (RUN_32BIT,
// fType = float, swe_code/main/include/swe_config.h:75-80).
//
// Device layouts as the FP64 path (swe_kernels.h) with float values:
//   wts32 f32 [nnbr][ld][4] (Dx, Dy, Dz, L of one slot and node: 16 B),
//   geo32 f32 [GEO_N][ld], state f32 [node][u v w h] (16 B per node).
// One lane per node: per stencil slot one 16-byte weight load, one 16-byte
// neighbour-state gather and one id load, the 16 sums in float.
//
// Arithmetic follows the reference's C types (the build uses
// -ffp-contract=off; EXACT kernels round every float product and sum
// separately, in the reference's order, and reproduce its FP32 build bit for
// bit): the stencil sums and the pointwise RHS are float
// (rk4_rbffd_swe.c:254-353); eval_K's coefficient dt/2.0 or dt is a float
// argument and K = H + c*F is float (:123-158); update_D is float
// (:160-186); update_H computes (dt/6.0)*D and the sum with H in double and
// rounds once to float on the store (:86-121: dt/6.0 is a double
// expression).  The fast kernels use fused multiply-adds.
#pragma once

#include "swe_kernels.h"

namespace swe {

struct DevMesh32 {
    const int *__restrict__ nbr;
    const float *__restrict__ wts;
    const float *__restrict__ geo;
    int n, ld, nnbr, sblk;
    float gh0;
};

struct DevStage32 {
    const float *src;   // gather source: H (stage 0) or K
    const float *h;     // H, own node
    float *d;           // D
    float *out;         // next K, H (stage 3) or F (ST_RHS)
    float c;            // K = H + c*F (float dt/2.0, dt/2.0, dt)
    double c6;          // stage 3: H += (dt/6.0)*D in double
    float dcoef;        // D += dcoef*F (2, 2, 1)
    int nblk;
};

template <bool EXACT>
__device__ __forceinline__ float madd32(float acc, float w, float s)
{
    if constexpr (EXACT) {
        const float p = w * s;
        return acc + p;
    } else {
        return __builtin_fmaf(w, s, acc);
    }
}

// y + (a*x) in float.
template <bool EXACT>
__device__ __forceinline__ float axpy32(float y, float a, float x)
{
    if constexpr (EXACT) {
        const float p = a * x;
        return y + p;
    } else {
        return __builtin_fmaf(a, x, y);
    }
}

// update_H: (float)((double)H + (dt/6.0) * (double)D).
template <bool EXACT>
__device__ __forceinline__ float update_h32(float h, double c6, float d)
{
    if constexpr (EXACT) {
        const double p = c6 * (double)d;
        return (float)((double)h + p);
    } else {
        return (float)__builtin_fma(c6, (double)d, (double)h);
    }
}

struct Sums32 {
    float dudx, dvdx, dwdx, dhdx;
    float dudy, dvdy, dwdy, dhdy;
    float dudz, dvdz, dwdz, dhdz;
    float lu, lv, lw, lh;
};

// Pointwise right-hand side of one node in float, operation order of
// rk4_rbffd_swe.c:331-353.
__device__ __forceinline__ void rhs_point32(const Sums32 &S, float u, float v, float w, float h,
                                            const float *g, float gh0, float F[4])
{
    const float x = g[G_X], y = g[G_Y], z = g[G_Z], f = g[G_F], ghm = g[G_GHM];
    const float ru = -((u * S.dudx) + (v * S.dudy) + (w * S.dudz) + (f * ((y * w) - (z * v)))
                       + S.dhdx);
    const float rv = -((u * S.dvdx) + (v * S.dvdy) + (w * S.dvdz) + (f * ((z * u) - (x * w)))
                       + S.dhdy);
    const float rw = -((u * S.dwdx) + (v * S.dwdy) + (w * S.dwdz) + (f * ((x * v) - (y * u)))
                       + S.dhdz);
    F[0] = (g[G_PU] * ru) + (g[G_PU + 1] * rv) + (g[G_PU + 2] * rw) + S.lu;
    F[1] = (g[G_PV] * ru) + (g[G_PV + 1] * rv) + (g[G_PV + 2] * rw) + S.lv;
    F[2] = (g[G_PW] * ru) + (g[G_PW + 1] * rv) + (g[G_PW + 2] * rw) + S.lw;
    F[3] = -((u * (S.dhdx - g[G_GG])) + (v * (S.dhdy - g[G_GG + 1]))
             + (w * (S.dhdz - g[G_GG + 2])) + ((h + gh0 - ghm) * (S.dudx + S.dvdy + S.dwdz)))
           + S.lh;
}

// One node's fused stage update of its four values (16-byte accesses).
template <int STAGE, bool EXACT>
__device__ __forceinline__ void stage_update32(const DevStage32 &st, long e, const float F[4])
{
    float4 *out = reinterpret_cast<float4 *>(st.out + e);
    if constexpr (STAGE == ST_RHS) {
        *out = make_float4(F[0], F[1], F[2], F[3]);
    } else {
        const float4 hv = *reinterpret_cast<const float4 *>(st.h + e);
        float4 *dp = reinterpret_cast<float4 *>(st.d + e);
        if constexpr (STAGE == ST_0) {
            *dp = make_float4(F[0], F[1], F[2], F[3]);
            *out = make_float4(axpy32<EXACT>(hv.x, st.c, F[0]), axpy32<EXACT>(hv.y, st.c, F[1]),
                               axpy32<EXACT>(hv.z, st.c, F[2]), axpy32<EXACT>(hv.w, st.c, F[3]));
        } else {
            const float4 d0 = *dp;
            const float4 dv = make_float4(axpy32<EXACT>(d0.x, st.dcoef, F[0]),
                                          axpy32<EXACT>(d0.y, st.dcoef, F[1]),
                                          axpy32<EXACT>(d0.z, st.dcoef, F[2]),
                                          axpy32<EXACT>(d0.w, st.dcoef, F[3]));
            if constexpr (STAGE == ST_3) {
                *out = make_float4(update_h32<EXACT>(hv.x, st.c6, dv.x),
                                   update_h32<EXACT>(hv.y, st.c6, dv.y),
                                   update_h32<EXACT>(hv.z, st.c6, dv.z),
                                   update_h32<EXACT>(hv.w, st.c6, dv.w));
            } else {
                *dp = dv;
                *out = make_float4(axpy32<EXACT>(hv.x, st.c, F[0]), axpy32<EXACT>(hv.y, st.c, F[1]),
                                   axpy32<EXACT>(hv.z, st.c, F[2]), axpy32<EXACT>(hv.w, st.c, F[3]));
            }
        }
    }
}

// Single state, one lane per node, FP32.  NNBR > 0: compile-time stencil
// width; 0: run-time width.
template <int STAGE, bool EXACT, int NNBR>
__global__ __launch_bounds__(256) void k_stage_lane32(DevMesh32 m, DevStage32 st)
{
    const int t = xcd_block(st.nblk) * (int)blockDim.x + (int)threadIdx.x;
    const bool live = t < m.n;
    const int node = live ? t : m.n - 1;
    const int nn = NNBR > 0 ? NNBR : m.nnbr;
    const long ld = m.sblk ? 16L : (long)m.ld;
    const long W = NNBR > 0 ? NNBR : m.nnbr;
    const long sb = m.sblk ? ((long)(node >> 4) * W << 4) + (node & 15) : (long)node;
    const int *__restrict__ nb = m.nbr + sb;
    const float *__restrict__ wp = m.wts + 4L * sb;
    const float *__restrict__ src = st.src;
    Sums32 S = {};
    auto slot = [&](int j) {
        const int q = nb[j * ld];
        const float4 w = *reinterpret_cast<const float4 *>(wp + 4L * j * ld);
        const float4 s = *reinterpret_cast<const float4 *>(src + 4L * q);
        S.dudx = madd32<EXACT>(S.dudx, w.x, s.x); S.dvdx = madd32<EXACT>(S.dvdx, w.x, s.y);
        S.dwdx = madd32<EXACT>(S.dwdx, w.x, s.z); S.dhdx = madd32<EXACT>(S.dhdx, w.x, s.w);
        S.dudy = madd32<EXACT>(S.dudy, w.y, s.x); S.dvdy = madd32<EXACT>(S.dvdy, w.y, s.y);
        S.dwdy = madd32<EXACT>(S.dwdy, w.y, s.z); S.dhdy = madd32<EXACT>(S.dhdy, w.y, s.w);
        S.dudz = madd32<EXACT>(S.dudz, w.z, s.x); S.dvdz = madd32<EXACT>(S.dvdz, w.z, s.y);
        S.dwdz = madd32<EXACT>(S.dwdz, w.z, s.z); S.dhdz = madd32<EXACT>(S.dhdz, w.z, s.w);
        S.lu = madd32<EXACT>(S.lu, w.w, s.x); S.lv = madd32<EXACT>(S.lv, w.w, s.y);
        S.lw = madd32<EXACT>(S.lw, w.w, s.z); S.lh = madd32<EXACT>(S.lh, w.w, s.w);
    };
    // a rolled loop of 4-slot bodies keeps ~4 slots of loads in flight per
    // lane (a fully unrolled one front-loads all 31 and spills)
#pragma unroll 4
    for (int j = 0; j < nn; ++j)
        slot(j);
    const float4 own = *reinterpret_cast<const float4 *>(src + 4L * node);
    float g[GEO_N];
#pragma unroll
    for (int f = 0; f < GEO_N; ++f)
        g[f] = m.geo[(long)f * m.ld + node];
    float F[4];
    rhs_point32(S, own.x, own.y, own.z, own.w, g, m.gh0, F);
    if (live)
        stage_update32<STAGE, EXACT>(st, 4L * node, F);
}


// A float moved across lanes by DPP (ctrl: quad_perm or row_half_mirror).
template <int CTRL>
__device__ __forceinline__ float dpp32(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Adds the partial sums of the S lanes of a group (S = 2, 4 or 8 aligned
// lanes): pairs, then quads (quad_perm), then halves of 8 (row_half_mirror).
// Every round adds two partial sums in one lane and the same two in its
// partner (x + y and y + x round alike), so all lanes of a group end with
// the same total.
template <int S>
__device__ __forceinline__ float group_sum32(float x)
{
    x += dpp32<0xB1>(x);            // quad_perm [1, 0, 3, 2]
    if constexpr (S >= 4)
        x += dpp32<0x4E>(x);        // quad_perm [2, 3, 0, 1]
    if constexpr (S >= 8)
        x += dpp32<0x141>(x);       // row_half_mirror: lane i <-> 7 - i within 8
    return x;
}

// Single state, S lanes per node splitting the stencil slots (fast mode):
// lane h of a node's group sums slots h, h + S, ... in order with FMAs, the
// group adds its S partial sums (group_sum32), and the group's first lane
// evaluates the RHS and writes the stage update.  S times the waves of
// k_stage_lane32 with S times fewer slots each, for meshes too small to fill
// the chip with one lane per node; the same loads per node (one 16-byte
// weight load, one 16-byte gather and one id per slot).
template <int STAGE, int NNBR, int S>
__global__ __launch_bounds__(256) void k_stage_split32(DevMesh32 m, DevStage32 st)
{
    static_assert(S == 2 || S == 4 || S == 8, "groups of 2, 4 or 8 lanes");
    const int t = xcd_block(st.nblk) * (int)blockDim.x + (int)threadIdx.x;
    const int h = (int)threadIdx.x & (S - 1);
    const bool live = t / S < m.n;
    const int node = live ? t / S : m.n - 1;
    const int nn = NNBR > 0 ? NNBR : m.nnbr;
    const long ld = m.sblk ? 16L : (long)m.ld;
    const long W = NNBR > 0 ? NNBR : m.nnbr;
    const long sb = m.sblk ? ((long)(node >> 4) * W << 4) + (node & 15) : (long)node;
    const int *__restrict__ nb = m.nbr + sb;
    const float *__restrict__ wp = m.wts + 4L * sb;
    const float *__restrict__ src = st.src;
    Sums32 P = {};
    auto slot = [&](int j) {
        const int q = nb[j * ld];
        const float4 w = *reinterpret_cast<const float4 *>(wp + 4L * j * ld);
        const float4 v = *reinterpret_cast<const float4 *>(src + 4L * q);
        P.dudx = madd32<false>(P.dudx, w.x, v.x); P.dvdx = madd32<false>(P.dvdx, w.x, v.y);
        P.dwdx = madd32<false>(P.dwdx, w.x, v.z); P.dhdx = madd32<false>(P.dhdx, w.x, v.w);
        P.dudy = madd32<false>(P.dudy, w.y, v.x); P.dvdy = madd32<false>(P.dvdy, w.y, v.y);
        P.dwdy = madd32<false>(P.dwdy, w.y, v.z); P.dhdy = madd32<false>(P.dhdy, w.y, v.w);
        P.dudz = madd32<false>(P.dudz, w.z, v.x); P.dvdz = madd32<false>(P.dvdz, w.z, v.y);
        P.dwdz = madd32<false>(P.dwdz, w.z, v.z); P.dhdz = madd32<false>(P.dhdz, w.z, v.w);
        P.lu = madd32<false>(P.lu, w.w, v.x); P.lv = madd32<false>(P.lv, w.w, v.y);
        P.lw = madd32<false>(P.lw, w.w, v.z); P.lh = madd32<false>(P.lh, w.w, v.w);
    };
    if constexpr (NNBR > 0) {
        constexpr int K = (NNBR + S - 1) / S;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int j = h + S * k;
            if (NNBR % S == 0 || j < NNBR)
                slot(j);
        }
    } else {
#pragma unroll 2
        for (int j = h; j < nn; j += S)
            slot(j);
    }
    Sums32 T;
    T.dudx = group_sum32<S>(P.dudx); T.dvdx = group_sum32<S>(P.dvdx);
    T.dwdx = group_sum32<S>(P.dwdx); T.dhdx = group_sum32<S>(P.dhdx);
    T.dudy = group_sum32<S>(P.dudy); T.dvdy = group_sum32<S>(P.dvdy);
    T.dwdy = group_sum32<S>(P.dwdy); T.dhdy = group_sum32<S>(P.dhdy);
    T.dudz = group_sum32<S>(P.dudz); T.dvdz = group_sum32<S>(P.dvdz);
    T.dwdz = group_sum32<S>(P.dwdz); T.dhdz = group_sum32<S>(P.dhdz);
    T.lu = group_sum32<S>(P.lu); T.lv = group_sum32<S>(P.lv);
    T.lw = group_sum32<S>(P.lw); T.lh = group_sum32<S>(P.lh);
    if (h != 0 || !live)
        return;
    const float4 own = *reinterpret_cast<const float4 *>(src + 4L * node);
    float g[GEO_N];
#pragma unroll
    for (int f = 0; f < GEO_N; ++f)
        g[f] = m.geo[(long)f * m.ld + node];
    float F[4];
    rhs_point32(T, own.x, own.y, own.z, own.w, g, m.gh0, F);
    stage_update32<STAGE, false>(st, 4L * node, F);
}

// Single state over LDS tiles (k_stage_ptile's scheme in FP32, one lane
// per node): a workgroup of TN lanes takes a tile of <= TN consecutive
// (Hilbert-ordered) nodes, copies the tile's <= 4 TN distinct neighbour
// states (16-byte rows, one direct-to-LDS load per row) into LDS, then sums
// every stencil slot in order from LDS through the 16-bit tile positions
// (node-major, 64 B per node) -- the same sums in the same order as
// k_stage_lane32 (bitwise in exact mode), with the 31 scattered gathers and
// 31 ids per node replaced by LDS reads and 64 B of positions.
template <int STAGE, bool EXACT, int NNBR, int TN>
__global__ __launch_bounds__(TN) void k_stage_tile32(DevMesh32 m, DevTiles tl, DevStage32 st)
{
    static_assert(NNBR > 0 && NNBR <= 31, "positions: 31 slots + the own row");
    static_assert(TN % 64 == 0, "whole waves");
    __shared__ float4 rows[4 * TN];
    const int tile = xcd_block(st.nblk);
    if (tile >= tl.ntiles)
        return;   // uniform over the workgroup
    const int t0 = tl.tile_start[tile];
    const int tcount = tl.tile_start[tile + 1] - t0;
    const int u0 = tl.ustart[tile];
    const int ucount = tl.ustart[tile + 1] - u0;
    const float *__restrict__ src = st.src;
    const int lane = (int)threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    // wave w copies rows [64 c, 64 c + 64) for c = w, w + waves, ...
    for (int c = wv; 64 * c < ucount; c += TN / 64) {
        const int r = 64 * c + lane;
        const int u = tl.ulist[u0 + (r < ucount ? r : ucount - 1)];
        if (r < ucount)
            dma16(src + 4L * u, reinterpret_cast<double *>(rows + 64 * c));
    }
    const int nl = (int)threadIdx.x;
    const bool live = nl < tcount;
    const int node = t0 + (live ? nl : tcount - 1);
    const long ld = m.sblk ? 16L : (long)m.ld;
    const long sb = m.sblk ? ((long)(node >> 4) * NNBR << 4) + (node & 15) : (long)node;
    const float *__restrict__ wp = m.wts + 4L * sb;
    const uint4 *pp = reinterpret_cast<const uint4 *>(tl.lidx_node + 32L * node);
    __builtin_amdgcn_s_waitcnt(0);   // this wave's row loads are in LDS
    __syncthreads();
    Sums32 S = {};
    // groups of 8 slots (one 16-byte load of their positions), rolled so
    // that ~8 slots of weight loads are in flight per lane: unrolled over
    // all 31 slots the weight loads are hoisted and take 266 registers
    unsigned ownw = 0xffffu;
#pragma unroll 1
    for (int g = 0; g < 4; ++g) {
        const uint4 pv = pp[g];
        const unsigned pw[4] = {pv.x, pv.y, pv.z, pv.w};
        if (g == 3)
            ownw = pv.w >> 16;   // spare slot 31: the node's own row
        static_for<8>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const int j = 8 * g + k;
            if (NNBR % 8 == 0 || j < NNBR) {
                const int pos = (int)((k & 1) ? (pw[k >> 1] >> 16) : (pw[k >> 1] & 0xffffu));
                const float4 w = *reinterpret_cast<const float4 *>(wp + 4L * j * ld);
                const float4 v = rows[pos];
                S.dudx = madd32<EXACT>(S.dudx, w.x, v.x); S.dvdx = madd32<EXACT>(S.dvdx, w.x, v.y);
                S.dwdx = madd32<EXACT>(S.dwdx, w.x, v.z); S.dhdx = madd32<EXACT>(S.dhdx, w.x, v.w);
                S.dudy = madd32<EXACT>(S.dudy, w.y, v.x); S.dvdy = madd32<EXACT>(S.dvdy, w.y, v.y);
                S.dwdy = madd32<EXACT>(S.dwdy, w.y, v.z); S.dhdy = madd32<EXACT>(S.dhdy, w.y, v.w);
                S.dudz = madd32<EXACT>(S.dudz, w.z, v.x); S.dvdz = madd32<EXACT>(S.dvdz, w.z, v.y);
                S.dwdz = madd32<EXACT>(S.dwdz, w.z, v.z); S.dhdz = madd32<EXACT>(S.dhdz, w.z, v.w);
                S.lu = madd32<EXACT>(S.lu, w.w, v.x); S.lv = madd32<EXACT>(S.lv, w.w, v.y);
                S.lw = madd32<EXACT>(S.lw, w.w, v.z); S.lh = madd32<EXACT>(S.lh, w.w, v.w);
            }
        });
    }
    // own state: the node's row, else a global load (nodes outside their
    // own stencil)
    float4 own;
    if (ownw != 0xffffu) {
        own = rows[ownw];
    } else {
        own = *reinterpret_cast<const float4 *>(src + 4L * node);
        asm volatile("" : "+v"(own.x), "+v"(own.y), "+v"(own.z), "+v"(own.w));
    }
    float g[GEO_N];
#pragma unroll
    for (int f = 0; f < GEO_N; ++f)
        g[f] = m.geo[(long)f * m.ld + node];
    float F[4];
    rhs_point32(S, own.x, own.y, own.z, own.w, g, m.gh0, F);
    if (live)
        stage_update32<STAGE, EXACT>(st, 4L * node, F);
}

}  // namespace swe
