// Device kernels of the RBF-FD shallow-water RK4 path for gfx950 (CDNA4).
//
// One RK stage = one launch: the 31-point stencil sums of Dx, Dy, Dz and the
// hyperviscosity operator L over the neighbour states, the pointwise SWE
// right-hand side (advection, Coriolis, pressure gradient, tangent
// projection, orography, continuity; swe_code/main/rk4_rbffd_swe.c:285-353)
// and the RK4 stage update of that node (rk4_rbffd_swe.c:30-186) fused into
// the epilogue, so the RHS is never written to memory.
//
// Device layouts (HBM):
//   nbr  int32  [nnbr][ld]        neighbour ids, stencil-slot major
//   wts  f64    [nnbr][ld][4]     (Dx, Dy, Dz, L) of one (slot, node): 32 B
//   geo  f64    [GEO_N][ld]       x y z f ghm p_u[3] p_v[3] p_w[3] gradghm[3]
//   state f64   [node][var][B]    var = u v w h, member fastest
// Every wave reads contiguous runs of the weight and index streams; the
// state gathers hit L2 because nodes are RCM-ordered and each XCD owns one
// contiguous node range (xcd_block).
//
// The build uses -ffp-contract=off.  EXACT kernels therefore round every
// product and sum separately, in the reference's order (sequential stencil
// slot j, rk4_rbffd_swe.c:285-316, expressions of :331-353, axpys of
// :108-183), and reproduce the reference CPU build bit for bit; the fast
// kernels use fused multiply-adds for the stencil sums and stage updates.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

namespace swe {

constexpr int GEO_N = 17;
enum GeoField {
    G_X = 0, G_Y, G_Z, G_F, G_GHM,
    G_PU = 5, G_PV = 8, G_PW = 11, G_GG = 14
};

// Kernel variants: the four RK stages and a plain RHS evaluation.
enum StageKind { ST_0 = 0, ST_1 = 1, ST_2 = 2, ST_3 = 3, ST_RHS = 4 };

struct DevMesh {
    const int *__restrict__ nbr;
    const double *__restrict__ wts;
    const double *__restrict__ geo;
    int n;        // nodes
    int ld;       // leading dimension of geo (and of nbr/wts when sblk == 0)
    int nnbr;     // stencil width (loop bound)
    int sblk;     // 0: nbr/wts slot-major [nnbr][ld]; 16: node blocks [n/16][nnbr][16]
    double gh0;
};

// Position of (slot 0, node) in nbr (and /4 in wts), and the distance
// between consecutive slots of one node, for both stencil layouts.
template <int NNBR>
__device__ __forceinline__ long slot_base(const DevMesh &m, int node)
{
    const long W = NNBR > 0 ? NNBR : m.nnbr;
    return m.sblk ? ((long)(node >> 4) * W << 4) + (node & 15) : (long)node;
}

__device__ __forceinline__ long slot_stride(const DevMesh &m)
{
    return m.sblk ? 16L : (long)m.ld;
}

struct DevStage {
    const double *src;  // gather source: H (stage 0) or K
    const double *h;    // H, own node (stage update)
    double *d;          // D, running RK sum
    double *out;        // next K, H (stage 3) or F (ST_RHS)
    double c;           // K = H + c*F (dt/2, dt/2, dt); stage 3: H += c*D (dt/6)
    double dcoef;       // D += dcoef*F (2, 2, 1)
    const int *active;  // ensemble: member slots [0, *active) are advanced
    int members;        // B: state stride of a (node, var) row
    int nblk;           // grid.x, a multiple of 8
};

// Workgroups are dispatched round-robin over the 8 XCDs.  Remap the
// hardware block id so that XCD x processes the contiguous logical blocks
// [x*nblk/8, (x+1)*nblk/8): its L2 then holds the state of one node band.
__device__ __forceinline__ int xcd_block(int nblk)
{
    const int b = blockIdx.x;
    return (b & 7) * (nblk >> 3) + (b >> 3);
}

template <bool EXACT>
__device__ __forceinline__ double madd(double acc, double w, double s)
{
    if constexpr (EXACT) {
        const double p = w * s;
        return acc + p;
    } else {
        return __builtin_fma(w, s, acc);
    }
}

// y + (a*x), the reference's axpy form.
template <bool EXACT>
__device__ __forceinline__ double axpy(double y, double a, double x)
{
    if constexpr (EXACT) {
        const double p = a * x;
        return y + p;
    } else {
        return __builtin_fma(a, x, y);
    }
}

struct Sums {
    double dudx, dvdx, dwdx, dhdx;
    double dudy, dvdy, dwdy, dhdy;
    double dudz, dvdz, dwdz, dhdz;
    double lu, lv, lw, lh;
};

struct Geo {
    double x, y, z, f, ghm;
    double pu0, pu1, pu2, pv0, pv1, pv2, pw0, pw1, pw2, gx, gy, gz;
};

__device__ __forceinline__ Geo load_geo(const double *__restrict__ geo, int ld, int i)
{
    Geo g;
    g.x = geo[G_X * (long)ld + i];
    g.y = geo[G_Y * (long)ld + i];
    g.z = geo[G_Z * (long)ld + i];
    g.f = geo[G_F * (long)ld + i];
    g.ghm = geo[G_GHM * (long)ld + i];
    g.pu0 = geo[(G_PU + 0) * (long)ld + i];
    g.pu1 = geo[(G_PU + 1) * (long)ld + i];
    g.pu2 = geo[(G_PU + 2) * (long)ld + i];
    g.pv0 = geo[(G_PV + 0) * (long)ld + i];
    g.pv1 = geo[(G_PV + 1) * (long)ld + i];
    g.pv2 = geo[(G_PV + 2) * (long)ld + i];
    g.pw0 = geo[(G_PW + 0) * (long)ld + i];
    g.pw1 = geo[(G_PW + 1) * (long)ld + i];
    g.pw2 = geo[(G_PW + 2) * (long)ld + i];
    g.gx = geo[(G_GG + 0) * (long)ld + i];
    g.gy = geo[(G_GG + 1) * (long)ld + i];
    g.gz = geo[(G_GG + 2) * (long)ld + i];
    return g;
}

// Pointwise right-hand side of one node, operation order of
// rk4_rbffd_swe.c:331-353 (no contraction in this build).
__device__ __forceinline__ void rhs_point(const Sums &S, double u, double v, double w,
                                          double h, const Geo &g, double gh0, double F[4])
{
    const double ru = -((u * S.dudx) + (v * S.dudy) + (w * S.dudz)
                        + (g.f * ((g.y * w) - (g.z * v))) + S.dhdx);
    const double rv = -((u * S.dvdx) + (v * S.dvdy) + (w * S.dvdz)
                        + (g.f * ((g.z * u) - (g.x * w))) + S.dhdy);
    const double rw = -((u * S.dwdx) + (v * S.dwdy) + (w * S.dwdz)
                        + (g.f * ((g.x * v) - (g.y * u))) + S.dhdz);
    F[0] = (g.pu0 * ru) + (g.pu1 * rv) + (g.pu2 * rw) + S.lu;
    F[1] = (g.pv0 * ru) + (g.pv1 * rv) + (g.pv2 * rw) + S.lv;
    F[2] = (g.pw0 * ru) + (g.pw1 * rv) + (g.pw2 * rw) + S.lw;
    F[3] = -((u * (S.dhdx - g.gx)) + (v * (S.dhdy - g.gy)) + (w * (S.dhdz - g.gz))
             + ((h + gh0 - g.ghm) * (S.dudx + S.dvdy + S.dwdz)))
           + S.lh;
}

// Stage update of one state value at linear index e (RK_substep,
// rk4_rbffd_swe.c:36-69):
//   stage 0: D = F;             K = H + (dt/2)*F
//   stage 1: D = D + 2*F;       K = H + (dt/2)*F
//   stage 2: D = D + 2*F;       K = H + dt*F
//   stage 3: D = D + 1*F;       H = H + (dt/6)*D   (D is dead afterwards)
template <int STAGE, bool EXACT>
__device__ __forceinline__ void stage_update(const DevStage &st, long e, double Fv)
{
    if constexpr (STAGE == ST_RHS) {
        st.out[e] = Fv;
    } else if constexpr (STAGE == ST_0) {
        const double hv = st.h[e];
        st.d[e] = Fv;
        st.out[e] = axpy<EXACT>(hv, st.c, Fv);
    } else if constexpr (STAGE == ST_1 || STAGE == ST_2) {
        const double hv = st.h[e];
        const double dv = axpy<EXACT>(st.d[e], st.dcoef, Fv);
        st.d[e] = dv;
        st.out[e] = axpy<EXACT>(hv, st.c, Fv);
    } else {
        const double hv = st.h[e];
        const double dv = axpy<EXACT>(st.d[e], st.dcoef, Fv);
        st.out[e] = axpy<EXACT>(hv, st.c, dv);
    }
}

// Stage update of two adjacent elements (e, e + 1; e even: two members, or
// two variables of one node) with 16-byte loads and stores; per element the
// arithmetic of stage_update.
template <int STAGE, bool EXACT>
__device__ __forceinline__ void stage_update_pair(const DevStage &st, long e, double F0, double F1)
{
    double2 *out = reinterpret_cast<double2 *>(st.out + e);
    if constexpr (STAGE == ST_RHS) {
        *out = make_double2(F0, F1);
    } else if constexpr (STAGE == ST_0) {
        const double2 hv = *reinterpret_cast<const double2 *>(st.h + e);
        *reinterpret_cast<double2 *>(st.d + e) = make_double2(F0, F1);
        *out = make_double2(axpy<EXACT>(hv.x, st.c, F0), axpy<EXACT>(hv.y, st.c, F1));
    } else if constexpr (STAGE == ST_1 || STAGE == ST_2) {
        const double2 hv = *reinterpret_cast<const double2 *>(st.h + e);
        double2 *dp = reinterpret_cast<double2 *>(st.d + e);
        const double2 d0 = *dp;
        *dp = make_double2(axpy<EXACT>(d0.x, st.dcoef, F0), axpy<EXACT>(d0.y, st.dcoef, F1));
        *out = make_double2(axpy<EXACT>(hv.x, st.c, F0), axpy<EXACT>(hv.y, st.c, F1));
    } else {
        const double2 hv = *reinterpret_cast<const double2 *>(st.h + e);
        const double2 d0 = *reinterpret_cast<const double2 *>(st.d + e);
        const double dv0 = axpy<EXACT>(d0.x, st.dcoef, F0);
        const double dv1 = axpy<EXACT>(d0.y, st.dcoef, F1);
        *out = make_double2(axpy<EXACT>(hv.x, st.c, dv0), axpy<EXACT>(hv.y, st.c, dv1));
    }
}

// Broadcast lane K of every aligned group of 4 lanes (DPP quad_perm).
template <int K>
__device__ __forceinline__ double quad_bcast(double v)
{
    constexpr int ctrl = K * 0x55;  // quad_perm [K, K, K, K]
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

template <int K>
__device__ __forceinline__ int quad_bcast_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, K * 0x55, 0xF, 0xF, false);
}

// f(std::integral_constant<int, 0>) ... f(std::integral_constant<int, N-1>),
// in order: a loop whose index is a compile-time constant in the body.
template <typename F, int... J>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, J...>)
{
    (f(std::integral_constant<int, J>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Geometry of a node, read once per quad: lane k loads fields k, k+4, ...
// and the quad shares them by DPP (17 loads per node instead of 4 x 17).
__device__ __forceinline__ Geo load_geo_quad(const double *__restrict__ geo, int ld, int i, int k)
{
    double gv[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        const int f = 4 * r + k;
        gv[r] = geo[(long)(f < GEO_N ? f : 0) * ld + i];
    }
    double gf[GEO_N];
    static_for<GEO_N>([&](auto fc) {
        constexpr int f = decltype(fc)::value;
        gf[f] = quad_bcast<f & 3>(gv[f >> 2]);
    });
    Geo g;
    g.x = gf[G_X]; g.y = gf[G_Y]; g.z = gf[G_Z]; g.f = gf[G_F]; g.ghm = gf[G_GHM];
    g.pu0 = gf[G_PU]; g.pu1 = gf[G_PU + 1]; g.pu2 = gf[G_PU + 2];
    g.pv0 = gf[G_PV]; g.pv1 = gf[G_PV + 1]; g.pv2 = gf[G_PV + 2];
    g.pw0 = gf[G_PW]; g.pw1 = gf[G_PW + 1]; g.pw2 = gf[G_PW + 2];
    g.gx = gf[G_GG]; g.gy = gf[G_GG + 1]; g.gz = gf[G_GG + 2];
    return g;
}

// Stencil-slot loop shared by the stage kernels.  PF == 0 leaves the load
// schedule to the compiler.  PF = D > 0 writes the loop as a ring of D slots:
// all neighbour ids are loaded first, each iteration consumes the oldest
// slot and issues the loads of slot j + D, with scheduling barriers between
// iterations.  The barriers carry no data dependency, so they shape the
// compiler's schedule without fixing it (in exact mode they keep it from
// precomputing all 31x16 products ahead of the serial sums); the kernels
// that need a firm rolling window (k_stage_mpair, k_stage_tspec) instead tie
// each next address to the running sums with an empty asm.  Accumulation
// is sequential in j for every sum in all forms.
template <int NNBR, int PF, typename LoadFn, typename AccFn, typename Slot>
__device__ __forceinline__ void stencil_loop(const int *__restrict__ nb, long ld, int nn,
                                             LoadFn load, AccFn acc, Slot *)
{
    if constexpr (NNBR > 0 && PF > 0) {
        int q[NNBR];
#pragma unroll
        for (int j = 0; j < NNBR; ++j)
            q[j] = nb[j * ld];
        Slot ring[PF];
#pragma unroll
        for (int j = 0; j < PF && j < NNBR; ++j)
            ring[j] = load(j, q[j]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < NNBR; ++j) {
            acc(ring[j % PF]);
            if (j + PF < NNBR)
                ring[j % PF] = load(j + PF, q[j + PF]);
            __builtin_amdgcn_sched_barrier(0);
        }
    } else if constexpr (NNBR > 0) {
#pragma unroll
        for (int j = 0; j < NNBR; ++j)
            acc(load(j, nb[j * ld]));
    } else {
#pragma unroll 4
        for (int j = 0; j < nn; ++j)
            acc(load(j, nb[j * ld]));
    }
}

struct QuadWSlot {
    double w;
    double s;
};

// ---------------------------------------------------------------------------
// Single state (B == 1), quad mapping with shared loads (any stencil width):
// four lanes per node, lane k reading variable k of every neighbour and
// keeping the Dx, Dy, Dz, L sums of that variable (each a sequential sum
// over the stencil).  Lane k loads only weight k of each slot -- one 8-byte
// load per lane, a wave's 16 nodes x 32 B read once -- and the quad
// exchanges the four weights by DPP; the stencil ids, the node's own state
// and its geometry are likewise read once per quad and shared by DPP.  A
// wave64 memory instruction costs the CU's address unit the same time
// whatever it fetches, so every load removed counts.
template <int STAGE, bool EXACT, int NNBR, int PF>
__global__ __launch_bounds__(256) void k_stage_quadw(DevMesh m, DevStage st)
{
    const int t = xcd_block(st.nblk) * (int)blockDim.x + (int)threadIdx.x;
    const int node_raw = t >> 2;
    const int k = t & 3;
    const bool live = node_raw < m.n;
    const int node = live ? node_raw : m.n - 1;
    const int nn = NNBR > 0 ? NNBR : m.nnbr;
    const long ld = slot_stride(m);
    const long sb = slot_base<NNBR>(m, node);
    const int *__restrict__ nb = m.nbr + sb;
    const double *__restrict__ wp = m.wts + 4L * sb + k;
    const double *__restrict__ src = st.src;

    double sx = 0.0, sy = 0.0, sz = 0.0, sl = 0.0;
    auto load = [&](int j, int q) {
        QuadWSlot r;
        r.w = wp[4L * j * ld];
        r.s = src[4L * q + k];
        return r;
    };
    auto acc = [&](const QuadWSlot &r) {
        const double wx = quad_bcast<0>(r.w), wy = quad_bcast<1>(r.w);
        const double wz = quad_bcast<2>(r.w), wl = quad_bcast<3>(r.w);
        sx = madd<EXACT>(sx, wx, r.s);
        sy = madd<EXACT>(sy, wy, r.s);
        sz = madd<EXACT>(sz, wz, r.s);
        sl = madd<EXACT>(sl, wl, r.s);
    };
    if constexpr (NNBR > 0) {
        // Stencil ids once per quad too: lane k loads slots k, k+4, ...;
        // slot j's id comes from lane j & 3 by DPP.
        constexpr int NQ = (NNBR + 3) / 4;
        int qv[NQ];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const int j = 4 * i + k;
            qv[i] = nb[(j < NNBR ? j : 0) * ld];
        }
        static_for<NNBR>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            acc(load(j, quad_bcast_i<j & 3>(qv[j >> 2])));
        });
    } else {
        stencil_loop<NNBR, PF>(nb, ld, nn, load, acc, (QuadWSlot *)nullptr);
    }
    Sums S;
    S.dudx = quad_bcast<0>(sx); S.dudy = quad_bcast<0>(sy);
    S.dudz = quad_bcast<0>(sz); S.lu = quad_bcast<0>(sl);
    S.dvdx = quad_bcast<1>(sx); S.dvdy = quad_bcast<1>(sy);
    S.dvdz = quad_bcast<1>(sz); S.lv = quad_bcast<1>(sl);
    S.dwdx = quad_bcast<2>(sx); S.dwdy = quad_bcast<2>(sy);
    S.dwdz = quad_bcast<2>(sz); S.lw = quad_bcast<2>(sl);
    S.dhdx = quad_bcast<3>(sx); S.dhdy = quad_bcast<3>(sy);
    S.dhdz = quad_bcast<3>(sz); S.lh = quad_bcast<3>(sl);

    const double own = src[4L * node + k];
    const Geo g = load_geo_quad(m.geo, m.ld, node, k);
    double F[4];
    rhs_point(S, quad_bcast<0>(own), quad_bcast<1>(own), quad_bcast<2>(own), quad_bcast<3>(own),
              g, m.gh0, F);
    const double Fk = k == 0 ? F[0] : k == 1 ? F[1] : k == 2 ? F[2] : F[3];
    if (live)
        stage_update<STAGE, EXACT>(st, 4L * node + k, Fk);
}

// DPP within aligned lane pairs: swap the two lanes, or broadcast lane K.
__device__ __forceinline__ double pair_swap(double v)
{
    constexpr int ctrl = 0xB1;  // quad_perm [1, 0, 3, 2]
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int K>
__device__ __forceinline__ int pair_bcast_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, K ? 0xF5 : 0xA0, 0xF, 0xF, false);  // [K, K, 2+K, 2+K]
}
template <int K>
__device__ __forceinline__ double pair_bcast(double v)
{
    constexpr int ctrl = K ? 0xF5 : 0xA0;
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// DPP row_newbcast (64-bit moves): every lane of a 16-lane row takes lane
// K's value of that row.
template <int K>
__device__ __forceinline__ double row_bcast64(double v)
{
    static_assert(K >= 0 && K < 16, "lane within a row");
    return __builtin_amdgcn_mov_dpp(v, 0x150 + K, 0xF, 0xF, false);
}

// ---------------------------------------------------------------------------
// Single state (B == 1), pair mapping: two lanes per node.  Lane k owns
// variables 2k, 2k+1 and weights 2k, 2k+1 of every slot: per slot it issues
// one 16-byte weight load, one 16-byte gather (the two variables of the
// neighbour) and every other slot's stencil id, and takes the other two
// weights from its partner by DPP -- 2.5 global load instructions per slot
// for 32 nodes, where the quad mapping needs 4.5.  A lane keeps 8 sequential
// sums: its two variables against its own two weights and against the
// partner's; the 16 sums, the own state and the geometry are exchanged by
// DPP once per node, each lane evaluates the RHS and updates its two
// variables with one 16-byte access.  Same sums, same order as the other
// single-state kernels (bitwise in exact mode).
template <int STAGE, bool EXACT, int NNBR>
__global__ __launch_bounds__(256) void k_stage_pair(DevMesh m, DevStage st)
{
    static_assert(NNBR > 0, "compile-time stencil width");
    const int t = xcd_block(st.nblk) * (int)blockDim.x + (int)threadIdx.x;
    const int node_raw = t >> 1;
    const int k = t & 1;
    const bool live = node_raw < m.n;
    const int node = live ? node_raw : m.n - 1;
    const long ld = slot_stride(m);
    const long sb = slot_base<NNBR>(m, node);
    const int *__restrict__ nb = m.nbr + sb;
    const double *__restrict__ wp = m.wts + 4L * sb + 2 * k;
    const double *__restrict__ src = st.src;

    constexpr int NQ = (NNBR + 1) / 2;
    int qv[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int j = 2 * i + k;
        qv[i] = nb[(j < NNBR ? j : 0) * ld];
    }
    // own0/own1: this lane's weights (2k, 2k+1); oth0/oth1: the partner's
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;   // variable 2k   x (own0, own1, oth0, oth1)
    double b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;   // variable 2k+1 x (own0, own1, oth0, oth1)
    static_for<NNBR>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const int q = pair_bcast_i<j & 1>(qv[j >> 1]);
        const double2 w = *reinterpret_cast<const double2 *>(wp + 4L * j * ld);
        const double2 s = *reinterpret_cast<const double2 *>(src + 4L * q + 2 * k);
        const double o0 = pair_swap(w.x), o1 = pair_swap(w.y);
        a0 = madd<EXACT>(a0, w.x, s.x); a1 = madd<EXACT>(a1, w.y, s.x);
        a2 = madd<EXACT>(a2, o0, s.x);  a3 = madd<EXACT>(a3, o1, s.x);
        b0 = madd<EXACT>(b0, w.x, s.y); b1 = madd<EXACT>(b1, w.y, s.y);
        b2 = madd<EXACT>(b2, o0, s.y);  b3 = madd<EXACT>(b3, o1, s.y);
    });
    // (dx, dy, dz, l) of this lane's variables: lane 0 owns (Dx, Dy), lane 1 (Dz, L)
    const double ax = k ? a2 : a0, ay = k ? a3 : a1, az = k ? a0 : a2, al = k ? a1 : a3;
    const double bx = k ? b2 : b0, by = k ? b3 : b1, bz = k ? b0 : b2, bl = k ? b1 : b3;
    // partner's variables
    const double pax = pair_swap(ax), pay = pair_swap(ay), paz = pair_swap(az), pal = pair_swap(al);
    const double pbx = pair_swap(bx), pby = pair_swap(by), pbz = pair_swap(bz), pbl = pair_swap(bl);
    Sums S;
    // variables u, v live on lane 0 (a, b); w, h on lane 1
    S.dudx = k ? pax : ax; S.dudy = k ? pay : ay; S.dudz = k ? paz : az; S.lu = k ? pal : al;
    S.dvdx = k ? pbx : bx; S.dvdy = k ? pby : by; S.dvdz = k ? pbz : bz; S.lv = k ? pbl : bl;
    S.dwdx = k ? ax : pax; S.dwdy = k ? ay : pay; S.dwdz = k ? az : paz; S.lw = k ? al : pal;
    S.dhdx = k ? bx : pbx; S.dhdy = k ? by : pby; S.dhdz = k ? bz : pbz; S.lh = k ? bl : pbl;

    const double2 own = *reinterpret_cast<const double2 *>(src + 4L * node + 2 * k);
    const double po0 = pair_swap(own.x), po1 = pair_swap(own.y);
    const double u = k ? po0 : own.x, v = k ? po1 : own.y;
    const double w = k ? own.x : po0, h = k ? own.y : po1;
    double gv[(GEO_N + 1) / 2];
#pragma unroll
    for (int r = 0; r < (GEO_N + 1) / 2; ++r) {
        const int f = 2 * r + k;
        gv[r] = m.geo[(long)(f < GEO_N ? f : 0) * m.ld + node];
    }
    double gf[GEO_N];
    static_for<GEO_N>([&](auto fc) {
        constexpr int f = decltype(fc)::value;
        gf[f] = pair_bcast<f & 1>(gv[f >> 1]);
    });
    Geo g;
    g.x = gf[G_X]; g.y = gf[G_Y]; g.z = gf[G_Z]; g.f = gf[G_F]; g.ghm = gf[G_GHM];
    g.pu0 = gf[G_PU]; g.pu1 = gf[G_PU + 1]; g.pu2 = gf[G_PU + 2];
    g.pv0 = gf[G_PV]; g.pv1 = gf[G_PV + 1]; g.pv2 = gf[G_PV + 2];
    g.pw0 = gf[G_PW]; g.pw1 = gf[G_PW + 1]; g.pw2 = gf[G_PW + 2];
    g.gx = gf[G_GG]; g.gy = gf[G_GG + 1]; g.gz = gf[G_GG + 2];
    double F[4];
    rhs_point(S, u, v, w, h, g, m.gh0, F);
    if (live)
        stage_update_pair<STAGE, EXACT>(st, 4L * node + 2 * k, k ? F[2] : F[0], k ? F[3] : F[1]);
}

// ---------------------------------------------------------------------------
// Ensemble (B >= 2): lane = (node, member), MB members of one node per lane
// group, member block blockIdx.y.  Each lane keeps all 16 sums of its
// member; the stencil weights and ids are shared by the MB lanes of a node
// (wave-uniform scalar loads when MB == 64), so one weight read from HBM
// serves every member of the block.  The neighbour state of a variable is a
// contiguous run of MB doubles.  Member blocks below 64, and 64 with a
// run-time stencil width (31-point stencils at 64 use k_stage_mpair).
struct MemberSlot {
    double2 w01, w23;
    double su, sv, sw, sh;
};

template <int STAGE, bool EXACT, int MB, int NNBR, int PF>
__global__ __launch_bounds__(256) void k_stage_members(DevMesh m, DevStage st)
{
    static_assert(MB >= 2 && MB <= 64 && (MB & (MB - 1)) == 0, "MB power of two");
    if ((int)blockIdx.y * MB >= *st.active)
        return;   // whole member block idle (uniform over the workgroup)
    const int t = xcd_block(st.nblk) * 256 + (int)threadIdx.x;
    int node_raw = t / MB;
    if constexpr (MB == 64)
        node_raw = __builtin_amdgcn_readfirstlane(node_raw);
    const int mem = (int)blockIdx.y * MB + (t & (MB - 1));
    const bool live = node_raw < m.n;
    const int node = live ? node_raw : m.n - 1;
    const int nn = NNBR > 0 ? NNBR : m.nnbr;
    const long ld = slot_stride(m);
    const long sb = slot_base<NNBR>(m, node);
    const long B = st.members;
    const int *__restrict__ nb = m.nbr + sb;
    const double *__restrict__ wp = m.wts + 4L * sb;
    const double *__restrict__ src = st.src + mem;

    Sums S = {};
    auto load = [&](int j, int q) {
        if constexpr (MB == 64)
            q = __builtin_amdgcn_readfirstlane(q);
        MemberSlot r;
        r.w01 = *reinterpret_cast<const double2 *>(wp + 4L * j * ld);
        r.w23 = *reinterpret_cast<const double2 *>(wp + 4L * j * ld + 2);
        const double *sq = src + 4L * q * B;
        r.su = sq[0];
        r.sv = sq[B];
        r.sw = sq[2 * B];
        r.sh = sq[3 * B];
        return r;
    };
    auto acc = [&](const MemberSlot &r) {
        const double su = r.su, sv = r.sv, sw = r.sw, sh = r.sh;
        S.dudx = madd<EXACT>(S.dudx, r.w01.x, su); S.dvdx = madd<EXACT>(S.dvdx, r.w01.x, sv);
        S.dwdx = madd<EXACT>(S.dwdx, r.w01.x, sw); S.dhdx = madd<EXACT>(S.dhdx, r.w01.x, sh);
        S.dudy = madd<EXACT>(S.dudy, r.w01.y, su); S.dvdy = madd<EXACT>(S.dvdy, r.w01.y, sv);
        S.dwdy = madd<EXACT>(S.dwdy, r.w01.y, sw); S.dhdy = madd<EXACT>(S.dhdy, r.w01.y, sh);
        S.dudz = madd<EXACT>(S.dudz, r.w23.x, su); S.dvdz = madd<EXACT>(S.dvdz, r.w23.x, sv);
        S.dwdz = madd<EXACT>(S.dwdz, r.w23.x, sw); S.dhdz = madd<EXACT>(S.dhdz, r.w23.x, sh);
        S.lu = madd<EXACT>(S.lu, r.w23.y, su); S.lv = madd<EXACT>(S.lv, r.w23.y, sv);
        S.lw = madd<EXACT>(S.lw, r.w23.y, sw); S.lh = madd<EXACT>(S.lh, r.w23.y, sh);
    };
    stencil_loop<NNBR, PF>(nb, ld, nn, load, acc, (MemberSlot *)nullptr);
    const double *own = src + 4L * node * B;
    const double u = own[0], v = own[B], w = own[2 * B], h = own[3 * B];
    const Geo g = load_geo(m.geo, m.ld, node);
    double F[4];
    rhs_point(S, u, v, w, h, g, m.gh0, F);
    if (live) {
        const long e = 4L * node * B + mem;
#pragma unroll
        for (int var = 0; var < 4; ++var)
            stage_update<STAGE, EXACT>(st, e + var * B, F[var]);
    }
}

// One stencil slot into the 16 sums (the reference's term order).
template <bool EXACT>
__device__ __forceinline__ void acc_slot(Sums &S, double2 w01, double2 w23, double su, double sv,
                                         double sw, double sh)
{
    S.dudx = madd<EXACT>(S.dudx, w01.x, su); S.dvdx = madd<EXACT>(S.dvdx, w01.x, sv);
    S.dwdx = madd<EXACT>(S.dwdx, w01.x, sw); S.dhdx = madd<EXACT>(S.dhdx, w01.x, sh);
    S.dudy = madd<EXACT>(S.dudy, w01.y, su); S.dvdy = madd<EXACT>(S.dvdy, w01.y, sv);
    S.dwdy = madd<EXACT>(S.dwdy, w01.y, sw); S.dhdy = madd<EXACT>(S.dhdy, w01.y, sh);
    S.dudz = madd<EXACT>(S.dudz, w23.x, su); S.dvdz = madd<EXACT>(S.dvdz, w23.x, sv);
    S.dwdz = madd<EXACT>(S.dwdz, w23.x, sw); S.dhdz = madd<EXACT>(S.dhdz, w23.x, sh);
    S.lu = madd<EXACT>(S.lu, w23.y, su); S.lv = madd<EXACT>(S.lv, w23.y, sv);
    S.lw = madd<EXACT>(S.lw, w23.y, sw); S.lh = madd<EXACT>(S.lh, w23.y, sh);
}

// ---------------------------------------------------------------------------
// Ensemble, member pairs (member block 64, B a multiple of 64).  A wave64
// memory instruction costs the CU's texture addresser a fixed ~16 cycles
// whatever its width, so an 8-byte-per-lane gather (one member per lane)
// moves only 32 B per addresser cycle.  Here lane (half h, pair p) of a wave
// holds members 2p, 2p+1 of node 2*wave + h and gathers them with one
// 16-byte load per variable: 4 loads per slot for 2 nodes x 64 members,
// half the instructions of one member per lane.  Each half-wave's node
// weights sit in a wave-private LDS row read back with uniform-per-half
// broadcasts; its neighbour ids come from one vector load and v_readlane.
// Per member the sums and the stage update are those of k_stage_members.
template <int STAGE, bool EXACT, int NNBR, int DEPTH>
__global__ __launch_bounds__(256) void k_stage_mpair(DevMesh m, DevStage st)
{
    static_assert(NNBR > 0 && NNBR <= 32, "one slot per half-wave lane");
    static_assert(DEPTH >= 1, "pipeline depth");
    if ((int)blockIdx.y * 64 >= *st.active)
        return;   // whole member block idle (uniform over the workgroup)
    const int lane = (int)threadIdx.x & 63;
    const int wid = (int)threadIdx.x >> 6;
    const int half = lane >> 5;
    const int p = lane & 31;
    const int t = xcd_block(st.nblk) * 256 + (int)threadIdx.x;
    const int node_raw = t >> 5;
    const bool live = node_raw < m.n;
    const int node = live ? node_raw : m.n - 1;
    const long ld = slot_stride(m);
    const long sb = slot_base<NNBR>(m, node);
    const long B = st.members;
    const int mem = (int)blockIdx.y * 64 + 2 * p;
    const double *__restrict__ src = st.src + mem;
    const int *__restrict__ nb = m.nbr + sb;
    const double *__restrict__ wp = m.wts + 4L * sb;

    __shared__ double2 wrow[4][2][2 * NNBR];
    const int jq = p < NNBR ? p : 0;
    const int qv = nb[jq * ld];
    {
        const double2 w01 = *reinterpret_cast<const double2 *>(wp + 4L * jq * ld);
        const double2 w23 = *reinterpret_cast<const double2 *>(wp + 4L * jq * ld + 2);
        if (p < NNBR) {
            wrow[wid][half][2 * p] = w01;
            wrow[wid][half][2 * p + 1] = w23;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    // Rolling pipeline DEPTH slots deep: slot j + DEPTH is issued once slot j
    // is consumed (the empty asm ties its id to all 32 sums after slot j).
    auto id_of = [&](int j) {
        const int qa = __builtin_amdgcn_readlane(qv, j);
        const int qb = __builtin_amdgcn_readlane(qv, 32 + j);
        return half ? qb : qa;
    };
    double2 ru[DEPTH], rv[DEPTH], rw[DEPTH], rh[DEPTH];
    auto issue = [&](int j, int q) {
        const double *sq = src + 4L * q * B;
        ru[j % DEPTH] = *reinterpret_cast<const double2 *>(sq);
        rv[j % DEPTH] = *reinterpret_cast<const double2 *>(sq + B);
        rw[j % DEPTH] = *reinterpret_cast<const double2 *>(sq + 2 * B);
        rh[j % DEPTH] = *reinterpret_cast<const double2 *>(sq + 3 * B);
    };
    Sums S0 = {}, S1 = {};
#pragma unroll
    for (int j = 0; j < DEPTH && j < NNBR; ++j)
        issue(j, id_of(j));
#pragma unroll
    for (int j = 0; j < NNBR; ++j) {
        const double2 su = ru[j % DEPTH], sv = rv[j % DEPTH], sw = rw[j % DEPTH],
                      sh = rh[j % DEPTH];
        const double2 w01 = wrow[wid][half][2 * j];
        const double2 w23 = wrow[wid][half][2 * j + 1];
        acc_slot<EXACT>(S0, w01, w23, su.x, sv.x, sw.x, sh.x);
        acc_slot<EXACT>(S1, w01, w23, su.y, sv.y, sw.y, sh.y);
        if (j + DEPTH < NNBR) {
            int q = id_of(j + DEPTH);
            asm volatile(""
                         : "+v"(q)
                         : "v"(S0.dudx), "v"(S0.dvdx), "v"(S0.dwdx), "v"(S0.dhdx),
                           "v"(S0.dudy), "v"(S0.dvdy), "v"(S0.dwdy), "v"(S0.dhdy),
                           "v"(S0.dudz), "v"(S0.dvdz), "v"(S0.dwdz), "v"(S0.dhdz),
                           "v"(S0.lu), "v"(S0.lv), "v"(S0.lw), "v"(S0.lh));
            asm volatile(""
                         : "+v"(q)
                         : "v"(S1.dudx), "v"(S1.dvdx), "v"(S1.dwdx), "v"(S1.dhdx),
                           "v"(S1.dudy), "v"(S1.dvdy), "v"(S1.dwdy), "v"(S1.dhdy),
                           "v"(S1.dudz), "v"(S1.dvdz), "v"(S1.dwdz), "v"(S1.dhdz),
                           "v"(S1.lu), "v"(S1.lv), "v"(S1.lw), "v"(S1.lh));
            issue(j + DEPTH, q);
        }
    }
    const double *own = src + 4L * node * B;
    const double2 u = *reinterpret_cast<const double2 *>(own);
    const double2 v = *reinterpret_cast<const double2 *>(own + B);
    const double2 w = *reinterpret_cast<const double2 *>(own + 2 * B);
    const double2 h = *reinterpret_cast<const double2 *>(own + 3 * B);
    const Geo g = load_geo(m.geo, m.ld, node);
    double F0[4], F1[4];
    rhs_point(S0, u.x, v.x, w.x, h.x, g, m.gh0, F0);
    rhs_point(S1, u.y, v.y, w.y, h.y, g, m.gh0, F1);
    if (live) {
        const long e = 4L * node * B + mem;
#pragma unroll
        for (int var = 0; var < 4; ++var)
            stage_update_pair<STAGE, EXACT>(st, e + var * B, F0[var], F1[var]);
    }
}

// ---------------------------------------------------------------------------
// LDS tiles: the device mesh is stored in a space-filling-curve node order
// and cut into tiles of consecutive nodes whose stencils share few
// neighbours; a tile's distinct neighbour states are copied into LDS once
// and every stencil entry is read from there.  Used by the single-state tile
// kernels (k_stage_ptile, k_stage_tile32) and the member-lane ensemble
// kernel (k_stage_tspec).
struct DevTiles {
    const int *tile_start;          // [ntiles + 1] first node of each tile
    const int *ustart;              // [ntiles + 1] offsets into ulist
    const int *ulist;               // distinct stencil nodes of each tile, ascending
    const int *ulist_tile;          // the same at a fixed stride: [ntiles][kTmemStride],
                                    // zero-padded (member-lane kernel only)
    const unsigned short *lidx;     // [nnbr][ld] position of nbr(j, node) in its tile's list
    const unsigned short *lidx_node;  // [n][32] the same, node-major (64 B per node)
    const double *wts_node;         // [n][32][4] weights, node-major (member-lane kernel only)
    const double *geo_node;         // [n][32] geometry 0..16, positions in 24..31 (same)
    int ntiles;
    int nmb;                        // member blocks per tile (B / MB)
};

// 16 bytes per lane from global memory straight into LDS (gfx950
// global_load_lds_dwordx4): lane l's bytes land at lds + 16 l, lds being
// wave-uniform.  No VGPR holds the data and no LDS write instruction is
// issued; the wave's vmcnt counts the load until the bytes are in LDS.
__device__ __forceinline__ void dma16(const void *src, double *lds)
{
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)lds, 16, 0,
                                     0);
}

// ---------------------------------------------------------------------------
// Single state (B == 1), pair mapping over LDS tiles.  The device mesh is in
// a space-filling-curve order cut into tiles of <= TN consecutive nodes
// whose stencils touch <= 4 TN distinct nodes (~3 per node at 128-node
// Hilbert tiles, against 31 stencil entries per node).  A workgroup (2 TN
// lanes) copies its tile's distinct neighbour states (32 B rows) into LDS,
// then runs the pair mapping of k_stage_pair with every neighbour read from
// LDS through the 16-bit tile positions (node-major, 64 B per node, two
// 16-byte loads per lane): the 31 scattered global gathers per node, which
// keep the texture path of k_stage_pair busy at large N, become LDS reads.
// Own state: the node's row (position in the spare slot 31), else a global
// load.  Same sums, same order as k_stage_pair (bitwise in exact mode).
// LOAD selects the staging schedule: 0 rows through registers; 1 rows by
// direct-to-LDS loads; 2 the same plus the weights of the first kPtilePre
// slots loaded before the barrier that waits for the rows.
constexpr int kPtileNodes = 128;   // default nodes per tile (4 waves x 32 nodes)
constexpr int kPtilePre = 8;

template <int TN>
constexpr int ptile_rows() { return 4 * TN; }

template <int STAGE, bool EXACT, int NNBR, int TN, int LOAD>
__global__ __launch_bounds__(2 * TN) void k_stage_ptile(DevMesh m, DevTiles tl, DevStage st)
{
    static_assert(NNBR > 0 && NNBR <= 31, "positions: 31 slots + the own row");
    static_assert(TN % 32 == 0, "whole waves");
    __shared__ double2 rows[2 * ptile_rows<TN>()];   // [row][half] = (u, v), (w, h)
    const int tile = xcd_block(st.nblk);
    if (tile >= tl.ntiles)
        return;   // uniform over the workgroup
    const int t0 = tl.tile_start[tile];
    const int tcount = tl.tile_start[tile + 1] - t0;
    const int u0 = tl.ustart[tile];
    const int ucount = tl.ustart[tile + 1] - u0;
    const double *__restrict__ src = st.src;
    if constexpr (LOAD == 0) {
        for (int i = (int)threadIdx.x; i < 2 * ucount; i += (int)blockDim.x) {
            const int u = tl.ulist[u0 + (i >> 1)];
            rows[i] = *reinterpret_cast<const double2 *>(src + 4L * u + 2 * (i & 1));
        }
    } else {
        // wave w copies rows [32 c, 32 c + 32) for c = w, w + waves, ...: lane
        // l's 16 bytes are half (l & 1) of row 32 c + l / 2, landing at
        // rows + 64 c + l (double2)
        const int lane = (int)threadIdx.x & 63;
        const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
        for (int c = wv; 32 * c < ucount; c += TN / 32) {
            const int r = 32 * c + (lane >> 1);
            const int u = tl.ulist[u0 + (r < ucount ? r : ucount - 1)];
            if (r < ucount)
                dma16(src + 4L * u + 2 * (lane & 1), reinterpret_cast<double *>(rows + 64 * c));
        }
    }
    const int nl = (int)threadIdx.x >> 1;
    const int k = (int)threadIdx.x & 1;
    const bool live = nl < tcount;
    const int node = t0 + (live ? nl : tcount - 1);   // dead pairs stay whole for DPP
    const long ld = slot_stride(m);
    const double *__restrict__ wp = m.wts + 4L * slot_base<NNBR>(m, node) + 2 * k;
    // positions of slots 16k .. 16k + 15 (two per dword)
    const uint4 *pp = reinterpret_cast<const uint4 *>(tl.lidx_node + 32L * node) + 2 * k;
    const uint4 pa = pp[0], pb = pp[1];
    const unsigned pw[8] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w};
    constexpr int NPRE = LOAD == 2 ? kPtilePre : 0;
    double2 wpre[NPRE > 0 ? NPRE : 1];
    static_for<NPRE>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        wpre[j] = *reinterpret_cast<const double2 *>(wp + 4L * j * ld);
    });
    if constexpr (LOAD != 0)
        __builtin_amdgcn_s_waitcnt(0);   // this wave's row loads are in LDS
    __syncthreads();
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;   // variable 2k   x (own0, own1, oth0, oth1)
    double b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;   // variable 2k+1 x (own0, own1, oth0, oth1)
    static_for<NNBR>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const unsigned wv = (unsigned)pair_bcast_i<j / 16>((int)pw[(j % 16) / 2]);
        const int pos = (int)((j & 1) ? (wv >> 16) : (wv & 0xffffu));
        double2 w;
        if constexpr (j < NPRE)
            w = wpre[j];
        else
            w = *reinterpret_cast<const double2 *>(wp + 4L * j * ld);
        const double2 s = rows[2 * pos + k];
        const double o0 = pair_swap(w.x), o1 = pair_swap(w.y);
        a0 = madd<EXACT>(a0, w.x, s.x); a1 = madd<EXACT>(a1, w.y, s.x);
        a2 = madd<EXACT>(a2, o0, s.x);  a3 = madd<EXACT>(a3, o1, s.x);
        b0 = madd<EXACT>(b0, w.x, s.y); b1 = madd<EXACT>(b1, w.y, s.y);
        b2 = madd<EXACT>(b2, o0, s.y);  b3 = madd<EXACT>(b3, o1, s.y);
    });
    const double ax = k ? a2 : a0, ay = k ? a3 : a1, az = k ? a0 : a2, al = k ? a1 : a3;
    const double bx = k ? b2 : b0, by = k ? b3 : b1, bz = k ? b0 : b2, bl = k ? b1 : b3;
    const double pax = pair_swap(ax), pay = pair_swap(ay), paz = pair_swap(az), pal = pair_swap(al);
    const double pbx = pair_swap(bx), pby = pair_swap(by), pbz = pair_swap(bz), pbl = pair_swap(bl);
    Sums S;
    S.dudx = k ? pax : ax; S.dudy = k ? pay : ay; S.dudz = k ? paz : az; S.lu = k ? pal : al;
    S.dvdx = k ? pbx : bx; S.dvdy = k ? pby : by; S.dvdz = k ? pbz : bz; S.lv = k ? pbl : bl;
    S.dwdx = k ? ax : pax; S.dwdy = k ? ay : pay; S.dwdz = k ? az : paz; S.lw = k ? al : pal;
    S.dhdx = k ? bx : pbx; S.dhdy = k ? by : pby; S.dhdz = k ? bz : pbz; S.lh = k ? bl : pbl;
    // own row: position in the spare slot 31 (lane 1's last dword, high half)
    const unsigned ownw = (unsigned)pair_bcast_i<1>((int)pw[7]) >> 16;
    double2 own;
    if (ownw != 0xffffu) {
        own = rows[2 * (int)ownw + k];
    } else {
        own = *reinterpret_cast<const double2 *>(src + 4L * node + 2 * k);
        asm volatile("" : "+v"(own.x), "+v"(own.y));   // no merged flat load (see tmem_compute)
    }
    const double po0 = pair_swap(own.x), po1 = pair_swap(own.y);
    const double u = k ? po0 : own.x, v = k ? po1 : own.y;
    const double w = k ? own.x : po0, h = k ? own.y : po1;
    double gv[(GEO_N + 1) / 2];
#pragma unroll
    for (int r = 0; r < (GEO_N + 1) / 2; ++r) {
        const int f = 2 * r + k;
        gv[r] = m.geo[(long)(f < GEO_N ? f : 0) * m.ld + node];
    }
    double gf[GEO_N];
    static_for<GEO_N>([&](auto fc) {
        constexpr int f = decltype(fc)::value;
        gf[f] = pair_bcast<f & 1>(gv[f >> 1]);
    });
    Geo g;
    g.x = gf[G_X]; g.y = gf[G_Y]; g.z = gf[G_Z]; g.f = gf[G_F]; g.ghm = gf[G_GHM];
    g.pu0 = gf[G_PU]; g.pu1 = gf[G_PU + 1]; g.pu2 = gf[G_PU + 2];
    g.pv0 = gf[G_PV]; g.pv1 = gf[G_PV + 1]; g.pv2 = gf[G_PV + 2];
    g.pw0 = gf[G_PW]; g.pw1 = gf[G_PW + 1]; g.pw2 = gf[G_PW + 2];
    g.gx = gf[G_GG]; g.gy = gf[G_GG + 1]; g.gz = gf[G_GG + 2];
    double F[4];
    rhs_point(S, u, v, w, h, g, m.gh0, F);
    if (live)
        stage_update_pair<STAGE, EXACT>(st, 4L * node + 2 * k, k ? F[2] : F[0], k ? F[3] : F[1]);
}

// Wave-uniform load through the constant address space: read-only, so it
// stays a scalar load even after stores the compiler cannot rule out as
// aliases.
__device__ __forceinline__ int c_i32(const int *q)
{
    return *(const __attribute__((address_space(4))) int *)q;
}

// ---------------------------------------------------------------------------
// Ensemble, LDS tiles with member lanes (member block 32).  An item is one
// tile (<= kTmemTile consecutive Morton nodes whose stencils touch
// <= kTmemRows distinct nodes) for 32 members.  A wave64 global load costs
// the CU's address unit 16 cycles whatever its width, an LDS read ~5, so
// the kernels move every byte they can through LDS:
//  * the tile's neighbour rows are copied into LDS once -- 16-byte global
//    loads, one row (4 variables x 32 members, 1 KB) per wave instruction --
//    instead of being gathered from L2 once per stencil that names them;
//    LDS rows keep the global layout [var][member] (16-byte writes, 8-byte
//    reads, contiguous across lanes);
//  * the weights (31 x 32 B), geometry and 16-bit stencil positions of each
//    wave's two nodes are staged in a wave-private LDS slice;
//  * lane = member, two nodes per wave: per slot the lane reads its member's
//    (u, v, w, h) (two ds_read2_b64); its node's four weights come from
//    registers by DPP broadcasts within each 16-lane row, read from LDS
//    once per item.
// Per member the sums, the RHS and the stage update are those of
// k_stage_members (bitwise in exact mode).
constexpr int kTmemTile = 8;    // nodes per tile: 4 waves x 2 nodes
constexpr int kTmemRows = 70;   // distinct rows per tile: 70 KB (+10 KB): two item images fill 160 KB
constexpr int kTmemWtsDoubles = 4 * 2 * 32 * 4;   // weights [wave][node][slot][4]
constexpr int kTmemGeoDoubles = 4 * 2 * 32;       // geometry [wave][node][32]
constexpr int kTmemFill = (kTmemRows + 3) / 4;    // rows per wave
constexpr int kTmemStride = 4 * kTmemFill;        // per-tile stride of the fixed-stride row ids

// One (tile, member block) item as wave slice `wid` (0..3) sees it: the
// wave owns tile nodes 2 wid, 2 wid + 1 (nA, nB) and copies tile rows
// [wid * kTmemFill, (wid + 1) * kTmemFill).  All fields wave-uniform.
struct TmemItem {
    int tile, t0, tcount, ucount, col0, nA, nB;
    bool live, hasB;
};

__device__ __forceinline__ TmemItem tmem_item(const DevTiles &tl, long it, int wid,
                                              int active_members)
{
    TmemItem I;
    const int tile = (int)(it / tl.nmb);
    const int mb = (int)(it - (long)tile * tl.nmb);
    I.tile = tile;
    I.t0 = c_i32(tl.tile_start + tile);
    I.tcount = c_i32(tl.tile_start + tile + 1) - I.t0;
    I.ucount = c_i32(tl.ustart + tile + 1) - c_i32(tl.ustart + tile);
    I.col0 = mb * 32;
    const int i0 = 2 * wid;
    I.live = i0 < I.tcount && I.col0 < active_members;
    I.hasB = i0 + 1 < I.tcount;
    I.nA = I.t0 + (i0 < I.tcount ? i0 : 0);
    I.nB = I.hasB ? I.nA + 1 : I.nA;
    return I;
}

// LDS image of one item (doubles): weights [4][2][32][4], geometry
// [4][2][32], rows [kTmemRows][2][32][2].
constexpr int kTmemItemDoubles = kTmemWtsDoubles + kTmemGeoDoubles + 128 * kTmemRows;

// Geometry-row slot of the 32 16-bit stencil positions (8 doubles).
constexpr int kTmemPosSlot = 24;
static_assert(GEO_N <= kTmemPosSlot, "geometry fields and positions share a 32-double row");

// Doubles 17 and 18 of wave slice w's first geometry row (unused by the
// geometry) carry the header of the workgroup's next item for compute wave
// w: (nA, nB) and (col0, live | hasB << 1), so a compute wave learns its
// next nodes from one LDS read instead of a scalar-memory round trip.
constexpr int kTmemHdrSlot = 17;
static_assert(GEO_N <= kTmemHdrSlot && kTmemHdrSlot + 2 <= kTmemPosSlot, "header slots");

__device__ __forceinline__ double tmem_hdr_word(const TmemItem &I, int which)
{
    const unsigned lo = which ? (unsigned)I.col0 : (unsigned)I.nA;
    const unsigned hi = which ? (unsigned)((I.live ? 1 : 0) | (I.hasB ? 2 : 0)) : (unsigned)I.nB;
    return __builtin_bit_cast(double, (unsigned long long)lo | (unsigned long long)hi << 32);
}

__device__ __forceinline__ TmemItem tmem_hdr_item(const double *img, int wid)
{
    const double *h = img + kTmemWtsDoubles + wid * 64 + kTmemHdrSlot;
    const unsigned long long a = __builtin_bit_cast(unsigned long long, h[0]);
    const unsigned long long b = __builtin_bit_cast(unsigned long long, h[1]);
    TmemItem I = {};
    I.nA = (int)(unsigned)a;
    I.nB = (int)(unsigned)(a >> 32);
    I.col0 = (int)(unsigned)b;
    I.live = (b >> 32) & 1u;
    I.hasB = (b >> 33) & 1u;
    return I;
}


// An item and its row ids as a loader wave needs them: every scalar load
// is issued before any is used (one round trip instead of one per group).
struct TmemLoaderItem {
    TmemItem I;
    int ids[kTmemFill];
};

__device__ __forceinline__ TmemLoaderItem tmem_loader_item(const DevTiles &tl, long it, int wid,
                                                           int active_members)
{
    TmemLoaderItem L;
    const int tile = (int)(it / tl.nmb);
    const int *ul = tl.ulist_tile + (long)tile * kTmemStride + wid * kTmemFill;
#pragma unroll
    for (int q = 0; q < kTmemFill; ++q)
        L.ids[q] = c_i32(ul + q);
    int t0 = c_i32(tl.tile_start + tile), t1 = c_i32(tl.tile_start + tile + 1);
    int u0 = c_i32(tl.ustart + tile), u1 = c_i32(tl.ustart + tile + 1);
    asm volatile("" : "+s"(t0), "+s"(t1), "+s"(u0), "+s"(u1));
    const int mb = (int)(it - (long)tile * tl.nmb);
    TmemItem &I = L.I;
    I.tile = tile;
    I.t0 = t0;
    I.tcount = t1 - t0;
    I.ucount = u1 - u0;
    I.col0 = mb * 32;
    const int i0 = 2 * wid;
    I.live = i0 < I.tcount && I.col0 < active_members;
    I.hasB = i0 + 1 < I.tcount;
    I.nA = I.t0 + (i0 < I.tcount ? i0 : 0);
    I.nB = I.hasB ? I.nA + 1 : I.nA;
    return L;
}

// Issues wave slice wid's share of an item's LDS image as direct-to-LDS
// loads: its kTmemFill tile rows, the weights of its two nodes and their
// geometry rows (lanes 0..31, 16 bytes each).  The header slots of the
// geometry row are overwritten by tmem_dma_header once these loads are
// complete.
__device__ __forceinline__ void tmem_dma_issue(double *img, const DevTiles &tl,
                                               const double *__restrict__ src, long B,
                                               const TmemItem &I, const int *ids, int wid,
                                               int lane)
{
    double *wl = img;
    double *gl = img + kTmemWtsDoubles;
    double *rows = gl + kTmemGeoDoubles;
    dma16(tl.wts_node + 128L * I.nA + 2 * lane, wl + wid * 256);
    dma16(tl.wts_node + 128L * I.nB + 2 * lane, wl + wid * 256 + 128);
    if (lane < 32)
        dma16(tl.geo_node + 32L * (lane < 16 ? I.nA : I.nB) + 2 * (lane & 15), gl + wid * 64);
    // rows [row][var][32 members]: lane (var fv, fp) loads members 2 fp, 2 fp + 1
    const double *__restrict__ s0 = src + I.col0 + 2 * (lane & 15) + (long)(lane >> 4) * B;
#pragma unroll
    for (int q = 0; q < kTmemFill; ++q) {
        const int rr = wid * kTmemFill + q;
        if (rr < I.ucount)   // wave-uniform
            dma16(s0 + 4L * ids[q] * B, rows + rr * 128);
    }
}

// LDS of one workgroup: two item images.
constexpr int kTmemLdsDoubles = 2 * kTmemItemDoubles;

// The next-item header (geometry slots kTmemHdrSlot, +1 of wave slice wid's
// first row), written after the slice's direct-to-LDS loads completed.
__device__ __forceinline__ void tmem_dma_header(double *img, const TmemItem &next, int wid,
                                                int lane)
{
    double *gl = img + kTmemWtsDoubles;
    if (lane == kTmemHdrSlot || lane == kTmemHdrSlot + 1)
        gl[wid * 64 + lane] = tmem_hdr_word(next, lane == kTmemHdrSlot + 1);
}

// The epilogue's global operands H and D (the own state comes from the
// tile's LDS rows) in the epilogue's pair layout: lanes 2p, 2p + 1 of a
// half-wave hold members 2p, 2p + 1 of variables (0, 1) and (2, 3)
// respectively, so every load and store is 16 bytes.
struct TmemOwn {
    double2 hv[2], dv[2];
};

__device__ __forceinline__ long tmem_pair_elem(long B, int node, int col0, int lane)
{
    const int ml = lane & 31;
    return 4L * node * B + col0 + (ml & ~1) + 2L * (ml & 1) * B;
}

template <int STAGE>
__device__ __forceinline__ TmemOwn tmem_own_load(const DevStage &st, const double *__restrict__ src,
                                                 long B, const TmemItem &I, int lane)
{
    TmemOwn O;
    const long e = tmem_pair_elem(B, (lane >> 5) ? I.nB : I.nA, I.col0, lane);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        // stage 0 gathers H itself: the epilogue takes H from the node's
        // own row in the tile (tmem_compute)
        O.hv[q] = (STAGE != ST_RHS && STAGE != ST_0)
                      ? *reinterpret_cast<const double2 *>(st.h + e + q * B)
                      : make_double2(0.0, 0.0);
        O.dv[q] = (STAGE != ST_RHS && STAGE != ST_0)
                      ? *reinterpret_cast<const double2 *>(st.d + e + q * B)
                      : make_double2(0.0, 0.0);
    }
    return O;
}

// The fused stage update (D, and K or H) of two consecutive elements from
// prefetched H and D operands (16-byte stores).
template <int STAGE, bool EXACT>
__device__ __forceinline__ void stage_update_pre_pair(const DevStage &st, long e, double2 F,
                                                      double2 hv, double2 dv)
{
    double2 *out = reinterpret_cast<double2 *>(st.out + e);
    if constexpr (STAGE == ST_RHS) {
        *out = F;
    } else if constexpr (STAGE == ST_0) {
        *reinterpret_cast<double2 *>(st.d + e) = F;
        *out = make_double2(axpy<EXACT>(hv.x, st.c, F.x), axpy<EXACT>(hv.y, st.c, F.y));
    } else if constexpr (STAGE == ST_1 || STAGE == ST_2) {
        *reinterpret_cast<double2 *>(st.d + e) =
            make_double2(axpy<EXACT>(dv.x, st.dcoef, F.x), axpy<EXACT>(dv.y, st.dcoef, F.y));
        *out = make_double2(axpy<EXACT>(hv.x, st.c, F.x), axpy<EXACT>(hv.y, st.c, F.y));
    } else {
        const double d0 = axpy<EXACT>(dv.x, st.dcoef, F.x);
        const double d1 = axpy<EXACT>(dv.y, st.dcoef, F.y);
        *out = make_double2(axpy<EXACT>(hv.x, st.c, d0), axpy<EXACT>(hv.y, st.c, d1));
    }
}

// x's upper 32 lanes <-> y's lower 32 lanes (v_permlane32_swap, per dword).
__device__ __forceinline__ void swap_rows32(double &x, double &y)
{
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(x),
                                                     (unsigned)__double2loint(y), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(x),
                                                     (unsigned)__double2hiint(y), false, false);
    x = __hiloint2double((int)hi[0], (int)lo[0]);
    y = __hiloint2double((int)hi[1], (int)lo[1]);
}

// x's odd 16-lane rows <-> y's even rows (v_permlane16_swap, per dword).
__device__ __forceinline__ void swap_rows16(double &x, double &y)
{
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x),
                                                     (unsigned)__double2loint(y), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x),
                                                     (unsigned)__double2hiint(y), false, false);
    x = __hiloint2double((int)hi[0], (int)lo[0]);
    y = __hiloint2double((int)hi[1], (int)lo[1]);
}

// Stencil sums of wave slice wid's two nodes on the FP64 matrix cores
// (fast mode).  Per node and group of 4 slots, v_mfma_f64_4x4x4f64 (4
// blocks b = (lane >> 2) & 3; A[i][k] in lane 16 k + 4 b + i, B[k][j] in
// lane 16 k + 4 b + j, D[i][j] in lane 16 i + 4 b + j) multiplies the 4 x 4
// weights (operator i, slot k) by 16 (variable, member) columns of the
// neighbour states: instruction g takes variable g >> 1, members 16 (g & 1)
// + 4 b + j.  8 slot groups x 8 instructions per node; slot 31 has zero
// weights (its position, the node's own row, is any valid row).  Lane row r
// then holds operator r's sums for 2 nodes x 2 member halves x 4 variables;
// a 4 x 4 transpose over the lane rows (row pairs 32 apart, then 16 apart)
// leaves lane (node half, member ml) with the 16 sums of its member.  The
// instruction sums 4 products per slot group, not the sequential sums of
// the vector path.
template <int NNBR>
__device__ __forceinline__ Sums tmem_sums_mfma(const double *img, int wid, int lane)
{
    static_assert(NNBR <= 31, "8 groups of 4 slots, slot 31 weighted zero");
    const double *wl = img + wid * 256;                         // [node][slot][op]
    const double *gl = img + kTmemWtsDoubles + wid * 64;        // [node][32]
    const double *rows = img + kTmemWtsDoubles + kTmemGeoDoubles;
    const int k = lane >> 4, i = lane & 3, c = lane & 15;
    double acc[2][8];
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        const unsigned short *pos16 =
            reinterpret_cast<const unsigned short *>(gl + n * 32 + kTmemPosSlot);
#pragma unroll
        for (int g = 0; g < 8; ++g)
            acc[n][g] = 0.0;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int slot = 4 * s + k;
            const double a = (slot < NNBR) ? wl[n * 128 + slot * 4 + i] : 0.0;
            unsigned p = pos16[slot];
            p = p == 0xffffu ? 0u : p;   // no own row (padding): any row, zero weight
            const double *rb = rows + p * 128 + c;
            double b[8];
#pragma unroll
            for (int g = 0; g < 8; ++g)
                b[g] = rb[(g >> 1) * 32 + 16 * (g & 1)];
#pragma unroll
            for (int g = 0; g < 8; ++g)
                acc[n][g] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b[g], acc[n][g], 0, 0, 0);
        }
    }
    // per variable v: M[row r][q] = acc[q >> 1][2 v + (q & 1)] -> lane row q
    // gets M[0..3][q]
    double T[4][4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double x0 = acc[0][2 * v], x1 = acc[0][2 * v + 1];
        double x2 = acc[1][2 * v], x3 = acc[1][2 * v + 1];
        swap_rows32(x0, x2);
        swap_rows32(x1, x3);
        swap_rows16(x0, x1);
        swap_rows16(x2, x3);
        T[v][0] = x0; T[v][1] = x1; T[v][2] = x2; T[v][3] = x3;
    }
    Sums S;
    S.dudx = T[0][0]; S.dudy = T[0][1]; S.dudz = T[0][2]; S.lu = T[0][3];
    S.dvdx = T[1][0]; S.dvdy = T[1][1]; S.dvdz = T[1][2]; S.lv = T[1][3];
    S.dwdx = T[2][0]; S.dwdy = T[2][1]; S.dwdz = T[2][2]; S.lw = T[2][3];
    S.dhdx = T[3][0]; S.dhdy = T[3][1]; S.dhdz = T[3][2]; S.lh = T[3][3];
    return S;
}

// Stencil sums of wave slice wid's two nodes from an item's LDS image, then
// the pointwise RHS and the stage update of every (node, member) lane.
template <int STAGE, bool EXACT, int NNBR, int DEPTH, bool MFMA = false>
__device__ __forceinline__ void tmem_compute(const DevMesh &m, const DevTiles &tl,
                                             const DevStage &st, long B, const double *img,
                                             const TmemItem &I, const TmemOwn &O, int wid,
                                             int lane)
{
    const int ml = lane & 31;
    const int half = lane >> 5;
    const double *wl = img;
    const double *gl = img + kTmemWtsDoubles;
    const double *rows = gl + kTmemGeoDoubles;
    // the half-wave's node positions (16-bit, 2 per dword), broadcast reads
    const double *gp = gl + wid * 64 + half * 32;
    unsigned p[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 v = *reinterpret_cast<const uint4 *>(gp + kTmemPosSlot + 2 * q);
        p[4 * q] = v.x; p[4 * q + 1] = v.y; p[4 * q + 2] = v.z; p[4 * q + 3] = v.w;
    }
    auto row_off = [&](int j) {
        const unsigned a = p[j >> 1];
        return (int)((j & 1) ? (a >> 16) : (a & 0xffffu)) * 128;
    };
    const double *lrow = rows + ml;
    // Weights: lane k of each 16-lane row holds its node's four weights of
    // slots k and k + 16 (rows 0, 1: node A; rows 2, 3: node B), four LDS
    // reads per item; slot j's weights reach the whole row by DPP
    // (row_newbcast from lane j & 15), so the stencil loop reads only the
    // member rows from LDS.
    const double2 *wq =
        reinterpret_cast<const double2 *>(wl) + wid * 128 + half * 64 + 2 * (lane & 15);
    const double2 wlo01 = wq[0], wlo23 = wq[1], whi01 = wq[32], whi23 = wq[33];
    // The LDS reads of a slot's member row run DEPTH slots ahead, in issue
    // order, so waiting for slot j leaves the later slots in flight; slot
    // j + DEPTH's reads are tied to the sums after slot j (bounded live
    // ranges in exact mode too).
    double2 ruv[DEPTH], rwh[DEPTH];
    auto issue = [&](int j, int off) {
        ruv[j % DEPTH] = make_double2(lrow[off], lrow[off + 32]);        // u, v
        rwh[j % DEPTH] = make_double2(lrow[off + 64], lrow[off + 96]);   // w, h
    };
    Sums S = {};
    if constexpr (MFMA) {
        S = tmem_sums_mfma<NNBR>(img, wid, lane);
    } else {
    static_for<DEPTH>(
        [&](auto jc) { issue(decltype(jc)::value, row_off(decltype(jc)::value)); });
    static_for<NNBR>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const double2 uv = ruv[j % DEPTH], wh = rwh[j % DEPTH];
        const double2 wa = j < 16 ? wlo01 : whi01, wb = j < 16 ? wlo23 : whi23;
        const double2 w01 = make_double2(row_bcast64<j & 15>(wa.x), row_bcast64<j & 15>(wa.y));
        const double2 w23 = make_double2(row_bcast64<j & 15>(wb.x), row_bcast64<j & 15>(wb.y));
        acc_slot<EXACT>(S, w01, w23, uv.x, uv.y, wh.x, wh.y);
        if constexpr (j + DEPTH < NNBR) {
            int off = row_off(j + DEPTH);
            asm volatile(""
                         : "+v"(off)
                         : "v"(S.dudx), "v"(S.dvdx), "v"(S.dwdx), "v"(S.dhdx),
                           "v"(S.dudy), "v"(S.dvdy), "v"(S.dwdy), "v"(S.dhdy),
                           "v"(S.dudz), "v"(S.dvdz), "v"(S.dwdz), "v"(S.dhdz),
                           "v"(S.lu), "v"(S.lv), "v"(S.lw), "v"(S.lh));
            issue(j + DEPTH, off);
        }
    });
    }
    Geo g;
    g.x = gp[G_X]; g.y = gp[G_Y]; g.z = gp[G_Z]; g.f = gp[G_F]; g.ghm = gp[G_GHM];
    g.pu0 = gp[G_PU]; g.pu1 = gp[G_PU + 1]; g.pu2 = gp[G_PU + 2];
    g.pv0 = gp[G_PV]; g.pv1 = gp[G_PV + 1]; g.pv2 = gp[G_PV + 2];
    g.pw0 = gp[G_PW]; g.pw1 = gp[G_PW + 1]; g.pw2 = gp[G_PW + 2];
    g.gx = gp[G_GG]; g.gy = gp[G_GG + 1]; g.gz = gp[G_GG + 2];
    // own state: the node's row in this tile (position in the spare slot 31
    // of its position list), else a global load (nodes outside their own
    // stencil, e.g. padding)
    const int node = half ? I.nB : I.nA;
    const long e = 4L * node * B + I.col0 + ml;
    const unsigned own = p[15] >> 16;
    double ou, ov, ow, oh;
    if (own != 0xffffu) {
        const double *r = lrow + own * 128;
        ou = r[0]; ov = r[32]; ow = r[64]; oh = r[96];
    } else {
        ou = st.src[e]; ov = st.src[e + B]; ow = st.src[e + 2 * B]; oh = st.src[e + 3 * B];
        // keeps the two branches apart: merged, their loads become flat loads
        // of a selected pointer, and a flat load waits for every outstanding
        // global access of the wave (the previous item's stores included)
        asm volatile("" : "+v"(ou), "+v"(ov), "+v"(ow), "+v"(oh));
    }
    double F[4];
    rhs_point(S, ou, ov, ow, oh, g, m.gh0, F);
    // to the pair layout: lane 2p keeps variables 0, 1 and takes member
    // 2p + 1's; lane 2p + 1 keeps 2, 3 and takes member 2p's (one DPP swap
    // of two values, every lane active)
    const int k = ml & 1;
    const double r0 = pair_swap(k ? F[0] : F[2]);
    const double r1 = pair_swap(k ? F[1] : F[3]);
    const double2 Fa = k ? make_double2(r0, F[2]) : make_double2(F[0], r0);
    const double2 Fb = k ? make_double2(r1, F[3]) : make_double2(F[1], r1);
    const long e2 = tmem_pair_elem(B, node, I.col0, lane);
    double2 hv0 = O.hv[0], hv1 = O.hv[1];
    if constexpr (STAGE == ST_0) {
        // H is the gathered state: this lane's pair-layout values from the
        // own row (variables 2k, 2k + 1 of members ml & ~1 and the next)
        if (own != 0xffffu) {
            const double *r = rows + own * 128 + 64 * k + (ml & ~1);
            hv0 = *reinterpret_cast<const double2 *>(r);
            hv1 = *reinterpret_cast<const double2 *>(r + 32);
        } else {
            hv0 = *reinterpret_cast<const double2 *>(st.h + e2);
            hv1 = *reinterpret_cast<const double2 *>(st.h + e2 + B);
            asm volatile("" : "+v"(hv0.x), "+v"(hv0.y), "+v"(hv1.x), "+v"(hv1.y));
        }
    }
    if (!half || I.hasB) {
        stage_update_pre_pair<STAGE, EXACT>(st, e2, Fa, hv0, O.dv[0]);
        stage_update_pre_pair<STAGE, EXACT>(st, e2 + B, Fb, hv1, O.dv[1]);
    }
}

// Items of a persistent launch: XCD x (hardware blocks are dealt
// round-robin over the 8 XCDs) owns items [x R, (x+1) R); its W workgroups
// interleave over them (j, j + W, ...), so the items in flight on one XCD
// are adjacent tiles that share halo rows and weights in that XCD's L2.
struct TmemRange {
    long it0, it1, W;
};

__device__ __forceinline__ TmemRange tmem_range(const DevTiles &tl, int nblk)
{
    const long items = (long)tl.ntiles * tl.nmb;
    const long R = (items + 7) / 8;
    const int xcd = (int)blockIdx.x & 7;
    TmemRange r;
    r.W = nblk >> 3;
    r.it0 = (long)xcd * R + ((long)blockIdx.x >> 3);
    const long xend = (long)(xcd + 1) * R;
    r.it1 = xend < items ? xend : items;
    return r;
}

// The same items with specialised waves and two LDS images (160 KB, one
// workgroup per CU): waves 4..7 fill one image while waves 0..3 compute
// from the other, one barrier per item.  Loader wave 4 + w fills the share
// of compute wave w with direct-to-LDS loads (dma16: no register staging,
// no LDS write instructions), one item ahead: right after the barrier that
// frees an image it issues the next item's loads into it, waits for them
// and writes the header naming the item after.  A compute wave issues its
// own epilogue loads (H, D) one item ahead, before its stencil loop.
template <int STAGE, bool EXACT, int NNBR, bool MFMA = false>
__global__ __launch_bounds__(512) void k_stage_tspec(DevMesh m, DevTiles tl, DevStage st)
{
    static_assert(!(MFMA && EXACT), "the matrix-core sums are fast mode only");
    static_assert(NNBR > 0 && NNBR <= 31, "node-major positions: 31 slots + the own row");
    extern __shared__ double lds[];   // two item images
    const TmemRange r = tmem_range(tl, st.nblk);
    const long B = st.members;
    const int lane = (int)threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int wid = wv & 3;
    const int active_members = c_i32(st.active);
    if (r.it0 >= r.it1)
        return;   // uniform over the workgroup
    // a clamped, valid item past the end keeps every scalar load in bounds
    auto clamp = [&](long i) { return i < r.it1 ? i : r.it0; };
    if (wv >= 4) {
        TmemLoaderItem L = tmem_loader_item(tl, r.it0, wid, active_members);
        tmem_dma_issue(lds, tl, st.src, B, L.I, L.ids, wid, lane);
        TmemLoaderItem Ln = tmem_loader_item(tl, clamp(r.it0 + r.W), wid, active_members);
        __builtin_amdgcn_s_waitcnt(0);
        tmem_dma_header(lds, Ln.I, wid, lane);
        __syncthreads();
        int b = 0;
        for (long it = r.it0; it < r.it1; it += r.W) {
            // image b ^ 1 is free: fill it with item it + W
            const bool more = it + r.W < r.it1;
            if (more)
                tmem_dma_issue(lds + (b ^ 1) * kTmemItemDoubles, tl, st.src, B, Ln.I, Ln.ids,
                               wid, lane);
            const TmemLoaderItem L2 = tmem_loader_item(tl, clamp(it + 2 * r.W), wid,
                                                       active_members);
            __builtin_amdgcn_s_waitcnt(0);
            if (more)
                tmem_dma_header(lds + (b ^ 1) * kTmemItemDoubles, L2.I, wid, lane);
            __syncthreads();   // image b is free, image b ^ 1 is filled
            Ln = L2;
            b ^= 1;
        }
    } else {
        // the epilogue operands of item i + 1 are loaded while item i is
        // computed (their HBM latency exceeds one stencil loop)
        TmemItem I = tmem_item(tl, r.it0, wid, active_members);
        TmemOwn o = tmem_own_load<STAGE>(st, st.src, B, I, lane);
        __syncthreads();
        int b = 0;
        for (long it = r.it0; it < r.it1; it += r.W) {
            // the next item's nodes from the header its loader left in this image
            const TmemItem In = tmem_hdr_item(lds + b * kTmemItemDoubles, wid);
            const TmemOwn on = tmem_own_load<STAGE>(st, st.src, B, In, lane);
            if (I.live)
                tmem_compute<STAGE, EXACT, NNBR, 4, MFMA>(m, tl, st, B,
                                                          lds + b * kTmemItemDoubles, I, o,
                                                          wid, lane);
            __syncthreads();
            I = In;
            o = on;
            b ^= 1;
        }
    }
}

}  // namespace swe
