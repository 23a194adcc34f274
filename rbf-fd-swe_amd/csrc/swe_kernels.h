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

// Broadcast lane K of every aligned group of 4 lanes (DPP quad_perm).
template <int K>
__device__ __forceinline__ double quad_bcast(double v)
{
    constexpr int ctrl = K * 0x55;  // quad_perm [K, K, K, K]
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Stencil-slot loop shared by the stage kernels.  PF == 0 leaves the load
// schedule to the compiler.  PF = D > 0 software-pipelines it: all neighbour
// ids are loaded first, then a ring of D slots is kept in flight -- each
// iteration consumes the oldest slot and issues the loads of slot j + D --
// with scheduling barriers between iterations, so a wave always has D
// gathers and 2*D weight loads outstanding and the compiler can neither
// hoist further loads nor (in exact mode) precompute all 31x16 products
// ahead of the serial sums, which would exhaust the registers.
// Accumulation is sequential in j for every sum in all forms.
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

struct QuadSlot {
    double2 w01, w23;
    double s;
};

// ---------------------------------------------------------------------------
// Single state (B == 1), quad mapping: four lanes per node.  Lane k of the
// quad reads variable k of every neighbour and keeps the Dx, Dy, Dz, L sums
// of that variable (each a sequential sum over the stencil); the 16 sums are
// then shared across the quad by DPP, every lane evaluates the pointwise RHS
// and lane k updates variable k.  256 threads = 64 nodes per workgroup.
template <int STAGE, bool EXACT, int NNBR, int PF>
__global__ __launch_bounds__(256) void k_stage_quad(DevMesh m, DevStage st)
{
    const int t = xcd_block(st.nblk) * 256 + (int)threadIdx.x;
    const int node_raw = t >> 2;
    const int k = t & 3;
    const bool live = node_raw < m.n;
    const int node = live ? node_raw : m.n - 1;  // keep quads whole for DPP
    const int nn = NNBR > 0 ? NNBR : m.nnbr;
    const long ld = slot_stride(m);
    const long sb = slot_base<NNBR>(m, node);
    const int *__restrict__ nb = m.nbr + sb;
    const double *__restrict__ wp = m.wts + 4L * sb;
    const double *__restrict__ src = st.src;

    double sx = 0.0, sy = 0.0, sz = 0.0, sl = 0.0;
    auto load = [&](int j, int q) {
        QuadSlot r;
        r.w01 = *reinterpret_cast<const double2 *>(wp + 4L * j * ld);
        r.w23 = *reinterpret_cast<const double2 *>(wp + 4L * j * ld + 2);
        r.s = src[4L * q + k];
        return r;
    };
    auto acc = [&](const QuadSlot &r) {
        sx = madd<EXACT>(sx, r.w01.x, r.s);
        sy = madd<EXACT>(sy, r.w01.y, r.s);
        sz = madd<EXACT>(sz, r.w23.x, r.s);
        sl = madd<EXACT>(sl, r.w23.y, r.s);
    };
    stencil_loop<NNBR, PF>(nb, ld, nn, load, acc, (QuadSlot *)nullptr);
    Sums S;
    S.dudx = quad_bcast<0>(sx); S.dudy = quad_bcast<0>(sy);
    S.dudz = quad_bcast<0>(sz); S.lu = quad_bcast<0>(sl);
    S.dvdx = quad_bcast<1>(sx); S.dvdy = quad_bcast<1>(sy);
    S.dvdz = quad_bcast<1>(sz); S.lv = quad_bcast<1>(sl);
    S.dwdx = quad_bcast<2>(sx); S.dwdy = quad_bcast<2>(sy);
    S.dwdz = quad_bcast<2>(sz); S.lw = quad_bcast<2>(sl);
    S.dhdx = quad_bcast<3>(sx); S.dhdy = quad_bcast<3>(sy);
    S.dhdz = quad_bcast<3>(sz); S.lh = quad_bcast<3>(sl);

    const double2 s01 = *reinterpret_cast<const double2 *>(src + 4L * node);
    const double2 s23 = *reinterpret_cast<const double2 *>(src + 4L * node + 2);
    const Geo g = load_geo(m.geo, m.ld, node);
    double F[4];
    rhs_point(S, s01.x, s01.y, s23.x, s23.y, g, m.gh0, F);
    const double Fk = k == 0 ? F[0] : k == 1 ? F[1] : k == 2 ? F[2] : F[3];
    if (live)
        stage_update<STAGE, EXACT>(st, 4L * node + k, Fk);
}

struct QuadWSlot {
    double w;
    double s;
};

// Single state, quad mapping with shared weight loads: lane k of the quad
// loads only weight k (Dx, Dy, Dz or L) of each slot -- one 8-byte load per
// lane, a wave's 16 nodes x 32 B read once instead of four times -- and the
// quad exchanges the four weights by DPP.  Same sums, same order as
// k_stage_quad.
template <int STAGE, bool EXACT, int NNBR, int PF>
__global__ __launch_bounds__(256) void k_stage_quadw(DevMesh m, DevStage st)
{
    const int t = xcd_block(st.nblk) * 256 + (int)threadIdx.x;
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
    stencil_loop<NNBR, PF>(nb, ld, nn, load, acc, (QuadWSlot *)nullptr);
    Sums S;
    S.dudx = quad_bcast<0>(sx); S.dudy = quad_bcast<0>(sy);
    S.dudz = quad_bcast<0>(sz); S.lu = quad_bcast<0>(sl);
    S.dvdx = quad_bcast<1>(sx); S.dvdy = quad_bcast<1>(sy);
    S.dvdz = quad_bcast<1>(sz); S.lv = quad_bcast<1>(sl);
    S.dwdx = quad_bcast<2>(sx); S.dwdy = quad_bcast<2>(sy);
    S.dwdz = quad_bcast<2>(sz); S.lw = quad_bcast<2>(sl);
    S.dhdx = quad_bcast<3>(sx); S.dhdy = quad_bcast<3>(sy);
    S.dhdz = quad_bcast<3>(sz); S.lh = quad_bcast<3>(sl);

    const double2 s01 = *reinterpret_cast<const double2 *>(src + 4L * node);
    const double2 s23 = *reinterpret_cast<const double2 *>(src + 4L * node + 2);
    const Geo g = load_geo(m.geo, m.ld, node);
    double F[4];
    rhs_point(S, s01.x, s01.y, s23.x, s23.y, g, m.gh0, F);
    const double Fk = k == 0 ? F[0] : k == 1 ? F[1] : k == 2 ? F[2] : F[3];
    if (live)
        stage_update<STAGE, EXACT>(st, 4L * node + k, Fk);
}

struct LaneSlot {
    double2 w01, w23, s01, s23;
};

// Single state (B == 1), lane mapping: one lane per node with all 16 sums.
// A quarter of the quad kernel's load instructions per node (each lane
// reads its own 32 B of weights and the neighbour's whole 32 B state), for
// meshes large enough to fill the chip with one lane per node.
template <int STAGE, bool EXACT, int NNBR, int PF>
__global__ __launch_bounds__(256) void k_stage_lane(DevMesh m, DevStage st)
{
    const int t = xcd_block(st.nblk) * 256 + (int)threadIdx.x;
    const bool live = t < m.n;
    const int node = live ? t : m.n - 1;
    const int nn = NNBR > 0 ? NNBR : m.nnbr;
    const long ld = slot_stride(m);
    const long sb = slot_base<NNBR>(m, node);
    const int *__restrict__ nb = m.nbr + sb;
    const double *__restrict__ wp = m.wts + 4L * sb;
    const double *__restrict__ src = st.src;

    Sums S = {};
    auto load = [&](int j, int q) {
        LaneSlot r;
        r.w01 = *reinterpret_cast<const double2 *>(wp + 4L * j * ld);
        r.w23 = *reinterpret_cast<const double2 *>(wp + 4L * j * ld + 2);
        r.s01 = *reinterpret_cast<const double2 *>(src + 4L * q);
        r.s23 = *reinterpret_cast<const double2 *>(src + 4L * q + 2);
        return r;
    };
    auto acc = [&](const LaneSlot &r) {
        const double su = r.s01.x, sv = r.s01.y, sw = r.s23.x, sh = r.s23.y;
        S.dudx = madd<EXACT>(S.dudx, r.w01.x, su); S.dvdx = madd<EXACT>(S.dvdx, r.w01.x, sv);
        S.dwdx = madd<EXACT>(S.dwdx, r.w01.x, sw); S.dhdx = madd<EXACT>(S.dhdx, r.w01.x, sh);
        S.dudy = madd<EXACT>(S.dudy, r.w01.y, su); S.dvdy = madd<EXACT>(S.dvdy, r.w01.y, sv);
        S.dwdy = madd<EXACT>(S.dwdy, r.w01.y, sw); S.dhdy = madd<EXACT>(S.dhdy, r.w01.y, sh);
        S.dudz = madd<EXACT>(S.dudz, r.w23.x, su); S.dvdz = madd<EXACT>(S.dvdz, r.w23.x, sv);
        S.dwdz = madd<EXACT>(S.dwdz, r.w23.x, sw); S.dhdz = madd<EXACT>(S.dhdz, r.w23.x, sh);
        S.lu = madd<EXACT>(S.lu, r.w23.y, su); S.lv = madd<EXACT>(S.lv, r.w23.y, sv);
        S.lw = madd<EXACT>(S.lw, r.w23.y, sw); S.lh = madd<EXACT>(S.lh, r.w23.y, sh);
    };
    stencil_loop<NNBR, PF>(nb, ld, nn, load, acc, (LaneSlot *)nullptr);
    const double2 s01 = *reinterpret_cast<const double2 *>(src + 4L * node);
    const double2 s23 = *reinterpret_cast<const double2 *>(src + 4L * node + 2);
    const Geo g = load_geo(m.geo, m.ld, node);
    double F[4];
    rhs_point(S, s01.x, s01.y, s23.x, s23.y, g, m.gh0, F);
    if (live) {
#pragma unroll
        for (int var = 0; var < 4; ++var)
            stage_update<STAGE, EXACT>(st, 4L * node + var, F[var]);
    }
}

// ---------------------------------------------------------------------------
// Ensemble (B >= 2): lane = (node, member), MB members of one node per lane
// group, member block blockIdx.y.  Each lane keeps all 16 sums of its
// member; the stencil weights and ids are shared by the MB lanes of a node
// (wave-uniform scalar loads when MB == 64), so one weight read from HBM
// serves every member of the block.  The neighbour state of a variable is a
// contiguous run of MB doubles.
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
    if constexpr (MB == 64 && !EXACT && NNBR > 0 && NNBR <= 32) {
        // One node per wave: lane j fetches the id of slot j and lanes 2j,
        // 2j+1 fetch its (Dx, Dy) and (Dz, L), one vector load each, up
        // front.  The weights go to a wave-private LDS row and are read back
        // per slot with uniform-address (broadcast) LDS loads; the id moves
        // to a scalar register with v_readlane.  No scalar-memory round trip
        // sits between the neighbour gathers, so several slots of gathers
        // stay in flight.
        __shared__ double2 wrow[4][64];
        const int lane = (int)threadIdx.x & 63;
        const int wid = (int)threadIdx.x >> 6;
        const int jq = lane < NNBR ? lane : 0;
        const int jw = (lane >> 1) < NNBR ? (lane >> 1) : 0;
        const int qv = nb[jq * ld];
        wrow[wid][lane] = *reinterpret_cast<const double2 *>(wp + 4L * jw * ld + 2 * (lane & 1));
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int j = 0; j < NNBR; ++j) {
            const int q = __builtin_amdgcn_readlane(qv, j);
            const double2 w01 = wrow[wid][2 * j];
            const double2 w23 = wrow[wid][2 * j + 1];
            const double *sq = src + 4L * q * B;
            const double su = sq[0], sv = sq[B], sw = sq[2 * B], sh = sq[3 * B];
            S.dudx = madd<EXACT>(S.dudx, w01.x, su); S.dvdx = madd<EXACT>(S.dvdx, w01.x, sv);
            S.dwdx = madd<EXACT>(S.dwdx, w01.x, sw); S.dhdx = madd<EXACT>(S.dhdx, w01.x, sh);
            S.dudy = madd<EXACT>(S.dudy, w01.y, su); S.dvdy = madd<EXACT>(S.dvdy, w01.y, sv);
            S.dwdy = madd<EXACT>(S.dwdy, w01.y, sw); S.dhdy = madd<EXACT>(S.dhdy, w01.y, sh);
            S.dudz = madd<EXACT>(S.dudz, w23.x, su); S.dvdz = madd<EXACT>(S.dvdz, w23.x, sv);
            S.dwdz = madd<EXACT>(S.dwdz, w23.x, sw); S.dhdz = madd<EXACT>(S.dhdz, w23.x, sh);
            S.lu = madd<EXACT>(S.lu, w23.y, su); S.lv = madd<EXACT>(S.lv, w23.y, sv);
            S.lw = madd<EXACT>(S.lw, w23.y, sw); S.lh = madd<EXACT>(S.lh, w23.y, sh);
            if constexpr (PF > 0) {
                if (j % PF == PF - 1)
                    __builtin_amdgcn_sched_barrier(0);
            }
        }
    } else {
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
    }
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

// ---------------------------------------------------------------------------
// Ensemble, tiled (B >= 2): the mesh is stored in a space-filling-curve node
// order and cut into tiles of consecutive nodes whose stencils share few
// neighbours (a 64-node tile touches ~220 distinct nodes, 9x reuse).  One
// workgroup = one tile x MB members: it first copies the tile's distinct
// neighbour rows (4 variables x MB members each) from HBM/L2 into LDS, then
// lane (node, member) gathers its 31 neighbours from LDS.  Member blocks of
// one tile are consecutive workgroups on the same XCD, so the tile's weights
// come from HBM once and from L2 for the other member blocks.
struct DevTiles {
    const int *tile_start;          // [ntiles + 1] first node of each tile
    const int *ustart;              // [ntiles + 1] offsets into ulist
    const int *ulist;               // distinct stencil nodes of each tile, ascending
    const unsigned short *lidx;     // [nnbr][ld] position of nbr(j, node) in its tile's list
    int ntiles;
    int nmb;                        // member blocks per tile (B / MB)
};

template <int STAGE, bool EXACT, int MB, int NNBR>
__global__ __launch_bounds__(512) void k_stage_tiled(DevMesh m, DevTiles tl, DevStage st)
{
    static_assert(MB >= 2 && MB <= 64 && (MB & (MB - 1)) == 0, "MB power of two");
    extern __shared__ double rows[];   // [ucount][4][MB]
    const int lb = xcd_block(st.nblk);
    const int tile = lb / tl.nmb;
    const int mb = lb - tile * tl.nmb;
    if (tile >= tl.ntiles || mb * MB >= *st.active)
        return;   // uniform over the workgroup
    const int t0 = tl.tile_start[tile];
    const int tcount = tl.tile_start[tile + 1] - t0;
    const int u0 = tl.ustart[tile];
    const int ucount = tl.ustart[tile + 1] - u0;
    const long B = st.members;
    const int col0 = mb * MB;
    const double *__restrict__ src = st.src;
    const int total = ucount * 4 * MB;
    for (int i = (int)threadIdx.x; i < total; i += (int)blockDim.x) {
        const int r = i / MB;
        const int c = i & (MB - 1);
        const int u = tl.ulist[u0 + (r >> 2)];
        rows[i] = src[(4L * u + (r & 3)) * B + col0 + c];
    }
    __syncthreads();
    const int nl = (int)threadIdx.x / MB;
    const int ml = (int)threadIdx.x & (MB - 1);
    if (nl >= tcount)
        return;   // no barrier follows
    const int node = t0 + nl;
    const int nn = NNBR > 0 ? NNBR : m.nnbr;
    const long ld = slot_stride(m);
    const long lxs = m.ld;   // tile-local ids are slot-major [nnbr][n]
    const unsigned short *__restrict__ lx = tl.lidx + node;
    const double *__restrict__ wp = m.wts + 4L * slot_base<NNBR>(m, node);
    Sums S = {};
    auto slot = [&](int j) {
        const double *sp = rows + (int)lx[j * lxs] * 4 * MB + ml;
        const double su = sp[0], sv = sp[MB], sw = sp[2 * MB], sh = sp[3 * MB];
        const double2 w01 = *reinterpret_cast<const double2 *>(wp + 4L * j * ld);
        const double2 w23 = *reinterpret_cast<const double2 *>(wp + 4L * j * ld + 2);
        S.dudx = madd<EXACT>(S.dudx, w01.x, su); S.dvdx = madd<EXACT>(S.dvdx, w01.x, sv);
        S.dwdx = madd<EXACT>(S.dwdx, w01.x, sw); S.dhdx = madd<EXACT>(S.dhdx, w01.x, sh);
        S.dudy = madd<EXACT>(S.dudy, w01.y, su); S.dvdy = madd<EXACT>(S.dvdy, w01.y, sv);
        S.dwdy = madd<EXACT>(S.dwdy, w01.y, sw); S.dhdy = madd<EXACT>(S.dhdy, w01.y, sh);
        S.dudz = madd<EXACT>(S.dudz, w23.x, su); S.dvdz = madd<EXACT>(S.dvdz, w23.x, sv);
        S.dwdz = madd<EXACT>(S.dwdz, w23.x, sw); S.dhdz = madd<EXACT>(S.dhdz, w23.x, sh);
        S.lu = madd<EXACT>(S.lu, w23.y, su); S.lv = madd<EXACT>(S.lv, w23.y, sv);
        S.lw = madd<EXACT>(S.lw, w23.y, sw); S.lh = madd<EXACT>(S.lh, w23.y, sh);
    };
    if constexpr (NNBR > 0) {
#pragma unroll 8
        for (int j = 0; j < NNBR; ++j)
            slot(j);
    } else {
#pragma unroll 4
        for (int j = 0; j < nn; ++j)
            slot(j);
    }
    const long e = 4L * node * B + col0 + ml;
    const double u = src[e], v = src[e + B], w = src[e + 2 * B], h = src[e + 3 * B];
    const Geo g = load_geo(m.geo, m.ld, node);
    double F[4];
    rhs_point(S, u, v, w, h, g, m.gh0, F);
#pragma unroll
    for (int var = 0; var < 4; ++var)
        stage_update<STAGE, EXACT>(st, e + var * B, F[var]);
}

}  // namespace swe
