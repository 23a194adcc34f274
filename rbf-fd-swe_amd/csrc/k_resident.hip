// Whole RK4 steps of one single-state mesh in one cooperative launch: each
// workgroup owns one tile of consecutive nodes for the whole launch and
// keeps everything static about them -- 31 x 4 weights, stencil positions,
// geometry -- and its running H and D in registers; per stage it waits for
// the tiles it exchanges rows with, copies its tile's distinct neighbour
// states into LDS, evaluates the stencil sums, the RHS and the stage update
// of its nodes, publishes their new state and its own stage count.  The
// static stencil data (1252 of the 1388 bytes a node costs one stage launch
// of the stage kernels) is read once per launch instead of once per stage.
//
// Same sums, same order and same stage arithmetic as k_stage_pair and
// k_stage_ptile (bitwise in exact mode and in fast mode); the RK stages and
// constants of RK_substep (rk4_rbffd_swe.c:30-72) and the fused stage
// kernels: stage 0 gathers H and writes Ka, stage 1 Ka -> Kb, stage 2
// Kb -> Ka, stage 3 Ka -> H.
//
// Cross-workgroup visibility: the L2 caches of the 8 XCDs are not coherent
// with each other, so every state access another workgroup depends on is
// agent-coherent (direct-to-LDS loads and buffer stores with the sc1 cache
// policy; the published states are small and live in the last-level
// cache), and the
// stage counts are agent-scope atomics written after the state stores
// have completed.  Every wait is bounded (spin_limit polls): a workgroup
// that exceeds it raises the abort flag, which every waiting workgroup
// sees, and all of them leave the kernel; the host reports the launch as
// failed.
#include "swe_ctx.h"

#include <algorithm>

namespace swe {

// Synchronisation state (unsigned words), zeroed before the first launch of
// a context and when the counts near 2^31: word 0 is the abort flag, then
// one 128-byte line per tile holding the number of stages the tile has
// completed, counted on across launches.  Instead of a grid-wide barrier a
// tile waits, before each stage, for the tiles it exchanges rows with (the
// owners of the rows it reads and the tiles that read its rows): they have
// then finished the previous stage, so the rows it reads are final and
// nobody still reads the buffer it is about to overwrite.  Neighbouring
// tiles are at most one stage apart and no wait depends on a later stage.
constexpr int kResLine = 32;

struct DevResident {
    double *H, *Ka, *Kb;
    unsigned *sync;     // [1 + ntiles][kResLine]
    unsigned *host_abort;   // pinned host word, set on a time-out
    const int *nbrs;    // [ntiles][kResMaxNbrs] tiles to wait for, -1 padded
    double c[4];        // K = H + c F (dt/2, dt/2, dt); stage 3: H += c D (dt/6)
    double dcoef[4];    // D += dcoef F (0: D = F, 2, 2, 1)
    int nsteps;
    int bytes;          // bytes of each state buffer (buffer-resource bound)
    int lds_rows;       // row halves the LDS holds (2 x the largest tile's rows)
    unsigned stage0;    // stages completed before this launch
    unsigned spin_limit;
};

__device__ __forceinline__ unsigned ld_agent(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_agent(unsigned *p, unsigned v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave 0 polls the neighbour tiles' stage counts (one per lane) until every
// one has reached `need`; false once any wait in the grid was aborted.
__device__ bool res_wait(const DevResident &r, int tile, unsigned need)
{
    __shared__ int ok;
    if (threadIdx.x < 64) {
        const int u = r.nbrs[tile * kResMaxNbrs + (int)threadIdx.x];
        const unsigned *f = r.sync + (long)(1 + (u >= 0 ? u : tile)) * kResLine;
        int good = 1;
        unsigned polls = 0;
        while (!__all(u < 0 || ld_agent(f) >= need)) {
            if (++polls > r.spin_limit || ld_agent(r.sync)) {
                if (threadIdx.x == 0) {
                    st_agent(r.sync, 1u);
                    __hip_atomic_store(r.host_abort, 1u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_SYSTEM);
                }
                good = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (threadIdx.x == 0)
            ok = good;
    }
    __syncthreads();
    return ok != 0;
}

// Every wave's state stores are complete, then the tile's count moves on.
__device__ __forceinline__ void res_publish(const DevResident &r, int tile, unsigned done)
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0)
        st_agent(r.sync + (long)(1 + tile) * kResLine, done);
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t state_rsrc(const double *p, int bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(p), (short)0, bytes,
                                             0x00020000);
}

constexpr int kSc1 = 16;   // cache policy: agent-coherent access

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// 16 bytes per lane from global memory straight into LDS, agent-coherent
// (global_load_lds_dwordx4 ... sc1): lane l's bytes land at lds + 16 l, lds
// being wave-uniform; the wave's vmcnt counts the load until they are there.
__device__ __forceinline__ void dma16_sc1(const void *src, void *lds)
{
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)lds, 16, 0,
                                     kSc1);
}

// Copies rows [0, count) of 16 bytes from byte offsets offs[] of `base`
// into rows[], one direct-to-LDS load per wave and 64 rows, then waits for
// this wave's copies.
__device__ __forceinline__ void copy_rows_sc1(const void *base, const int *offs, void *rows,
                                              int count)
{
    const int lane = (int)threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const char *b = static_cast<const char *>(base);
    char *r = static_cast<char *>(rows);
    for (int i0 = 64 * wv; i0 < count; i0 += (int)blockDim.x) {
        const int i = i0 + lane;
        if (i < count)
            dma16_sc1(b + offs[i], r + 16L * i0);
    }
    __builtin_amdgcn_s_waitcnt(0);
}

__device__ __forceinline__ void store_sc1(__amdgpu_buffer_rsrc_t r, int off_bytes, double2 v)
{
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, off_bytes, 0, kSc1);
}

template <bool EXACT, int NNBR>
__global__ __launch_bounds__(64 * kResMaxWaves) void k_steps_resident(DevMesh m, DevTiles tl,
                                                                      DevResident r)
{
    static_assert(NNBR > 0 && NNBR <= 31, "positions: 31 slots + the own row");
    // LDS: the tile's rows [row][half] = (u, v), (w, h), then the byte
    // offsets of those row halves in a state buffer
    extern __shared__ double2 rows[];
    const int tile = (int)blockIdx.x;   // one workgroup per tile
    const int t0 = tl.tile_start[tile];
    const int tcount = tl.tile_start[tile + 1] - t0;
    const int u0 = tl.ustart[tile];
    const int ucount = tl.ustart[tile + 1] - u0;
    const int tid = (int)threadIdx.x;
    const int nl = tid >> 1;
    const int k = tid & 1;
    const bool live = nl < tcount;
    const int node = t0 + (live ? nl : tcount - 1);   // dead pairs stay whole for DPP

    int *offs = reinterpret_cast<int *>(rows + r.lds_rows);
    for (int i = tid; i < 2 * ucount; i += (int)blockDim.x)
        offs[i] = 8 * (4 * tl.ulist[u0 + (i >> 1)] + 2 * (i & 1));
    // this lane's weights (2k, 2k + 1) of every slot
    const long ld = slot_stride(m);
    const double *__restrict__ wp = m.wts + 4L * slot_base<NNBR>(m, node) + 2 * k;
    double2 w[NNBR];
    static_for<NNBR>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        w[j] = *reinterpret_cast<const double2 *>(wp + 4L * j * ld);
    });
    // tile positions of slots 16k .. 16k + 15 (two per dword)
    const uint4 *pp = reinterpret_cast<const uint4 *>(tl.lidx_node + 32L * node) + 2 * k;
    const uint4 pa = pp[0], pb = pp[1];
    const unsigned pw[8] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w};
    double gv[(GEO_N + 1) / 2];
#pragma unroll
    for (int q = 0; q < (GEO_N + 1) / 2; ++q) {
        const int f = 2 * q + k;
        gv[q] = m.geo[(long)(f < GEO_N ? f : 0) * m.ld + node];
    }
    const int e = 8 * (4 * node + 2 * k);   // byte offset of this lane's two variables
    double2 H = *reinterpret_cast<const double2 *>(r.H + 4L * node + 2 * k);
    double2 D = make_double2(0.0, 0.0);
    double2 K = H;   // the state the next stage gathers, this lane's part
    unsigned stage = r.stage0;
    bool alive = true;

    for (int step = 0; step < r.nsteps && alive; ++step) {
        static_for<4>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            if (!alive)
                return;
            const double *src = s == 0 ? r.H : (s == 2 ? r.Kb : r.Ka);
            double *out = s == 0 ? r.Ka : (s == 1 ? r.Kb : (s == 2 ? r.Ka : r.H));
            if (!(alive = res_wait(r, tile, stage)))   // also: offsets written, rows free
                return;
            copy_rows_sc1(src, offs, rows, 2 * ucount);
            __syncthreads();
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;   // variable 2k   x (own0, own1, oth0, oth1)
            double b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;   // variable 2k+1 x (own0, own1, oth0, oth1)
            static_for<NNBR>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const unsigned wv = (unsigned)pair_bcast_i<j / 16>((int)pw[(j % 16) / 2]);
                const int pos = (int)((j & 1) ? (wv >> 16) : (wv & 0xffffu));
                const double2 sv = rows[2 * pos + k];
                const double o0 = pair_swap(w[j].x), o1 = pair_swap(w[j].y);
                a0 = madd<EXACT>(a0, w[j].x, sv.x); a1 = madd<EXACT>(a1, w[j].y, sv.x);
                a2 = madd<EXACT>(a2, o0, sv.x);     a3 = madd<EXACT>(a3, o1, sv.x);
                b0 = madd<EXACT>(b0, w[j].x, sv.y); b1 = madd<EXACT>(b1, w[j].y, sv.y);
                b2 = madd<EXACT>(b2, o0, sv.y);     b3 = madd<EXACT>(b3, o1, sv.y);
            });
            const double ax = k ? a2 : a0, ay = k ? a3 : a1, az = k ? a0 : a2, al = k ? a1 : a3;
            const double bx = k ? b2 : b0, by = k ? b3 : b1, bz = k ? b0 : b2, bl = k ? b1 : b3;
            const double pax = pair_swap(ax), pay = pair_swap(ay), paz = pair_swap(az),
                         pal = pair_swap(al);
            const double pbx = pair_swap(bx), pby = pair_swap(by), pbz = pair_swap(bz),
                         pbl = pair_swap(bl);
            Sums S;
            S.dudx = k ? pax : ax; S.dudy = k ? pay : ay; S.dudz = k ? paz : az; S.lu = k ? pal : al;
            S.dvdx = k ? pbx : bx; S.dvdy = k ? pby : by; S.dvdz = k ? pbz : bz; S.lv = k ? pbl : bl;
            S.dwdx = k ? ax : pax; S.dwdy = k ? ay : pay; S.dwdz = k ? az : paz; S.lw = k ? al : pal;
            S.dhdx = k ? bx : pbx; S.dhdy = k ? by : pby; S.dhdz = k ? bz : pbz; S.lh = k ? bl : pbl;
            const double po0 = pair_swap(K.x), po1 = pair_swap(K.y);
            const double u = k ? po0 : K.x, v = k ? po1 : K.y;
            const double ww = k ? K.x : po0, h = k ? K.y : po1;
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
            rhs_point(S, u, v, ww, h, g, m.gh0, F);
            const double f0 = k ? F[2] : F[0], f1 = k ? F[3] : F[1];
            // the stage update of stage_update_pair, on registers
            if constexpr (s == 0) {
                D = make_double2(f0, f1);
                K = make_double2(axpy<EXACT>(H.x, r.c[0], f0), axpy<EXACT>(H.y, r.c[0], f1));
            } else if constexpr (s == 1 || s == 2) {
                D = make_double2(axpy<EXACT>(D.x, r.dcoef[s], f0), axpy<EXACT>(D.y, r.dcoef[s], f1));
                K = make_double2(axpy<EXACT>(H.x, r.c[s], f0), axpy<EXACT>(H.y, r.c[s], f1));
            } else {
                const double d0 = axpy<EXACT>(D.x, r.dcoef[3], f0);
                const double d1 = axpy<EXACT>(D.y, r.dcoef[3], f1);
                H = make_double2(axpy<EXACT>(H.x, r.c[3], d0), axpy<EXACT>(H.y, r.c[3], d1));
                K = H;
            }
            if (live)
                store_sc1(state_rsrc(out, r.bytes), e, K);
            res_publish(r, tile, ++stage);
        });
    }
}

// FP32 (the reference's RUN_32BIT type), one lane per node: the sums,
// RHS and stage update of k_stage_lane32 (bitwise equal to it in both
// modes, and in exact mode to the reference's FP32 build), the tile's rows
// as 16-byte float4 in LDS, the same neighbour waits.
struct DevResident32 {
    float *H, *Ka, *Kb;
    unsigned *sync;
    unsigned *host_abort;
    const int *nbrs;
    float c[3];         // K = H + c F: (float)(dt/2.0), (float)(dt/2.0), dt
    double c6;          // stage 3: H += (dt/6.0) D in double
    float dcoef[4];     // D += dcoef F (0: D = F, 2, 2, 1)
    int nsteps;
    int bytes;
    int lds_rows;       // rows the LDS holds (the largest tile's)
    unsigned stage0;
    unsigned spin_limit;
};

__device__ __forceinline__ void store_sc1_f4(__amdgpu_buffer_rsrc_t r, int off_bytes, float4 v)
{
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, off_bytes, 0, kSc1);
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t state_rsrc32(const float *p, int bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), (short)0, bytes, 0x00020000);
}

template <bool EXACT, int NNBR>
__global__ __launch_bounds__(64 * kResMaxWaves) void k_steps_resident32(DevMesh32 m, DevTiles tl,
                                                                        DevResident32 r)
{
    static_assert(NNBR > 0 && NNBR <= 32, "positions: 32 16-bit slots per node");
    extern __shared__ float4 rows32[];   // the tile's rows, then their byte offsets
    const int tile = (int)blockIdx.x;
    const int t0 = tl.tile_start[tile];
    const int tcount = tl.tile_start[tile + 1] - t0;
    const int u0 = tl.ustart[tile];
    const int ucount = tl.ustart[tile + 1] - u0;
    const int tid = (int)threadIdx.x;
    const bool live = tid < tcount;
    const int node = t0 + (live ? tid : tcount - 1);
    int *offs = reinterpret_cast<int *>(rows32 + r.lds_rows);
    for (int i = tid; i < ucount; i += (int)blockDim.x)
        offs[i] = 16 * tl.ulist[u0 + i];
    const long ld = m.sblk ? 16L : (long)m.ld;
    const long sb = m.sblk ? ((long)(node >> 4) * NNBR << 4) + (node & 15) : (long)node;
    const float *__restrict__ wp = m.wts + 4L * sb;
    float4 w[NNBR];
    static_for<NNBR>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        w[j] = *reinterpret_cast<const float4 *>(wp + 4L * j * ld);
    });
    unsigned pw[16];
    {
        const uint4 *pp = reinterpret_cast<const uint4 *>(tl.lidx_node + 32L * node);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 v = pp[q];
            pw[4 * q] = v.x; pw[4 * q + 1] = v.y; pw[4 * q + 2] = v.z; pw[4 * q + 3] = v.w;
        }
    }
    float g[GEO_N];
#pragma unroll
    for (int f = 0; f < GEO_N; ++f)
        g[f] = m.geo[(long)f * m.ld + node];
    const int e = 16 * node;
    float4 H = *reinterpret_cast<const float4 *>(r.H + 4L * node);
    float4 D = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 K = H;
    unsigned stage = r.stage0;
    bool alive = true;
    DevResident sync_view{};   // the fields res_wait and res_publish read
    sync_view.sync = r.sync;
    sync_view.host_abort = r.host_abort;
    sync_view.nbrs = r.nbrs;
    sync_view.spin_limit = r.spin_limit;

    for (int step = 0; step < r.nsteps && alive; ++step) {
        static_for<4>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            if (!alive)
                return;
            const float *src = s == 0 ? r.H : (s == 2 ? r.Kb : r.Ka);
            float *out = s == 0 ? r.Ka : (s == 1 ? r.Kb : (s == 2 ? r.Ka : r.H));
            if (!(alive = res_wait(sync_view, tile, stage)))   // also: offsets written, rows free
                return;
            copy_rows_sc1(src, offs, rows32, ucount);
            __syncthreads();
            Sums32 S = {};
            static_for<NNBR>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const unsigned pv = pw[j >> 1];
                const float4 sv = rows32[(j & 1) ? (pv >> 16) : (pv & 0xffffu)];
                const float4 wj = w[j];
                S.dudx = madd32<EXACT>(S.dudx, wj.x, sv.x); S.dvdx = madd32<EXACT>(S.dvdx, wj.x, sv.y);
                S.dwdx = madd32<EXACT>(S.dwdx, wj.x, sv.z); S.dhdx = madd32<EXACT>(S.dhdx, wj.x, sv.w);
                S.dudy = madd32<EXACT>(S.dudy, wj.y, sv.x); S.dvdy = madd32<EXACT>(S.dvdy, wj.y, sv.y);
                S.dwdy = madd32<EXACT>(S.dwdy, wj.y, sv.z); S.dhdy = madd32<EXACT>(S.dhdy, wj.y, sv.w);
                S.dudz = madd32<EXACT>(S.dudz, wj.z, sv.x); S.dvdz = madd32<EXACT>(S.dvdz, wj.z, sv.y);
                S.dwdz = madd32<EXACT>(S.dwdz, wj.z, sv.z); S.dhdz = madd32<EXACT>(S.dhdz, wj.z, sv.w);
                S.lu = madd32<EXACT>(S.lu, wj.w, sv.x); S.lv = madd32<EXACT>(S.lv, wj.w, sv.y);
                S.lw = madd32<EXACT>(S.lw, wj.w, sv.z); S.lh = madd32<EXACT>(S.lh, wj.w, sv.w);
            });
            float F[4];
            rhs_point32(S, K.x, K.y, K.z, K.w, g, m.gh0, F);
            // the stage update of stage_update32, on registers
            if constexpr (s == 0) {
                D = make_float4(F[0], F[1], F[2], F[3]);
                K = make_float4(axpy32<EXACT>(H.x, r.c[0], F[0]), axpy32<EXACT>(H.y, r.c[0], F[1]),
                                axpy32<EXACT>(H.z, r.c[0], F[2]), axpy32<EXACT>(H.w, r.c[0], F[3]));
            } else {
                D = make_float4(axpy32<EXACT>(D.x, r.dcoef[s], F[0]),
                                axpy32<EXACT>(D.y, r.dcoef[s], F[1]),
                                axpy32<EXACT>(D.z, r.dcoef[s], F[2]),
                                axpy32<EXACT>(D.w, r.dcoef[s], F[3]));
                if constexpr (s == 3) {
                    H = make_float4(update_h32<EXACT>(H.x, r.c6, D.x), update_h32<EXACT>(H.y, r.c6, D.y),
                                    update_h32<EXACT>(H.z, r.c6, D.z), update_h32<EXACT>(H.w, r.c6, D.w));
                    K = H;
                } else {
                    K = make_float4(axpy32<EXACT>(H.x, r.c[s], F[0]), axpy32<EXACT>(H.y, r.c[s], F[1]),
                                    axpy32<EXACT>(H.z, r.c[s], F[2]), axpy32<EXACT>(H.w, r.c[s], F[3]));
                }
            }
            if (live)
                store_sc1_f4(state_rsrc32(out, r.bytes), e, K);
            res_publish(sync_view, tile, ++stage);
        });
    }
}

__global__ void k_res_clear(unsigned *p, int words)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < words)
        p[i] = 0u;
}

namespace {

template <bool EX>
void *resident_kernel(bool fp32)
{
    return fp32 ? reinterpret_cast<void *>(&k_steps_resident32<EX, 31>)
                : reinterpret_cast<void *>(&k_steps_resident<EX, 31>);
}

}  // namespace

long resident_sync_bytes(int ntiles) { return (1L + ntiles) * kResLine * sizeof(unsigned); }

// LDS budget of both kernel instances and a check that one workgroup per
// tile can be resident at once (the cooperative launch needs it).  The
// budget is a property of the kernel, shared by every context: it is set to
// what the largest tile any context may build needs (kResMaxRows rows), so
// that a context created later never lowers it below an earlier one's use.
int resident_prepare(swe_hip_ctx *c, std::string *err)
{
    const int lds_max = kResMaxRows * (c->fp32 ? (int)(sizeof(float4) + sizeof(int))
                                               : 2 * (int)(sizeof(double2) + sizeof(int)));
    if (c->res_lds > lds_max) {
        *err = "resident kernel: LDS budget";
        return SWE_HIP_EDEVICE;
    }
    for (bool ex : {false, true}) {
        void *fn = ex ? resident_kernel<true>(c->fp32) : resident_kernel<false>(c->fp32);
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max)
            != hipSuccess) {
            *err = "resident kernel: LDS budget";
            return SWE_HIP_EDEVICE;
        }
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, c->res_block, c->res_lds)
                != hipSuccess
            || (long)per_cu * c->ncu < c->tiles.ntiles) {
            *err = "resident kernel: the launch cannot be co-resident on this device";
            return SWE_HIP_EDEVICE;
        }
    }
    return 0;
}

int resident_launch(swe_hip_ctx *c, int nsteps, std::string *err);

// n fused RK steps in cooperative launches of at most 2^20 steps each.  The
// launches are not waited for: a time-out sets a pinned host word, read by
// resident_check once the stream has drained.
int resident_steps(swe_hip_ctx *c, int nsteps, std::string *err)
{
    if (c->res_pending && hipStreamQuery(c->stream) == hipSuccess)
        if (int rc = resident_check(c, err))
            return rc;
    for (int done = 0; done < nsteps;) {
        const int n = std::min(nsteps - done, 1 << 20);
        if (int rc = resident_launch(c, n, err))
            return rc;
        done += n;
    }
    return 0;
}

int resident_check(swe_hip_ctx *c, std::string *err)
{
    if (!c->res_pending)
        return 0;
    c->res_pending = false;
    if (*c->h_res_abort) {
        *c->h_res_abort = 0;
        c->res_stages = 0;   // the counts are inconsistent: clear them next time
        *err = "resident kernel: a wait for neighbouring tiles exceeded its poll limit (the "
               "state is invalid)";
        return SWE_HIP_EDEVICE;
    }
    return 0;
}

int resident_launch(swe_hip_ctx *c, int nsteps, std::string *err)
{
    const long long stages = 4LL * nsteps;
    if (c->res_stages == 0 || c->res_stages + stages > (1LL << 31)) {
        const int words = (int)(resident_sync_bytes(c->tiles.ntiles) / sizeof(unsigned));
        k_res_clear<<<(words + 255) / 256, 256, 0, c->stream>>>(
            reinterpret_cast<unsigned *>(c->d_res_sync), words);
        c->res_stages = 0;
    }
    DevTiles tl = tiles_of(c, 1);
    void *fn = exact(c) ? resident_kernel<true>(c->fp32) : resident_kernel<false>(c->fp32);
    hipError_t e = hipSuccess;
    if (c->fp32) {
        DevMesh32 m = mesh32_of(c);
        DevResident32 r{};
        r.H = c->f_H;
        r.Ka = c->f_Ka;
        r.Kb = c->f_Kb;
        r.sync = reinterpret_cast<unsigned *>(c->d_res_sync);
        r.host_abort = c->d_res_abort;
        r.nbrs = c->d_res_nbrs;
        // the constants of the FP32 stage kernels (rk4_rbffd_swe.c:123-186)
        const float dt = (float)c->dt;
        r.c[0] = (float)(dt / 2.0); r.c[1] = (float)(dt / 2.0); r.c[2] = dt;
        r.c6 = dt / 6.0;
        r.dcoef[0] = 0.0f; r.dcoef[1] = 2.0f; r.dcoef[2] = 2.0f; r.dcoef[3] = 1.0f;
        r.nsteps = nsteps;
        r.bytes = (int)(state_count(c) * (long)sizeof(float));
        r.spin_limit = c->res_spin;
        r.lds_rows = c->tiles.umax_used;
        r.stage0 = (unsigned)c->res_stages;
        void *args[] = {&m, &tl, &r};
        e = hipLaunchCooperativeKernel(fn, dim3(c->tiles.ntiles), dim3(c->res_block), args,
                                       (unsigned)c->res_lds, c->stream);
    } else {
        DevMesh m = mesh_of(c);
        DevResident r{};
        r.H = c->d_H;
        r.Ka = c->d_Ka;
        r.Kb = c->d_Kb;
        r.sync = reinterpret_cast<unsigned *>(c->d_res_sync);
        r.host_abort = c->d_res_abort;
        r.nbrs = c->d_res_nbrs;
        const double dt = c->dt;
        r.c[0] = dt / 2.0; r.c[1] = dt / 2.0; r.c[2] = dt; r.c[3] = dt / 6.0;
        r.dcoef[0] = 0.0; r.dcoef[1] = 2.0; r.dcoef[2] = 2.0; r.dcoef[3] = 1.0;
        r.nsteps = nsteps;
        r.bytes = (int)(state_count(c) * (long)sizeof(double));
        r.spin_limit = c->res_spin;
        r.lds_rows = 2 * c->tiles.umax_used;
        r.stage0 = (unsigned)c->res_stages;
        void *args[] = {&m, &tl, &r};
        e = hipLaunchCooperativeKernel(fn, dim3(c->tiles.ntiles), dim3(c->res_block), args,
                                       (unsigned)c->res_lds, c->stream);
    }
    if (e != hipSuccess) {
        c->res_stages = 0;
        *err = std::string("resident kernel launch: ") + hipGetErrorString(e);
        return SWE_HIP_EDEVICE;
    }
    c->res_stages += stages;
    c->res_pending = true;
    return 0;
}

}  // namespace swe
