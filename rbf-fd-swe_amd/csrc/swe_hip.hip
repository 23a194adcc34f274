// C-ABI implementation of libswe_hip.so (include/swe_hip.h): context
// set-up and device repacking of the static model data, the drop-in stage
// driver (RK_substep semantics, swe_code/main/rk4_rbffd_swe.c:30-186), the
// fused and graph-captured fast path, state upload/read-back in the
// reference's host layouts (swe_code/main/include/swe_config.h:21-54).
#include "swe_ctx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace swe;

namespace swe {

// ---------------------------------------------------------------------------
// Element-wise kernels of the drop-in stage driver (rk4_rbffd_swe.c:74-186),
// over `count` contiguous values.
__global__ __launch_bounds__(256) void k_copy(double *__restrict__ dst,
                                              const double *__restrict__ src, long count)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        dst[i] = src[i];
}

template <bool EXACT>
__global__ __launch_bounds__(256) void k_axpy_inplace(double *__restrict__ y, double a,
                                                      const double *__restrict__ x, long count)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        y[i] = axpy<EXACT>(y[i], a, x[i]);
}

template <bool EXACT>
__global__ __launch_bounds__(256) void k_axpy_to(double *__restrict__ out,
                                                 const double *__restrict__ y, double a,
                                                 const double *__restrict__ x, long count)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        out[i] = axpy<EXACT>(y[i], a, x[i]);
}

// Member slot <-> node-major [node][var] staging copy.
__global__ __launch_bounds__(256) void k_put_member(double *__restrict__ state,
                                                    const double *__restrict__ flat,
                                                    int member, int B, long count)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        state[i * B + member] = flat[i];
}

__global__ void k_set_int(int *dst, int value)
{
    if (threadIdx.x == 0)
        dst[threadIdx.x] = value;
}

// Copies member slot `from` to slot `to` for every (node, var) row.
__global__ __launch_bounds__(256) void k_move_member(double *__restrict__ state, int from,
                                                     int to, int B, long rows)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows)
        state[i * B + to] = state[i * B + from];
}

__global__ __launch_bounds__(256) void k_get_member(double *__restrict__ flat,
                                                    const double *__restrict__ state,
                                                    int member, int B, long count)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        flat[i] = state[i * B + member];
}

// Element-wise kernels of the FP32 drop-in stage driver over `count` values.
__global__ __launch_bounds__(256) void k_copy32(float *__restrict__ dst,
                                                const float *__restrict__ src, long count)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        dst[i] = src[i];
}

template <bool EXACT>
__global__ __launch_bounds__(256) void k_axpy_inplace32(float *__restrict__ y, float a,
                                                        const float *__restrict__ x, long count)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        y[i] = axpy32<EXACT>(y[i], a, x[i]);
}

template <bool EXACT>
__global__ __launch_bounds__(256) void k_axpy_to32(float *__restrict__ out,
                                                   const float *__restrict__ y, float a,
                                                   const float *__restrict__ x, long count)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        out[i] = axpy32<EXACT>(y[i], a, x[i]);
}

template <bool EXACT>
__global__ __launch_bounds__(256) void k_update_h32(float *__restrict__ H, double c6,
                                                    const float *__restrict__ D, long count)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        H[i] = update_h32<EXACT>(H[i], c6, D[i]);
}

}  // namespace swe

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

}  // namespace

int swe::set_error(int code, const std::string &msg) { return fail(code, msg); }

namespace {

// After the context's stream has drained: a grid-barrier time-out of the
// last resident launch (its state is invalid) becomes this call's error.
int check_resident(swe_hip_ctx *c)
{
    std::string why;
    if (int rc = swe::resident_check(c, &why))
        return fail(rc, why);
    return 0;
}

}  // namespace

namespace {

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(SWE_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// RK steps per captured graph: a call of n steps replays the 64-step graph
// n / 64 times and then at most three of each shorter one (a replay
// boundary costs a few microseconds)
constexpr int kGraphSteps[kGraphCount] = {64, 16, 4, 1};

}  // namespace

namespace {

// Fused stage s of one RK step: gathers from H (s=0), Ka (s=1), Kb (s=2),
// Ka (s=3); writes Ka, Kb, Ka, H.
// FP32 stage s of one RK step: the same buffers and constants in float;
// stage 3 keeps dt/6.0 in double (update_H, rk4_rbffd_swe.c:86-121).
void launch_fused32(const swe_hip_ctx *c, int s)
{
    DevStage32 st{};
    st.h = c->f_H;
    st.d = c->f_D;
    const float dt = (float)c->dt;
    switch (s) {
    case 0:
        st.src = c->f_H; st.out = c->f_Ka; st.c = (float)(dt / 2.0); st.dcoef = 0.0f;
        break;
    case 1:
        st.src = c->f_Ka; st.out = c->f_Kb; st.c = (float)(dt / 2.0); st.dcoef = 2.0f;
        break;
    case 2:
        st.src = c->f_Kb; st.out = c->f_Ka; st.c = dt; st.dcoef = 2.0f;
        break;
    default:
        st.src = c->f_Ka; st.out = c->f_H; st.c6 = dt / 6.0; st.dcoef = 1.0f;
        break;
    }
    launch_fp32_stage(c, s, st);
}

void launch_fused(const swe_hip_ctx *c, int s)
{
    if (c->fp32) {
        launch_fused32(c, s);
        return;
    }
    DevStage st{};
    st.members = c->B;
    st.active = c->d_active;
    st.h = c->d_H;
    st.d = c->d_D;
    const double dt = c->dt;
    switch (s) {
    case 0:
        st.src = c->d_H; st.out = c->d_Ka; st.c = dt / 2.0; st.dcoef = 0.0;
        launch_stage(c, ST_0, st);
        break;
    case 1:
        st.src = c->d_Ka; st.out = c->d_Kb; st.c = dt / 2.0; st.dcoef = 2.0;
        launch_stage(c, ST_1, st);
        break;
    case 2:
        st.src = c->d_Kb; st.out = c->d_Ka; st.c = dt; st.dcoef = 2.0;
        launch_stage(c, ST_2, st);
        break;
    default:
        st.src = c->d_Ka; st.out = c->d_H; st.c = dt / 6.0; st.dcoef = 1.0;
        launch_stage(c, ST_3, st);
        break;
    }
}

void launch_rhs(const swe_hip_ctx *c, const double *src, double *F)
{
    DevStage st{};
    st.members = c->B;
    st.active = c->d_active;
    st.src = src;
    st.h = src;
    st.out = F;
    launch_stage(c, ST_RHS, st);
}

int elementwise_blocks(long count) { return (int)((count + kBlock - 1) / kBlock); }

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(SWE_HIP_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return 0;
}

// Runs `fn` and, when timing is on, synchronises and adds the wall time to
// *slot (clFinish-per-launch timing of RK_ocl.c).
template <typename Fn>
int timed(swe_hip_ctx *c, double *slot, const char *what, Fn fn)
{
    const auto t0 = std::chrono::steady_clock::now();
    fn();
    int rc = check_launch(what);
    if (rc)
        return rc;
    if (c->timing && slot) {
        HIPCHK(hipStreamSynchronize(c->stream));
        *slot += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return 0;
}

int capture_steps(swe_hip_ctx *c, int steps, hipGraphExec_t *exec)
{
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < steps; ++i)
        for (int s = 0; s < 4; ++s)
            launch_fused(c, s);
    hipError_t le = hipGetLastError();
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (le != hipSuccess || e != hipSuccess) {
        if (graph)
            (void)hipGraphDestroy(graph);
        return fail(SWE_HIP_EDEVICE, std::string("graph capture: ")
                                         + hipGetErrorString(le != hipSuccess ? le : e));
    }
    e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess)
        return fail(SWE_HIP_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return 0;
}

void release_mesh(SweMesh *m)
{
    if (!m || --m->refs > 0)
        return;
    (void)hipSetDevice(m->dev);
    if (m->d_nbr) (void)hipFree(m->d_nbr);
    if (m->d_wts) (void)hipFree(m->d_wts);
    if (m->d_geo) (void)hipFree(m->d_geo);
    if (m->d_wts32) (void)hipFree(m->d_wts32);
    if (m->d_geo32) (void)hipFree(m->d_geo32);
    delete m;
}

void free_ctx(swe_hip_ctx *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->dev);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipGraphExec_t g : c->graphs)
        if (g) (void)hipGraphExecDestroy(g);
    void *ptrs[] = {c->d_H, c->d_Ka, c->d_Kb, c->d_D, c->d_F, c->d_stage, c->d_active,
                    c->f_H, c->f_Ka, c->f_Kb, c->f_D, c->f_F,
                    c->tiles.d_tile_start, c->tiles.d_ustart, c->tiles.d_ulist,
                    c->tiles.d_lidx, c->tiles.d_lidx_node, c->tiles.d_ulist_tile,
                    c->tiles.d_wts_node, c->tiles.d_geo_node, c->d_res_sync, c->d_res_nbrs};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->h_res_abort) (void)hipHostFree(c->h_res_abort);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    release_mesh(c->mesh);
    delete c;
}

int dev_alloc(long long *counter, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 8);
    if (e != hipSuccess)
        return fail(SWE_HIP_ENOMEM, std::string("hipMalloc(") + std::to_string(bytes)
                                        + "): " + hipGetErrorString(e));
    *counter += (long long)bytes;
    return 0;
}

int member_block(int B, int want)
{
    if (B == 1)
        return 1;
    if (want > 0)
        return (want >= 2 && want <= 64 && (want & (want - 1)) == 0 && B % want == 0) ? want : -1;
    int mb = 64;
    while (mb > 2 && (B % mb) != 0)
        mb >>= 1;
    return (B % mb) == 0 ? mb : -1;
}

// Checks sizes, layout and pointers of the static data (no device use).
int validate_static(const swe_hip_static *s)
{
    const int n = s->compute_size;
    if (n <= 0 || s->Nnbr < 1 || s->Nnbr > 64 || s->padded_Nnbr < s->Nnbr || s->simd < 1)
        return fail(SWE_HIP_EINVAL, "bad sizes (compute_size, Nnbr 1..64, padded_Nnbr, simd)");
    if (s->compute_pid_s != 0 || s->patch_size != n)
        return fail(SWE_HIP_EINVAL,
                    "only the serial single-patch layout is supported (compute_pid_s == 0, "
                    "patch_size == compute_size): the whole sphere runs on one GPU");
    if ((!s->cfdl || s->sfdl) && n % s->simd != 0)
        return fail(SWE_HIP_EINVAL, "compute_size must be a multiple of SIMD_LENGTH");
    if (!s->x || !s->y || !s->z || !s->f || !s->ghm || !s->p_u || !s->p_v || !s->p_w
        || !s->gradghm || !s->idx || !s->Dx || !s->Dy || !s->Dz || !s->L)
        return fail(SWE_HIP_EINVAL, "a static array pointer is NULL");
    return 0;
}

// Builds the device copy of validated static data; every stencil id is
// range-checked on the host before anything reaches the device.
// 3-D Morton code of a point of the unit cube [-1, 1]^3 (21 bits per axis).
uint64_t morton3(double x, double y, double z)
{
    auto spread = [](uint64_t v) {
        v &= 0x1fffff;
        v = (v | v << 32) & 0x1f00000000ffffULL;
        v = (v | v << 16) & 0x1f0000ff0000ffULL;
        v = (v | v << 8) & 0x100f00f00f00f00fULL;
        v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
        v = (v | v << 2) & 0x1249249249249249ULL;
        return v;
    };
    auto q = [](double a) {
        double t = (a + 1.0) * 0.5 * 2097151.0;
        t = t < 0.0 ? 0.0 : (t > 2097151.0 ? 2097151.0 : t);
        return (uint64_t)t;
    };
    return spread(q(x)) | spread(q(y)) << 1 | spread(q(z)) << 2;
}

// 3-D Hilbert index of a point of [-1, 1]^3 (21 bits per axis): Skilling's
// transform of the axes to the transposed index (J. Skilling, "Programming
// the Hilbert curve", AIP Conf. Proc. 707, 2004), then bit interleaving.
// Consecutive indices are always neighbouring cells, so a run of nodes in
// this order is more compact than in Morton order (fewer, fuller tiles).
uint64_t hilbert3(double x, double y, double z)
{
    auto q = [](double a) {
        double t = (a + 1.0) * 0.5 * 2097151.0;
        t = t < 0.0 ? 0.0 : (t > 2097151.0 ? 2097151.0 : t);
        return (uint32_t)t;
    };
    uint32_t X[3] = {q(x), q(y), q(z)};
    const uint32_t M = 1u << 20;
    for (uint32_t Q = M; Q > 1; Q >>= 1) {   // inverse undo
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < 3; ++i)   // Gray encode
        X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q)
            t ^= Q - 1;
    for (int i = 0; i < 3; ++i)
        X[i] ^= t;
    uint64_t key = 0;
    for (int b = 20; b >= 0; --b)
        for (int i = 0; i < 3; ++i)
            key = key << 1 | ((X[i] >> b) & 1u);
    return key;
}

int build_mesh(const swe_hip_static *s, int device, int order, int sblk, bool fp32,
               SweMesh **out)
{
    *out = nullptr;
    const int n = s->compute_size;
    const HostLayout lay{s->cfdl ? 1 : 0, s->sfdl ? 1 : 0, s->simd};
    const int W = s->Nnbr, pW = s->padded_Nnbr;
    // Device node order: the host's (RCM) order, or a Morton or Hilbert
    // curve through the node coordinates (compact tiles for the LDS-staged
    // kernels).  The arithmetic of every node is the same in any order.
    std::vector<int> perm, inv;
    if (order == SWE_HIP_ORDER_MORTON || order == SWE_HIP_ORDER_HILBERT) {
        std::vector<std::pair<uint64_t, int>> key(n);
        for (int i = 0; i < n; ++i)
            key[i] = {order == SWE_HIP_ORDER_MORTON ? morton3(s->x[i], s->y[i], s->z[i])
                                                    : hilbert3(s->x[i], s->y[i], s->z[i]),
                      i};
        std::sort(key.begin(), key.end());
        perm.resize(n);
        inv.resize(n);
        for (int d = 0; d < n; ++d) {
            perm[d] = key[d].second;
            inv[key[d].second] = d;
        }
    }
    auto host_of = [&](int d) { return perm.empty() ? d : perm[d]; };
    auto dev_of = [&](int h) { return inv.empty() ? h : inv[h]; };
    // Repack into the device layouts, checking every stencil id.  nbr keeps
    // the slot-major form [nnbr][n] (the tile builder reads it); the device
    // stencil arrays are slot-major or in node blocks of 16.
    const size_t nround = sblk ? ((size_t)n + 15) / 16 * 16 : (size_t)n;
    auto pos = [&](int j, int dnode) -> size_t {
        return sblk ? (((size_t)(dnode >> 4) * W + j) << 4) + (dnode & 15)
                    : (size_t)j * n + dnode;
    };
    std::vector<int> nbr((size_t)W * n);
    std::vector<int> dnbr(sblk ? (size_t)W * nround : 0);
    std::vector<double> wts((size_t)W * nround * 4);
    std::vector<double> geo((size_t)GEO_N * n);
    for (int dnode = 0; dnode < n; ++dnode) {
        const int i = host_of(dnode);
        for (int j = 0; j < W; ++j) {
            const long d = lay.dm(i, j, pW);
            const int q = s->idx[d];
            if (q < 0 || q >= n)
                return fail(SWE_HIP_ERANGE, "idx[" + std::to_string(d) + "] = " + std::to_string(q)
                                                + " outside [0, " + std::to_string(n) + ")");
            nbr[(size_t)j * n + dnode] = dev_of(q);
            if (sblk)
                dnbr[pos(j, dnode)] = dev_of(q);
            double *w = &wts[pos(j, dnode) * 4];
            w[0] = s->Dx[d];
            w[1] = s->Dy[d];
            w[2] = s->Dz[d];
            w[3] = s->L[d];
        }
        geo[(size_t)G_X * n + dnode] = s->x[i];
        geo[(size_t)G_Y * n + dnode] = s->y[i];
        geo[(size_t)G_Z * n + dnode] = s->z[i];
        geo[(size_t)G_F * n + dnode] = s->f[i];
        geo[(size_t)G_GHM * n + dnode] = s->ghm[i];
        for (int c3 = 0; c3 < 3; ++c3) {
            const long v = lay.vect(i, c3);
            geo[(size_t)(G_PU + c3) * n + dnode] = s->p_u[v];
            geo[(size_t)(G_PV + c3) * n + dnode] = s->p_v[v];
            geo[(size_t)(G_PW + c3) * n + dnode] = s->p_w[v];
            geo[(size_t)(G_GG + c3) * n + dnode] = s->gradghm[v];
        }
    }
    if (sblk) {   // rows of the last block beyond n repeat the last node
        for (int dnode = n; dnode < (int)nround; ++dnode)
            for (int j = 0; j < W; ++j) {
                dnbr[pos(j, dnode)] = dnbr[pos(j, n - 1)];
                for (int c4 = 0; c4 < 4; ++c4)
                    wts[pos(j, dnode) * 4 + c4] = wts[pos(j, n - 1) * 4 + c4];
            }
    }
    const std::vector<int> &upload_nbr = sblk ? dnbr : nbr;
    SweMesh *m = new SweMesh();
    m->order = order;
    m->sblk = sblk;
    m->dev = device;
    m->n = n;
    m->ld = n;
    m->nnbr = W;
    m->gh0 = s->gh0;
    m->dt = s->dt;
    m->lay = lay;
    m->fp32 = fp32;
    int rc = 0;
    hipError_t e = hipSuccess;
    if (!rc) rc = dev_alloc(&m->bytes, (void **)&m->d_nbr, sizeof(int) * upload_nbr.size());
    if (!rc)
        e = hipMemcpy(m->d_nbr, upload_nbr.data(), sizeof(int) * upload_nbr.size(),
                      hipMemcpyHostToDevice);
    if (fp32) {   // values rounded to float (exact for data read as FP32)
        const std::vector<float> w32(wts.begin(), wts.end()), g32(geo.begin(), geo.end());
        if (!rc) rc = dev_alloc(&m->bytes, (void **)&m->d_wts32, sizeof(float) * w32.size());
        if (!rc) rc = dev_alloc(&m->bytes, (void **)&m->d_geo32, sizeof(float) * g32.size());
        if (!rc && e == hipSuccess)
            e = hipMemcpy(m->d_wts32, w32.data(), sizeof(float) * w32.size(),
                          hipMemcpyHostToDevice);
        if (!rc && e == hipSuccess)
            e = hipMemcpy(m->d_geo32, g32.data(), sizeof(float) * g32.size(),
                          hipMemcpyHostToDevice);
    } else {
        if (!rc) rc = dev_alloc(&m->bytes, (void **)&m->d_wts, sizeof(double) * wts.size());
        if (!rc) rc = dev_alloc(&m->bytes, (void **)&m->d_geo, sizeof(double) * geo.size());
        if (!rc && e == hipSuccess)
            e = hipMemcpy(m->d_wts, wts.data(), sizeof(double) * wts.size(),
                          hipMemcpyHostToDevice);
        if (!rc && e == hipSuccess)
            e = hipMemcpy(m->d_geo, geo.data(), sizeof(double) * geo.size(),
                          hipMemcpyHostToDevice);
    }
    if (!rc && e != hipSuccess)
        rc = fail(SWE_HIP_EDEVICE, std::string("static data upload: ") + hipGetErrorString(e));
    if (rc) {
        release_mesh(m);
        return rc;
    }
    m->perm = std::move(perm);
    m->h_nbr = std::move(nbr);
    *out = m;
    return 0;
}

// Greedy tiles along the device node order: a tile grows while it has fewer
// than tmax nodes and its distinct stencil nodes fit umax LDS rows.
// Node-major copies of the per-node data a member-lane tile item stages,
// built once per context: wts_node [n][32][4] (slots past the stencil
// zero) and geo_node [n][32] (fields 0..16, zeros, the node's 32 16-bit
// tile positions as doubles 24..31).  Thread = (node, slot).
__global__ void k_pack_tile_nodes(DevMesh m, int n, const unsigned short *lidx_node,
                                  double *wts_node, double *geo_node)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 32L * n)
        return;
    const int node = (int)(t >> 5), j = (int)(t & 31);
    double w[4] = {0.0, 0.0, 0.0, 0.0};
    if (j < m.nnbr) {
        const double *p = m.wts + 4L * slot_base<0>(m, node) + 4L * j * slot_stride(m);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            w[k] = p[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        wts_node[4 * t + k] = w[k];
    double g = 0.0;
    if (j < GEO_N)
        g = m.geo[(long)j * m.ld + node];
    else if (j >= kTmemPosSlot)
        g = reinterpret_cast<const double *>(lidx_node + 32L * node)[j - kTmemPosSlot];
    geo_node[t] = g;
}

int build_tiles(swe_hip_ctx *c, int tmax, int umax, std::vector<int> *h_tstart = nullptr,
                std::vector<int> *h_ustart = nullptr, std::vector<int> *h_ulist = nullptr)
{
    const int n = c->n, W = c->nnbr;
    const std::vector<int> &nbr = c->mesh->h_nbr;
    std::vector<int> tstart{0}, ustart{0}, ulist;
    std::vector<unsigned short> lidx((size_t)W * n), lidx_node((size_t)32 * n, 0);
    std::vector<int> mark(n, -1), pos(n, 0);
    std::vector<int> cur;
    int tile = 0, node = 0;
    while (node < n) {
        cur.clear();
        const int first = node;
        while (node < n && node - first < tmax) {
            int fresh = 0;
            for (int j = 0; j < W; ++j) {
                const int q = nbr[(size_t)j * n + node];
                if (mark[q] != tile) {
                    mark[q] = tile;
                    cur.push_back(q);
                    ++fresh;
                }
            }
            if ((int)cur.size() > umax) {
                if (node == first)
                    return fail(SWE_HIP_EINVAL, "one stencil exceeds the LDS tile budget");
                for (int k = 0; k < fresh; ++k)
                    mark[cur.back()] = -1, cur.pop_back();
                break;
            }
            ++node;
        }
        std::sort(cur.begin(), cur.end());
        for (int k = 0; k < (int)cur.size(); ++k)
            pos[cur[k]] = k;
        for (int i = first; i < node; ++i) {
            for (int j = 0; j < W; ++j) {
                const unsigned short l = (unsigned short)pos[nbr[(size_t)j * n + i]];
                lidx[(size_t)j * n + i] = l;
                if (j < 32)
                    lidx_node[(size_t)32 * i + j] = l;
            }
            // spare slot 31: the node's own row in its tile (0xffff if its
            // stencil does not contain it)
            if (W < 32)
                lidx_node[(size_t)32 * i + 31] = mark[i] == tile ? (unsigned short)pos[i] : 0xffff;
        }
        ulist.insert(ulist.end(), cur.begin(), cur.end());
        tstart.push_back(node);
        ustart.push_back((int)ulist.size());
        ++tile;
    }
    // fixed-stride copy for the member-lane kernel (tiles of <= kTmemRows rows)
    std::vector<int> ulist_tile;
    if (umax <= kTmemRows) {
        ulist_tile.assign((size_t)tile * kTmemStride, 0);
        for (int k = 0; k < tile; ++k)
            std::copy(ulist.begin() + ustart[k], ulist.begin() + ustart[k + 1],
                      ulist_tile.begin() + (size_t)k * kTmemStride);
    }
    SweTiles &t = c->tiles;
    t.ntiles = tile;
    t.tmax = tmax;
    t.umax = umax;
    t.umax_used = 0;
    for (int k = 0; k < tile; ++k)
        t.umax_used = std::max(t.umax_used, ustart[k + 1] - ustart[k]);
    int rc = 0;
    if (!rc) rc = dev_alloc(&c->bytes, (void **)&t.d_tile_start, sizeof(int) * tstart.size());
    if (!rc) rc = dev_alloc(&c->bytes, (void **)&t.d_ustart, sizeof(int) * ustart.size());
    if (!rc) rc = dev_alloc(&c->bytes, (void **)&t.d_ulist, sizeof(int) * ulist.size());
    if (!rc) rc = dev_alloc(&c->bytes, (void **)&t.d_lidx, sizeof(unsigned short) * lidx.size());
    if (!rc) rc = dev_alloc(&c->bytes, (void **)&t.d_lidx_node, sizeof(unsigned short) * lidx_node.size());
    if (!rc && !ulist_tile.empty())
        rc = dev_alloc(&c->bytes, (void **)&t.d_ulist_tile, sizeof(int) * ulist_tile.size());
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpy(t.d_tile_start, tstart.data(), sizeof(int) * tstart.size(), hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess) e = hipMemcpy(t.d_ustart, ustart.data(), sizeof(int) * ustart.size(), hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess) e = hipMemcpy(t.d_ulist, ulist.data(), sizeof(int) * ulist.size(), hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess) e = hipMemcpy(t.d_lidx, lidx.data(), sizeof(unsigned short) * lidx.size(), hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess) e = hipMemcpy(t.d_lidx_node, lidx_node.data(), sizeof(unsigned short) * lidx_node.size(), hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess && !ulist_tile.empty())
        e = hipMemcpy(t.d_ulist_tile, ulist_tile.data(), sizeof(int) * ulist_tile.size(), hipMemcpyHostToDevice);
    if (!rc && e != hipSuccess)
        rc = fail(SWE_HIP_EDEVICE, std::string("tile upload: ") + hipGetErrorString(e));
    if (!rc && !ulist_tile.empty()) {
        rc = dev_alloc(&c->bytes, (void **)&t.d_wts_node, sizeof(double) * 128 * (size_t)n);
        if (!rc) rc = dev_alloc(&c->bytes, (void **)&t.d_geo_node, sizeof(double) * 32 * (size_t)n);
        if (!rc) {
            const long threads = 32L * n;
            k_pack_tile_nodes<<<(unsigned)((threads + 255) / 256), 256, 0, c->stream>>>(
                mesh_of(c), n, t.d_lidx_node, t.d_wts_node, t.d_geo_node);
            rc = check_launch("k_pack_tile_nodes");
        }
    }
    if (rc)
        t.ntiles = 0;
    if (!rc && h_tstart) {
        *h_tstart = tstart;
        *h_ustart = ustart;
        *h_ulist = ulist;
    }
    return rc;
}

// Resident kernel: for every tile, the tiles it must wait for before a
// stage -- owners of the rows it reads and tiles that read its rows.
int resident_neighbours(swe_hip_ctx *c, const std::vector<int> &tstart,
                        const std::vector<int> &ustart, const std::vector<int> &ulist)
{
    const int nt = (int)tstart.size() - 1;
    std::vector<int> owner(c->n);
    for (int t = 0; t < nt; ++t)
        for (int i = tstart[t]; i < tstart[t + 1]; ++i)
            owner[i] = t;
    std::vector<std::vector<int>> nb(nt);
    for (int t = 0; t < nt; ++t)
        for (int k = ustart[t]; k < ustart[t + 1]; ++k) {
            const int u = owner[ulist[k]];
            if (u != t) {
                nb[t].push_back(u);
                nb[u].push_back(t);
            }
        }
    std::vector<int> flat((size_t)nt * kResMaxNbrs, -1);
    for (int t = 0; t < nt; ++t) {
        std::sort(nb[t].begin(), nb[t].end());
        nb[t].erase(std::unique(nb[t].begin(), nb[t].end()), nb[t].end());
        if ((int)nb[t].size() > kResMaxNbrs)
            return fail(SWE_HIP_EINVAL, "the resident kernel's tiles exchange rows with more than "
                                        + std::to_string(kResMaxNbrs) + " other tiles");
        std::copy(nb[t].begin(), nb[t].end(), flat.begin() + (size_t)t * kResMaxNbrs);
    }
    int rc = dev_alloc(&c->bytes, (void **)&c->d_res_nbrs, sizeof(int) * flat.size());
    if (!rc && hipMemcpy(c->d_res_nbrs, flat.data(), sizeof(int) * flat.size(),
                         hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(SWE_HIP_EDEVICE, "resident kernel: neighbour upload");
    return rc;
}

}  // namespace

namespace {

// Milliseconds per fused RK step of one candidate configuration on the
// static data of `base`: one warm-up step (graph capture), then `steps`
// timed with HIP events on the candidate's stream.
int time_options(swe_hip_ctx *base, const swe_hip_options &opt, int steps, double *per_step)
{
    swe_hip_options o = opt;
    o.share = base;
    o.device = base->dev;
    swe_hip_ctx *c = nullptr;
    int rc = swe_hip_create_ex(nullptr, &o, &c);
    if (rc)
        return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.0f;
    rc = swe_hip_rk_steps(c, 1);
    hipError_t e = hipSuccess;
    if (!rc) e = hipEventCreate(&e0);
    if (!rc && e == hipSuccess) e = hipEventCreate(&e1);
    if (!rc && e == hipSuccess) e = hipEventRecord(e0, c->stream);
    if (!rc && e == hipSuccess) rc = swe_hip_rk_steps(c, steps);
    if (!rc && e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
    if (!rc && e == hipSuccess) rc = check_resident(c);
    if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    free_ctx(c);
    if (!rc && e != hipSuccess)
        rc = fail(SWE_HIP_EDEVICE, std::string("tune timing: ") + hipGetErrorString(e));
    if (!rc)
        *per_step = (double)ms / steps;
    return rc;
}

// Every admissible member block (powers of two 2..64 dividing members) on
// the member-block kernels, and with `with_tiled` also the LDS-tiled ones:
// member blocks 4, 8, 16 (one member per lane) and 32 (member-lane tiles,
// 31-point stencils only).  Tiles follow the node order of base's device
// copy.
int tune_candidates(swe_hip_ctx *base, int members, unsigned flags, int steps, bool with_tiled,
                    int *best_mb, int *best_tiled, double *best_ms)
{
    *best_mb = 0;
    *best_tiled = 0;
    double best = 1e300;
    for (int tiled = 0; tiled <= (with_tiled ? 1 : 0); ++tiled)
        for (int mb = 64; mb >= 2; mb >>= 1) {
            if (members % mb)
                continue;
            if (tiled && !(mb == 32 && members >= 32 && base->nnbr == 31))
                continue;
            double per = 0.0;
            const swe_hip_options o{members, flags, base->dev, mb, base, 0, 0, 0, tiled, 0, 0, 0};
            const int rc = time_options(base, o, steps, &per);
            if (rc)
                return rc;
            if (per < best) {
                best = per;
                *best_mb = mb;
                *best_tiled = tiled;
            }
        }
    if (best_ms)
        *best_ms = best;
    return *best_mb ? 0 : fail(SWE_HIP_EINVAL, "members has no admissible member block");
}

}  // namespace

extern "C" {

const char *swe_hip_strerror(int code)
{
    switch (code) {
    case SWE_HIP_OK: return "success";
    case SWE_HIP_EINVAL: return "invalid argument";
    case SWE_HIP_ENOMEM: return "out of memory";
    case SWE_HIP_EDEVICE: return "HIP device or runtime error";
    case SWE_HIP_ERANGE: return "stencil index outside the patch";
    default: return "unknown error";
    }
}

const char *swe_hip_last_error(void) { return g_last_error.c_str(); }

int swe_hip_create_ex(const swe_hip_static *s, const swe_hip_options *o, swe_hip_ctx **out)
{
    if (!out)
        return fail(SWE_HIP_EINVAL, "out is NULL");
    *out = nullptr;
    if (!o)
        return fail(SWE_HIP_EINVAL, "options are NULL");
    if (!s && !o->share)
        return fail(SWE_HIP_EINVAL, "static data is NULL");
    if (o->members < 1 || o->members > 4096)
        return fail(SWE_HIP_EINVAL, "members must be in 1..4096");
    if (o->kernel < SWE_HIP_KERNEL_AUTO || o->kernel > SWE_HIP_KERNEL_RESIDENT)
        return fail(SWE_HIP_EINVAL, "kernel must be one of SWE_HIP_KERNEL_*");
    if (o->tiled < 0 || o->tiled > 1)
        return fail(SWE_HIP_EINVAL, "tiled must be 0 or 1");
    if (o->prefetch != 0 && o->prefetch != 4 && o->prefetch != 8)
        return fail(SWE_HIP_EINVAL, "prefetch must be 0, 4 or 8");
    if (o->node_order < SWE_HIP_ORDER_HOST || o->node_order > SWE_HIP_ORDER_HILBERT)
        return fail(SWE_HIP_EINVAL, "node_order must be SWE_HIP_ORDER_HOST, _MORTON or _HILBERT");
    if (o->slot_block != 0 && o->slot_block != 16)
        return fail(SWE_HIP_EINVAL, "slot_block must be 0 or 16");
    if (o->flags & ~(SWE_HIP_EXACT | SWE_HIP_FP32))
        return fail(SWE_HIP_EINVAL, "unknown flag bits");
    const bool fp32 = (o->flags & SWE_HIP_FP32) != 0;
    if (fp32 && (o->members != 1 || o->tiled))
        return fail(SWE_HIP_EINVAL, "SWE_HIP_FP32 contexts hold one state (members == 1, not "
                                    "tiled)");
    if (o->share && o->share->fp32 != fp32)
        return fail(SWE_HIP_EINVAL, "share: the shared context has the other model precision");
    if (fp32 && o->kernel == SWE_HIP_KERNEL_QUADW)
        return fail(SWE_HIP_EINVAL, "SWE_HIP_FP32 kernels are AUTO, LANE, PAIR, QUAD, OCT and TILE");
    if (fp32 && (o->flags & SWE_HIP_EXACT) && o->kernel != SWE_HIP_KERNEL_AUTO
        && o->kernel != SWE_HIP_KERNEL_LANE && o->kernel != SWE_HIP_KERNEL_TILE
        && o->kernel != SWE_HIP_KERNEL_RESIDENT)
        return fail(SWE_HIP_EINVAL, "exact SWE_HIP_FP32 contexts sum each stencil in order: "
                                    "kernel AUTO, LANE, TILE or RESIDENT");
    if (!fp32 && (o->kernel == SWE_HIP_KERNEL_OCT || o->kernel == SWE_HIP_KERNEL_QUAD
                  || o->kernel == SWE_HIP_KERNEL_LANE))
        return fail(SWE_HIP_EINVAL, "FP64 single-state kernels are AUTO, PAIR, QUADW and TILE "
                                    "(QUAD, LANE and OCT are FP32 mappings)");
    if (o->kernel == SWE_HIP_KERNEL_RESIDENT && (o->members != 1 || o->tiled))
        return fail(SWE_HIP_EINVAL, "the resident kernel takes single-state contexts");
    if (o->members == 1 && o->tile_nodes != 0 && o->tile_nodes != 64 && o->tile_nodes != 128
        && o->tile_nodes != 256)
        return fail(SWE_HIP_EINVAL, "tile_nodes of the single-state tile kernel must be 0, 64, "
                                    "128 or 256");
    if (o->members > 1 && (o->tile_nodes < 0 || o->tile_nodes > kTmemTile))
        return fail(SWE_HIP_EINVAL, "tile_nodes of the member-lane tiled kernel must be 0..8");
    if (o->tile_rows != 0 && (o->tile_rows < 16 || o->tile_rows > kTmemRows))
        return fail(SWE_HIP_EINVAL, "tile_rows must be 0 or 16..70");
    if (!o->share) {
        const int rc = validate_static(s);
        if (rc)
            return rc;
    }
    const int mb = member_block(o->members, o->member_block);
    if (mb < 1)
        return fail(SWE_HIP_EINVAL, "member_block must be a power of two in 2..64 dividing "
                                    "members (members > 1 must be even)");
    if (o->tiled && (mb != 32 || o->members < 32))
        return fail(SWE_HIP_EINVAL, "the tiled ensemble kernel takes member block 32 (member "
                                    "lanes over LDS tiles)");
    int device = o->device;
    if (o->share)
        device = o->share->dev;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(SWE_HIP_EDEVICE, "no HIP device visible");
    if (device < 0 || device >= ndev)
        return fail(SWE_HIP_EINVAL, "device ordinal out of range");
    if (hipSetDevice(device) != hipSuccess)
        return fail(SWE_HIP_EDEVICE, "hipSetDevice failed");
    {
        // device address of the no-allocation entry point's scratch, once
        // per device; a failure is reported by that entry point only
        std::string why;
        (void)noalloc_setup(device, &why);
    }
    SweMesh *m = nullptr;
    if (o->share) {
        m = o->share->mesh;
        m->refs++;
    } else {
        // AUTO on a mesh the resident kernel will take: its tiles are
        // compact in Hilbert order (N = 40962, FP32: 3.7 against 4.6 us per
        // stage in the host's RCM order), and the order changes no result
        int order = o->node_order;
        int ncu = 0;
        if (o->kernel == SWE_HIP_KERNEL_AUTO && order == SWE_HIP_ORDER_HOST && o->members == 1
            && !o->tiled && s->Nnbr == 31 && s->compute_size >= kResAutoMinNodes
            && hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device)
                   == hipSuccess
            && s->compute_size <= 32 * kResMaxWaves * ncu)
            order = SWE_HIP_ORDER_HILBERT;
        int rc = build_mesh(s, device, order, o->slot_block, (o->flags & SWE_HIP_FP32) != 0, &m);
        if (rc)
            return rc;
    }
    swe_hip_ctx *c = new swe_hip_ctx();
    c->mesh = m;
    c->dev = device;
    c->n = m->n;
    c->ld = m->ld;
    c->nnbr = m->nnbr;
    c->gh0 = m->gh0;
    c->dt = m->dt;
    c->lay = m->lay;
    c->d_nbr = m->d_nbr;
    c->sblk = m->sblk;
    c->d_wts = m->d_wts;
    c->d_geo = m->d_geo;
    c->B = o->members;
    c->MB = mb;
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess || c->ncu <= 0)
        c->ncu = 256;
    // auto: the pair mapping (half the address-unit work per node of the
    // quad mapping) with one-wave workgroups is the fastest stage kernel
    // measured at N = 40962, 163842 and 655362; it needs a compile-time
    // stencil width.  Up to 256 nodes per compute unit the resident
    // whole-step kernel beats it at every size measured, its ~20 us launch
    // included (per stage over 20-step calls: N = 642 3.4 against 4.6 us,
    // 2562 3.6 against 4.6, 10242 3.7 against 5.1, 40962 4.8 against 10.4).
    if (o->kernel == SWE_HIP_KERNEL_AUTO) {
        c->kernel = m->nnbr == 31 ? SWE_HIP_KERNEL_PAIR : SWE_HIP_KERNEL_QUADW;
        if (!fp32 && o->members == 1 && !o->tiled && m->nnbr == 31 && m->n >= kResAutoMinNodes
            && m->n <= 32 * kResMaxWaves * c->ncu) {
            c->kernel = SWE_HIP_KERNEL_RESIDENT;
            c->res_auto = true;
        }
    }
    else if ((o->kernel == SWE_HIP_KERNEL_PAIR || o->kernel == SWE_HIP_KERNEL_TILE)
             && m->nnbr != 31)
        c->kernel = SWE_HIP_KERNEL_QUADW;   // both need a compile-time stencil width
    else
        c->kernel = o->kernel;
    c->fp32 = fp32;
    if (fp32) {
        // lanes per node of the FP32 kernels; auto in fast mode: two lanes
        // per node below 262144 nodes, where one lane per node leaves fewer
        // than 4 waves per SIMD (N = 40962: 5.7 against 10.8 us per launch),
        // one lane above (N = 655362: 108.5 against 112.8 us)
        const int fast_auto = c->n <= 262144 ? SWE_HIP_KERNEL_PAIR : SWE_HIP_KERNEL_LANE;
        int k = o->kernel == SWE_HIP_KERNEL_AUTO
                    ? ((o->flags & SWE_HIP_EXACT) ? SWE_HIP_KERNEL_LANE : fast_auto)
                    : o->kernel;
        // the resident kernel (both modes; its single stages are the lane
        // kernel's) over the same node range as in FP64
        if (o->kernel == SWE_HIP_KERNEL_AUTO && m->nnbr == 31 && m->n >= kResAutoMinNodes
            && m->n <= 32 * kResMaxWaves * c->ncu) {
            k = SWE_HIP_KERNEL_RESIDENT;
            c->res_auto = true;
        }
        c->kernel = k;
        c->split32 = k == SWE_HIP_KERNEL_PAIR ? 2 : k == SWE_HIP_KERNEL_QUAD ? 4
                   : k == SWE_HIP_KERNEL_OCT ? 8 : 1;
    }
    c->prefetch = o->prefetch;
    c->tile_nodes = o->tile_nodes;
    c->tile_rows = o->tile_rows;
    c->flags = o->flags;
    int rc = 0;
    auto check = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && !rc)
            rc = fail(SWE_HIP_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
    };
    check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    // F (the RHS array of the drop-in path) is allocated on first use: the
    // fused stages never store the RHS
    const size_t sbytes = (fp32 ? sizeof(float) : sizeof(double)) * (size_t)state_count(c);
    void **bufs[] = {(void **)&c->d_H, (void **)&c->d_Ka, (void **)&c->d_Kb, (void **)&c->d_D};
    void **bufs32[] = {(void **)&c->f_H, (void **)&c->f_Ka, (void **)&c->f_Kb, (void **)&c->f_D};
    for (int k = 0; k < 4; ++k)
        if (!rc) rc = dev_alloc(&c->bytes, fp32 ? bufs32[k] : bufs[k], sbytes);
    if (!rc && !fp32)
        rc = dev_alloc(&c->bytes, (void **)&c->d_stage, sizeof(double) * 4 * (size_t)c->n);
    if (!rc) rc = dev_alloc(&c->bytes, (void **)&c->d_active, sizeof(int));
    for (int k = 0; k < 4; ++k)
        if (!rc)
            check(hipMemsetAsync(*(fp32 ? bufs32[k] : bufs[k]), 0, sbytes, c->stream),
                  "hipMemsetAsync");
    if (!rc) {
        k_set_int<<<1, 64, 0, c->stream>>>(c->d_active, c->B);
        rc = check_launch("k_set_int");
    }
    if (!rc && c->B == 1 && c->kernel == SWE_HIP_KERNEL_TILE && (!fp32 || c->nnbr == 31)) {
        if (!c->tile_nodes)
            c->tile_nodes = kPtileNodes;
        rc = build_tiles(c, c->tile_nodes, 4 * c->tile_nodes);
    }
    if (!rc && c->kernel == SWE_HIP_KERNEL_RESIDENT) {
        // one tile of <= ceil(n / CUs) consecutive nodes per compute unit;
        // two lanes per node in FP64, one in FP32
        const int tmax = (c->n + c->ncu - 1) / c->ncu;
        const int waves = fp32 ? (tmax + 63) / 64 : (tmax + 31) / 32;
        if (c->nnbr != 31 || tmax > 32 * kResMaxWaves)
            rc = fail(SWE_HIP_EINVAL, "the resident kernel takes 31-point stencils and at most "
                                      + std::to_string(32 * kResMaxWaves)
                                      + " nodes per compute unit");
        std::vector<int> ts, us, ul;
        if (!rc) {
            c->res_block = 64 * waves;
            rc = build_tiles(c, tmax, kResMaxRows, &ts, &us, &ul);
        }
        if (!rc && c->tiles.ntiles > c->ncu)
            rc = fail(SWE_HIP_EINVAL, "the resident kernel's tiles exceed one per compute unit");
        if (!rc)
            rc = resident_neighbours(c, ts, us, ul);
        if (!rc) {
            c->res_lds = c->tiles.umax_used * (fp32 ? (int)(sizeof(float4) + sizeof(int))
                                                    : 2 * (int)(sizeof(double2) + sizeof(int)));
            rc = dev_alloc(&c->bytes, &c->d_res_sync,
                           (size_t)resident_sync_bytes(c->tiles.ntiles));
        }
        if (!rc)
            check(hipHostMalloc((void **)&c->h_res_abort, sizeof(unsigned)), "hipHostMalloc");
        // SWE_HIP_RESIDENT_SPIN_LIMIT: polls a wait may take (default 2^22,
        // seconds); a small value makes waits time out, which exercises the
        // abort path without any wait that does not end
        if (const char *v = std::getenv("SWE_HIP_RESIDENT_SPIN_LIMIT")) {
            const long lim = std::strtol(v, nullptr, 10);
            if (lim >= 1 && lim <= (1L << 30))
                c->res_spin = (unsigned)lim;
        }
        if (!rc) {
            *c->h_res_abort = 0;
            check(hipHostGetDevicePointer((void **)&c->d_res_abort, c->h_res_abort, 0),
                  "hipHostGetDevicePointer");
        }
        if (!rc) {
            std::string why;
            if (int e = resident_prepare(c, &why))
                rc = fail(e == SWE_HIP_EDEVICE ? SWE_HIP_EINVAL : e, why);
        }
        // AUTO: a mesh whose tiles do not fit the resident scheme (too many
        // rows or neighbouring tiles, e.g. long bands of a caller's node
        // order) or that cannot be co-resident runs the stage kernels
        if (rc == SWE_HIP_EINVAL && c->res_auto) {
            c->res_auto = false;
            c->tiles.ntiles = 0;
            if (fp32) {
                c->kernel = (o->flags & SWE_HIP_EXACT) || c->n > 262144 ? SWE_HIP_KERNEL_LANE
                                                                         : SWE_HIP_KERNEL_PAIR;
                c->split32 = c->kernel == SWE_HIP_KERNEL_PAIR ? 2 : 1;
            } else {
                c->kernel = SWE_HIP_KERNEL_PAIR;
            }
            rc = 0;
        }
    }
    if (!rc && o->tiled) {
        int lds_max = 0;
        {
            if (c->nnbr != 31)
                rc = fail(SWE_HIP_EINVAL, "the member-lane tiled kernel (member block 32) needs "
                                          "a 31-point stencil");
            else if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                           device) != hipSuccess ||
                     (size_t)lds_max < sizeof(double) * (size_t)kTmemLdsDoubles)
                rc = fail(SWE_HIP_EDEVICE, "the member-lane tiled kernel needs 160 KB of LDS "
                                           "per workgroup");
            else
                rc = build_tiles(c, c->tile_nodes ? c->tile_nodes : kTmemTile,
                                 c->tile_rows ? c->tile_rows : kTmemRows);
        }
    }
    if (!rc)
        check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (rc) {
        free_ctx(c);
        return rc;
    }
    *out = c;
    return 0;
}

int swe_hip_create(const swe_hip_static *s, int members, unsigned flags, int device,
                   swe_hip_ctx **out)
{
    swe_hip_options o{members, flags, device, 0, nullptr, 0, 0, 0, 0, 0, 0, 0};
    return swe_hip_create_ex(s, &o, out);
}

void swe_hip_destroy(swe_hip_ctx *ctx) { free_ctx(ctx); }

int swe_hip_create32(const swe_hip_static32 *s, int members, unsigned flags, int device,
                     swe_hip_ctx **out)
{
    if (!out)
        return fail(SWE_HIP_EINVAL, "out is NULL");
    *out = nullptr;
    if (!s)
        return fail(SWE_HIP_EINVAL, "static data is NULL");
    const long n = s->compute_size, nw = n * (long)s->padded_Nnbr;
    if (n <= 0 || s->padded_Nnbr <= 0 || !s->x || !s->y || !s->z || !s->f || !s->ghm || !s->p_u
        || !s->p_v || !s->p_w || !s->gradghm || !s->idx || !s->Dx || !s->Dy || !s->Dz || !s->L)
        return fail(SWE_HIP_EINVAL, "bad sizes or a NULL static array");
    // widen to the double descriptor (exact), repacked and rounded back to
    // float on upload by the FP32 context
    std::vector<double> buf((size_t)(5 * n + 12 * n + 4 * nw));
    double *p = buf.data();
    auto widen = [&](const float *a, long count) {
        double *dst = p;
        for (long k = 0; k < count; ++k)
            dst[k] = (double)a[k];
        p += count;
        return (const double *)dst;
    };
    swe_hip_static d{};
    d.compute_size = s->compute_size;
    d.patch_size = s->patch_size;
    d.compute_pid_s = s->compute_pid_s;
    d.Nnbr = s->Nnbr;
    d.padded_Nnbr = s->padded_Nnbr;
    d.cfdl = s->cfdl;
    d.sfdl = s->sfdl;
    d.simd = s->simd;
    d.gh0 = s->gh0;
    d.dt = s->dt;
    d.x = widen(s->x, n); d.y = widen(s->y, n); d.z = widen(s->z, n);
    d.f = widen(s->f, n); d.ghm = widen(s->ghm, n);
    d.p_u = widen(s->p_u, 3 * n); d.p_v = widen(s->p_v, 3 * n); d.p_w = widen(s->p_w, 3 * n);
    d.gradghm = widen(s->gradghm, 3 * n);
    d.idx = s->idx;
    d.Dx = widen(s->Dx, nw); d.Dy = widen(s->Dy, nw); d.Dz = widen(s->Dz, nw);
    d.L = widen(s->L, nw);
    swe_hip_options o{members, flags | SWE_HIP_FP32, device, 0, nullptr, 0, 0, 0, 0, 0, 0, 0};
    return swe_hip_create_ex(&d, &o, out);
}

int swe_hip_tune_member_block(swe_hip_ctx *base, int members, unsigned flags, int steps,
                              int *best_mb, double *best_ms)
{
    if (!base || members < 2 || steps < 1 || !best_mb)
        return fail(SWE_HIP_EINVAL, "tune_member_block: bad arguments");
    int tiled = 0;
    return tune_candidates(base, members, flags, steps, false, best_mb, &tiled, best_ms);
}

int swe_hip_tune_ensemble(swe_hip_ctx *base, int members, unsigned flags, int steps,
                          int *best_mb, int *best_tiled, double *best_ms)
{
    if (!base || members < 2 || steps < 1 || !best_mb || !best_tiled)
        return fail(SWE_HIP_EINVAL, "tune_ensemble: bad arguments");
    return tune_candidates(base, members, flags, steps, true, best_mb, best_tiled, best_ms);
}

int swe_hip_tune_ex(swe_hip_ctx *base, int members, unsigned flags, int steps,
                    swe_hip_options *best, double *best_ms)
{
    if (!base || members < 1 || steps < 1 || !best
        || ((flags & SWE_HIP_FP32) != 0) != base->fp32 || (base->fp32 && members != 1))
        return fail(SWE_HIP_EINVAL, "tune_ex: bad arguments (flags must carry base's "
                                    "precision; FP32 contexts hold one member)");
    std::vector<swe_hip_options> cand;
    auto add = [&](int mb, int kernel, int prefetch, int tiled, int tn, int tr) {
        cand.push_back(swe_hip_options{members, flags, base->dev, mb, base, kernel, prefetch, 0,
                                       tiled, 0, tn, tr});
    };
    if (base->fp32) {   // lanes per node (exact mode: one, in order) and LDS tiles
        add(0, SWE_HIP_KERNEL_LANE, 0, 0, 0, 0);
        if (!(flags & SWE_HIP_EXACT))
            for (int k : {SWE_HIP_KERNEL_PAIR, SWE_HIP_KERNEL_QUAD, SWE_HIP_KERNEL_OCT})
                add(0, k, 0, 0, 0, 0);
        if (base->nnbr == 31)
            for (int tn : {64, 128, 256})
                add(0, SWE_HIP_KERNEL_TILE, 0, 0, tn, 0);
        if (base->nnbr == 31 && base->n <= 32 * kResMaxWaves * base->ncu)
            add(0, SWE_HIP_KERNEL_RESIDENT, 0, 0, 0, 0);
    } else if (members == 1) {
        add(0, SWE_HIP_KERNEL_PAIR, 0, 0, 0, 0);
        if (base->nnbr == 31 && base->n <= 32 * kResMaxWaves * base->ncu)
            add(0, SWE_HIP_KERNEL_RESIDENT, 0, 0, 0, 0);
        if (base->nnbr == 31)
            for (int tn : {64, 128, 256})
                for (int pf : {0, 4, 8})
                    add(0, SWE_HIP_KERNEL_TILE, pf, 0, tn, 0);
    } else {
        for (int mb = 64; mb >= 2; mb >>= 1) {
            if (members % mb)
                continue;
            add(mb, 0, 0, 0, 0, 0);
            if (mb == 32 && base->nnbr == 31)   // member-lane tiles: nodes x rows per tile
                for (int tn : {8, 6})
                    for (int tr : {kTmemRows, 56})
                        add(mb, 0, 0, 1, tn, tr);
        }
    }
    double best_t = 1e300;
    int found = -1;
    for (size_t k = 0; k < cand.size(); ++k) {
        double per = 0.0;
        const int rc = time_options(base, cand[k], steps, &per);
        if (rc == SWE_HIP_EINVAL)
            continue;   // not possible on this mesh and node order (e.g. the
                        // resident kernel's tiles of a scattered order)
        if (rc)
            return rc;
        if (per < best_t) {
            best_t = per;
            found = (int)k;
        }
    }
    if (found < 0)
        return fail(SWE_HIP_EINVAL, "tune_ex: members has no admissible configuration");
    *best = cand[found];
    if (best_ms)
        *best_ms = best_t;
    return 0;
}

int swe_hip_info(const swe_hip_ctx *c, int *n_nodes, int *members, int *nnbr, unsigned *flags)
{
    if (!c)
        return fail(SWE_HIP_EINVAL, "ctx is NULL");
    if (n_nodes) *n_nodes = c->n;
    if (members) *members = c->B;
    if (nnbr) *nnbr = c->nnbr;
    if (flags) *flags = c->flags;
    return 0;
}

void *swe_hip_state_device_ptr(swe_hip_ctx *c)
{
    return c ? (c->fp32 ? (void *)c->f_H : (void *)c->d_H) : nullptr;
}
void *swe_hip_stream(swe_hip_ctx *c) { return c ? (void *)c->stream : nullptr; }
long long swe_hip_device_bytes(const swe_hip_ctx *c) { return c ? c->bytes : 0; }
long long swe_hip_static_bytes(const swe_hip_ctx *c) { return c ? c->mesh->bytes : 0; }

int swe_hip_query(swe_hip_ctx *c)
{
    if (!c)
        return fail(SWE_HIP_EINVAL, "ctx is NULL");
    hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) {
        if (int rc = check_resident(c))
            return rc;
        return 1;
    }
    if (e == hipErrorNotReady)
        return 0;
    return fail(SWE_HIP_EDEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int swe_hip_set_active(swe_hip_ctx *c, int active)
{
    if (!c || active < 0 || active > c->B)
        return fail(SWE_HIP_EINVAL, "set_active: bad ctx or count");
    HIPCHK(hipSetDevice(c->dev));
    k_set_int<<<1, 64, 0, c->stream>>>(c->d_active, active);
    return check_launch("k_set_int");
}

int swe_hip_move_member(swe_hip_ctx *c, int from, int to)
{
    if (!c || from < 0 || to < 0 || from >= c->B || to >= c->B)
        return fail(SWE_HIP_EINVAL, "move_member: bad ctx or slot");
    if (from == to)
        return 0;
    HIPCHK(hipSetDevice(c->dev));
    const long rows = 4L * c->n;
    k_move_member<<<elementwise_blocks(rows), kBlock, 0, c->stream>>>(c->d_H, from, to, c->B, rows);
    return check_launch("k_move_member");
}

static int to_device_member(swe_hip_ctx *c, int member, const double *Hhost, double *dst)
{
    const int n = c->n;
    c->host_tmp.resize(4 * (size_t)n);
    double *t = c->host_tmp.data();
    const std::vector<int> &perm = c->mesh->perm;
    for (int d = 0; d < n; ++d) {
        const int i = perm.empty() ? d : perm[d];
        for (int v = 0; v < 4; ++v)
            t[4 * (size_t)d + v] = Hhost[c->lay.svm(i, v)];
    }
    HIPCHK(hipSetDevice(c->dev));
    if (c->B == 1) {
        HIPCHK(hipMemcpyAsync(dst, t, sizeof(double) * 4 * n, hipMemcpyHostToDevice, c->stream));
    } else {
        HIPCHK(hipMemcpyAsync(c->d_stage, t, sizeof(double) * 4 * n, hipMemcpyHostToDevice,
                              c->stream));
        k_put_member<<<elementwise_blocks(4L * n), kBlock, 0, c->stream>>>(dst, c->d_stage, member,
                                                                         c->B, 4L * n);
        int rc = check_launch("k_put_member");
        if (rc)
            return rc;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

static int from_device_member(swe_hip_ctx *c, int member, const double *src, double *Hhost)
{
    const int n = c->n;
    c->host_tmp.resize(4 * (size_t)n);
    double *t = c->host_tmp.data();
    HIPCHK(hipSetDevice(c->dev));
    if (c->B == 1) {
        HIPCHK(hipMemcpyAsync(t, src, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, c->stream));
    } else {
        k_get_member<<<elementwise_blocks(4L * n), kBlock, 0, c->stream>>>(c->d_stage, src, member,
                                                                         c->B, 4L * n);
        int rc = check_launch("k_get_member");
        if (rc)
            return rc;
        HIPCHK(hipMemcpyAsync(t, c->d_stage, sizeof(double) * 4 * n, hipMemcpyDeviceToHost,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (int rc = check_resident(c))
        return rc;
    const std::vector<int> &perm = c->mesh->perm;
    for (int d = 0; d < n; ++d) {
        const int i = perm.empty() ? d : perm[d];
        for (int v = 0; v < 4; ++v)
            Hhost[c->lay.svm(i, v)] = t[4 * (size_t)d + v];
    }
    return 0;
}

// FP32 single state: host layout (SVM_LIN_ID) <-> device [node][4] floats.
static int to_device32(swe_hip_ctx *c, const float *Hhost, float *dst)
{
    const int n = c->n;
    std::vector<float> t(4 * (size_t)n);
    const std::vector<int> &perm = c->mesh->perm;
    for (int d = 0; d < n; ++d) {
        const int i = perm.empty() ? d : perm[d];
        for (int v = 0; v < 4; ++v)
            t[4 * (size_t)d + v] = Hhost[c->lay.svm(i, v)];
    }
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipMemcpyAsync(dst, t.data(), sizeof(float) * 4 * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

static int from_device32(swe_hip_ctx *c, const float *src, float *Hhost)
{
    const int n = c->n;
    std::vector<float> t(4 * (size_t)n);
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipMemcpyAsync(t.data(), src, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (int rc = check_resident(c))
        return rc;
    const std::vector<int> &perm = c->mesh->perm;
    for (int d = 0; d < n; ++d) {
        const int i = perm.empty() ? d : perm[d];
        for (int v = 0; v < 4; ++v)
            Hhost[c->lay.svm(i, v)] = t[4 * (size_t)d + v];
    }
    return 0;
}

int swe_hip_set_state(swe_hip_ctx *c, int member, const double *H)
{
    if (!c || !H || member < 0 || member >= (c ? c->B : 0))
        return fail(SWE_HIP_EINVAL, "set_state: bad ctx, pointer or member");
    if (c->fp32)
        return fail(SWE_HIP_EINVAL, "set_state: FP32 context, use swe_hip_set_state32");
    return to_device_member(c, member, H, c->d_H);
}

int swe_hip_get_state(swe_hip_ctx *c, int member, double *H)
{
    if (!c || !H || member < 0 || member >= (c ? c->B : 0))
        return fail(SWE_HIP_EINVAL, "get_state: bad ctx, pointer or member");
    if (c->fp32)
        return fail(SWE_HIP_EINVAL, "get_state: FP32 context, use swe_hip_get_state32");
    return from_device_member(c, member, c->d_H, H);
}

int swe_hip_set_state32(swe_hip_ctx *c, int member, const float *H)
{
    if (!c || !H || member != 0 || !c->fp32)
        return fail(SWE_HIP_EINVAL, "set_state32: bad ctx, pointer or member (FP32 contexts "
                                    "only, member 0)");
    return to_device32(c, H, c->f_H);
}

int swe_hip_get_state32(swe_hip_ctx *c, int member, float *H)
{
    if (!c || !H || member != 0 || !c->fp32)
        return fail(SWE_HIP_EINVAL, "get_state32: bad ctx, pointer or member (FP32 contexts "
                                    "only, member 0)");
    return from_device32(c, c->f_H, H);
}

// The RHS buffer F of the drop-in path, allocated (and zeroed) on its first
// use.
static int ensure_F(swe_hip_ctx *c)
{
    void **F = c->fp32 ? (void **)&c->f_F : (void **)&c->d_F;
    if (*F)
        return 0;
    HIPCHK(hipSetDevice(c->dev));
    const size_t sbytes = (c->fp32 ? sizeof(float) : sizeof(double)) * (size_t)state_count(c);
    int rc = dev_alloc(&c->bytes, F, sbytes);
    if (rc)
        return rc;
    HIPCHK(hipMemsetAsync(*F, 0, sbytes, c->stream));
    return 0;
}

static void launch_rhs32(const swe_hip_ctx *c, const float *src, float *F)
{
    DevStage32 st{};
    st.src = src;
    st.h = src;
    st.out = F;
    launch_fp32_stage(c, ST_RHS, st);
}

int swe_hip_eval_rhs32(swe_hip_ctx *c, int member, const float *K, float *F)
{
    if (!c || !K || !F || member != 0 || !c->fp32)
        return fail(SWE_HIP_EINVAL, "eval_rhs32: bad ctx, pointer or member (FP32 contexts "
                                    "only, member 0)");
    int rc = ensure_F(c);
    if (!rc)
        rc = to_device32(c, K, c->f_Kb);
    if (rc)
        return rc;
    launch_rhs32(c, c->f_Kb, c->f_F);
    rc = check_launch("eval_RHS kernel (FP32)");
    if (rc)
        return rc;
    return from_device32(c, c->f_F, F);
}

// The drop-in substep of an FP32 context: RK_substep with the reference's
// FP32 types (rk4_rbffd_swe.c:30-186; eval_K's coefficient dt/2.0 rounded
// to float, update_H's dt/6.0 kept in double).
static int rk_substep32(swe_hip_ctx *c, int s)
{
    const long cnt = state_count(c);
    const int eb = elementwise_blocks(cnt);
    const float dt = (float)c->dt;
    const bool ex = exact(c);
    float *H = c->f_H, *K = c->f_Ka, *F = c->f_F, *D = c->f_D;
    hipStream_t q = c->stream;
    int rc = timed(c, &c->acc.t_eval_rhs, "eval_RHS kernel (FP32)",
                   [&] { launch_rhs32(c, s == 0 ? H : K, F); });
    if (rc)
        return rc;
    if (s == 0) {
        rc = timed(c, nullptr, "copy kernel", [&] { k_copy32<<<eb, kBlock, 0, q>>>(D, F, cnt); });
    } else {
        const float coeff = s == 3 ? 1.0f : 2.0f;
        rc = timed(c, &c->acc.t_update_D, "update_D kernel", [&] {
            if (ex) k_axpy_inplace32<true><<<eb, kBlock, 0, q>>>(D, coeff, F, cnt);
            else k_axpy_inplace32<false><<<eb, kBlock, 0, q>>>(D, coeff, F, cnt);
        });
    }
    if (rc)
        return rc;
    if (s < 3) {
        const float cK = s == 2 ? dt : (float)(dt / 2.0);
        rc = timed(c, &c->acc.t_eval_K, "eval_K kernel", [&] {
            if (ex) k_axpy_to32<true><<<eb, kBlock, 0, q>>>(K, H, cK, F, cnt);
            else k_axpy_to32<false><<<eb, kBlock, 0, q>>>(K, H, cK, F, cnt);
        });
    } else {
        const double c6 = dt / 6.0;
        rc = timed(c, &c->acc.t_update_H, "update_H kernel", [&] {
            if (ex) k_update_h32<true><<<eb, kBlock, 0, q>>>(H, c6, D, cnt);
            else k_update_h32<false><<<eb, kBlock, 0, q>>>(H, c6, D, cnt);
        });
    }
    return rc;
}

int swe_hip_eval_rhs(swe_hip_ctx *c, int member, const double *K, double *F)
{
    if (!c || !K || !F || member < 0 || member >= (c ? c->B : 0))
        return fail(SWE_HIP_EINVAL, "eval_rhs: bad ctx, pointer or member");
    if (c->fp32)
        return fail(SWE_HIP_EINVAL, "eval_rhs: FP32 context, use swe_hip_eval_rhs32");
    int rc = ensure_F(c);
    if (rc)
        return rc;
    rc = to_device_member(c, member, K, c->d_Kb);
    if (rc)
        return rc;
    launch_rhs(c, c->d_Kb, c->d_F);
    rc = check_launch("eval_RHS kernel");
    if (rc)
        return rc;
    return from_device_member(c, member, c->d_F, F);
}

int swe_hip_rk_substep(swe_hip_ctx *c, int s)
{
    if (!c || s < 0 || s > 3)
        return fail(SWE_HIP_EINVAL, "rk_substep: bad ctx or substep_id");
    HIPCHK(hipSetDevice(c->dev));
    if (int rc0 = ensure_F(c))
        return rc0;
    if (c->fp32)
        return rk_substep32(c, s);
    const long cnt = state_count(c);
    const int eb = elementwise_blocks(cnt);
    const double dt = c->dt;
    const bool ex = exact(c);
    double *H = c->d_H, *K = c->d_Ka, *F = c->d_F, *D = c->d_D;
    hipStream_t q = c->stream;
    int rc = timed(c, &c->acc.t_eval_rhs, "eval_RHS kernel",
                   [&] { launch_rhs(c, s == 0 ? H : K, F); });
    if (rc)
        return rc;
    if (s == 0) {
        rc = timed(c, nullptr, "copy kernel", [&] { k_copy<<<eb, kBlock, 0, q>>>(D, F, cnt); });
    } else {
        const double coeff = s == 3 ? 1.0 : 2.0;
        rc = timed(c, &c->acc.t_update_D, "update_D kernel", [&] {
            if (ex) k_axpy_inplace<true><<<eb, kBlock, 0, q>>>(D, coeff, F, cnt);
            else k_axpy_inplace<false><<<eb, kBlock, 0, q>>>(D, coeff, F, cnt);
        });
    }
    if (rc)
        return rc;
    if (s < 3) {
        const double cK = s == 2 ? dt : dt / 2.0;
        rc = timed(c, &c->acc.t_eval_K, "eval_K kernel", [&] {
            if (ex) k_axpy_to<true><<<eb, kBlock, 0, q>>>(K, H, cK, F, cnt);
            else k_axpy_to<false><<<eb, kBlock, 0, q>>>(K, H, cK, F, cnt);
        });
    } else {
        const double c6 = dt / 6.0;
        rc = timed(c, &c->acc.t_update_H, "update_H kernel", [&] {
            if (ex) k_axpy_inplace<true><<<eb, kBlock, 0, q>>>(H, c6, D, cnt);
            else k_axpy_inplace<false><<<eb, kBlock, 0, q>>>(H, c6, D, cnt);
        });
    }
    return rc;
}

int swe_hip_rk_steps(swe_hip_ctx *c, int nsteps)
{
    if (!c || nsteps < 0)
        return fail(SWE_HIP_EINVAL, "rk_steps: bad ctx or nsteps");
    HIPCHK(hipSetDevice(c->dev));
    // a resident context chosen by AUTO replays the stage graphs for calls
    // too short to repay a launch (same arithmetic, bitwise)
    if (c->kernel == SWE_HIP_KERNEL_RESIDENT && (!c->res_auto || nsteps >= kResAutoMinSteps)) {
        std::string why;
        const long long before = c->res_stages;
        int rc = resident_steps(c, nsteps, &why);
        if (rc == SWE_HIP_EDEVICE && c->res_auto && c->res_stages == 0 && before == 0
            && !c->res_pending) {
            // AUTO, and the device refused the first cooperative launch
            // (e.g. compute units reserved by another process): the stage
            // kernels from here on
            c->res_auto = false;
            c->kernel = c->fp32 ? SWE_HIP_KERNEL_LANE : SWE_HIP_KERNEL_PAIR;
            c->split32 = 1;
            (void)hipGetLastError();
        } else if (rc) {
            return fail(rc, why);
        } else {
            return 0;
        }
    }
    // every graph is captured on the first call, so no later (timed) call
    // pays for a capture
    for (int g = 0; g < kGraphCount; ++g)
        if (!c->graphs[g])
            if (int rc = capture_steps(c, kGraphSteps[g], &c->graphs[g]))
                return rc;
    int left = nsteps;
    for (int g = 0; g < kGraphCount; ++g)
        for (; left >= kGraphSteps[g]; left -= kGraphSteps[g])
            HIPCHK(hipGraphLaunch(c->graphs[g], c->stream));
    return 0;
}

int swe_hip_set_timing(swe_hip_ctx *c, int on)
{
    if (!c)
        return fail(SWE_HIP_EINVAL, "ctx is NULL");
    c->timing = on != 0;
    return 0;
}

int swe_hip_get_timers(swe_hip_ctx *c, swe_hip_timers *t, int reset)
{
    if (!c || !t)
        return fail(SWE_HIP_EINVAL, "get_timers: NULL argument");
    *t = c->acc;
    if (reset)
        c->acc = swe_hip_timers{};
    return 0;
}

int swe_hip_synchronize(swe_hip_ctx *c)
{
    if (!c)
        return fail(SWE_HIP_EINVAL, "ctx is NULL");
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipStreamSynchronize(c->stream));
    return check_resident(c);
}

}  // extern "C"
