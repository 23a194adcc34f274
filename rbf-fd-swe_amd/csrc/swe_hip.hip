// C-ABI implementation of libswe_hip.so (include/swe_hip.h): context
// set-up and device repacking of the static model data, the drop-in stage
// driver (RK_substep semantics, swe_code/main/rk4_rbffd_swe.c:30-186), the
// fused and graph-captured fast path, state upload/read-back in the
// reference's host layouts (swe_code/main/include/swe_config.h:21-54).
#include "swe_hip.h"
#include "swe_kernels.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace swe;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(SWE_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kBlock = 256;
constexpr int kGraphSteps = 8;   // RK steps per captured graph

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

}  // namespace

// Device copy of one patch's static data, shared by every context created
// with `share` (reference-counted; owned by the contexts).
struct SweMesh {
    int refs = 1;
    int dev = 0;
    int n = 0, ld = 0, nnbr = 0;
    double gh0 = 0.0, dt = 0.0;
    HostLayout lay{1, 0, 1};
    int *d_nbr = nullptr;
    double *d_wts = nullptr, *d_geo = nullptr;
    long long bytes = 0;
};

struct swe_hip_ctx {
    SweMesh *mesh = nullptr;
    int dev = 0;
    hipStream_t stream = nullptr;
    // copies of the mesh fields used on every launch
    int n = 0, ld = 0, nnbr = 0;
    double gh0 = 0.0, dt = 0.0;
    HostLayout lay{1, 0, 1};
    int *d_nbr = nullptr;
    double *d_wts = nullptr, *d_geo = nullptr;
    int B = 1, MB = 1;
    int kernel = SWE_HIP_KERNEL_QUAD;   // single-state mapping
    int prefetch = 0;                   // stencil slots per load batch (0, 4, 8)
    unsigned flags = 0;
    double *d_H = nullptr, *d_Ka = nullptr, *d_Kb = nullptr, *d_D = nullptr, *d_F = nullptr;
    double *d_stage = nullptr;   // [n*4] staging for member upload/read-back
    int *d_active = nullptr;     // member slots advanced by the ensemble kernel
    hipGraphExec_t graph1 = nullptr, graphN = nullptr;
    bool timing = false;
    swe_hip_timers acc{};
    long long bytes = 0;         // state buffers
    std::vector<double> host_tmp;
};

namespace {

bool exact(const swe_hip_ctx *c) { return (c->flags & SWE_HIP_EXACT) != 0; }

long state_count(const swe_hip_ctx *c) { return 4L * c->n * c->B; }

int blocks_x(const swe_hip_ctx *c, long threads)
{
    long nb = (threads + kBlock - 1) / kBlock;
    nb = (nb + 7) / 8 * 8;   // multiple of 8 for the XCD remap
    return (int)nb;
}

DevMesh mesh_of(const swe_hip_ctx *c)
{
    DevMesh m;
    m.nbr = c->d_nbr;
    m.wts = c->d_wts;
    m.geo = c->d_geo;
    m.n = c->n;
    m.ld = c->ld;
    m.nnbr = c->nnbr;
    m.gh0 = c->gh0;
    return m;
}

template <int STAGE, bool EX, int NN>
void launch_quad(const swe_hip_ctx *c, const DevMesh &m, DevStage st)
{
    st.nblk = blocks_x(c, 4L * c->n);
    const dim3 g(st.nblk), b(kBlock);
    if (NN > 0 && c->prefetch == 8)
        k_stage_quad<STAGE, EX, NN, 8><<<g, b, 0, c->stream>>>(m, st);
    else if (NN > 0 && c->prefetch == 4)
        k_stage_quad<STAGE, EX, NN, 4><<<g, b, 0, c->stream>>>(m, st);
    else
        k_stage_quad<STAGE, EX, NN, 0><<<g, b, 0, c->stream>>>(m, st);
}

template <int STAGE, bool EX, int NN>
void launch_lane(const swe_hip_ctx *c, const DevMesh &m, DevStage st)
{
    st.nblk = blocks_x(c, (long)c->n);
    const dim3 g(st.nblk), b(kBlock);
    if (NN > 0 && (EX || c->prefetch == 4))
        k_stage_lane<STAGE, EX, NN, 4><<<g, b, 0, c->stream>>>(m, st);
    else if (NN > 0 && c->prefetch == 8)
        k_stage_lane<STAGE, EX, NN, 8><<<g, b, 0, c->stream>>>(m, st);
    else
        k_stage_lane<STAGE, EX, NN, 0><<<g, b, 0, c->stream>>>(m, st);
}

template <int STAGE, bool EX, int MBv, int NN>
void launch_members(const swe_hip_ctx *c, const DevMesh &m, DevStage st)
{
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

template <int STAGE, bool EX, int NN>
void launch_stage_nn(const swe_hip_ctx *c, const DevStage &st)
{
    const DevMesh m = mesh_of(c);
    switch (c->MB) {
    case 1:
        if (c->kernel == SWE_HIP_KERNEL_LANE)
            launch_lane<STAGE, EX, NN>(c, m, st);
        else
            launch_quad<STAGE, EX, NN>(c, m, st);
        break;
    case 2: launch_members<STAGE, EX, 2, NN>(c, m, st); break;
    case 4: launch_members<STAGE, EX, 4, NN>(c, m, st); break;
    case 8: launch_members<STAGE, EX, 8, NN>(c, m, st); break;
    case 16: launch_members<STAGE, EX, 16, NN>(c, m, st); break;
    case 32: launch_members<STAGE, EX, 32, NN>(c, m, st); break;
    default: launch_members<STAGE, EX, 64, NN>(c, m, st); break;
    }
}

template <int STAGE>
void launch_stage(const swe_hip_ctx *c, const DevStage &st)
{
    const bool nn31 = c->nnbr == 31;
    if (exact(c)) {
        if (nn31) launch_stage_nn<STAGE, true, 31>(c, st);
        else launch_stage_nn<STAGE, true, 0>(c, st);
    } else {
        if (nn31) launch_stage_nn<STAGE, false, 31>(c, st);
        else launch_stage_nn<STAGE, false, 0>(c, st);
    }
}

// Fused stage s of one RK step: gathers from H (s=0), Ka (s=1), Kb (s=2),
// Ka (s=3); writes Ka, Kb, Ka, H.
void launch_fused(const swe_hip_ctx *c, int s)
{
    DevStage st{};
    st.members = c->B;
    st.active = c->d_active;
    st.h = c->d_H;
    st.d = c->d_D;
    const double dt = c->dt;
    switch (s) {
    case 0:
        st.src = c->d_H; st.out = c->d_Ka; st.c = dt / 2.0; st.dcoef = 0.0;
        launch_stage<ST_0>(c, st);
        break;
    case 1:
        st.src = c->d_Ka; st.out = c->d_Kb; st.c = dt / 2.0; st.dcoef = 2.0;
        launch_stage<ST_1>(c, st);
        break;
    case 2:
        st.src = c->d_Kb; st.out = c->d_Ka; st.c = dt; st.dcoef = 2.0;
        launch_stage<ST_2>(c, st);
        break;
    default:
        st.src = c->d_Ka; st.out = c->d_H; st.c = dt / 6.0; st.dcoef = 1.0;
        launch_stage<ST_3>(c, st);
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
    launch_stage<ST_RHS>(c, st);
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
    delete m;
}

void free_ctx(swe_hip_ctx *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->dev);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph1) (void)hipGraphExecDestroy(c->graph1);
    if (c->graphN) (void)hipGraphExecDestroy(c->graphN);
    void *ptrs[] = {c->d_H, c->d_Ka, c->d_Kb, c->d_D, c->d_F, c->d_stage, c->d_active};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
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
int build_mesh(const swe_hip_static *s, int device, SweMesh **out)
{
    *out = nullptr;
    const int n = s->compute_size;
    const HostLayout lay{s->cfdl ? 1 : 0, s->sfdl ? 1 : 0, s->simd};
    const int W = s->Nnbr, pW = s->padded_Nnbr;
    // Repack into the device layouts, checking every stencil id.
    std::vector<int> nbr((size_t)W * n);
    std::vector<double> wts((size_t)W * n * 4);
    std::vector<double> geo((size_t)GEO_N * n);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < W; ++j) {
            const long d = lay.dm(i, j, pW);
            const int q = s->idx[d];
            if (q < 0 || q >= n)
                return fail(SWE_HIP_ERANGE, "idx[" + std::to_string(d) + "] = " + std::to_string(q)
                                                + " outside [0, " + std::to_string(n) + ")");
            nbr[(size_t)j * n + i] = q;
            double *w = &wts[((size_t)j * n + i) * 4];
            w[0] = s->Dx[d];
            w[1] = s->Dy[d];
            w[2] = s->Dz[d];
            w[3] = s->L[d];
        }
        geo[(size_t)G_X * n + i] = s->x[i];
        geo[(size_t)G_Y * n + i] = s->y[i];
        geo[(size_t)G_Z * n + i] = s->z[i];
        geo[(size_t)G_F * n + i] = s->f[i];
        geo[(size_t)G_GHM * n + i] = s->ghm[i];
        for (int c3 = 0; c3 < 3; ++c3) {
            const long v = lay.vect(i, c3);
            geo[(size_t)(G_PU + c3) * n + i] = s->p_u[v];
            geo[(size_t)(G_PV + c3) * n + i] = s->p_v[v];
            geo[(size_t)(G_PW + c3) * n + i] = s->p_w[v];
            geo[(size_t)(G_GG + c3) * n + i] = s->gradghm[v];
        }
    }
    SweMesh *m = new SweMesh();
    m->dev = device;
    m->n = n;
    m->ld = n;
    m->nnbr = W;
    m->gh0 = s->gh0;
    m->dt = s->dt;
    m->lay = lay;
    int rc = 0;
    if (!rc) rc = dev_alloc(&m->bytes, (void **)&m->d_nbr, sizeof(int) * nbr.size());
    if (!rc) rc = dev_alloc(&m->bytes, (void **)&m->d_wts, sizeof(double) * wts.size());
    if (!rc) rc = dev_alloc(&m->bytes, (void **)&m->d_geo, sizeof(double) * geo.size());
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpy(m->d_nbr, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess)
        e = hipMemcpy(m->d_wts, wts.data(), sizeof(double) * wts.size(), hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess)
        e = hipMemcpy(m->d_geo, geo.data(), sizeof(double) * geo.size(), hipMemcpyHostToDevice);
    if (!rc && e != hipSuccess)
        rc = fail(SWE_HIP_EDEVICE, std::string("static data upload: ") + hipGetErrorString(e));
    if (rc) {
        release_mesh(m);
        return rc;
    }
    *out = m;
    return 0;
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
    if (o->prefetch != 0 && o->prefetch != 4 && o->prefetch != 8)
        return fail(SWE_HIP_EINVAL, "prefetch must be 0, 4 or 8");
    if (!o->share) {
        const int rc = validate_static(s);
        if (rc)
            return rc;
    }
    const int mb = member_block(o->members, o->member_block);
    if (mb < 1)
        return fail(SWE_HIP_EINVAL, "member_block must be a power of two in 2..64 dividing "
                                    "members (members > 1 must be even)");
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
    SweMesh *m = nullptr;
    if (o->share) {
        m = o->share->mesh;
        m->refs++;
    } else {
        int rc = build_mesh(s, device, &m);
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
    c->d_wts = m->d_wts;
    c->d_geo = m->d_geo;
    c->B = o->members;
    c->MB = mb;
    c->kernel = o->kernel == SWE_HIP_KERNEL_LANE ? SWE_HIP_KERNEL_LANE : SWE_HIP_KERNEL_QUAD;
    c->prefetch = o->prefetch;
    c->flags = o->flags;
    int rc = 0;
    auto check = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && !rc)
            rc = fail(SWE_HIP_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
    };
    check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    const size_t sbytes = sizeof(double) * (size_t)state_count(c);
    double **bufs[] = {&c->d_H, &c->d_Ka, &c->d_Kb, &c->d_D, &c->d_F};
    for (double **b : bufs)
        if (!rc) rc = dev_alloc(&c->bytes, (void **)b, sbytes);
    if (!rc) rc = dev_alloc(&c->bytes, (void **)&c->d_stage, sizeof(double) * 4 * (size_t)c->n);
    if (!rc) rc = dev_alloc(&c->bytes, (void **)&c->d_active, sizeof(int));
    for (double **b : bufs)
        if (!rc)
            check(hipMemsetAsync(*b, 0, sbytes, c->stream), "hipMemsetAsync");
    if (!rc) {
        k_set_int<<<1, 64, 0, c->stream>>>(c->d_active, c->B);
        rc = check_launch("k_set_int");
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
    swe_hip_options o{members, flags, device, 0, nullptr, 0, 0};
    return swe_hip_create_ex(s, &o, out);
}

void swe_hip_destroy(swe_hip_ctx *ctx) { free_ctx(ctx); }

int swe_hip_tune_member_block(swe_hip_ctx *base, int members, unsigned flags, int steps,
                              int *best_mb, double *best_ms)
{
    if (!base || members < 2 || steps < 1 || !best_mb)
        return fail(SWE_HIP_EINVAL, "tune_member_block: bad arguments");
    *best_mb = 0;
    double best = 1e300;
    for (int mb = 64; mb >= 2; mb >>= 1) {
        if (members % mb)
            continue;
        swe_hip_options o{members, flags, base->dev, mb, base, 0, 0};
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
        if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        free_ctx(c);
        if (!rc && e != hipSuccess)
            rc = fail(SWE_HIP_EDEVICE, std::string("tune timing: ") + hipGetErrorString(e));
        if (rc)
            return rc;
        const double per = (double)ms / steps;
        if (per < best) {
            best = per;
            *best_mb = mb;
        }
    }
    if (best_ms)
        *best_ms = best;
    return *best_mb ? 0 : fail(SWE_HIP_EINVAL, "members has no admissible member block");
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

void *swe_hip_state_device_ptr(swe_hip_ctx *c) { return c ? (void *)c->d_H : nullptr; }
void *swe_hip_stream(swe_hip_ctx *c) { return c ? (void *)c->stream : nullptr; }
long long swe_hip_device_bytes(const swe_hip_ctx *c) { return c ? c->bytes : 0; }
long long swe_hip_static_bytes(const swe_hip_ctx *c) { return c ? c->mesh->bytes : 0; }

int swe_hip_query(swe_hip_ctx *c)
{
    if (!c)
        return fail(SWE_HIP_EINVAL, "ctx is NULL");
    hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess)
        return 1;
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
    for (int i = 0; i < n; ++i)
        for (int v = 0; v < 4; ++v)
            t[4 * (size_t)i + v] = Hhost[c->lay.svm(i, v)];
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
    for (int i = 0; i < n; ++i)
        for (int v = 0; v < 4; ++v)
            Hhost[c->lay.svm(i, v)] = t[4 * (size_t)i + v];
    return 0;
}

int swe_hip_set_state(swe_hip_ctx *c, int member, const double *H)
{
    if (!c || !H || member < 0 || member >= (c ? c->B : 0))
        return fail(SWE_HIP_EINVAL, "set_state: bad ctx, pointer or member");
    return to_device_member(c, member, H, c->d_H);
}

int swe_hip_get_state(swe_hip_ctx *c, int member, double *H)
{
    if (!c || !H || member < 0 || member >= (c ? c->B : 0))
        return fail(SWE_HIP_EINVAL, "get_state: bad ctx, pointer or member");
    return from_device_member(c, member, c->d_H, H);
}

int swe_hip_eval_rhs(swe_hip_ctx *c, int member, const double *K, double *F)
{
    if (!c || !K || !F || member < 0 || member >= (c ? c->B : 0))
        return fail(SWE_HIP_EINVAL, "eval_rhs: bad ctx, pointer or member");
    int rc = to_device_member(c, member, K, c->d_Kb);
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
    int rc = 0;
    if (!c->graph1 && (rc = capture_steps(c, 1, &c->graph1)))
        return rc;
    if (!c->graphN && (rc = capture_steps(c, kGraphSteps, &c->graphN)))
        return rc;
    int left = nsteps;
    for (; left >= kGraphSteps; left -= kGraphSteps)
        HIPCHK(hipGraphLaunch(c->graphN, c->stream));
    for (; left > 0; --left)
        HIPCHK(hipGraphLaunch(c->graph1, c->stream));
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
    return 0;
}

}  // extern "C"
