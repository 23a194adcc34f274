// swe_hip_rk_steps_noalloc (include/swe_hip.h): fused RK4 steps of a
// caller-owned device state, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops).  The RK scratch of a
// step -- the two stage states K and the running sum D, 3 x 4 x n doubles --
// lives in the file-scope device array `scratch` below instead of the
// context's buffers, and the stages are launched directly (no graph, whose
// instantiation allocates).  Nothing here calls hipMalloc or hipFree.
//
// Every GPU the library runs on holds its own copy of `scratch` (64 MiB)
// until the process ends.  Host code never uses the array's name as an
// address: swe_hip_create_ex looks up the device address of the copy on
// each device once (noalloc_setup) and the entry point passes that address
// to the kernels.  All calls on one GPU share its copy, so calls that
// overlap on one GPU race; each call writes every scratch value before
// reading it (stage 0 writes K and D), so nothing needs clearing between
// calls.
#include "swe_ctx.h"

#include <mutex>
#include <string>

namespace swe {

// 3 x 4 x n doubles for n <= kNoallocMaxNodes (655362-node meshes fit).
constexpr long kNoallocDoubles = 64L * 1024 * 1024 / sizeof(double);
constexpr long kNoallocMaxNodes = kNoallocDoubles / 12;

}  // namespace swe

static __device__ __align__(256) double scratch[swe::kNoallocDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};
std::string g_setup_error[kMaxDevices];   // why a device has no scratch address

}  // namespace

namespace swe {

// Device address of this device's `scratch` (the current device must be
// `dev`).  Called at context set-up; a second call on a device returns at
// once.
int noalloc_setup(int dev, std::string *err)
{
    if (dev < 0 || dev >= kMaxDevices) {
        *err = "device ordinal beyond the no-allocation scratch table";
        return SWE_HIP_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[dev])
        return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        g_setup_error[dev] = *err;
        return SWE_HIP_EDEVICE;
    }
    g_scratch[dev] = static_cast<double *>(p);
    return 0;
}

}  // namespace swe

extern "C" int swe_hip_rk_steps_noalloc(swe_hip_ctx *c, double *H, int nsteps)
{
    using namespace swe;
    if (!c || !H || nsteps < 0)
        return set_error(SWE_HIP_EINVAL, "rk_steps_noalloc: bad ctx, state or nsteps");
    if (c->B != 1 || c->fp32)
        return set_error(SWE_HIP_EINVAL, "rk_steps_noalloc: single-state FP64 contexts only "
                                         "(members == 1)");
    if ((long)c->n > kNoallocMaxNodes)
        return set_error(
            SWE_HIP_EINVAL, "rk_steps_noalloc: " + std::to_string(c->n) + " nodes exceed the "
                                + std::to_string(kNoallocMaxNodes) + "-node scratch");
    double *base = nullptr;
    std::string why;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (c->dev >= 0 && c->dev < kMaxDevices) {
            base = g_scratch[c->dev];
            why = g_setup_error[c->dev];
        }
    }
    if (!base)
        return set_error(SWE_HIP_EDEVICE,
                         "rk_steps_noalloc: no scratch address on this device: " + why);
    if (hipSetDevice(c->dev) != hipSuccess)
        return set_error(SWE_HIP_EDEVICE, "rk_steps_noalloc: hipSetDevice failed");
    const long cnt = 4L * c->n;
    double *Ka = base, *Kb = base + cnt, *D = base + 2 * cnt;
    const double dt = c->dt;
    for (int i = 0; i < nsteps; ++i) {
        // the stages of swe_hip_rk_steps (gathers from H, Ka, Kb, Ka)
        DevStage st{};
        st.members = 1;
        st.active = c->d_active;
        st.h = H;
        st.d = D;
        st.src = H; st.out = Ka; st.c = dt / 2.0; st.dcoef = 0.0;
        launch_stage(c, ST_0, st);
        st.src = Ka; st.out = Kb; st.c = dt / 2.0; st.dcoef = 2.0;
        launch_stage(c, ST_1, st);
        st.src = Kb; st.out = Ka; st.c = dt; st.dcoef = 2.0;
        launch_stage(c, ST_2, st);
        st.src = Ka; st.out = H; st.c = dt / 6.0; st.dcoef = 1.0;
        launch_stage(c, ST_3, st);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(SWE_HIP_EDEVICE,
                         std::string("rk_steps_noalloc: ") + hipGetErrorString(e));
    return 0;
}
