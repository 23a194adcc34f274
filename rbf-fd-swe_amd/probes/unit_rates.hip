// unit_rates: per-CU throughput of the memory paths the stage kernels choose
// between on gfx950 -- LDS b128 reads (distinct and broadcast addresses),
// ds_bpermute, L1-resident global loads of 16 and 8 bytes per lane, and the
// FP64 FMA and FP64 MFMA (16x16x4) rates.  Prints bytes per clock per CU and clocks per wave64
// instruction per CU (per SIMD for the FMA).
// All indices stay inside their buffers: LDS offsets are masked to 32 KB,
// global offsets to 16 KB per workgroup.
// Build: make -C rbf-fd-swe_amd probes; run: rbf-fd-swe_amd/probes/unit_rates
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)

constexpr int ITERS = 4096;

__global__ __launch_bounds__(256) void k_lds(double *out, int bcast)
{
    __shared__ double2 buf[2048];   // 32 KB
    for (int i = threadIdx.x; i < 2048; i += 256) buf[i] = make_double2(i, -i);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    double2 acc = make_double2(0, 0);
    int off = bcast ? 0 : lane;
#pragma unroll 8
    for (int it = 0; it < ITERS; ++it) {
        const double2 v = buf[(off + it * 64) & 2047];
        acc.x += v.x; acc.y += v.y;
    }
    out[blockIdx.x * 256 + threadIdx.x] = acc.x + acc.y;
}

__global__ __launch_bounds__(256) void k_bperm(double *out)
{
    const int lane = threadIdx.x & 63;
    int v = lane * 3 + blockIdx.x;
    int acc = 0;
#pragma unroll 8
    for (int it = 0; it < ITERS; ++it) {
        acc += __builtin_amdgcn_ds_bpermute(((lane + it) & 63) * 4, v + it);
    }
    out[blockIdx.x * 256 + threadIdx.x] = acc;
}

__global__ __launch_bounds__(256) void k_fma64(double *out, double a)
{
    double x[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = threadIdx.x + i;
    for (int it = 0; it < ITERS; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = __builtin_fma(x[i], a, 1.0e-9);
    }
    double acc = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) acc += x[i];
    out[blockIdx.x * 256 + threadIdx.x] = acc;
}

typedef double mfma_d4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_mfma64(double *out, double a)
{
    // 4 independent 16x16x4 FP64 accumulations per wave
    mfma_d4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
    const double x = a + threadIdx.x * 1e-9, y = a - threadIdx.x * 1e-9;
    for (int it = 0; it < ITERS; ++it) {
        c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(y, x, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(x, x, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(y, y, c3, 0, 0, 0);
    }
    const mfma_d4 s = c0 + c1 + c2 + c3;
    out[blockIdx.x * 256 + threadIdx.x] = s.x + s.y + s.z + s.w;
}

template <int W>
__global__ __launch_bounds__(256) void k_glob(const double *__restrict__ src, double *out)
{
    // src holds 16 KB per block (L1-resident); each lane reads W doubles
    const int lane = threadIdx.x & 63;
    const double *base = src + (blockIdx.x & 1023) * 2048;
    double acc = 0;
#pragma unroll 8
    for (int it = 0; it < ITERS; ++it) {
        const int o = ((lane + it * 64) * W) & 2047;
        if constexpr (W == 2) {
            const double2 v = *reinterpret_cast<const double2 *>(base + o);
            acc += v.x + v.y;
        } else {
            acc += base[o];
        }
    }
    out[blockIdx.x * 256 + threadIdx.x] = acc;
}

// FP64 FMA issue rate at a given occupancy: 16 independent accumulators per
// lane (the stencil sums of one slot), the weight either an ordinary VGPR or
// taken from lane k of the 16-lane row by 64-bit DPP row_newbcast.  Each
// wave records the shader-clock cycles of its loop in out[wave].
template <int K>
__device__ __forceinline__ void fmac_bc(double &acc, double w, double x)
{
    asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(w), "v"(x), "i"(K));
}
template <bool DPP>
__global__ __launch_bounds__(256) void k_fma64_acc16(double *out, double x0)
{
    double s[16];
    const double w = 1.0 + threadIdx.x * 1e-6;   // written long before any DPP read
    double x = x0;
#pragma unroll
    for (int i = 0; i < 16; ++i) s[i] = 0.0;
    const unsigned long long c0 = __builtin_readcyclecounter();
    for (int it = 0; it < ITERS; ++it) {
        if constexpr (DPP) {
            fmac_bc<0>(s[0], w, x); fmac_bc<1>(s[1], w, x); fmac_bc<2>(s[2], w, x); fmac_bc<3>(s[3], w, x);
            fmac_bc<4>(s[4], w, x); fmac_bc<5>(s[5], w, x); fmac_bc<6>(s[6], w, x); fmac_bc<7>(s[7], w, x);
            fmac_bc<8>(s[8], w, x); fmac_bc<9>(s[9], w, x); fmac_bc<10>(s[10], w, x); fmac_bc<11>(s[11], w, x);
            fmac_bc<12>(s[12], w, x); fmac_bc<13>(s[13], w, x); fmac_bc<14>(s[14], w, x); fmac_bc<15>(s[15], w, x);
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i) s[i] = __builtin_fma(w, x, s[i]);
        }
        asm volatile("" : "+v"(x));
    }
    const unsigned long long c1 = __builtin_readcyclecounter();
    double acc = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc += s[i];
    out[blockIdx.x * 256 + threadIdx.x] = acc;
    if ((threadIdx.x & 63) == 0)
        out[gridDim.x * 256 + blockIdx.x * 4 + (threadIdx.x >> 6)] = (double)(c1 - c0);
}

int main()
{
    int dev = 0; hipDeviceProp_t p; CHECK(hipGetDeviceProperties(&p, dev));
    const int cus = p.multiProcessorCount;
    const double ghz = p.clockRate / 1e6;
    printf("CUs %d clock %.3f GHz LDS/block %zu\n", cus, ghz, p.sharedMemPerBlock);
    const int blocks = cus * 8;
    double *out, *src;
    CHECK(hipMalloc(&out, sizeof(double) * blocks * 256));
    CHECK(hipMalloc(&src, sizeof(double) * 1024 * 2048));
    CHECK(hipMemset(src, 0, sizeof(double) * 1024 * 2048));
    hipEvent_t a, b; CHECK(hipEventCreate(&a)); CHECK(hipEventCreate(&b));
    auto run = [&](const char *name, double bytes_per_lane_iter, auto launch) {
        launch(); CHECK(hipDeviceSynchronize());
        CHECK(hipEventRecord(a)); for (int r = 0; r < 5; ++r) launch(); CHECK(hipEventRecord(b));
        CHECK(hipEventSynchronize(b)); float ms; CHECK(hipEventElapsedTime(&ms, a, b)); ms /= 5;
        const double bytes = (double)blocks * 256 * ITERS * bytes_per_lane_iter;
        const double per_cu_clk = bytes / (ms * 1e-3) / cus / (ghz * 1e9);
        const double instr = (double)blocks * 4 * ITERS;   // wave instructions
        printf("%-22s %8.3f ms  %8.1f TB/s  %6.1f B/clk/CU  %6.2f clk/wave-instr/CU\n", name, ms,
               bytes / (ms * 1e-3) / 1e12, per_cu_clk, (ms * 1e-3) * ghz * 1e9 * cus / instr);
    };
    run("lds_b128_distinct", 16, [&] { k_lds<<<blocks, 256>>>(out, 0); });
    run("lds_b128_broadcast", 16, [&] { k_lds<<<blocks, 256>>>(out, 1); });
    run("ds_bpermute_b32", 4, [&] { k_bperm<<<blocks, 256>>>(out); });
    run("global_dwordx4_L1", 16, [&] { k_glob<2><<<blocks, 256>>>(src, out); });
    run("global_dwordx2_L1", 8, [&] { k_glob<1><<<blocks, 256>>>(src, out); });
    // FP64 FMA: 8 independent chains per lane, ITERS x 8 FMAs = 16 flop each
    {
        k_fma64<<<blocks, 256>>>(out, 0.999999); CHECK(hipDeviceSynchronize());
        CHECK(hipEventRecord(a)); for (int r = 0; r < 5; ++r) k_fma64<<<blocks, 256>>>(out, 0.999999);
        CHECK(hipEventRecord(b)); CHECK(hipEventSynchronize(b));
        float ms; CHECK(hipEventElapsedTime(&ms, a, b)); ms /= 5;
        const double fmas = (double)blocks * 256 * ITERS * 8;
        const double instr = (double)blocks * 4 * ITERS * 8;
        printf("%-22s %8.3f ms  %8.1f TFLOP/s  %6.2f clk/wave-instr/SIMD\n", "v_fma_f64", ms,
               2 * fmas / (ms * 1e-3) / 1e12, (ms * 1e-3) * ghz * 1e9 * cus * 4 / instr);
    }
    // FP64 MFMA 16x16x4: 2*16*16*4 flop per wave instruction
    {
        k_mfma64<<<blocks, 256>>>(out, 0.5); CHECK(hipDeviceSynchronize());
        CHECK(hipEventRecord(a)); for (int r = 0; r < 5; ++r) k_mfma64<<<blocks, 256>>>(out, 0.5);
        CHECK(hipEventRecord(b)); CHECK(hipEventSynchronize(b));
        float ms; CHECK(hipEventElapsedTime(&ms, a, b)); ms /= 5;
        const double instr = (double)blocks * 4 * ITERS * 4;
        printf("%-22s %8.3f ms  %8.1f TFLOP/s  %6.2f clk/wave-instr/SIMD\n", "v_mfma_f64_16x16x4", ms,
               instr * 2 * 16 * 16 * 4 / (ms * 1e-3) / 1e12, (ms * 1e-3) * ghz * 1e9 * cus * 4 / instr);
    }
    // FP64 FMA issue rate with 16 accumulators, 1 and 2 waves per SIMD,
    // plain and DPP row-broadcast weight: shader-clock cycles per wave
    // instruction, from each wave's own cycle counter
    for (int wps = 1; wps <= 2; ++wps)
        for (int dpp = 0; dpp <= 1; ++dpp) {
            const int nb = cus * wps;   // 4 waves per block: wps waves per SIMD
            auto launch = [&] {
                if (dpp) k_fma64_acc16<true><<<nb, 256>>>(out, 0.5);
                else k_fma64_acc16<false><<<nb, 256>>>(out, 0.5);
            };
            launch(); CHECK(hipDeviceSynchronize());
            std::vector<double> h((size_t)nb * 4);
            CHECK(hipMemcpy(h.data(), out + (size_t)nb * 256, sizeof(double) * nb * 4,
                            hipMemcpyDeviceToHost));
            double cyc = 0;
            for (double v : h) cyc += v;
            cyc /= h.size();
            printf("v_fmac_f64%-11s %d wave(s)/SIMD  %6.2f clk/wave-instr (wave's own counter)\n",
                   dpp ? "_dpp bcast" : "", wps, cyc / ((double)ITERS * 16));
        }
    CHECK(hipFree(out)); CHECK(hipFree(src));
    return 0;
}
