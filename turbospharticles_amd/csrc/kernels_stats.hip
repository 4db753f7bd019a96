// kernels_stats.hip -- per-scene statistics on the device (SimBackend_GPU::FillStatisticsRecord,
// GPU/backend/gpu_backend.cpp:1067-1208, statistics_kernels.cu): neighbor-count minimum, maximum
// and 99th percentile from an exact histogram of the counts, the averages of the relative density
// error rho/rho0 - 1 (raw, positive part, absolute), and the velocity magnitude's sum, minimum and
// maximum. A served scene is monitored without copying its particles to the host.
#include <mutex>

#include "tsp_device.h"

namespace tsp
{
    constexpr uint32_t kHistBins = 65536;   // neighbor counts are stored in 16 bits

    // counts of one relation of one scene: u16 (FF, FB, FR) or u32 (RF, RB, RR)
    __global__ void __launch_bounds__(kBlock) kStatsHist(const uint16_t* __restrict__ c16, const uint32_t* __restrict__ c32,
        uint32_t mask, uint32_t n, uint32_t* __restrict__ hist)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) { return; }
        const uint32_t v = c16 ? (static_cast<uint32_t>(c16[i]) & mask) : c32[i];
        atomicAdd(&hist[v < kHistBins ? v : kHistBins - 1u], 1u);
    }

    // minimum, maximum and the value at sorted position (99 (n - 1)) / 100 of a histogram of n
    // values (ComputeStatsFromHistogram, statistics_kernels.cu:82-127), and their sum; one block
    __global__ void __launch_bounds__(kBlock) kStatsHistScan(const uint32_t* __restrict__ hist, uint32_t n,
        uint64_t* __restrict__ out)
    {
        constexpr uint32_t kPer = kHistBins / kBlock;
        __shared__ uint32_t chunk[kBlock];
        __shared__ uint64_t wsum[kBlock];
        const uint32_t t = threadIdx.x;
        uint32_t s = 0;
        uint64_t ws = 0;
        for (uint32_t b = 0; b < kPer; b++)
        {
            const uint32_t h = hist[t * kPer + b];
            s += h;
            ws += static_cast<uint64_t>(h) * (t * kPer + b);
        }
        chunk[t] = s;
        wsum[t] = ws;
        __syncthreads();
        if (t != 0) { return; }
        uint64_t total = 0;
        for (uint32_t c = 0; c < kBlock; c++) { total += wsum[c]; }
        uint32_t mn = 0, mx = 0, p99 = 0;
        if (n > 0)
        {
            const uint64_t i99 = (99ull * static_cast<uint64_t>(n - 1u)) / 100ull;
            uint64_t before = 0;
            bool haveMin = false, haveP99 = false;
            uint32_t lastChunk = 0;
            for (uint32_t c = 0; c < kBlock; c++)
            {
                if (chunk[c] == 0u) { continue; }
                const uint32_t lo = c * kPer;
                if (!haveMin)
                {
                    for (uint32_t v = lo; v < lo + kPer; v++) { if (hist[v]) { mn = v; break; } }
                    haveMin = true;
                }
                lastChunk = c;
                const uint64_t after = before + chunk[c];
                if (!haveP99 && i99 < after)
                {
                    uint64_t b = before;
                    for (uint32_t v = lo; v < lo + kPer; v++)
                    {
                        b += hist[v];
                        if (i99 < b) { p99 = v; break; }
                    }
                    haveP99 = true;
                }
                before = after;
            }
            for (uint32_t v = lastChunk * kPer + kPer; v > lastChunk * kPer; v--) { if (hist[v - 1]) { mx = v - 1; break; } }
        }
        out[0] = mn; out[1] = mx; out[2] = p99; out[3] = total;
    }

    // per block of the scene's fluid: FP64 sums of the density error terms and of |v|, min and max |v|
    __global__ void __launch_bounds__(kBlock) kStatsFluid(const float4* __restrict__ vel, uint32_t off, uint32_t n,
        float rho0, double* __restrict__ partial)
    {
        __shared__ double red[4][kBlock / 64];
        __shared__ float vmin[kBlock / 64], vmax[kBlock / 64];
        const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
        double raw = 0.0, pos = 0.0, ab = 0.0, vs = 0.0;
        float mn = 3.4e38f, mx = 0.0f;
        if (q < n)
        {
            const float4 v = vel[off + q];
            const float e = v.w / rho0 - 1.0f;
            raw = e; pos = e > 0.0f ? e : 0.0f; ab = __builtin_fabsf(e);
            const float m = __builtin_sqrtf(__builtin_fmaf(v.x, v.x, __builtin_fmaf(v.y, v.y, v.z * v.z)));
            vs = m; mn = m; mx = m;
        }
        for (int o = 32; o > 0; o >>= 1)
        {
            raw += __shfl_xor(raw, o, 64); pos += __shfl_xor(pos, o, 64);
            ab += __shfl_xor(ab, o, 64); vs += __shfl_xor(vs, o, 64);
            mn = fminf(mn, __shfl_xor(mn, o, 64)); mx = fmaxf(mx, __shfl_xor(mx, o, 64));
        }
        const uint32_t w = threadIdx.x >> 6;
        if ((threadIdx.x & 63u) == 0) { red[0][w] = raw; red[1][w] = pos; red[2][w] = ab; red[3][w] = vs; vmin[w] = mn; vmax[w] = mx; }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            double a[4] = { 0, 0, 0, 0 };
            float lo = 3.4e38f, hi = 0.0f;
            for (uint32_t k = 0; k < kBlock / 64; k++)
            {
                for (int r = 0; r < 4; r++) { a[r] += red[r][k]; }
                lo = fminf(lo, vmin[k]); hi = fmaxf(hi, vmax[k]);
            }
            double* p = partial + 6ull * blockIdx.x;
            p[0] = a[0]; p[1] = a[1]; p[2] = a[2]; p[3] = a[3]; p[4] = lo; p[5] = hi;
        }
    }

    __global__ void __launch_bounds__(kBlock) kStatsFluidSum(const double* __restrict__ partial, uint32_t nBlocks,
        double* __restrict__ out)
    {
        if (threadIdx.x != 0) { return; }
        double a[4] = { 0, 0, 0, 0 };
        double lo = nBlocks ? 3.4e38 : 0.0, hi = 0.0;
        for (uint32_t b = 0; b < nBlocks; b++)
        {
            for (int r = 0; r < 4; r++) { a[r] += partial[6ull * b + r]; }
            lo = partial[6ull * b + 4] < lo ? partial[6ull * b + 4] : lo;
            hi = partial[6ull * b + 5] > hi ? partial[6ull * b + 5] : hi;
        }
        for (int r = 0; r < 4; r++) { out[r] = a[r]; }
        out[4] = lo; out[5] = hi;
    }

    __global__ void __launch_bounds__(kBlock) kStatsClear(uint32_t* __restrict__ hist)
    {
        hist[blockIdx.x * kBlock + threadIdx.x] = 0u;
    }

    static inline uint32_t blocksN(uint64_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

    // hist: kHistBins words (cleared by a kernel, not a memset node, so the sequence can be
    // captured); out4: min, max, p99, sum
    void launchStatsCounts(hipStream_t s, const uint16_t* c16, const uint32_t* c32, uint32_t mask, uint32_t n,
                           uint32_t* hist, uint64_t* out4)
    {
        kStatsClear<<<kHistBins / kBlock, kBlock, 0, s>>>(hist);
        if (n) { kStatsHist<<<blocksN(n), kBlock, 0, s>>>(c16, c32, mask, n, hist); }
        kStatsHistScan<<<1, kBlock, 0, s>>>(hist, n, out4);
    }

    // partial: 6 doubles per block of n; out6: sums of raw, positive, absolute error and |v|, min and max |v|
    void launchStatsFluid(hipStream_t s, const float4* vel, uint32_t off, uint32_t n, float rho0, double* partial, double* out6)
    {
        if (n) { kStatsFluid<<<blocksN(n), kBlock, 0, s>>>(vel, off, n, rho0, partial); }
        kStatsFluidSum<<<1, kBlock, 0, s>>>(partial, blocksN(n), out6);
    }

    // --- the statistics scratch of the allocation-free entry point (tsp_get_statistics_noalloc):
    //     one static array per GPU, for callers that cannot allocate device memory. Layout: the
    //     count histogram, the relations' (min, max, p99, sum), the fluid sums, then 6 FP64 block
    //     partials per 256 fluid particles of a scene of at most kStatsMaxFluid particles.
    //     Every call on one GPU shares that GPU's array: calls that overlap on one GPU (contexts
    //     on different streams) race, and each call leaves its values in it. ---
    constexpr uint32_t kStatsMaxFluid = 1u << 24;
    constexpr size_t kStatsHistOff = 0;
    constexpr size_t kStatsCountsOff = kStatsHistOff + sizeof(uint32_t) * kHistBins;
    constexpr size_t kStatsFluidOff = kStatsCountsOff + 256;
    constexpr size_t kStatsPartialOff = kStatsFluidOff + 256;
    constexpr size_t kStatsScratchBytes = kStatsPartialOff + sizeof(double) * 6 * (kStatsMaxFluid / kBlock);
    static_assert(sizeof(uint64_t) * 24 <= 256 && sizeof(double) * 6 <= 256, "scratch layout");

    static __device__ __align__(256) unsigned char scratch[kStatsScratchBytes];

    constexpr int kStatsMaxDevices = 64;
    static void* scratchPtr[kStatsMaxDevices];
    static bool scratchReady[kStatsMaxDevices];
    static std::mutex scratchMutex;

    // once per device, with that device current: the device address of its copy of the array
    // (the array's name in host code is a host-side stand-in, not a device address)
    hipError_t statsScratchInit(int device)
    {
        if (device < 0 || device >= kStatsMaxDevices) { return hipErrorInvalidDevice; }
        std::lock_guard<std::mutex> lock(scratchMutex);
        if (scratchReady[device]) { return hipSuccess; }
        void* ptr = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (rc != hipSuccess) { return rc; }
        scratchPtr[device] = ptr;
        scratchReady[device] = true;
        return hipSuccess;
    }

    // the scratch buffers of one device (statsScratchInit ran for it), or false
    bool statsScratch(int device, uint32_t** hist, uint64_t** counts, double** fluid, double** partial,
                      uint32_t* maxFluid)
    {
        if (device < 0 || device >= kStatsMaxDevices) { return false; }
        std::lock_guard<std::mutex> lock(scratchMutex);
        if (!scratchReady[device]) { return false; }
        unsigned char* base = static_cast<unsigned char*>(scratchPtr[device]);
        *hist = reinterpret_cast<uint32_t*>(base + kStatsHistOff);
        *counts = reinterpret_cast<uint64_t*>(base + kStatsCountsOff);
        *fluid = reinterpret_cast<double*>(base + kStatsFluidOff);
        *partial = reinterpret_cast<double*>(base + kStatsPartialOff);
        *maxFluid = kStatsMaxFluid;
        return true;
    }
}
