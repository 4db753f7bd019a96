// math_selftest.hip -- checks on the GPU that the short correctly rounded square root and quotient
// of the pair kernels (sqrtCR, divCR in tsp_device.h) give the same bits as the compiler's general
// correctly rounded lowering (built with -fhip-fp32-correctly-rounded-divide-sqrt):
//   1. sqrtCR over 0 and every float from 2^-96 to +inf (its domain, 1.88e9 inputs);
//   2. divCR(koutScale(h), dist) over every float dist in [h, 2h] for particle sizes h from
//      0.0008 to 1 (the operands of the gradient factor's outer branch);
//   3. divCR over 2^32 pseudo-random pairs with |a| in [2^-60, 2^61) and |b| in [2^-30, 2^32).
// Prints the counts and exits 0 when every case matches, 1 on a mismatch, 3 without a device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "tsp_device.h"

using namespace tsp;

namespace
{
    constexpr int kThreads = 256;
    constexpr int kBlocks = 2048;

    struct Result
    {
        unsigned long long tested;
        unsigned long long bad;
        uint32_t firstBad;
        uint32_t pad;
    };

    __device__ void record(Result* r, unsigned long long tested, unsigned long long bad, uint32_t first)
    {
        atomicAdd(&r->tested, tested);
        if (bad)
        {
            atomicAdd(&r->bad, bad);
            atomicMin(&r->firstBad, first);
        }
    }

    __global__ void __launch_bounds__(kThreads) kSqrt(uint64_t n, Result* r)
    {
        const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
        unsigned long long tested = 0, bad = 0;
        uint32_t first = 0xffffffffu;
        for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            // 0, then every float from 2^-96 (the domain of sqrtCR) to +inf
            const uint32_t bits = i == 0 ? 0u : 0x0f800000u + static_cast<uint32_t>(i - 1);
            const float x = __uint_as_float(bits);
            const float a = sqrtCR(x);
            const float b = __builtin_sqrtf(x);
            tested++;
            if (__float_as_uint(a) != __float_as_uint(b)) { bad++; first = min(first, bits); }
        }
        record(r, tested, bad, first);
    }

    __global__ void __launch_bounds__(kThreads) kDivRange(float num, uint32_t lo, uint32_t hi, Result* r)
    {
        const uint32_t stride = gridDim.x * blockDim.x;
        unsigned long long tested = 0, bad = 0;
        uint32_t first = 0xffffffffu;
        for (uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x; i <= hi && i >= lo; i += stride)
        {
            const float den = __uint_as_float(i);
            const float a = divCR(num, den);
            const float b = num / den;
            tested++;
            if (__float_as_uint(a) != __float_as_uint(b)) { bad++; first = min(first, i); }
            if (hi - i < stride) { break; }
        }
        record(r, tested, bad, first);
    }

    __device__ uint64_t mix64(uint64_t z)
    {
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }

    __global__ void __launch_bounds__(kThreads) kDivRandom(uint64_t n, uint64_t seed, Result* r)
    {
        const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
        unsigned long long tested = 0, bad = 0;
        uint32_t first = 0xffffffffu;
        for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            const uint64_t z = mix64(seed + 0x9e3779b97f4a7c15ull * (i + 1));
            const uint32_t ea = static_cast<uint32_t>(127 - 60 + (z % 121));        // |a| in [2^-60, 2^61)
            const uint32_t eb = static_cast<uint32_t>(127 - 30 + ((z >> 8) % 62));  // |b| in [2^-30, 2^32)
            const uint32_t ma = static_cast<uint32_t>(z >> 16) & 0x7fffffu;
            const uint32_t mb = static_cast<uint32_t>(z >> 39) & 0x7fffffu;
            const uint32_t sa = static_cast<uint32_t>(z >> 62) & 1u, sb = static_cast<uint32_t>(z >> 63) & 1u;
            const float a = __uint_as_float((sa << 31) | (ea << 23) | ma);
            const float b = __uint_as_float((sb << 31) | (eb << 23) | mb);
            const float q1 = divCR(a, b);
            const float q2 = a / b;
            tested++;
            if (__float_as_uint(q1) != __float_as_uint(q2)) { bad++; first = min(first, static_cast<uint32_t>(i)); }
        }
        record(r, tested, bad, first);
    }

    bool check(hipError_t e, const char* what)
    {
        if (e != hipSuccess) { std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e)); return false; }
        return true;
    }

    bool report(const char* name, Result* dr)
    {
        Result h{};
        if (!check(hipDeviceSynchronize(), name) || !check(hipMemcpy(&h, dr, sizeof(h), hipMemcpyDeviceToHost), "D2H")) { return false; }
        std::printf("%-44s tested %llu mismatches %llu%s\n", name, h.tested, h.bad, h.bad ? " FAIL" : "");
        if (h.bad) { std::printf("  first mismatch at input %u (0x%08x)\n", h.firstBad, h.firstBad); }
        std::fflush(stdout);
        return h.bad == 0 && h.tested > 0;
    }

    bool reset(Result* dr)
    {
        Result z{};
        z.firstBad = 0xffffffffu;
        return check(hipMemcpy(dr, &z, sizeof(z), hipMemcpyHostToDevice), "H2D");
    }

    uint32_t bitsOf(float f)
    {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
    }
}

int main()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) { std::fprintf(stderr, "no HIP device\n"); return 3; }
    Result* dr = nullptr;
    if (!check(hipMalloc(&dr, sizeof(Result)), "hipMalloc")) { return 1; }
    bool ok = true;

    ok = reset(dr) && ok;
    kSqrt<<<kBlocks, kThreads>>>(1ull + (0x7f800000ull - 0x0f800000ull + 1ull), dr);
    ok = report("sqrt, 0 and every float in [2^-96, inf]", dr) && ok;

    const float hs[] = { 0.0008f, 0.001f, 0.002f, 0.004f, 0.005f, 0.01f, 0.02f, 0.05f, 0.1f, 0.5f, 1.0f };
    for (float h : hs)
    {
        // koutScale as the kernel constants compute it: -3 alpha hinv^3, alpha = hinv^3 / (4 pi) in FP64
        const double hinvD = 1.0 / static_cast<double>(h);
        const float hinv = static_cast<float>(hinvD);
        const float alpha = static_cast<float>(hinvD * hinvD * hinvD / (4.0 * 3.14159265358979323846));
        const float kout = -3.0f * alpha * hinv * hinv * hinv;
        const float ksize = 1.0f / hinv, ktwo = 2.0f / hinv;
        ok = reset(dr) && ok;
        kDivRange<<<kBlocks, kThreads>>>(kout, bitsOf(ksize), bitsOf(ktwo), dr);
        char name[96];
        std::snprintf(name, sizeof(name), "div koutScale / dist, h = %g, dist in [h, 2h]", h);
        ok = report(name, dr) && ok;
    }

    ok = reset(dr) && ok;
    kDivRandom<<<kBlocks, kThreads>>>(1ull << 32, 0x5eed, dr);
    ok = report("div, 2^32 random pairs in range", dr) && ok;

    (void)hipFree(dr);
    std::printf("%s\n", ok ? "PASS" : "FAIL");
    return ok ? 0 : 1;
}
