// kernels_build.hip -- a generated scene written straight into a context's arrays: the fluid
// block with its SplitMix64 jitter and the container's boundary layers with their psi / V kernel
// sums, restating per particle the host builder's formulas (host/scene_plan.h: FluidPosition,
// BoundaryPosition, ComputeCoefficients; reference scene/builder.cpp:174-376, :415-642,
// :1400-1509). The tests hold the two builders to the same bits.
//
// The coefficient sums follow the host's order exactly: the boundary particles sorted by cell
// (z, y, x) and, within a cell, by index (a stable radix sort of packed cell keys), and for each
// particle the nine rows (dz, dy) of three cells each, visited in sorted order.
#include "tsp_device.h"

namespace tsp
{
    static inline uint32_t blocksN(uint64_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

    __device__ __forceinline__ BuildVec vadd(BuildVec a, BuildVec b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
    __device__ __forceinline__ BuildVec vmul(BuildVec a, float s) { return { a.x * s, a.y * s, a.z * s }; }

    // std::min / std::max / std::clamp as the host evaluates them
    __device__ __forceinline__ float smin(float a, float b) { return b < a ? b : a; }
    __device__ __forceinline__ float smaxf(float a, float b) { return a < b ? b : a; }
    __device__ __forceinline__ float sclamp(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }

    // Jitter3D (builder.cpp:1486-1501): SplitMix64 of seed + key * golden, uniform in [-amp, amp)
    __device__ __forceinline__ float jitter3D(uint64_t baseSeed, uint64_t key, float amp)
    {
        uint64_t x = baseSeed + key * 0x9E3779B97F4A7C15ull;
        x ^= x >> 30;
        x *= 0xBF58476D1CE4E5B9ull;
        x ^= x >> 27;
        x *= 0x94D049BB133111EBull;
        x ^= x >> 31;
        const double u01 = static_cast<double>(x >> 11) * (1.0 / 9007199254740992.0);
        return static_cast<float>((u01 * 2.0 - 1.0) * static_cast<double>(amp));
    }

    // the fluid block's records: position (pressure 0), velocity 0 with density rho0, the other
    // fields 0, ids = lattice index
    __global__ void __launch_bounds__(kBlock) kBuildFluid(Dev d, uint32_t off, uint32_t n, BuildFluid b)
    {
        const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
        if (idx >= n) { return; }
        const uint64_t k = idx;
        const uint32_t iu = static_cast<uint32_t>(k % b.nu);
        const uint32_t jv = static_cast<uint32_t>((k / b.nu) % b.nv);
        const uint32_t kw = static_cast<uint32_t>(k / (static_cast<uint64_t>(b.nu) * b.nv));
        float u = b.uOff + static_cast<float>(iu) * b.size;
        float v = b.vOff + static_cast<float>(jv) * b.size;
        float w = b.wOff + static_cast<float>(kw) * b.size;
        if (b.jitAmp > 0.0f)
        {
            u += jitter3D(b.seed, k * 3ull + 0ull, b.jitAmp);
            v += jitter3D(b.seed, k * 3ull + 1ull, b.jitAmp);
            w += jitter3D(b.seed, k * 3ull + 2ull, b.jitAmp);
        }
        u = sclamp(u, b.radius, b.uLen - b.radius);
        v = sclamp(v, b.radius, b.vLen - b.radius);
        w = sclamp(w, b.radius, b.wLen - b.radius);
        const BuildVec p = vadd(vadd(vadd(b.origin, vmul(b.u, u)), vmul(b.v, v)), vmul(b.w, w));
        const uint32_t i = off + idx;
        d.pos[i] = make_float4(p.x, p.y, p.z, 0.0f);
        d.vel[i] = make_float4(0.0f, 0.0f, 0.0f, b.rho0);
        d.vp[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        d.ap[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        d.source[i] = 0.0f;
        d.diag[i] = 0.0f;
        d.id[i] = k;
    }

    __device__ __forceinline__ int floorInt(float v) { return static_cast<int>(__builtin_floorf(v)); }

    __device__ __forceinline__ uint32_t packCell(const BuildBoundary& c, int x, int y, int z)
    {
        return (static_cast<uint32_t>(z - c.cellMin[2]) << (c.bits[0] + c.bits[1]))
             | (static_cast<uint32_t>(y - c.cellMin[1]) << c.bits[0]) | static_cast<uint32_t>(x - c.cellMin[0]);
    }

    // the boundary positions of every face, their ids and packed cell keys (the sort's input)
    __global__ void __launch_bounds__(kBlock) kBuildBoundaryPos(BuildBoundary c, uint32_t n, float* __restrict__ xyz,
                                                                uint64_t* __restrict__ ids, uint32_t* __restrict__ key,
                                                                uint32_t* __restrict__ idx)
    {
        const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
        if (k >= n) { return; }
        uint32_t q = 0;
        while (q + 1 < c.nFaces && k >= c.faces[q + 1].first) { q++; }
        const BuildFace& f = c.faces[q];
        const uint64_t rem = k - f.first;
        const uint32_t j = static_cast<uint32_t>(rem % f.nv);
        const uint32_t i = static_cast<uint32_t>((rem / f.nv) % f.nu);
        const uint32_t layer = static_cast<uint32_t>(rem / (static_cast<uint64_t>(f.nu) * f.nv));
        const float outOffset = c.size * 0.5f + static_cast<float>(layer) * c.size;
        const float su = smin(static_cast<float>(i) * c.size, f.uLen);
        const float sv = smin(static_cast<float>(j) * c.size, f.vLen);
        const BuildVec p = vadd(vadd(vadd(f.origin, vmul(f.u, su)), vmul(f.v, sv)), vmul(f.n, outOffset));
        xyz[k] = p.x; xyz[n + k] = p.y; xyz[2ull * n + k] = p.z;
        ids[k] = k;
        key[k] = packCell(c, floorInt(p.x * c.cellSizeInv), floorInt(p.y * c.cellSizeInv), floorInt(p.z * c.cellSizeInv));
        idx[k] = k;
    }

    // the scalar weight of the coefficient sums (builder.cpp:49-79)
    __device__ __forceinline__ float coeffWeight(float dx, float dy, float dz, const BuildBoundary& c)
    {
        const float distSqu = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        if (distSqu >= c.radiusSqu) { return 0.0f; }
        if (distSqu < 1e-7f) { return 4.0f * c.alpha; }
        const float dist = __builtin_sqrtf(distSqu);
        const float q = dist * c.sizeInv;
        const float t1 = smaxf(1.0f - q, 0.0f);
        const float t2 = smaxf(2.0f - q, 0.0f);
        const float t1_3 = (t1 * t1) * t1;
        const float t2_3 = (t2 * t2) * t2;
        return c.alpha * (t2_3 - 4.0f * t1_3);
    }

    // first sorted position with key >= v (lower) or > v (upper)
    __device__ __forceinline__ uint32_t bound(const uint32_t* __restrict__ key, uint32_t n, uint32_t v, bool upper)
    {
        uint32_t lo = 0, hi = n;
        while (lo < hi)
        {
            const uint32_t mid = lo + (hi - lo) / 2;
            const bool before = upper ? key[mid] <= v : key[mid] < v;
            if (before) { lo = mid + 1; } else { hi = mid; }
        }
        return lo;
    }

    // ComputeParticleCoefficients for one boundary particle (original index i): V = 1 / sum W over
    // the nine rows of its cell's neighborhood in sorted order, psi = rho0 V; the scene's records
    __global__ void __launch_bounds__(kBlock) kBuildBoundaryCoeff(BuildBoundary c, uint32_t n, const float* __restrict__ xyz,
                                                                  const uint32_t* __restrict__ key,
                                                                  const uint32_t* __restrict__ sorted,
                                                                  float4* __restrict__ bpos, float* __restrict__ bvol)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) { return; }
        const float xi = xyz[i], yi = xyz[n + i], zi = xyz[2ull * n + i];
        const int cx = floorInt(xi * c.cellSizeInv), cy = floorInt(yi * c.cellSizeInv), cz = floorInt(zi * c.cellSizeInv);
        float sumW = 0.0f;
        for (int dz = -1; dz <= 1; dz++)
        {
            const int z = cz + dz;
            if (z < c.cellMin[2] || z > c.cellMax[2]) { continue; }
            for (int dy = -1; dy <= 1; dy++)
            {
                const int y = cy + dy;
                if (y < c.cellMin[1] || y > c.cellMax[1]) { continue; }
                const int x0 = max(cx - 1, c.cellMin[0]), x1 = min(cx + 1, c.cellMax[0]);
                if (x0 > x1) { continue; }
                const uint32_t lo = bound(key, n, packCell(c, x0, y, z), false);
                const uint32_t hi = bound(key, n, packCell(c, x1, y, z), true);
                for (uint32_t t = lo; t < hi; t++)
                {
                    const uint32_t j = sorted[t];
                    sumW += coeffWeight(xi - xyz[j], yi - xyz[n + j], zi - xyz[2ull * n + j], c);
                }
            }
        }
        float vol, psi;
        if (sumW > 0.0f && __builtin_isfinite(sumW))
        {
            vol = 1.0f / sumW;
            psi = c.rho0 * vol;
        }
        else
        {
            psi = c.massB;
            vol = c.rho0 > 0.0f ? c.massB / c.rho0 : 0.0f;
        }
        bpos[i] = make_float4(xi, yi, zi, psi);
        bvol[i] = vol;
    }

    bool launchRadixSort(hipStream_t, bool, uint32_t*, uint32_t*, uint32_t*, uint32_t*, uint32_t, uint32_t, uint32_t*, uint32_t*,
                         uint32_t*, bool);

    void launchBuildFluid(hipStream_t s, const Dev& d, uint32_t off, uint32_t n, const BuildFluid& b)
    {
        if (n) { kBuildFluid<<<blocksN(n), kBlock, 0, s>>>(d, off, n, b); }
    }

    // scratch: xyz 3n floats; key, idx, key2, idx2 n words each; hist 256 per 2048-key tile;
    // tmp a zeroed scan state for the digit counts (its word 2 is set if a scan gave up)
    void launchBuildBoundary(hipStream_t s, const BuildBoundary& c, uint32_t n, float4* bpos, float* bvol, uint64_t* ids,
                             float* xyz, uint32_t* key, uint32_t* idx, uint32_t* key2, uint32_t* idx2, uint32_t* hist,
                             uint32_t* tmp)
    {
        if (!n) { return; }
        kBuildBoundaryPos<<<blocksN(n), kBlock, 0, s>>>(c, n, xyz, ids, key, idx);
        launchRadixSort(s, false, key, idx, key2, idx2, n, c.bits[0] + c.bits[1] + c.bits[2], hist, tmp, tmp + 2, true);
        kBuildBoundaryCoeff<<<blocksN(n), kBlock, 0, s>>>(c, n, xyz, key, idx, bpos, bvol);
    }
}
