// kernels_export.hip -- the export snapshot of one scene on the device
// (SimBackend_GPU::PackExportSnapshot, GPU/backend/gpu_backend.cpp:346-551, and the shared
// surface criterion IsNearSurfaceFluidParticle, core/backend_sim.h:109-117).
//
// A fluid particle is near the surface when it touches the boundary or a rigid body, or has
// fewer FF neighbors than the threshold; the exported fluid particles are the near-surface ones
// and their FF neighbors (ClassifyExportSurfaceParticles, CPU/backend/cpu_backend.cpp:937-981).
// They leave in the backend's particle order with their position and speed, compacted on the
// device, so a served scene is exported without copying its whole state to the host.
#include "tsp_device.h"

namespace tsp
{
    static inline uint32_t blocksN(uint64_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

    // flags[i - off] = 1 for every exported particle of the segment [off, off + n) (cleared before;
    // the marks of different particles overlap and are plain stores of the same value)
    __global__ void __launch_bounds__(kBlock) kSurfaceMark(Dev d, uint32_t off, uint32_t n, uint32_t threshold,
                                                           uint32_t* __restrict__ flags)
    {
        const uint32_t local = blockIdx.x * kBlock + threadIdx.x;
        if (local >= n) { return; }
        const uint32_t i = off + local;
        const uint32_t cFF = d.cnt[REL_FF][i];
        const uint32_t cFB = d.cnt[REL_FB][i], cFR = d.cnt[REL_FR][i];
        if (!(cFB > 0u || cFR > 0u || cFF < threshold)) { return; }
        flags[local] = 1u;
        const List ff = listOf(d, REL_FF, i);
        const uint32_t blk = i / kBlock;
        const uint32_t hc = d.haloCount[blk];
        const uint32_t* halo = d.haloIdx + static_cast<size_t>(blk) * kHaloMax;
        visitFF(ff, hc <= kHaloMax, [&](uint32_t q, uint32_t) { flags[halo[q >> kSlotShift] - off] = 1u; },
                [&](uint32_t j, uint32_t) { flags[j - off] = 1u; });
    }

    // the exported particles at their scanned slots: position triples and speed
    __global__ void __launch_bounds__(kBlock) kSurfaceCompact(Dev d, uint32_t off, uint32_t n,
                                                              const uint32_t* __restrict__ scan,
                                                              float* __restrict__ xyz, float* __restrict__ speed)
    {
        const uint32_t local = blockIdx.x * kBlock + threadIdx.x;
        if (local >= n) { return; }
        const uint32_t o = scan[local];
        if (scan[local + 1] == o) { return; }
        const float4 p = d.pos[off + local];
        const float4 v = d.vel[off + local];
        xyz[3ull * o] = p.x; xyz[3ull * o + 1] = p.y; xyz[3ull * o + 2] = p.z;
        speed[o] = __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    }

    void launchScan(hipStream_t s, uint32_t* v, uint32_t m, uint32_t* tmp, uint32_t* fail);

    // the scene's exported fluid particles: flags (n + 1 words) end as their exclusive scan,
    // flags[n] the count; xyz and speed receive them in particle order
    void launchExportSurface(hipStream_t s, const Dev& d, uint32_t off, uint32_t n, uint32_t threshold,
                             uint32_t* flags, uint32_t* tmp, float* xyz, float* speed)
    {
        (void)hipMemsetAsync(flags, 0, sizeof(uint32_t) * (n + 1ull), s);
        if (n) { kSurfaceMark<<<blocksN(n), kBlock, 0, s>>>(d, off, n, threshold, flags); }
        launchScan(s, flags, n + 1, tmp, &d.status->scanFail);
        if (n) { kSurfaceCompact<<<blocksN(n), kBlock, 0, s>>>(d, off, n, flags, xyz, speed); }
    }
}
