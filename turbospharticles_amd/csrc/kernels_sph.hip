// kernels_sph.hip -- IISPH solver kernels of the step on gfx950.
//
// One thread per fluid particle in Morton order; a wave64 reads two SELL-32 slices, so the list
// loads of a neighbor column are two 64-byte rows. FF sums use the reference CPU's dict-chunk
// form (one scale after the sum), FB/FR sums its normal form, each in the list order the fill
// wrote (CPU/backend/sph_solver.cpp):
//   kDensity         Density_ChunkedSELL                            :734-890
//   kNonPressure     NonPressureAcceleration_ChunkedSELL            :892-1149
//   kSourceDiag      SourceTermAndDiagonalElementAndPressureJacobiInit :1225-1504
//   kPressureAcc     PressureAcceleration_ChunkedSELL               :1507-1715
//   kJacobi          PressureRelaxedJacobiSolver_ChunkedSELL        :1781-2049, FP64 error partials
//   finalizeSweep    the error sums, averages and TerminationConditionSatisfied (:205-261, :686-732)
//                    decided on the device by block 0 of the next sweep's kPressureAcc (kFinalize
//                    after the last sweep of a group), so no host round trip is needed per sweep
//   kIntegration     Integration                                    :2051-2089
// Rigid coupling (CPU/backend/cpu_backend.cpp:2001-2303, sph_solver.cpp:1151-1223, :1717-1778):
// particle kinematics, reaction forces over the RF lists, FP64 per-body force/torque sums,
// prediction and semi-implicit integration with Gram-Schmidt re-orthonormalization.
#include "tsp_device.h"

namespace tsp
{
    // copies the FF halo records of a block into LDS (load(slot, particle index)); every thread
    // of the block calls it. A count above kHaloMax (kNoHalo) means the block has no halo.
    template <typename Load>
    __device__ __forceinline__ void stageHalo(const Dev& d, uint32_t blk, uint32_t hc, Load&& load)
    {
        if (hc <= kHaloMax)
        {
            const uint32_t* idx = d.haloIdx + static_cast<size_t>(blk) * kHaloMax;
            for (uint32_t q = threadIdx.x; q < hc; q += blockDim.x) { load(q, idx[q]); }
        }
        __syncthreads();
    }

    template <bool kFast>
    __global__ void __launch_bounds__(kBlock) kDensity(Dev d)
    {
        if (stopped(d)) { return; }
        __shared__ float4 hp[kHaloMax];
        const uint32_t blk = xcdBlock();
        const uint32_t p = blk * kBlock + threadIdx.x;   // list position; i its particle
        const uint32_t i = particleAt(d, p);
        const uint32_t sc = d.blockScene[blk];
        const Consts& k = d.consts[sc];
        const uint32_t hc = d.haloCount[blk];
        stageHalo(d, blk, hc, [&](uint32_t q, uint32_t j) { hp[q] = d.pos[j]; });
        if (i - d.scenes[sc].offF >= d.scenes[sc].nF) { return; }
        const float4 pi = d.pos[i];
        const float xi = pi.x, yi = pi.y, zi = pi.z;
        const List ff = listAt(d, REL_FF, p, i);
        float dF;
        if (ff.n <= kDictMax)
        {
            // dict chunk: unscaled weights, scaled once (sph_solver.cpp:797-810)
            float wsum = 0.0f;
            const auto ffTerm = [&](const float4& pj) { wsum += kwUnscaled<kFast>(dist2(xi - pj.x, yi - pj.y, zi - pj.z), k); };
            visitFF(ff, hc <= kHaloMax, [&](uint32_t q, uint32_t) { ffTerm(recAt(hp, q)); }, [&](uint32_t j, uint32_t) { ffTerm(d.pos[j]); });
            dF = wsum * k.weightScale;
        }
        else
        {
            // normal chunk: the alpha-scaled weight per neighbor (sph_solver.cpp:826-831)
            float s = 0.0f;
            const auto ffTerm = [&](const float4& pj) { s = __builtin_fmaf(k.mass, kwNormal<kFast>(xi - pj.x, yi - pj.y, zi - pj.z, k), s); };
            visitFF(ff, hc <= kHaloMax, [&](uint32_t q, uint32_t) { ffTerm(recAt(hp, q)); }, [&](uint32_t j, uint32_t) { ffTerm(d.pos[j]); });
            dF = s;
        }
        float dB = 0.0f, dR = 0.0f;
        const List fb = listAt(d, REL_FB, p, i);
        visit(fb, [&](uint32_t j, uint32_t)
        {
            const float4 bj = d.bpos[j];
            dB = __builtin_fmaf(bj.w, kwNormal<kFast>(xi - bj.x, yi - bj.y, zi - bj.z, k), dB);
        });
        const List fr = listAt(d, REL_FR, p, i);
        visit(fr, [&](uint32_t j, uint32_t)
        {
            const float4 rj = d.rpos[j];
            dR = __builtin_fmaf(rj.w, kwNormal(xi - rj.x, yi - rj.y, zi - rj.z, k), dR);
        });
        d.vel[i].w = dF + dB + dR;
    }

    template <bool kFast>
    __global__ void __launch_bounds__(kBlock) kNonPressure(Dev d)
    {
        if (stopped(d)) { return; }
        __shared__ float4 hp[kHaloMax], hv[kHaloMax];
        const uint32_t blk = xcdBlock();
        const uint32_t p = blk * kBlock + threadIdx.x;   // list position; i its particle
        const uint32_t i = particleAt(d, p);
        const uint32_t sc = d.blockScene[blk];
        const Consts& k = d.consts[sc];
        const uint32_t hc = d.haloCount[blk];
        stageHalo(d, blk, hc, [&](uint32_t q, uint32_t j) { hp[q] = d.pos[j]; hv[q] = d.vel[j]; });
        if (i - d.scenes[sc].offF >= d.scenes[sc].nF) { return; }
        const float4 pi = d.pos[i], vi = d.vel[i];
        const float xi = pi.x, yi = pi.y, zi = pi.z, vxi = vi.x, vyi = vi.y, vzi = vi.z;
        float aFx = 0.0f, aFy = 0.0f, aFz = 0.0f;
        const List ff = listAt(d, REL_FF, p, i);
        // dict chunks scale by the mass after the sum, normal chunks per neighbor (sph_solver.cpp:974-1049)
        const bool dict = ff.n <= kDictMax;
        // the two forms as two loops (a wave runs the normal one only for a list longer than kDictMax)
        const auto sumFF = [&](auto dictFlag)
        {
            constexpr bool kDict = decltype(dictFlag)::value;
            const auto ffTerm = [&](const float4& pj, const float4& vj)
            {
                const float dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
                const float dvx = vxi - vj.x, dvy = vyi - vj.y, dvz = vzi - vj.z;
                const float d2 = dist2(dx, dy, dz);
                const float kg = kgFactor<kFast>(d2, k);
                const float dot = __builtin_fmaf(dvx, dx, __builtin_fmaf(dvy, dy, dvz * dz));
                const float rj = vmax(vj.w, 1e-7f);
                const float coeff = (kDict ? dot : k.mass * dot) / (rj * (d2 + k.hsqEps)) * kg;
                aFx = __builtin_fmaf(coeff, dx, aFx);
                aFy = __builtin_fmaf(coeff, dy, aFy);
                aFz = __builtin_fmaf(coeff, dz, aFz);
            };
            visitFF(ff, hc <= kHaloMax, [&](uint32_t q, uint32_t) { ffTerm(recAt(hp, q), recAt(hv, q)); },
                    [&](uint32_t j, uint32_t) { ffTerm(d.pos[j], d.vel[j]); });
        };
        if (dict) { sumFF(Flag<true>{}); } else { sumFF(Flag<false>{}); }
        if (dict) { aFx *= k.mass; aFy *= k.mass; aFz *= k.mass; }
        float aBx = 0.0f, aBy = 0.0f, aBz = 0.0f;
        const List fb = listAt(d, REL_FB, p, i);
        visit(fb, [&](uint32_t j, uint32_t e)
        {
            const float4 bj = d.bpos[j];
            const float dx = xi - bj.x, dy = yi - bj.y, dz = zi - bj.z;
            const float d2 = dist2(dx, dy, dz);
            const float f = kgFactor<kFast>(d2, k);
            const float gx = f * dx, gy = f * dy, gz = f * dz;
            const float dvx = vxi - 0.0f, dvy = vyi - 0.0f, dvz = vzi - 0.0f;
            const float dot = __builtin_fmaf(dvx, dx, __builtin_fmaf(dvy, dy, dvz * dz));
            const float coeff = (d.bvol[j] * dot) / (d2 + k.hsqEps);
            aBx = __builtin_fmaf(coeff, gx, aBx);
            aBy = __builtin_fmaf(coeff, gy, aBy);
            aBz = __builtin_fmaf(coeff, gz, aBz);
        });
        float aRx = 0.0f, aRy = 0.0f, aRz = 0.0f;
        const List fr = listAt(d, REL_FR, p, i);
        visit(fr, [&](uint32_t j, uint32_t e)
        {
            const float4 rj = d.rpos[j], uj = d.rvel[j];
            const float dx = xi - rj.x, dy = yi - rj.y, dz = zi - rj.z;
            const float d2 = dist2(dx, dy, dz);
            const float f = kgFactor(d2, k);
            const float gx = f * dx, gy = f * dy, gz = f * dz;
            const float dvx = vxi - uj.x, dvy = vyi - uj.y, dvz = vzi - uj.z;
            const float dot = __builtin_fmaf(dvx, dx, __builtin_fmaf(dvy, dy, dvz * dz));
            const float coeff = (uj.w * dot) / (d2 + k.hsqEps);
            aRx = __builtin_fmaf(coeff, gx, aRx);
            aRy = __builtin_fmaf(coeff, gy, aRy);
            aRz = __builtin_fmaf(coeff, gz, aRz);
        });
        const float ax = aFx * k.visF + (aBx + aRx) * k.visB + k.gx;
        const float ay = aFy * k.visF + (aBy + aRy) * k.visB + k.gy;
        const float az = aFz * k.visF + (aBz + aRz) * k.visB + k.gz;
        d.vp[i] = make_float4(__builtin_fmaf(ax, k.dt, vxi), __builtin_fmaf(ay, k.dt, vyi), __builtin_fmaf(az, k.dt, vzi), 0.0f);
    }

    template <bool kFast>
    __global__ void __launch_bounds__(kBlock) kSourceDiag(Dev d)
    {
        if (stopped(d)) { return; }
        __shared__ float4 hp[kHaloMax], hv[kHaloMax];
        const uint32_t blk = xcdBlock();
        const uint32_t p = blk * kBlock + threadIdx.x;   // list position; i its particle
        const uint32_t i = particleAt(d, p);
        const uint32_t sc = d.blockScene[blk];
        const Consts& k = d.consts[sc];
        const uint32_t hc = d.haloCount[blk];
        stageHalo(d, blk, hc, [&](uint32_t q, uint32_t j) { hp[q] = d.pos[j]; hv[q] = d.vp[j]; });
        if (i - d.scenes[sc].offF >= d.scenes[sc].nF) { return; }
        const float4 pi = d.pos[i], vi = d.vp[i];
        const float xi = pi.x, yi = pi.y, zi = pi.z, vxi = vi.x, vyi = vi.y, vzi = vi.z;
        float sF = 0.0f, cFx = 0.0f, cFy = 0.0f, cFz = 0.0f, g2 = 0.0f;
        const List ff = listAt(d, REL_FF, p, i);
        const auto ffTerm = [&](const float4& pj, const float4& vj, uint32_t e)
        {
            const float dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
            const float dvx = vxi - vj.x, dvy = vyi - vj.y, dvz = vzi - vj.z;
            const float d2 = dist2(dx, dy, dz);
            const float kg = kgFactor<kFast>(d2, k);
#if TSP_TUNE_KG_STORE
            // the sweeps of this step read the pair's gradient factor instead of recomputing it
            d.kgFF[e] = kg;
#else
            (void)e;
#endif
            const float st = __builtin_fmaf(dvx, dx, __builtin_fmaf(dvy, dy, dvz * dz));
            sF = __builtin_fmaf(kg, st, sF);
            cFx = __builtin_fmaf(kg, dx, cFx);
            cFy = __builtin_fmaf(kg, dy, cFy);
            cFz = __builtin_fmaf(kg, dz, cFz);
            g2 = __builtin_fmaf(kg, kg * d2, g2);
        };
        visitFF(ff, hc <= kHaloMax, [&](uint32_t q, uint32_t e) { ffTerm(recAt(hp, q), recAt(hv, q), e); },
                [&](uint32_t j, uint32_t e) { ffTerm(d.pos[j], d.vp[j], e); });
        float sB = 0.0f, cBx = 0.0f, cBy = 0.0f, cBz = 0.0f;
        const List fb = listAt(d, REL_FB, p, i);
        visit(fb, [&](uint32_t j, uint32_t)
        {
            const float4 bj = d.bpos[j];
            const float dx = xi - bj.x, dy = yi - bj.y, dz = zi - bj.z;
            const float f = kgFactor<kFast>(dist2(dx, dy, dz), k);
            const float gx = f * dx, gy = f * dy, gz = f * dz;
            const float dvx = vxi - 0.0f, dvy = vyi - 0.0f, dvz = vzi - 0.0f;
            const float st = __builtin_fmaf(dvx, gx, __builtin_fmaf(dvy, gy, dvz * gz));
            const float psi = bj.w;
            sB = __builtin_fmaf(psi, st, sB);
            cBx = __builtin_fmaf(psi, gx, cBx);
            cBy = __builtin_fmaf(psi, gy, cBy);
            cBz = __builtin_fmaf(psi, gz, cBz);
        });
        float sR = 0.0f, cRx = 0.0f, cRy = 0.0f, cRz = 0.0f;
        const List fr = listAt(d, REL_FR, p, i);
        visit(fr, [&](uint32_t j, uint32_t)
        {
            const float4 rj = d.rpos[j], uj = d.rvp[j];
            const float dx = xi - rj.x, dy = yi - rj.y, dz = zi - rj.z;
            const float f = kgFactor(dist2(dx, dy, dz), k);
            const float gx = f * dx, gy = f * dy, gz = f * dz;
            const float dvx = vxi - uj.x, dvy = vyi - uj.y, dvz = vzi - uj.z;
            const float st = __builtin_fmaf(dvx, gx, __builtin_fmaf(dvy, gy, dvz * gz));
            const float psi = rj.w;
            sR = __builtin_fmaf(psi, st, sR);
            cRx = __builtin_fmaf(psi, gx, cRx);
            cRy = __builtin_fmaf(psi, gy, cRy);
            cRz = __builtin_fmaf(psi, gz, cRz);
        });
        sF *= k.mass;
        const float src = k.rho0 - d.vel[i].w - k.dt * (sF + sB + sR);
        const float ctx = cFx * k.cFacF + (cBx + cRx) * k.cFacB;
        const float cty = cFy * k.cFacF + (cBy + cRy) * k.cFacB;
        const float ctz = cFz * k.cFacF + (cBz + cRz) * k.cFacB;
        const float dA = k.mass * __builtin_fmaf(ctx, cFx, __builtin_fmaf(cty, cFy, ctz * cFz));
        const float dB = k.mass * k.dFacB * g2;
        const float dC = __builtin_fmaf(ctx, cBx + cRx, __builtin_fmaf(cty, cBy + cRy, ctz * (cBz + cRz)));
        const float diag = k.dtsq * (dA + dB + dC);
        const bool valid = __builtin_fabsf(diag) > 1e-7f;
        const float cold = k.omega * (src / diag);
        float init = cold;
        if (k.warm > 0.0f)
        {
            const float prev = pi.w;
            init = prev > 0.0f ? k.warm * prev : cold;
        }
        d.source[i] = src;
        d.diag[i] = diag;
        d.pos[i].w = vmax(0.0f, valid ? init : 0.0f);
    }

    // TerminationConditionSatisfied, sph_solver.cpp:686-732
    __device__ __forceinline__ bool terminated(const Solve& s, const Consts& k)
    {
        if (s.iters < k.minIters) { return false; }
        if (s.iters >= k.maxIters) { return true; }
        switch (k.termCond)
        {
        case 0: return s.iters >= k.fixedIters;
        case 1: return s.avgRaw <= k.densityDevTarget;
        case 2: return s.avgPos <= k.densityDevTarget;
        case 3: return s.avgAbs <= k.densityDevTarget;
        default: return s.iters >= 3;
        }
    }

    // the solver state before the first sweep: N_solverIters_ = 1, errors 0, then the
    // increment and error decision of the first sweep (sph_solver.cpp:207-235)
    __global__ void __launch_bounds__(kBlock) kSolverInit(Dev d)
    {
        if (stopped(d)) { return; }
        for (uint32_t sc = threadIdx.x; sc < d.nScenes; sc += kBlock)
        {
            Solve& s = d.solve[sc];
            const Consts& k = d.consts[sc];
            const bool any = d.scenes[sc].nF > 0;
            s.iters = any ? 1u : 0u;
            s.avgRaw = s.avgPos = s.avgAbs = 0.0f;
            s.sumRaw = s.sumPos = s.sumAbs = 0.0;
            s.converged = any ? 0u : 1u;
            s.sweeps = 0;
            s.pending = 0;
            s.iters++;
            s.compute = (s.iters >= k.minIters && (k.termCond != 0 || terminated(s, k))) ? 1u : 0u;
        }
    }

    // sum over the 64 lanes of a wave (a fixed xor tree: the same order on every run)
    __device__ __forceinline__ double waveSum(double v)
    {
        for (int o = 32; o > 0; o >>= 1) { v += __shfl_xor(v, o, 64); }
        return v;
    }

    // threads of a solver sweep block (one 256-particle tile per block)
    constexpr uint32_t kSweepThreads = TSP_TUNE_SWEEP_THREADS;
    static_assert(kBlock % kSweepThreads == 0 && kSweepThreads % 64 == 0, "sweep block size");

    // the bookkeeping of one scene after a Jacobi sweep (sph_solver.cpp:207-261): the scene's block
    // partials summed in a fixed order, the averages, TerminationConditionSatisfied and the error
    // decision of the next sweep. Run by one whole block of kSweepThreads; nTot is the grid size of
    // the sweep.
    __device__ void finalizeSweep(const Dev& d, uint32_t sc, uint32_t nTot)
    {
        __shared__ double red[3][kSweepThreads / 64];
        __shared__ uint32_t flags[2];
        Solve& s = d.solve[sc];
        const Consts& k = d.consts[sc];
        const uint32_t b0 = d.scenes[sc].blk0, nBlocks = d.scenes[sc].nBlk;
        // one read of the flags for the whole block, so every wave takes the same path (the first
        // barrier lets a previous call's readers finish before thread 0 writes them again)
        __syncthreads();
        if (threadIdx.x == 0) { flags[0] = s.pending; flags[1] = s.compute; }
        __syncthreads();
        if (!flags[0]) { return; }
        const bool compute = flags[1] != 0u;
        if (compute)
        {
            for (int r = 0; r < 3; r++)
            {
                const double* src = d.errPartial + static_cast<size_t>(r) * nTot + b0;
                double a = 0.0;
                uint32_t b = threadIdx.x;
                constexpr uint32_t kT = kSweepThreads;
                for (; b + 3 * kT < nBlocks; b += 4 * kT)
                {
                    const double v0 = src[b], v1 = src[b + kT], v2 = src[b + 2 * kT], v3 = src[b + 3 * kT];
                    a += v0; a += v1; a += v2; a += v3;
                }
                for (; b < nBlocks; b += kT) { a += src[b]; }
                a = waveSum(a);
                if ((threadIdx.x & 63u) == 0) { red[r][threadIdx.x >> 6] = a; }
            }
            __syncthreads();
            if (threadIdx.x == 0)
            {
                for (int r = 0; r < 3; r++)
                {
                    double a = red[r][0];
                    for (uint32_t w = 1; w < kSweepThreads / 64; w++) { a += red[r][w]; }
                    red[r][0] = a;
                }
            }
        }
        if (threadIdx.x == 0)
        {
            s.pending = 0;
            s.sweeps++;
            if (compute)
            {
                s.sumRaw = red[0][0]; s.sumPos = red[1][0]; s.sumAbs = red[2][0];
                const float n = static_cast<float>(d.scenes[sc].nF);
                s.avgRaw = static_cast<float>(red[0][0]) / n;
                s.avgPos = static_cast<float>(red[1][0]) / n;
                s.avgAbs = static_cast<float>(red[2][0]) / n;
            }
            if (terminated(s, k))
            {
                s.converged = 1u;
            }
            else
            {
                s.iters++;
                s.compute = (s.iters >= k.minIters && (k.termCond != 0 || terminated(s, k))) ? 1u : 0u;
            }
        }
        __syncthreads();
    }

    // a sweep's pressure accelerations; block 0 of a sweep launch first finishes the bookkeeping of
    // the previous Jacobi sweep, which the next Jacobi launch reads. Blocks that see the solve
    // converged skip their particles: the final launch (sweep = false) recomputes every a^p.
    // The sweeps walk the 256-particle tiles of all scenes. One block per tile, in the XCD-aware
    // order of xcdBlock(); or (TSP_TUNE_PERSISTENT) a grid of at most kSweepGrid resident blocks,
    // each walking its XCD's contiguous range of tiles with a stride of the blocks per XCD.
    constexpr uint32_t kSweepGrid = 256u * 6u;

    template <typename Fn>
    __device__ __forceinline__ void forTiles(uint32_t nTiles, Fn tile)
    {
#if TSP_TUNE_PERSISTENT
        if (gridDim.x >= nTiles)
        {
            tile(xcdBlock());
            return;
        }
        const uint32_t b = blockIdx.x, xcd = b & 7u, stride = gridDim.x >> 3;
        const uint32_t per = nTiles >> 3, rem = nTiles & 7u;
        const uint32_t start = xcd < rem ? xcd * (per + 1u) : rem * (per + 1u) + (xcd - rem) * per;
        const uint32_t end = start + per + (xcd < rem ? 1u : 0u);
        for (uint32_t t = start + (b >> 3); t < end; t += stride)
        {
            tile(t);
            // the next tile restages the LDS records and flags this one has read
            __syncthreads();
        }
#else
        (void)nTiles;
        tile(xcdBlock());
#endif
    }

    static inline uint32_t sweepGrid(uint32_t nTiles)
    {
#if TSP_TUNE_PERSISTENT
        return nTiles < kSweepGrid ? nTiles : kSweepGrid;
#else
        return nTiles;
#endif
    }

    // one tile of a sweep's pressure accelerations
    template <bool kFast>
    __device__ __forceinline__ void pressureAccTile(const Dev& d, uint32_t blk, bool sweep)
    {
        __shared__ float4 hp[kHaloMax];
        __shared__ uint32_t skip;
        const uint32_t sc = d.blockScene[blk];
        const Consts& k = d.consts[sc];
        if (sweep)
        {
            // one read per block, so the whole block leaves together
            if (threadIdx.x == 0) { skip = __hip_atomic_load(&d.solve[sc].converged, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
            __syncthreads();
            if (skip) { return; }
        }
        const uint32_t hc = d.haloCount[blk];
        stageHalo(d, blk, hc, [&](uint32_t q, uint32_t j) { hp[q] = d.pos[j]; });
        for (uint32_t r = 0; r < kBlock / kSweepThreads; r++)
        {
        const uint32_t p = blk * kBlock + r * kSweepThreads + threadIdx.x;   // list position; i its particle
        const uint32_t i = particleAt(d, p);
        if (i - d.scenes[sc].offF >= d.scenes[sc].nF) { continue; }
        const float4 p4 = d.pos[i];
        const float xi = p4.x, yi = p4.y, zi = p4.z, pi = p4.w;
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        const List ff = listAt(d, REL_FF, p, i);
        // dict chunks sum (p_i + p_j) kg and scale once; normal chunks scale each neighbor's
        // pressure term (sph_solver.cpp:1573-1640)
        const bool dict = ff.n <= kDictMax;
        const float outerF = k.mass * pi * k.presDis;
        // the two forms as two loops (a wave runs the normal one only for a list longer than kDictMax)
        const auto sumFF = [&](auto dictFlag)
        {
            constexpr bool kDict = decltype(dictFlag)::value;
            const auto ffTerm = [&](const float4& pj, uint32_t e)
            {
                const float dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
#if TSP_TUNE_KG_STORE
                const float kg = d.kgFF[e];
#else
                (void)e;
                const float kg = kgFactor<kFast>(dist2(dx, dy, dz), k);
#endif
                const float coeff = (kDict ? pi + pj.w : outerF + k.mass * pj.w * k.presDis) * kg;
                sx = __builtin_fmaf(coeff, dx, sx);
                sy = __builtin_fmaf(coeff, dy, sy);
                sz = __builtin_fmaf(coeff, dz, sz);
            };
            visitFFSplit(ff, hc <= kHaloMax, [&](uint32_t q) { return recAt(hp, q); }, ffTerm,
                         [&](uint32_t j, uint32_t e) { ffTerm(d.pos[j], e); });
        };
        if (dict) { sumFF(Flag<true>{}); } else { sumFF(Flag<false>{}); }
        float ax = dict ? sx * k.presOuterFF : sx, ay = dict ? sy * k.presOuterFF : sy, az = dict ? sz * k.presOuterFF : sz;
        const List fb = listAt(d, REL_FB, p, i);
        visit(fb, [&](uint32_t j, uint32_t)
        {
            const float4 bj = d.bpos[j];
            const float dx = xi - bj.x, dy = yi - bj.y, dz = zi - bj.z;
            const float f = kgFactor<kFast>(dist2(dx, dy, dz), k);
            const float coeff = k.presOuterB * bj.w * pi;
            ax = __builtin_fmaf(coeff, f * dx, ax);
            ay = __builtin_fmaf(coeff, f * dy, ay);
            az = __builtin_fmaf(coeff, f * dz, az);
        });
        const List fr = listAt(d, REL_FR, p, i);
        visit(fr, [&](uint32_t j, uint32_t)
        {
            const float4 rj = d.rpos[j];
            const float dx = xi - rj.x, dy = yi - rj.y, dz = zi - rj.z;
            const float f = kgFactor(dist2(dx, dy, dz), k);
            const float coeff = k.presOuterB * rj.w * pi;
            ax = __builtin_fmaf(coeff, f * dx, ax);
            ay = __builtin_fmaf(coeff, f * dy, ay);
            az = __builtin_fmaf(coeff, f * dz, az);
        });
        d.ap[i] = make_float4(0.0f - ax, 0.0f - ay, 0.0f - az, 0.0f);
        }
    }

    template <bool kFast>
    __global__ void __launch_bounds__(kSweepThreads) TSP_WAVES_ATTR(TSP_TUNE_PA_WAVES) kPressureAcc(Dev d, bool sweep, uint32_t nTiles)
    {
        if (stopped(d)) { return; }
        // block b of the launch finishes the previous sweep of scenes b, b + grid, ...
        if (sweep)
        {
            for (uint32_t q = blockIdx.x; q < d.nScenes; q += gridDim.x) { finalizeSweep(d, q, nTiles); }
        }
        forTiles(nTiles, [&](uint32_t blk) { pressureAccTile<kFast>(d, blk, sweep); });
    }


    // one tile of a Jacobi sweep; nTot is the number of tiles (the stride of the error partials)
    template <bool kFast>
    __device__ __forceinline__ void jacobiTile(const Dev& d, uint32_t blk, uint32_t nTot)
    {
        __shared__ double red[3][kSweepThreads / 64];
#if TSP_TUNE_JACOBI_REC24
        // 24-byte halo records (x, y, z, a^p x | a^p y, a^p z): 24 KB per block, six blocks per CU
        __shared__ float4 hp[kHaloMax];
        __shared__ float2 ha[kHaloMax];
#else
        __shared__ float4 hp[kHaloMax], ha[kHaloMax];
#endif
        const uint32_t sc = d.blockScene[blk];
        const Consts& k = d.consts[sc];
        const Scene& S = d.scenes[sc];
        // the scene's flags were written by an earlier launch: every thread reads the same values
        if (d.solve[sc].converged) { return; }
        const bool compute = d.solve[sc].compute != 0u;
        if (blk == S.blk0 && threadIdx.x == 0) { d.solve[sc].pending = 1u; }
        const uint32_t hc = d.haloCount[blk];
        // (positions only: the pressures this launch writes into pos.w are not read here)
#if TSP_TUNE_JACOBI_REC24
        stageHalo(d, blk, hc, [&](uint32_t q, uint32_t j)
        {
            const float4 pj = d.pos[j], aj = d.ap[j];
            hp[q] = make_float4(pj.x, pj.y, pj.z, aj.x);
            ha[q] = make_float2(aj.y, aj.z);
        });
#else
        stageHalo(d, blk, hc, [&](uint32_t q, uint32_t j) { hp[q] = d.pos[j]; ha[q] = d.ap[j]; });
#endif
        double eRaw = 0.0, ePos = 0.0, eAbs = 0.0;
        for (uint32_t r = 0; r < kBlock / kSweepThreads; r++)
        {
        const uint32_t p = blk * kBlock + r * kSweepThreads + threadIdx.x;   // list position; i its particle
        const uint32_t i = particleAt(d, p);
        if (i - S.offF < S.nF)
        {
            const float4 pi = d.pos[i], ai = d.ap[i];
            const float xi = pi.x, yi = pi.y, zi = pi.z, axi = ai.x, ayi = ai.y, azi = ai.z;
            float sf = 0.0f;
            const List ff = listAt(d, REL_FF, p, i);
            const auto ffTerm = [&](float xj, float yj, float zj, float axj, float ayj, float azj, uint32_t e)
            {
                const float dx = xi - xj, dy = yi - yj, dz = zi - zj;
#if TSP_TUNE_KG_STORE
                const float kg = d.kgFF[e];
#else
                (void)e;
                const float kg = kgFactor<kFast>(dist2(dx, dy, dz), k);
#endif
                const float sv = __builtin_fmaf(axi - axj, dx, __builtin_fmaf(ayi - ayj, dy, (azi - azj) * dz));
                sf = __builtin_fmaf(kg, sv, sf);
            };
            // a halo record: position and a^p of one neighbor
            struct Rec
            {
                float x, y, z, ax, ay, az;
            };
#if TSP_TUNE_JACOBI_REC24
            const auto haloLoad = [&](uint32_t q)
            {
                const float4 pj = recAt(hp, q);
                const float2 aj = recAt(ha, q);
                return Rec{ pj.x, pj.y, pj.z, pj.w, aj.x, aj.y };
            };
#else
            const auto haloLoad = [&](uint32_t q)
            {
                const float4 pj = recAt(hp, q), aj = recAt(ha, q);
                return Rec{ pj.x, pj.y, pj.z, aj.x, aj.y, aj.z };
            };
#endif
            visitFFSplit(ff, hc <= kHaloMax, haloLoad,
                         [&](const Rec& r, uint32_t e) { ffTerm(r.x, r.y, r.z, r.ax, r.ay, r.az, e); },
                         [&](uint32_t j, uint32_t e)
                         {
                             const float4 pj = d.pos[j], aj = d.ap[j];
                             ffTerm(pj.x, pj.y, pj.z, aj.x, aj.y, aj.z, e);
                         });
            sf *= k.mass;
            float sb = 0.0f;
            const List fb = listAt(d, REL_FB, p, i);
            visit(fb, [&](uint32_t j, uint32_t)
            {
                const float4 bj = d.bpos[j];
                const float dx = xi - bj.x, dy = yi - bj.y, dz = zi - bj.z;
                const float f = kgFactor<kFast>(dist2(dx, dy, dz), k);
                const float sv = __builtin_fmaf(axi, f * dx, __builtin_fmaf(ayi, f * dy, azi * (f * dz)));
                sb += bj.w * sv;
            });
            const List fr = listAt(d, REL_FR, p, i);
            visit(fr, [&](uint32_t j, uint32_t)
            {
                const float4 rj = d.rpos[j];
                const float dx = xi - rj.x, dy = yi - rj.y, dz = zi - rj.z;
                const float f = kgFactor(dist2(dx, dy, dz), k);
                const float sv = __builtin_fmaf(axi, f * dx, __builtin_fmaf(ayi, f * dy, azi * (f * dz)));
                sb += rj.w * sv;
            });
            const float sys = k.dtsq * (sf + sb);
            const float src = d.source[i], diag = d.diag[i];
            const float resid = src - sys;
            const float pred = sys - src;
            const bool valid = __builtin_fabsf(diag) > 1e-7f;
            d.pos[i].w = vmax(0.0f, valid ? __builtin_fmaf(k.omega, resid / diag, pi.w) : 0.0f);
            if (compute && valid)
            {
                const double e = static_cast<double>(pred / k.rho0);
                eRaw += e;
                ePos += 0.0 < e ? e : 0.0;
                eAbs += __builtin_fabs(e);
            }
        }
        }
        if (!compute) { return; }
        eRaw = waveSum(eRaw);
        ePos = waveSum(ePos);
        eAbs = waveSum(eAbs);
        const uint32_t w = threadIdx.x >> 6;
        if ((threadIdx.x & 63u) == 0) { red[0][w] = eRaw; red[1][w] = ePos; red[2][w] = eAbs; }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            for (int r = 0; r < 3; r++)
            {
                double s = 0.0;
                for (uint32_t q = 0; q < kSweepThreads / 64; q++) { s += red[r][q]; }
                d.errPartial[static_cast<size_t>(r) * nTot + blk] = s;
            }
        }
    }

    template <bool kFast>
    __global__ void __launch_bounds__(kSweepThreads) TSP_WAVES_ATTR(TSP_TUNE_JACOBI_WAVES) kJacobi(Dev d, uint32_t nTiles)
    {
        if (stopped(d)) { return; }
        forTiles(nTiles, [&](uint32_t blk) { jacobiTile<kFast>(d, blk, nTiles); });
    }

    // the bookkeeping of the last sweep of a group, one block per scene, before the host reads
    // the solver states; nTot is the grid size of the sweeps
    __global__ void __launch_bounds__(kSweepThreads) kFinalize(Dev d, uint32_t nTot)
    {
        if (stopped(d)) { return; }
        finalizeSweep(d, blockIdx.x, nTot);
    }

    __global__ void __launch_bounds__(kBlock) kIntegration(Dev d)
    {
        if (stopped(d)) { return; }
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        const uint32_t sc = d.blockScene[blockIdx.x];
        const Consts& k = d.consts[sc];
        if (i - d.scenes[sc].offF >= d.scenes[sc].nF) { return; }
        const float4 a = d.ap[i], v = d.vp[i], p = d.pos[i];
        const float vx = __builtin_fmaf(k.dt, a.x, v.x);
        const float vy = __builtin_fmaf(k.dt, a.y, v.y);
        const float vz = __builtin_fmaf(k.dt, a.z, v.z);
        d.vel[i] = make_float4(vx, vy, vz, d.vel[i].w);
        d.pos[i] = make_float4(__builtin_fmaf(k.dt, vx, p.x), __builtin_fmaf(k.dt, vy, p.y), __builtin_fmaf(k.dt, vz, p.z), p.w);
    }

    // ----------------------------------------------------------------------------------------
    // rigid bodies; the body state is field-major: rb[field * nRB + body] (rbv)

    // UpdateRigidParticles(predicted), cpu_backend.cpp:2009-2063
    __global__ void __launch_bounds__(kBlock) kRigidUpdate(Dev d, bool predicted)
    {
        if (stopped(d)) { return; }
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i < d.nR) { rigidUpdateOne(d, i, predicted); }
    }

    // ClearRigidForcesAndAddGravity, cpu_backend.cpp:2065-2087
    __global__ void kRigidClearGravity(Dev d)
    {
        if (stopped(d)) { return; }
        const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
        if (b >= d.nRB) { return; }
        const Consts& k = d.consts[d.bodyScene[b]];
        const uint32_t nb = d.nRB;
        for (int f = 0; f < 6; f++) { d.rbForce[f * nb + b] = 0.0f; }
        const float m = rbv(d, RB_MASS, b), mi = rbv(d, RB_MASS_INV, b);
        if (m <= 0.0f || mi <= 0.0f) { return; }
        d.rbForce[0 * nb + b] += m * k.gx;
        d.rbForce[1 * nb + b] += m * k.gy;
        d.rbForce[2 * nb + b] += m * k.gz;
    }

    // AccumulateRigidParticleNonPressureForces_ChunkedSELL / ...PressureForces (sph_solver.cpp:1151-1223,
    // :1717-1778): reaction of the fluid neighbors of each rigid particle
    __global__ void __launch_bounds__(kBlock) kRigidForces(Dev d, bool pressure)
    {
        if (stopped(d)) { return; }
        // one wave per rigid particle: the lanes compute 64 pairs' coefficients and gradients at
        // once, then every lane folds them in list order (the single-thread sum, term by term)
        const uint32_t i = (blockIdx.x * kBlock + threadIdx.x) >> 6;
        const uint32_t lane = threadIdx.x & 63u;
        if (i >= d.nR) { return; }
        const Consts& k = d.consts[d.bodyScene[d.rbody[i]]];
        const float4 ri = d.rpos[i], ui = d.rvel[i];
        const float xi = ri.x, yi = ri.y, zi = ri.z;
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        const uint32_t q0 = d.rfOff[i], q1 = d.rfOff[i + 1];
        for (uint32_t base = q0; base < q1; base += 64)
        {
            const uint32_t q = base + lane;
            float fc = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
            if (q < q1)
            {
                const uint32_t j = d.rfIdx[q];
                const float4 pj = d.pos[j];
                const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
                kgScalar(dx, dy, dz, k, gx, gy, gz);
                if (pressure)
                {
                    fc = k.mass * k.presOuterB * ri.w * pj.w;
                }
                else
                {
                    const float4 vj = d.vel[j];
                    const float dvx = vj.x - ui.x, dvy = vj.y - ui.y, dvz = vj.z - ui.z;
                    const float dot = __builtin_fmaf(dvx, dx, __builtin_fmaf(dvy, dy, dvz * dz));
                    const float d2 = dist2(dx, dy, dz);
                    const float coeff = (ui.w * dot) / (d2 + k.hsqEps);
                    fc = -k.mass * k.visB * coeff;
                }
            }
            const uint32_t n = min(64u, q1 - base);
            for (uint32_t l = 0; l < n; l++)
            {
                const float c = __shfl(fc, static_cast<int>(l), 64);
                fx = __builtin_fmaf(c, __shfl(gx, static_cast<int>(l), 64), fx);
                fy = __builtin_fmaf(c, __shfl(gy, static_cast<int>(l), 64), fy);
                fz = __builtin_fmaf(c, __shfl(gz, static_cast<int>(l), 64), fz);
            }
        }
        if (lane != 0) { return; }
        if (pressure) { d.rfx[i] += fx; d.rfy[i] += fy; d.rfz[i] += fz; }
        else { d.rfx[i] = fx; d.rfy[i] = fy; d.rfz[i] = fz; }
    }

    // ReduceRigidParticleForcesToBody, cpu_backend.cpp:2089-2172, with FP64 sums: one block per body
    // one block of kReduceThreads per body: a body's particles are few blocks' worth, so the
    // block is wide to keep the chains of dependent loads and sums short
    constexpr uint32_t kReduceThreads = 1024;

    __global__ void __launch_bounds__(kReduceThreads) kRigidReduce(Dev d)
    {
        if (stopped(d)) { return; }
        __shared__ double red[6][kReduceThreads];
        const uint32_t b = blockIdx.x;
        const uint32_t nb = d.nRB;
        const float a[9] = { rbv(d, RB_A00 + 0, b), rbv(d, RB_A00 + 1, b), rbv(d, RB_A00 + 2, b),
                             rbv(d, RB_A00 + 3, b), rbv(d, RB_A00 + 4, b), rbv(d, RB_A00 + 5, b),
                             rbv(d, RB_A00 + 6, b), rbv(d, RB_A00 + 7, b), rbv(d, RB_A00 + 8, b) };
        double acc[6] = { 0, 0, 0, 0, 0, 0 };
        for (uint32_t q = d.byBodyOff[b] + threadIdx.x; q < d.byBodyOff[b + 1]; q += kReduceThreads)
        {
            const uint32_t i = d.byBodyIdx[q];
            const float lx = d.rxrel[i], ly = d.ryrel[i], lz = d.rzrel[i];
            const float rx = __builtin_fmaf(a[0], lx, __builtin_fmaf(a[1], ly, a[2] * lz));
            const float ry = __builtin_fmaf(a[3], lx, __builtin_fmaf(a[4], ly, a[5] * lz));
            const float rz = __builtin_fmaf(a[6], lx, __builtin_fmaf(a[7], ly, a[8] * lz));
            const double fx = d.rfx[i], fy = d.rfy[i], fz = d.rfz[i];
            acc[0] += fx; acc[1] += fy; acc[2] += fz;
            acc[3] += static_cast<double>(ry) * fz - static_cast<double>(rz) * fy;
            acc[4] += static_cast<double>(rz) * fx - static_cast<double>(rx) * fz;
            acc[5] += static_cast<double>(rx) * fy - static_cast<double>(ry) * fx;
        }
        for (int f = 0; f < 6; f++) { red[f][threadIdx.x] = acc[f]; }
        __syncthreads();
        for (uint32_t o = kReduceThreads / 2; o > 0; o >>= 1)
        {
            if (threadIdx.x < o) { for (int f = 0; f < 6; f++) { red[f][threadIdx.x] += red[f][threadIdx.x + o]; } }
            __syncthreads();
        }
        if (threadIdx.x == 0)
        {
            for (int f = 0; f < 6; f++) { d.rbForce[f * nb + b] += static_cast<float>(red[f][0]); }
        }
    }

    // PredictRigidBodies, cpu_backend.cpp:2174-2235
    __global__ void kRigidPredict(Dev d)
    {
        if (stopped(d)) { return; }
        const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
        if (b < d.nRB) { rigidPredictOne(d, b); }
    }

    __device__ __forceinline__ bool normalize3(float& x, float& y, float& z)
    {
        const float len = __builtin_sqrtf(__builtin_fmaf(x, x, __builtin_fmaf(y, y, z * z)));
        if (len <= 1e-12f || !__builtin_isfinite(len)) { return false; }
        const float inv = 1.0f / len;
        x *= inv; y *= inv; z *= inv;
        return true;
    }

    __device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz)
    {
        return __builtin_fmaf(ax, bx, __builtin_fmaf(ay, by, az * bz));
    }

    // IntegrateRigidBodies, cpu_backend.cpp:2237-2303
    __global__ void kRigidIntegrate(Dev d)
    {
        if (stopped(d)) { return; }
        const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
        if (b >= d.nRB) { return; }
        const Consts& k = d.consts[d.bodyScene[b]];
        const uint32_t nb = d.nRB;
        float* rb = d.rb;
        for (int a = 0; a < 3; a++)
        {
            rb[(RB_VX + a) * nb + b] = d.rbPred[a * nb + b];
            rb[(RB_LX + a) * nb + b] = d.rbPred[(3 + a) * nb + b];
        }
        for (int a = 0; a < 3; a++) { rb[(RB_XCM + a) * nb + b] += k.dt * rb[(RB_VX + a) * nb + b]; }
        const float wx = d.rbPred[6 * nb + b], wy = d.rbPred[7 * nb + b], wz = d.rbPred[8 * nb + b];
        rb[(RB_WX + 0) * nb + b] = wx;
        rb[(RB_WX + 1) * nb + b] = wy;
        rb[(RB_WX + 2) * nb + b] = wz;
        float A[9];
        for (int q = 0; q < 9; q++) { A[q] = rb[(RB_A00 + q) * nb + b]; }
        const float dt = k.dt;
        float N[9];
        N[0] = A[0] + dt * (-wz * A[3] + wy * A[6]);
        N[1] = A[1] + dt * (-wz * A[4] + wy * A[7]);
        N[2] = A[2] + dt * (-wz * A[5] + wy * A[8]);
        N[3] = A[3] + dt * (wz * A[0] - wx * A[6]);
        N[4] = A[4] + dt * (wz * A[1] - wx * A[7]);
        N[5] = A[5] + dt * (wz * A[2] - wx * A[8]);
        N[6] = A[6] + dt * (-wy * A[0] + wx * A[3]);
        N[7] = A[7] + dt * (-wy * A[1] + wx * A[4]);
        N[8] = A[8] + dt * (-wy * A[2] + wx * A[5]);
        float b1x = N[0], b1y = N[3], b1z = N[6];
        float b2x = N[1], b2y = N[4], b2z = N[7];
        float b3x = N[2], b3y = N[5], b3z = N[8];
        bool ok = normalize3(b1x, b1y, b1z);
        const float d12 = dot3(b1x, b1y, b1z, b2x, b2y, b2z);
        b2x -= d12 * b1x; b2y -= d12 * b1y; b2z -= d12 * b1z;
        ok = ok && normalize3(b2x, b2y, b2z);
        const float d13 = dot3(b1x, b1y, b1z, b3x, b3y, b3z);
        const float d23 = dot3(b2x, b2y, b2z, b3x, b3y, b3z);
        b3x -= d13 * b1x + d23 * b2x;
        b3y -= d13 * b1y + d23 * b2y;
        b3z -= d13 * b1z + d23 * b2z;
        ok = ok && normalize3(b3x, b3y, b3z);
        if (!ok) { atomicOr(&d.solve[d.bodyScene[b]].rigidDegenerate, 1u); return; }
        const float out[9] = { b1x, b2x, b3x, b1y, b2y, b3y, b1z, b2z, b3z };
        for (int q = 0; q < 9; q++) { rb[(RB_A00 + q) * nb + b] = out[q]; }
    }

    // ----------------------------------------------------------------------------------------
    // host launchers

    static inline uint32_t blocksOf(uint64_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

    void launchDensity(hipStream_t s, const Dev& d)
    {
        if (!d.nF) { return; }
        if (d.fastPairs) { kDensity<true><<<blocksOf(d.nF), kBlock, 0, s>>>(d); }
        else { kDensity<false><<<blocksOf(d.nF), kBlock, 0, s>>>(d); }
    }
    void launchNonPressure(hipStream_t s, const Dev& d)
    {
        if (!d.nF) { return; }
        if (d.fastPairs) { kNonPressure<true><<<blocksOf(d.nF), kBlock, 0, s>>>(d); }
        else { kNonPressure<false><<<blocksOf(d.nF), kBlock, 0, s>>>(d); }
    }
    void launchSourceDiag(hipStream_t s, const Dev& d)
    {
        if (d.nF)
        {
            if (d.fastPairs) { kSourceDiag<true><<<blocksOf(d.nF), kBlock, 0, s>>>(d); }
            else { kSourceDiag<false><<<blocksOf(d.nF), kBlock, 0, s>>>(d); }
        }
        if (d.nScenes) { kSolverInit<<<1, kBlock, 0, s>>>(d); }
    }
    void launchPressureAcc(hipStream_t s, const Dev& d, bool sweep)
    {
        if (!d.nF) { return; }
        const uint32_t n = blocksOf(d.nF);
        if (d.fastPairs) { kPressureAcc<true><<<sweepGrid(n), kSweepThreads, 0, s>>>(d, sweep, n); }
        else { kPressureAcc<false><<<sweepGrid(n), kSweepThreads, 0, s>>>(d, sweep, n); }
    }
    void launchJacobi(hipStream_t s, const Dev& d)
    {
        if (!d.nF) { return; }
        const uint32_t n = blocksOf(d.nF);
        if (d.fastPairs) { kJacobi<true><<<sweepGrid(n), kSweepThreads, 0, s>>>(d, n); }
        else { kJacobi<false><<<sweepGrid(n), kSweepThreads, 0, s>>>(d, n); }
    }
    void launchFinalize(hipStream_t s, const Dev& d)
    {
        if (d.nF && d.nScenes) { kFinalize<<<d.nScenes, kSweepThreads, 0, s>>>(d, blocksOf(d.nF)); }
    }
    void launchIntegration(hipStream_t s, const Dev& d) { if (d.nF) { kIntegration<<<blocksOf(d.nF), kBlock, 0, s>>>(d); } }

    void launchRigidUpdate(hipStream_t s, const Dev& d, bool predicted) { if (d.nR) { kRigidUpdate<<<blocksOf(d.nR), kBlock, 0, s>>>(d, predicted); } }
    void launchRigidClearGravity(hipStream_t s, const Dev& d) { if (d.nRB) { kRigidClearGravity<<<blocksOf(d.nRB), kBlock, 0, s>>>(d); } }
    void launchRigidForces(hipStream_t s, const Dev& d, bool pressure)
    {
        if (d.nR) { kRigidForces<<<blocksOf(64ull * d.nR), kBlock, 0, s>>>(d, pressure); }
    }
    void launchRigidReduce(hipStream_t s, const Dev& d) { if (d.nRB) { kRigidReduce<<<d.nRB, kReduceThreads, 0, s>>>(d); } }
    void launchRigidPredict(hipStream_t s, const Dev& d) { if (d.nRB) { kRigidPredict<<<blocksOf(d.nRB), kBlock, 0, s>>>(d); } }
    void launchRigidIntegrate(hipStream_t s, const Dev& d) { if (d.nRB) { kRigidIntegrate<<<blocksOf(d.nRB), kBlock, 0, s>>>(d); } }
}
