// kernels_search.hip -- neighbor search of the IISPH step on gfx950.
//
// Re-order (DataReOrder_SolverStep), Construct and Query of the reference's compact
// cell-linked list, written for wave64:
//   keys      Morton cell id of every particle in the local cell frame, with the frame
//             hysteresis test of UpdateCellFrame (CPU/backend/neighbor_search.cpp:1902-1960)
//   sort      stable LSD radix sort of (key, index), 8-bit digits, keys clamped so that
//             out-of-frame particles sort last (neighbor_search.cpp:403-564); per tile the
//             stable ranks come from wave64 ballots, not from a serial scan
//   permute   one gather of the packed particle records per particle
//   cells     dense {begin, end} table over the frame, written at the runs of the sorted keys;
//             the previous step's runs are cleared first, so no global counter is needed
//   count     per particle: FF/FB/FR counts over the 27-cell stencil (inclusive cutoff,
//             distance fma(dx,dx,fma(dy,dy,dz*dz)), neighbor_search.cpp:1130-1249); per
//             32-particle slice: pitch, base and u16/u32 format by width-32 lane reductions
//   scan      multi-block exclusive scans (slice offsets, radix digit offsets, RF offsets)
//   fill      writes the entries in the reference's iteration order: z-y-x stencil raster,
//             ascending index within a cell; FF entries are staged in LDS and written as whole
//             slice rows. FF lists whose indices span more than 0xffff keep the DictDelta16x2
//             order of the reference's per-particle dict chunks (low group first, high group
//             reversed, CPU/backend/neighbor_lists.h:171-197)
//   halo      per 256-particle block: the distinct FF neighbors (LDS hash set, bitonic sort) and
//             the block's FF entries rewritten as slots of that list, so the solver kernels
//             stage the neighbor records in LDS once per block
#include "tsp_device.h"

namespace tsp
{
    // ----------------------------------------------------------------------------------------
    // multi-block exclusive scan of u32 values (tiles of 2048 = 256 threads x 8)

    constexpr uint32_t kScanTile = 2048;

    __device__ __forceinline__ uint32_t blockExclusiveScan(uint32_t v, uint32_t& total)
    {
        __shared__ uint32_t s[kBlock];
        s[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < kBlock; off <<= 1)
        {
            const uint32_t a = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += a;
            __syncthreads();
        }
        total = s[kBlock - 1];
        const uint32_t r = s[threadIdx.x] - v;
        __syncthreads();
        return r;
    }

    // One-pass exclusive scan (decoupled look-back, Merrill & Garland 2016): each block takes the
    // next tile number from a counter, so a tile waits only on tiles whose blocks started before
    // it; it publishes its sum, adds its predecessors' sums walking back to the first inclusive
    // prefix, publishes that prefix and writes its values. Status words carry the scan's epoch
    // (bit 33 up), a prefix flag (bit 32) and the value, so no clearing pass is needed: the block
    // of the last tile resets the counter and counts the scan in the epoch word. A scan's epoch is
    // 1 + (scans before it mod 2^31 - 1), never 0, so the zeroed state of a new buffer matches no
    // scan. A wait longer than kScanSpins polls gives up and sets the fail word (never expected;
    // the host reports it).
    constexpr uint32_t kScanSpins = 1u << 20;

    __global__ void __launch_bounds__(kBlock) kScanOnePass(uint32_t* __restrict__ v, uint32_t m, uint32_t* __restrict__ tmp,
        uint32_t* __restrict__ fail)
    {
        __shared__ uint32_t sTile, sEpoch, sPrefix;
        uint32_t* counter = tmp;
        uint32_t* epochWord = tmp + 1;
        unsigned long long* status = reinterpret_cast<unsigned long long*>(tmp + kScanHeaderWords);
        const uint32_t nb = (m + kScanTile - 1) / kScanTile;
        if (threadIdx.x == 0)
        {
            sEpoch = __hip_atomic_load(epochWord, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            sTile = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const uint32_t tile = sTile;
        const unsigned long long tag = static_cast<unsigned long long>(sEpoch % 0x7fffffffu + 1u) << 33;
        const uint32_t base = tile * kScanTile + threadIdx.x * 8u;
        uint32_t loc[8];
        uint32_t sum = 0;
        for (uint32_t q = 0; q < 8; q++) { loc[q] = base + q < m ? v[base + q] : 0u; sum += loc[q]; }
        uint32_t total;
        const uint32_t r = blockExclusiveScan(sum, total);
        if (threadIdx.x == 0)
        {
            constexpr unsigned long long kPrefix = 1ull << 32;
            uint32_t prefix = 0;
            if (tile > 0)
            {
                __hip_atomic_store(&status[tile], tag | total, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                uint32_t j = tile - 1, spins = 0;
                for (;;)
                {
                    const unsigned long long w = __hip_atomic_load(&status[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                    if ((w >> 33) != (tag >> 33))
                    {
                        if (++spins > kScanSpins) { atomicOr(fail, 1u); break; }
                        __builtin_amdgcn_s_sleep(2);
                        continue;
                    }
                    prefix += static_cast<uint32_t>(w);
                    if ((w & kPrefix) || j == 0) { break; }
                    j--;
                }
            }
            __hip_atomic_store(&status[tile], tag | kPrefix | (prefix + total), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            sPrefix = prefix;
            if (tile == nb - 1)
            {
                // every block has drawn its tile: ready the counter and epoch for the next scan
                __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(epochWord, sEpoch + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();
        uint32_t run = sPrefix + r;
        for (uint32_t q = 0; q < 8; q++)
        {
            if (base + q < m) { v[base + q] = run; }
            run += loc[q];
        }
    }

    // exclusive scan of m values in place; tmp is a scan state of scanTmpWords(m) words, zeroed
    // when allocated (fail: set if a look-back gave up)
    void launchScan(hipStream_t s, uint32_t* v, uint32_t m, uint32_t* tmp, uint32_t* fail)
    {
        if (m == 0) { return; }
        const uint32_t nb = (m + kScanTile - 1) / kScanTile;
        kScanOnePass<<<nb, kBlock, 0, s>>>(v, m, tmp, fail);
    }

    // clears the search status block and the scenes' frame-abort flags
    __global__ void __launch_bounds__(kBlock) kResetSearch(Dev d)
    {
        constexpr uint32_t words = static_cast<uint32_t>(sizeof(Status) / sizeof(uint32_t));
        static_assert(words <= kBlock, "the reset covers one block");
        constexpr uint32_t keep = static_cast<uint32_t>(offsetof(Status, scanFail) / sizeof(uint32_t));
        if (threadIdx.x < words && threadIdx.x != keep) { reinterpret_cast<uint32_t*>(d.status)[threadIdx.x] = 0u; }
        for (uint32_t s = threadIdx.x; s < d.nScenes; s += kBlock)
        {
            Solve& so = d.solve[s];
            so.abort = 0u;
            so.contactActive = 0u; so.contactDone = 0u; so.contactIters = 0u; so.contactConverged = 0u;
            so.contactAvg = 0.0f; so.contactBoundary = 0u; so.contactBad = 0u; so.contacts = 0ull;
        }
    }

    // ----------------------------------------------------------------------------------------
    // keys

    __device__ __forceinline__ bool insideDomain(const Frame& f, float x, float y, float z)
    {
        if (!f.domainSet) { return true; }
        return x >= f.domMin[0] && x <= f.domMax[0] && y >= f.domMin[1] && y <= f.domMax[1]
            && z >= f.domMin[2] && z <= f.domMax[2];
    }

    // cell id of one position in its scene's key range; sets the abort flags when an in-domain
    // particle left the inner frame
    __device__ __forceinline__ uint64_t cellKey(const Frame& f, float x, float y, float z, bool check, Status* st, Solve* so)
    {
        const int cx = floorToInt(x * f.cellSizeInv);
        const int cy = floorToInt(y * f.cellSizeInv);
        const int cz = floorToInt(z * f.cellSizeInv);
        if (check && insideDomain(f, x, y, z))
        {
            if (cx < f.innerLo[0] || cx > f.innerHi[0] || cy < f.innerLo[1] || cy > f.innerHi[1]
                || cz < f.innerLo[2] || cz > f.innerHi[2])
            {
                atomicOr(&st->abort, 1u);
                atomicOr(&so->abort, 1u);
            }
        }
        const int lx = cx - f.orig[0], ly = cy - f.orig[1], lz = cz - f.orig[2];
        if (lx < 0 || lx >= static_cast<int>(f.dim[0]) || ly < 0 || ly >= static_cast<int>(f.dim[1])
            || lz < 0 || lz >= static_cast<int>(f.dim[2]))
        {
            return f.emptyKey;
        }
        return f.keyBase + morton64(static_cast<uint32_t>(lx), static_cast<uint32_t>(ly), static_cast<uint32_t>(lz));
    }

    enum Pop { POP_F = 0, POP_B = 1, POP_R = 2 };

    __device__ __forceinline__ uint32_t sceneOfParticle(const Dev& d, int pop, uint32_t i)
    {
        if (pop == POP_F) { return d.blockScene[i / kBlock]; }
        if (pop == POP_B) { return d.sceneB[i]; }
        return d.bodyScene[d.rbody[i]];
    }

    // key of particle i of a population; the padding slots of a fluid segment take the dead key
    __device__ __forceinline__ uint64_t particleKey(const Dev& d, int pop, const float4* p, uint32_t i, bool check)
    {
        const uint32_t s = sceneOfParticle(d, pop, i);
        const Frame& f = d.frames[s];
        if (pop == POP_F && i - d.scenes[s].offF >= d.scenes[s].nF) { return f.deadKey; }
        const float4 q = p[i];
        return cellKey(f, q.x, q.y, q.z, check, d.status, &d.solve[s]);
    }

    template <typename K>
    __global__ void __launch_bounds__(kBlock) kKeys(Dev d, int pop, const float4* __restrict__ p, uint32_t n, bool check,
        K* __restrict__ key, uint32_t* __restrict__ idx)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) { return; }
        key[i] = static_cast<K>(particleKey(d, pop, p, i, check));
        idx[i] = i;
    }

    // boundary keys; flags a descent so the host knows whether the order needs a sort
    template <typename K>
    __global__ void __launch_bounds__(kBlock) kKeysSorted(Dev d, const float4* __restrict__ p, uint32_t n,
        K* __restrict__ key, uint32_t* __restrict__ idx)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) { return; }
        const uint64_t k = particleKey(d, POP_B, p, i, false);
        key[i] = static_cast<K>(k);
        idx[i] = i;
        if (i > 0 && particleKey(d, POP_B, p, i - 1, false) > k) { atomicOr(&d.status->bUnsorted, 1u); }
    }

    // ----------------------------------------------------------------------------------------
    // stable LSD radix sort, kRadixBits-bit digits (kRadixBins <= 256 buckets), tiles of 2048 keys

    constexpr uint32_t kTile = 2048;
    constexpr uint32_t kRounds = kTile / kBlock;

    template <typename K>
    __global__ void __launch_bounds__(kBlock) kRadixHist(const K* __restrict__ key, uint32_t n, uint32_t shift,
        uint32_t nTiles, uint32_t* __restrict__ hist)
    {
        __shared__ uint32_t cnt[kRadixBins];
        if (threadIdx.x < kRadixBins) { cnt[threadIdx.x] = 0; }
        __syncthreads();
        const uint32_t base = blockIdx.x * kTile;
        for (uint32_t r = 0; r < kRounds; r++)
        {
            const uint32_t i = base + r * kBlock + threadIdx.x;
            if (i < n) { atomicAdd(&cnt[static_cast<uint32_t>(key[i] >> shift) & (kRadixBins - 1u)], 1u); }
        }
        __syncthreads();
        if (threadIdx.x < kRadixBins) { hist[threadIdx.x * nTiles + blockIdx.x] = cnt[threadIdx.x]; }
    }

    template <typename K>
    __global__ void __launch_bounds__(kBlock) kRadixScatter(const K* __restrict__ keyIn,
        const uint32_t* __restrict__ idxIn, K* __restrict__ keyOut, uint32_t* __restrict__ idxOut, uint32_t n,
        uint32_t shift, uint32_t nTiles, const uint32_t* __restrict__ offsets)
    {
        __shared__ uint32_t run[kRadixBins];
        __shared__ uint32_t waveCnt[kBlock / 64][kRadixBins];
        const uint32_t t = threadIdx.x;
        const uint32_t wave = t >> 6, lane = t & 63u;
        if (t < kRadixBins)
        {
            run[t] = offsets[t * nTiles + blockIdx.x];
            for (uint32_t w = 0; w < kBlock / 64; w++) { waveCnt[w][t] = 0; }
        }
        __syncthreads();
        const uint64_t lanesBelow = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
        const uint32_t base = blockIdx.x * kTile;
        for (uint32_t r = 0; r < kRounds; r++)
        {
            const uint32_t i = base + r * kBlock + t;
            const bool valid = i < n;
            const K k = valid ? keyIn[i] : K(0);
            const uint32_t d = static_cast<uint32_t>(k >> shift) & (kRadixBins - 1u);
            // lanes of this wave holding the same digit
            uint64_t match = __ballot(valid);
            for (uint32_t bit = 0; bit < kRadixBits; bit++)
            {
                const bool set = (d >> bit) & 1u;
                const uint64_t bs = __ballot(set);
                match &= set ? bs : ~bs;
            }
            const uint32_t rankInWave = __popcll(match & lanesBelow);
            if (valid && rankInWave == 0) { waveCnt[wave][d] = __popcll(match); }
            __syncthreads();
            uint32_t rank = 0;
            if (valid)
            {
                rank = run[d] + rankInWave;
                for (uint32_t w = 0; w < wave; w++) { rank += waveCnt[w][d]; }
            }
            __syncthreads();
            if (t < kRadixBins)
            {
                uint32_t tot = 0;
                for (uint32_t w = 0; w < kBlock / 64; w++) { tot += waveCnt[w][t]; waveCnt[w][t] = 0; }
                run[t] += tot;
            }
            __syncthreads();
            if (valid)
            {
                keyOut[rank] = k;
                idxOut[rank] = idxIn[i];
            }
        }
    }

    // the whole stable LSD sort of up to kSmallSort keys in one block of 1024 threads (one launch
    // instead of five per digit). Per digit: each thread loads its 16 elements (i = c*1024 + t) at
    // once; the digit counts are scanned by the first 256 threads with wave shuffles; then the
    // elements are ranked chunk by chunk in element order: wave ballots give the rank inside a
    // wave, 256 digit threads turn the per-wave counts into per-wave offsets, and the elements
    // scatter. The passes alternate between the two buffers in global memory, and an odd pass
    // count copies back at the end.
    constexpr uint32_t kSmallThreads = 1024;
    constexpr uint32_t kSmallPer = 16;
    constexpr uint32_t kSmallSort = kSmallThreads * kSmallPer;

    template <typename K>
    __global__ void __launch_bounds__(kSmallThreads) kSortSmall(K* __restrict__ key, uint32_t* __restrict__ idx,
        K* __restrict__ key2, uint32_t* __restrict__ idx2, uint32_t n, uint32_t bits)
    {
        constexpr uint32_t kWaves = kSmallThreads / 64;
        __shared__ uint32_t run[kRadixBins];
        __shared__ uint32_t waveTot[4];
        __shared__ uint32_t cnt[kWaves][kRadixBins];
        __shared__ uint32_t pre[kWaves][kRadixBins];
        const uint32_t t = threadIdx.x;
        const uint32_t wave = t >> 6, lane = t & 63u;
        const uint64_t lanesBelow = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
        for (uint32_t w = 0; w < kWaves; w++) { if (t < kRadixBins) { cnt[w][t] = 0; } }
        K* kIn = key; uint32_t* iIn = idx; K* kOut = key2; uint32_t* iOut = idx2;
        for (uint32_t shift = 0; shift < bits; shift += kRadixBits)
        {
            K kr[kSmallPer];
            uint32_t ir[kSmallPer];
#pragma unroll
            for (uint32_t c = 0; c < kSmallPer; c++)
            {
                const uint32_t i = c * kSmallThreads + t;
                kr[c] = i < n ? kIn[i] : K(0);
                ir[c] = i < n ? iIn[i] : 0u;
            }
            if (t < kRadixBins) { run[t] = 0; }
            __syncthreads();
#pragma unroll
            for (uint32_t c = 0; c < kSmallPer; c++)
            {
                if (c * kSmallThreads + t < n) { atomicAdd(&run[static_cast<uint32_t>(kr[c] >> shift) & (kRadixBins - 1u)], 1u); }
            }
            __syncthreads();
            // exclusive scan of the 256 digit counts: inclusive scan per wave, then the wave totals
            uint32_t v = 0, incl = 0;
            if (t < kRadixBins)
            {
                v = run[t];
                incl = v;
                for (int o = 1; o < 64; o <<= 1)
                {
                    const uint32_t u = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), o, 64));
                    if (lane >= static_cast<uint32_t>(o)) { incl += u; }
                }
                if (lane == 63) { waveTot[wave] = incl; }
            }
            __syncthreads();
            if (t < kRadixBins)
            {
                uint32_t off = 0;
                for (uint32_t w = 0; w < wave; w++) { off += waveTot[w]; }
                run[t] = off + incl - v;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t c = 0; c < kSmallPer; c++)
            {
                if (c * kSmallThreads >= n) { break; }
                const uint32_t i = c * kSmallThreads + t;
                const bool valid = i < n;
                const K k = kr[c];
                const uint32_t dg = static_cast<uint32_t>(k >> shift) & (kRadixBins - 1u);
                uint64_t match = __ballot(valid);
                for (uint32_t bit = 0; bit < kRadixBits; bit++)
                {
                    const bool set = (dg >> bit) & 1u;
                    const uint64_t bs = __ballot(set);
                    match &= set ? bs : ~bs;
                }
                const uint32_t rankInWave = __popcll(match & lanesBelow);
                if (valid && rankInWave == 0) { cnt[wave][dg] = __popcll(match); }
                __syncthreads();
                // per digit: offsets of the waves' runs (and the counts cleared for the next chunk)
                if (t < kRadixBins)
                {
                    uint32_t cw[kWaves];
#pragma unroll
                    for (uint32_t w = 0; w < kWaves; w++) { cw[w] = cnt[w][t]; }
                    uint32_t acc = run[t];
#pragma unroll
                    for (uint32_t w = 0; w < kWaves; w++) { pre[w][t] = acc; acc += cw[w]; cnt[w][t] = 0; }
                    run[t] = acc;
                }
                __syncthreads();
                if (valid)
                {
                    const uint32_t rank = pre[wave][dg] + rankInWave;
                    kOut[rank] = k;
                    iOut[rank] = ir[c];
                }
            }
            __syncthreads();
            K* tk = kIn; kIn = kOut; kOut = tk;
            uint32_t* ti = iIn; iIn = iOut; iOut = ti;
        }
        if (kIn != key)
        {
            for (uint32_t i = t; i < n; i += kSmallThreads) { key[i] = kIn[i]; idx[i] = iIn[i]; }
        }
    }

    // ----------------------------------------------------------------------------------------
    // permutation (gather into the second buffers)

    __global__ void __launch_bounds__(kBlock) kPermute(Gather g, const uint32_t* __restrict__ idx)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= g.n) { return; }
        const uint32_t o = idx[i];
        for (uint32_t f = 0; f < g.n4; f++) { g.d4[f][i] = g.s4[f][o]; }
        for (uint32_t f = 0; f < g.n1; f++) { g.d1[f][i] = g.s1[f][o]; }
        if (g.idS) { g.idD[i] = g.idS[o]; }
    }

    // ----------------------------------------------------------------------------------------
    // cell tables

    // cell-table index of an in-frame key (local Morton id = key - keyBase)
    __device__ __forceinline__ uint32_t linearOfKey(const Frame& f, uint64_t k)
    {
        const uint64_t m = k - f.keyBase;
        return linearCell(f, compact21(m), compact21(m >> 1), compact21(m >> 2));
    }

    // clears the cells of the runs of the previous sorted keys
    __global__ void __launch_bounds__(kBlock) kClearCells(Dev d, const uint32_t* __restrict__ key, uint32_t n,
        uint2* __restrict__ cell)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) { return; }
        const uint64_t k = keyAt(d, key, i);
        const Frame& f = d.frames[sceneOfKey(d, k)];
        if (k >= f.emptyKey) { return; }
        if (i == 0 || keyAt(d, key, i - 1) != k) { cell[linearOfKey(f, k)] = make_uint2(0u, 0u); }
    }

    __global__ void __launch_bounds__(kBlock) kClearAll(uint2* __restrict__ cell, uint64_t n)
    {
        const uint64_t q = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
        if (q < n) { cell[q] = make_uint2(0u, 0u); }
    }

    __global__ void __launch_bounds__(kBlock) kBuildCells(Dev d, const uint32_t* __restrict__ key, uint32_t n,
        uint2* __restrict__ cell)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) { return; }
        const uint64_t k = keyAt(d, key, i);
        const Frame& f = d.frames[sceneOfKey(d, k)];
        if (k >= f.emptyKey) { return; }
        const bool first = i == 0 || keyAt(d, key, i - 1) != k;
        const bool last = i + 1 == n || keyAt(d, key, i + 1) != k;
        if (!first && !last) { return; }
        const uint32_t lin = linearOfKey(f, k);
        if (first) { cell[lin].x = i; }
        if (last) { cell[lin].y = i + 1; }
    }

    // ----------------------------------------------------------------------------------------
    // query: one thread per fluid particle (sorted order), 27-cell stencil in z-y-x raster order

    // the stencil's cell box, and the frame's cell-table layout kept in registers (the walk
    // stores lists, so the compiler would otherwise reload the frame at every row)
    struct Stencil
    {
        int lo[3], hi[3];
        uint32_t off, dx, dxy;   // first cell of the scene's tables, cells per y step, per z step

        __device__ __forceinline__ uint32_t cell(int x, int y, int z) const
        {
            return off + static_cast<uint32_t>(x) + dx * static_cast<uint32_t>(y) + dxy * static_cast<uint32_t>(z);
        }
    };

    __device__ __forceinline__ bool stencilOf(const Frame& f, uint64_t k, Stencil& s)
    {
        if (k >= f.emptyKey) { return false; }
        const uint64_t m = k - f.keyBase;
        const int l[3] = { static_cast<int>(compact21(m)), static_cast<int>(compact21(m >> 1)),
                           static_cast<int>(compact21(m >> 2)) };
        for (int a = 0; a < 3; a++)
        {
            s.lo[a] = max(l[a] - f.reach, 0);
            s.hi[a] = min(l[a] + f.reach, static_cast<int>(f.dim[a]) - 1);
        }
        s.off = f.cellOff;
        s.dx = f.dim[0];
        s.dxy = f.dim[0] * f.dim[1];
        return true;
    }

    // The fluid cell ranges of a block's stencil box -- the bounding box of the stencils of its
    // particles, a few hundred cells for a Morton-ordered block -- staged in LDS, so the rows of
    // the walk read them there instead of from the cell table. F is null when the box is off for
    // the block (larger than kBoxCells, or TSP_TUNE_BOX_CELLS = 0).
    constexpr uint32_t kBoxCells = TSP_TUNE_BOX_CELLS;
    struct CellBox
    {
        const uint2* F;
        int x0, y0, z0;
        uint32_t w, wh;

        __device__ __forceinline__ uint32_t at(int x, int y, int z) const
        {
            return static_cast<uint32_t>(x - x0) + w * static_cast<uint32_t>(y - y0) + wh * static_cast<uint32_t>(z - z0);
        }
    };

    // block-wide (every thread calls it): the box of the active stencils, staged into lds
    // (kBoxCells entries) when it fits
    __device__ __forceinline__ CellBox stageCellBox(const Dev& d, const Frame& f, const Stencil& s, bool active,
                                                    uint2* lds, int* red)
    {
        CellBox b{};
        if (kBoxCells == 0) { return b; }
        if (threadIdx.x < 3) { red[threadIdx.x] = 0x7fffffff; red[threadIdx.x + 3] = -1; }
        __syncthreads();
        if (active)
        {
            for (int a = 0; a < 3; a++) { atomicMin(&red[a], s.lo[a]); atomicMax(&red[a + 3], s.hi[a]); }
        }
        __syncthreads();
        b.x0 = red[0]; b.y0 = red[1]; b.z0 = red[2];
        const int ex = red[3] - red[0] + 1, ey = red[4] - red[1] + 1, ez = red[5] - red[2] + 1;
        __syncthreads();
        if (ex <= 0) { return b; }   // no active particle
        const uint64_t vol = static_cast<uint64_t>(ex) * static_cast<uint64_t>(ey) * static_cast<uint64_t>(ez);
        if (vol > kBoxCells) { return b; }
        b.w = static_cast<uint32_t>(ex);
        b.wh = static_cast<uint32_t>(ex) * static_cast<uint32_t>(ey);
        for (uint32_t q = threadIdx.x; q < vol; q += blockDim.x)
        {
            const int x = b.x0 + static_cast<int>(q % b.w), y = b.y0 + static_cast<int>((q / b.w) % static_cast<uint32_t>(ey));
            const int z = b.z0 + static_cast<int>(q / b.wh);
            lds[q] = d.cellF[linearCell(f, x, y, z)];
        }
        __syncthreads();
        b.F = lds;
        return b;
    }

    __device__ __forceinline__ bool within(float xi, float yi, float zi, const float4& pj, float r2)
    {
        const float dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
        return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)) <= r2;
    }

    // reductions over the lanes of one SELL slice (with 32-lane slices the two halves of a wave64
    // are separate)
    __device__ __forceinline__ uint32_t sliceMax(uint32_t v)
    {
        for (int o = kSlice / 2; o > 0; o >>= 1) { v = max(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o, kSlice))); }
        return v;
    }

    __device__ __forceinline__ uint32_t sliceMin(uint32_t v)
    {
        for (int o = kSlice / 2; o > 0; o >>= 1) { v = min(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o, kSlice))); }
        return v;
    }

    __device__ __forceinline__ uint32_t sliceSum(uint32_t v)
    {
        for (int o = kSlice / 2; o > 0; o >>= 1) { v += static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o, kSlice)); }
        return v;
    }

    // the particle index of the q-th fluid candidate of a row stream (the three ranges in order)
    __device__ __forceinline__ uint32_t rowCandidate(const uint2 (&r)[3], uint32_t q)
    {
        const uint32_t n0 = r[0].y - r[0].x, n01 = n0 + (r[1].y - r[1].x);
        return q < n0 ? r[0].x + q : (q < n01 ? r[1].x + (q - n0) : r[2].x + (q - n01));
    }

    // The candidates of the stencil, per relation in list order: cells in z-y-x raster order,
    // ascending index inside a cell. onF(j, pos), onB(j, bpos), onR(j, rpos) see each candidate of
    // their relation once; the relations are separate lists, so only the order inside each one
    // matters. withF / withB / withR = false skip a relation's candidates (and its cell ranges);
    // needF(row) = false skips a row's fluid candidates.
#if TSP_TUNE_SEARCH_ROW
    // Row form: the ranges of up to three x-adjacent cells load together, and each relation's
    // candidates along the row form one stream whose loads issue TSP_TUNE_SEARCH_BATCH at a time
    // before their tests.
    template <bool kFlat = false, typename Index, typename Load, typename Fn>
    __device__ __forceinline__ void rowStream(const uint2 (&r)[3], Index index, Load load, Fn fn)
    {
        constexpr uint32_t kW = TSP_TUNE_SEARCH_BATCH;
#if TSP_TUNE_SEARCH_ROW == 2
        // cell by cell along the row, kW candidates at a time (a cell's candidates are contiguous)
        uint32_t q0 = 0;
#pragma unroll
        for (int c = 0; c < 3; c++)
        {
            const uint32_t b = r[c].x, n = r[c].y - r[c].x;
            for (uint32_t t = 0; t < n; t += kW)
            {
                uint32_t j[kW];
                float4 p[kW];
#pragma unroll
                for (uint32_t u = 0; u < kW; u++)
                {
                    j[u] = 0u;
                    p[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    if (t + u < n) { j[u] = index(b + t + u); p[u] = load(j[u]); }
                }
#pragma unroll
                for (uint32_t u = 0; u < kW; u++)
                {
                    if (t + u < n) { fn(j[u], p[u], q0 + t + u); }
                }
            }
            q0 += n;
        }
#else
        // the row's three ranges as one stream: candidate q is q + the offset of its range
        const uint32_t n0 = r[0].y - r[0].x, n1 = r[1].y - r[1].x, n2 = r[2].y - r[2].x;
        const uint32_t n01 = n0 + n1, total = n01 + n2;
        const uint32_t o0 = r[0].x, o1 = r[1].x - n0, o2 = r[2].x - n01;
        for (uint32_t t = 0; t < total; t += kW)
        {
            uint32_t j[kW];
            float4 p[kW];
#pragma unroll
            for (uint32_t u = 0; u < kW; u++)
            {
                if (kFlat)
                {
                    // past the stream's end the last candidate loads again (fn does not see it):
                    // the loads need no branch
                    const uint32_t q = min(t + u, total - 1u);
                    const uint32_t e = q + (q < n0 ? o0 : (q < n01 ? o1 : o2));
                    j[u] = index(e);
                    p[u] = load(j[u]);
                }
                else
                {
                    const uint32_t q = t + u;
                    const uint32_t e = q + (q < n0 ? o0 : (q < n01 ? o1 : o2));
                    j[u] = 0u;
                    p[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    if (q < total) { j[u] = index(e); p[u] = load(j[u]); }
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < kW; u++)
            {
                if (t + u < total) { fn(j[u], p[u], t + u); }
            }
        }
#endif
    }

    // preF(row, rf) may take a row's fluid candidates itself (returns true) instead of the stream;
    // postF(row) follows the row's fluid part. row counts the rows of three cells in stencil order.
    template <bool kFlat = false, typename OnF, typename OnB, typename OnR, typename NeedF, typename PreF, typename PostF>
    __device__ __forceinline__ void forStencil(const Dev& d, const Frame& f, const Stencil& s, const CellBox& box, bool withF,
                                               bool withB, bool withR, OnF onF, OnB onB, OnR onR, NeedF needF, PreF preF,
                                               PostF postF)
    {
        const auto same = [](uint32_t e) { return e; };
        uint32_t row = 0;
        for (int cz = s.lo[2]; cz <= s.hi[2]; cz++)
        {
            for (int cy = s.lo[1]; cy <= s.hi[1]; cy++)
            {
                for (int cx = s.lo[0]; cx <= s.hi[0]; cx += 3, row++)
                {
                    const int nx = min(3, s.hi[0] - cx + 1);
                    const uint32_t lin = s.cell(cx, cy, cz);
                    const bool rowF = withF && needF(row);
                    uint2 rf[3], rb[3], rr[3];
#if TSP_TUNE_RANGE_LOADS
                    // every lane loads the row's nine ranges (cells past the row's end load its
                    // last cell) and then selects: no predicated loads into zeroed registers, so
                    // nothing waits on loads still in flight from the previous row
#pragma unroll
                    for (int c = 0; c < 3; c++)
                    {
                        const bool in = c < nx;
                        const uint32_t cc = lin + static_cast<uint32_t>(min(c, nx - 1));
                        const uint2 vf = box.F ? box.F[box.at(cx, cy, cz) + static_cast<uint32_t>(min(c, nx - 1))] : d.cellF[cc];
                        const uint2 vb = d.cellB[cc], vr = d.cellR[cc];
                        rf[c] = (in && rowF) ? vf : make_uint2(0u, 0u);
                        rb[c] = (in && withB) ? vb : make_uint2(0u, 0u);
                        rr[c] = (in && withR && d.nR) ? vr : make_uint2(0u, 0u);
                    }
#else
#pragma unroll
                    for (int c = 0; c < 3; c++)
                    {
                        const bool in = c < nx;
                        rf[c] = (in && rowF) ? (box.F ? box.F[box.at(cx, cy, cz) + c] : d.cellF[lin + c]) : make_uint2(0u, 0u);
                        rb[c] = (in && withB) ? d.cellB[lin + c] : make_uint2(0u, 0u);
                        rr[c] = (in && withR && d.nR) ? d.cellR[lin + c] : make_uint2(0u, 0u);
                    }
#endif
                    if (rowF)
                    {
                        if (!preF(row, rf))
                        {
                            rowStream<kFlat>(rf, same, [&](uint32_t j) { return d.pos[j]; },
                                             [&](uint32_t j, const float4& p, uint32_t q) { onF(j, p, row, q); });
                        }
                        postF(row);
                    }
                    if (withB)
                    {
                        rowStream<kFlat>(rb, same, [&](uint32_t j) { return d.bpos[j]; },
                                         [&](uint32_t j, const float4& p, uint32_t) { onB(j, p); });
                    }
                    if (withR)
                    {
                        rowStream<kFlat>(rr, [&](uint32_t e) { return d.idxR[e]; }, [&](uint32_t j) { return d.rpos[j]; },
                                         [&](uint32_t j, const float4& p, uint32_t) { onR(j, p); });
                    }
                }
            }
        }
    }
#else
    // Cell form: one cell at a time, its candidates one at a time (no hit masks: every row is
    // tested, row numbers are not passed)
    template <bool kFlat = false, typename OnF, typename OnB, typename OnR, typename NeedF, typename PreF, typename PostF>
    __device__ __forceinline__ void forStencil(const Dev& d, const Frame& f, const Stencil& s, const CellBox&, bool withF,
                                               bool withB, bool withR, OnF onF, OnB onB, OnR onR, NeedF, PreF, PostF)
    {
        for (int cz = s.lo[2]; cz <= s.hi[2]; cz++)
        {
            for (int cy = s.lo[1]; cy <= s.hi[1]; cy++)
            {
                for (int cx = s.lo[0]; cx <= s.hi[0]; cx++)
                {
                    const uint32_t lin = s.cell(cx, cy, cz);
                    if (withF)
                    {
                        const uint2 rf = d.cellF[lin];
                        for (uint32_t j = rf.x; j < rf.y; j++) { onF(j, d.pos[j], kMaskRows, 0u); }
                    }
                    if (withB)
                    {
                        const uint2 rb = d.cellB[lin];
                        for (uint32_t j = rb.x; j < rb.y; j++) { onB(j, d.bpos[j]); }
                    }
                    if (withR && d.nR)
                    {
                        const uint2 rr = d.cellR[lin];
                        for (uint32_t q = rr.x; q < rr.y; q++)
                        {
                            const uint32_t j = d.idxR[q];
                            onR(j, d.rpos[j]);
                        }
                    }
                }
            }
        }
    }
#endif

    __global__ void __launch_bounds__(kBlock) TSP_WAVES_ATTR(TSP_TUNE_QUERY_WAVES) kQueryCount(Dev d, uint32_t nSlices)
    {
        const uint32_t blk = xcdBlock();
        const uint32_t i = blk * kBlock + threadIdx.x;
        const uint32_t sc = d.blockScene[blk];
        const Frame& f = d.frames[sc];
        const float r2 = d.consts[sc].r2;
        uint32_t cnt[3] = { 0, 0, 0 };
        bool tiny = false;
        uint32_t mn[3] = { kEmpty32, kEmpty32, kEmpty32 };
        uint32_t mx[3] = { 0, 0, 0 };
        Stencil s;
        // rows the masks do not cover (all of them unless the row form walks a reach-1 stencil)
        uint32_t uncovered = (1u << kMaskRows) - 1u, rowMask = 0;
        if (i < d.nF && stencilOf(f, keyAt(d, d.keyF, i), s))
        {
            const float4 pi = d.pos[i];
            const bool masks = TSP_TUNE_SEARCH_ROW == 1 && f.reach == 1;
            if (masks) { uncovered = 0; }
#if TSP_TUNE_SEARCH_FLAT
            // a hit updates the counts, spans, tiny flag and masks through selects, without a branch
            // (nearly every candidate step of a wave has a hit in some lane)
            const auto hitOf = [&](const float4& pj, float& d2)
            {
                d2 = dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
                return d2 <= r2;
            };
            forStencil<TSP_TUNE_SEARCH_FLAT == 1>(d, f, s, CellBox{}, true, true, true,
                [&](uint32_t j, const float4& pj, uint32_t row, uint32_t q)
                {
                    float d2;
                    const bool hit = hitOf(pj, d2);
                    cnt[0] += hit ? 1u : 0u;
                    mn[0] = min(mn[0], hit ? j : kEmpty32);
                    mx[0] = max(mx[0], hit ? j : 0u);
                    tiny = tiny || (hit && tinyDist2(d2));
                    const uint32_t bit = hit ? 1u : 0u;
                    rowMask |= q < 32u ? bit << q : 0u;
                    uncovered |= (q >= 32u && row < kMaskRows) ? bit << row : 0u;
                },
                [&](uint32_t j, const float4& pj)
                {
                    float d2;
                    const bool hit = hitOf(pj, d2);
                    cnt[1] += hit ? 1u : 0u;
                    mn[1] = min(mn[1], hit ? j : kEmpty32);
                    mx[1] = max(mx[1], hit ? j : 0u);
                    tiny = tiny || (hit && tinyDist2(d2));
                },
                [&](uint32_t j, const float4& pj)
                {
                    float d2;
                    const bool hit = hitOf(pj, d2);
                    cnt[2] += hit ? 1u : 0u;
                    mn[2] = min(mn[2], hit ? j : kEmpty32);
                    mx[2] = max(mx[2], hit ? j : 0u);
                },
#else
            forStencil(d, f, s, CellBox{}, true, true, true,
                [&](uint32_t j, const float4& pj, uint32_t row, uint32_t q)
                {
                    if (within(pi.x, pi.y, pi.z, pj, r2))
                    {
                        cnt[0]++; mn[0] = min(mn[0], j); mx[0] = max(mx[0], j);
                        tiny = tiny || tinyDist2(dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z));
                        if (q < 32u) { rowMask |= 1u << q; } else if (row < kMaskRows) { uncovered |= 1u << row; }
                    }
                },
                [&](uint32_t j, const float4& pj)
                {
                    if (within(pi.x, pi.y, pi.z, pj, r2))
                    {
                        cnt[1]++; mn[1] = min(mn[1], j); mx[1] = max(mx[1], j);
                        tiny = tiny || tinyDist2(dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z));
                    }
                },
                [&](uint32_t j, const float4& pj)
                {
                    if (within(pi.x, pi.y, pi.z, pj, r2)) { cnt[2]++; mn[2] = min(mn[2], j); mx[2] = max(mx[2], j); }
                },
#endif
                [&](uint32_t) { return true; },
                [&](uint32_t, const uint2 (&)[3]) { return false; },
                [&](uint32_t row)
                {
                    if (masks && row < kMaskRows) { d.ffMask[static_cast<size_t>(row) * d.nF + i] = rowMask; }
                    rowMask = 0;
                });
        }
        if (i < d.nF)
        {
            d.ffMask[static_cast<size_t>(kMaskRows) * d.nF + i] = uncovered;
            const bool wide = cnt[0] > 0 && mx[0] - mn[0] > 0xffffu;
            d.cnt[REL_FF][i] = static_cast<uint16_t>(cnt[0]);
            d.cnt[REL_FB][i] = static_cast<uint16_t>(cnt[1]);
            d.cnt[REL_FR][i] = static_cast<uint16_t>(cnt[2]);
            d.ffMin[i] = (cnt[0] > 0 ? mn[0] : 0u) | (wide ? kWideBit : 0u);
            // u16 counts, as the reference's (neighbor_lists.h:43)
            if (cnt[0] > 0xffffu || cnt[1] > 0xffffu || cnt[2] > 0xffffu) { atomicOr(&d.status->overflow, 2u); }
            if (tiny) { atomicOr(&d.status->pairTiny, 1u); }
        }
        // the particles whose FF list spans more than 0xffff join the list the wide fill walks
        // (one counter update per wave)
        {
            const bool wideHere = i < d.nF && cnt[0] > 0 && mx[0] - mn[0] > 0xffffu;
            const uint64_t ballot = __ballot(wideHere);
            if (ballot != 0ull)
            {
                const uint32_t lane64 = threadIdx.x & 63u;
                const int first = __ffsll(static_cast<unsigned long long>(ballot)) - 1;
                uint32_t base = 0;
                if (static_cast<int>(lane64) == first) { base = atomicAdd(&d.status->wideCount, static_cast<uint32_t>(__popcll(ballot))); }
                base = static_cast<uint32_t>(__shfl(static_cast<int>(base), first, 64));
                if (wideHere) { d.wideList[base + static_cast<uint32_t>(__popcll(ballot & ((1ull << lane64) - 1ull)))] = i; }
            }
        }
#if TSP_TUNE_LANE_SORT
        // the block's list positions, longest lists first (ties in particle order): a thread's
        // rank is the number of the block's sort keys above its own, the keys made distinct by
        // the thread index in their low byte. The slice statistics below then run over positions.
        {
            __shared__ uint32_t sortKey[kBlock];
            __shared__ uint32_t byPos[9][kBlock];
            const uint32_t t = threadIdx.x;
            const uint32_t nFF = min(cnt[0], 255u), nFB = min(cnt[1], 255u);
            const uint32_t key = TSP_TUNE_LANE_SORT == 2 ? (nFB << 8) | nFF : nFF;
            const uint32_t own = (key << 8) | (kBlock - 1u - t);
            sortKey[t] = own;
            __syncthreads();
            uint32_t rank = 0;
            for (uint32_t u = 0; u < kBlock; u += 4)
            {
                const uint4 v = *reinterpret_cast<const uint4*>(&sortKey[u]);
                rank += (v.x > own ? 1u : 0u) + (v.y > own ? 1u : 0u) + (v.z > own ? 1u : 0u) + (v.w > own ? 1u : 0u);
            }
            if (i < d.nF)
            {
                d.laneRank[i] = static_cast<uint8_t>(rank);
                d.rankWho[blk * kBlock + rank] = static_cast<uint8_t>(t);
            }
            for (int r = 0; r < 3; r++) { byPos[r][rank] = cnt[r]; byPos[3 + r][rank] = mn[r]; byPos[6 + r][rank] = mx[r]; }
            __syncthreads();
            for (int r = 0; r < 3; r++) { cnt[r] = byPos[r][t]; mn[r] = byPos[3 + r][t]; mx[r] = byPos[6 + r][t]; }
        }
#endif
        // per slice (list positions i of the block): pitch, base (min index), format, entries
        const uint32_t slice = i / kSlice;
        for (int r = 0; r < 3; r++)
        {
            const uint32_t pitch = sliceMax(cnt[r]);
            const uint32_t smin = sliceMin(mn[r]);
            const uint32_t smax = sliceMax(mx[r]);
            const uint32_t sum = sliceSum(cnt[r]);
            if ((threadIdx.x & (kSlice - 1)) == 0 && slice < nSlices)
            {
                SliceMeta m;
                m.pitch = static_cast<uint16_t>(pitch);
                const bool delta = pitch == 0 || smax - smin <= 0xffffu;
                m.format = delta ? kFmtDelta16 : kFmtAbs32;
                m.base = (pitch == 0 || !delta) ? 0u : smin;
                m.offset = pitch * kSlice * (delta ? 1u : 2u);
                m.entries = sum;
                d.meta[r][slice] = m;
            }
        }
    }

    // storage units of every slice into a scan buffer of nSlices + 1 entries per relation, and
    // their total in 64 bits (the u32 scan of the offsets wraps beyond 2^32 units; the host
    // refuses such a pool before any list is written)
    __global__ void __launch_bounds__(kBlock) kSliceUnits(Dev d, uint32_t nSlices, uint32_t* __restrict__ units)
    {
        __shared__ unsigned long long red[kBlock];
        const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
        const uint32_t rel = blockIdx.y;
        const uint32_t u = q < nSlices ? d.meta[rel][q].offset : 0u;
        if (q <= nSlices) { units[rel * (nSlices + 1) + q] = u; }
        red[threadIdx.x] = u;
        __syncthreads();
        for (uint32_t o = kBlock / 2; o > 0; o >>= 1)
        {
            if (threadIdx.x < o) { red[threadIdx.x] += red[threadIdx.x + o]; }
            __syncthreads();
        }
        if (threadIdx.x == 0 && red[0]) { atomicAdd(reinterpret_cast<unsigned long long*>(&d.status->poolUnits[rel]), red[0]); }
    }

    // writes the scanned offsets back, sums the entries and checks the pool capacity
    __global__ void __launch_bounds__(kBlock) kSliceApply(Dev d, uint32_t nSlices, const uint32_t* __restrict__ offs)
    {
        __shared__ uint32_t red[kBlock];
        const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
        const uint32_t rel = blockIdx.y;
        uint32_t e = 0;
        if (q < nSlices)
        {
            SliceMeta& m = d.meta[rel][q];
            e = m.entries;
            m.offset = offs[rel * (nSlices + 1) + q];
        }
        red[threadIdx.x] = e;
        __syncthreads();
        for (uint32_t o = kBlock / 2; o > 0; o >>= 1)
        {
            if (threadIdx.x < o) { red[threadIdx.x] += red[threadIdx.x + o]; }
            __syncthreads();
        }
        if (threadIdx.x == 0)
        {
            atomicAdd(reinterpret_cast<unsigned long long*>(&d.status->entries[rel]), static_cast<unsigned long long>(red[0]));
        }
        if (q == 0)
        {
            const uint64_t total = d.status->poolUnits[rel];   // summed in 64 bits by kSliceUnits
            if (total > d.poolCap[rel]) { atomicOr(&d.status->overflow, 1u); }
        }
    }

    constexpr uint32_t kStage = 40; // FF entries per thread staged in LDS by the fill (u16 slices)

    __global__ void __launch_bounds__(kBlock) TSP_WAVES_ATTR(TSP_TUNE_QUERY_WAVES) kQueryFill(Dev d)
    {
        if (stopped(d)) { return; }
        __shared__ uint16_t stage[kStage][kBlock];
        const uint32_t blk = xcdBlock();
        const uint32_t t = threadIdx.x;
        const uint32_t i = blk * kBlock + t;
        const uint32_t sc = d.blockScene[blk];
        const Frame& f = d.frames[sc];
        const float r2 = d.consts[sc].r2;
        // this particle's list position (its lane and slice), from the count pass; col is its
        // column of the staging array
        const uint32_t pos = i < d.nF ? sellPos(d, i) : i;
        const uint32_t col = pos & (kBlock - 1u);
        const uint32_t lane = pos & (kSlice - 1), slice = pos / kSlice;
        Stencil s;
        const bool active = i < d.nF && stencilOf(f, keyAt(d, d.keyF, i), s);
        __shared__ uint2 boxF[kBoxCells > 0 ? kBoxCells : 1];
        __shared__ int boxRed[6];
        const CellBox box = stageCellBox(d, f, s, active, boxF, boxRed);
        const uint32_t cF = i < d.nF ? d.cnt[REL_FF][i] : 0u;
        const uint32_t fMin = i < d.nF ? d.ffMin[i] : 0u;
        const bool wide = (fMin & kWideBit) != 0u;
        SliceMeta mF{}, mB{}, mR{};
        if (i < d.nF) { mF = d.meta[REL_FF][slice]; mB = d.meta[REL_FB][slice]; mR = d.meta[REL_FR][slice]; }
        // the FF entries stage in LDS as u16 offsets from mF.base: the slice base in a u16 slice,
        // the particle's smallest neighbor in a u32 slice, whose stores ignore the base (the list
        // spans at most 0xffff unless wide)
        if (mF.format != kFmtDelta16) { mF.base = fMin; }
        const bool staged = !wide;
        if (active)
        {
            const float4 pi = d.pos[i];
            uint32_t kF = 0, kB = 0, kR = 0;
            const uint32_t uncovered = d.ffMask[static_cast<size_t>(kMaskRows) * d.nF + i];
            // the row's hit mask, read before its cell ranges (and the next row's a row ahead): a
            // covered row without hits loads no range; relations without entries are not walked
            uint32_t rowMask = 0, nextMask = d.ffMask[i];
            const auto ffHit = [&](uint32_t j)
            {
                if (staged && kF < kStage) { stage[kF][col] = static_cast<uint16_t>(j - mF.base); }
                else { sellWrite(d.pool[REL_FF], mF, lane, kF, j); }
                kF++;
            };
            forStencil(d, f, s, box, !wide, d.cnt[REL_FB][i] != 0, d.cnt[REL_FR][i] != 0,
                [&](uint32_t j, const float4& pj, uint32_t, uint32_t)
                {
                    if (within(pi.x, pi.y, pi.z, pj, r2)) { ffHit(j); }
                },
                [&](uint32_t j, const float4& pj)
                {
                    if (within(pi.x, pi.y, pi.z, pj, r2)) { sellWrite(d.pool[REL_FB], mB, lane, kB++, j); }
                },
                [&](uint32_t j, const float4& pj)
                {
                    if (within(pi.x, pi.y, pi.z, pj, r2)) { sellWrite(d.pool[REL_FR], mR, lane, kR++, j); }
                },
                [&](uint32_t row)
                {
                    const uint32_t cur = nextMask;
#if TSP_TUNE_RANGE_LOADS
                    // unconditional (the last row's word again past the rows): the load stays in flight
                    nextMask = d.ffMask[static_cast<size_t>(min(row + 1u, kMaskRows - 1u)) * d.nF + i];
#else
                    if (row + 1u < kMaskRows) { nextMask = d.ffMask[static_cast<size_t>(row + 1u) * d.nF + i]; }
#endif
                    if (row >= kMaskRows || ((uncovered >> row) & 1u)) { return true; }
                    rowMask = cur;
                    return rowMask != 0u;
                },
                // a row the count pass covered: its hits, in stream order, without loading a position
                [&](uint32_t row, const uint2 (&rf)[3])
                {
                    if (row >= kMaskRows || ((uncovered >> row) & 1u)) { return false; }
                    uint32_t m = rowMask;
                    while (m)
                    {
                        const uint32_t q = static_cast<uint32_t>(__ffs(static_cast<int>(m))) - 1u;
                        m &= m - 1u;
                        ffHit(rowCandidate(rf, q));
                    }
                    return true;
                },
                [&](uint32_t) {});
            // (the FF lists of wide particles: kQueryFillWide)
        }
        // the staged FF entries leave as whole slice rows: column k of the slice's lanes per store
#if TSP_TUNE_LANE_SORT
        // (thread t writes list position t of the block: the staged entries of the particle there,
        // whose count and base it reads back; a particle outside the frame has no FF entries)
        __syncthreads();
        const uint32_t p = blk * kBlock + t;
        const uint32_t w = particleAt(d, p);
        const uint32_t wMin = d.ffMin[w];
        const uint32_t own = (wMin & kWideBit) ? 0u : min(static_cast<uint32_t>(d.cnt[REL_FF][w]), kStage);
        SliceMeta mw{};
        if (own) { mw = d.meta[REL_FF][p / kSlice]; }
        const uint32_t base = mw.format == kFmtDelta16 ? mw.base : wMin;
        const uint32_t lw = p & (kSlice - 1);
#else
        __syncthreads();
        const uint32_t own = (active && staged) ? min(cF, kStage) : 0u, base = mF.base;
        const SliceMeta mw = mF;
        const uint32_t lw = lane;
#endif
        const uint32_t rows = sliceMax(own);
        if (mw.format == kFmtDelta16)
        {
            uint16_t* row = d.pool[REL_FF] + mw.offset + lw;
            for (uint32_t k = 0; k < rows; k++)
            {
                if (k < own) { row[k * kSlice] = stage[k][t]; }
            }
        }
        else
        {
            uint32_t* row = reinterpret_cast<uint32_t*>(d.pool[REL_FF] + mw.offset) + lw;
            for (uint32_t k = 0; k < rows; k++)
            {
                if (k < own) { row[k * kSlice] = base + stage[k][t]; }
            }
        }
    }

    // The FF lists of the particles whose list spans more than 0xffff (the count pass lists them):
    // the DictDelta16x2 split of a per-particle dict (neighbor_lists.cpp:838-858) -- the neighbors
    // within 0xffff of the smallest form the low group, the rest the high group, which is stored
    // from the back when it fits 0xffff itself, else the candidate order stays; a list of more than
    // kDictMax entries is no dict (candidate order). One thread per listed particle walks its
    // stencil with the count pass's hit masks, twice for a dict: the high group's smallest and the
    // largest neighbor, then the entries. Such a particle's slice stores u32 entries.
    __global__ void __launch_bounds__(kBlock) TSP_WAVES_ATTR(TSP_TUNE_QUERY_WAVES) kQueryFillWide(Dev d)
    {
        if (stopped(d)) { return; }
        const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
        if (q >= d.status->wideCount) { return; }
        const uint32_t i = d.wideList[q];
        const uint32_t sc = d.blockScene[i / kBlock];
        const Frame& f = d.frames[sc];
        const float r2 = d.consts[sc].r2;
        Stencil s;
        if (!stencilOf(f, keyAt(d, d.keyF, i), s)) { return; }
        const uint32_t pos = sellPos(d, i);
        const uint32_t lane = pos & (kSlice - 1);
        const SliceMeta mF = d.meta[REL_FF][pos / kSlice];
        const float4 pi = d.pos[i];
        const uint32_t n = d.cnt[REL_FF][i];
        const uint32_t minRef = d.ffMin[i] & ~kWideBit;
        const uint32_t uncovered = d.ffMask[static_cast<size_t>(kMaskRows) * d.nF + i];
        const auto walk = [&](auto&& onHit)
        {
            uint32_t rowMask = 0, nextMask = d.ffMask[i];
            forStencil(d, f, s, CellBox{}, true, false, false,
                [&](uint32_t j, const float4& pj, uint32_t, uint32_t)
                {
                    if (within(pi.x, pi.y, pi.z, pj, r2)) { onHit(j); }
                },
                [&](uint32_t, const float4&) {},
                [&](uint32_t, const float4&) {},
                [&](uint32_t row)
                {
                    const uint32_t cur = nextMask;
                    if (row + 1u < kMaskRows) { nextMask = d.ffMask[static_cast<size_t>(row + 1u) * d.nF + i]; }
                    if (row >= kMaskRows || ((uncovered >> row) & 1u)) { return true; }
                    rowMask = cur;
                    return rowMask != 0u;
                },
                [&](uint32_t row, const uint2 (&rf)[3])
                {
                    if (row >= kMaskRows || ((uncovered >> row) & 1u)) { return false; }
                    uint32_t m = rowMask;
                    while (m)
                    {
                        const uint32_t b = static_cast<uint32_t>(__ffs(static_cast<int>(m))) - 1u;
                        m &= m - 1u;
                        onHit(rowCandidate(rf, b));
                    }
                    return true;
                },
                [&](uint32_t) {});
        };
        const bool dict = n <= kDictMax;
        uint32_t base2 = kEmpty32, maxRef = 0;
        if (dict)
        {
            walk([&](uint32_t j)
            {
                maxRef = max(maxRef, j);
                if (j - minRef > 0xffffu) { base2 = min(base2, j); }
            });
        }
        const bool split = dict && maxRef - base2 <= 0xffffu;
        uint32_t kLow = 0, kHigh = n;
        walk([&](uint32_t j)
        {
            if (!split || j - minRef <= 0xffffu) { sellWrite(d.pool[REL_FF], mF, lane, kLow++, j); }
            else { sellWrite(d.pool[REL_FF], mF, lane, --kHigh, j); }
        });
    }

    // ----------------------------------------------------------------------------------------
    // FF halo of each 256-particle block: the distinct neighbors of the block's particles,
    // ascending, and the block's FF entries rewritten as slots of that list. One block per
    // particle block; a halo larger than kHaloMax leaves the block's entries as particle indices.

    constexpr uint32_t kHashSlots = 2048;   // > kHaloMax + kBlock: probes always end

    // With TSP_TUNE_SLOT16 the halo slots of a u32 slice are stored as u16 in the first half of
    // its storage (entry k of lane l at u16 position k * kSlice + l, which the lanes reach only
    // after reading the u32 entries there), and the slice turns u16 once its entries are written.
    __device__ __forceinline__ void slotsToU16(const Dev& d, uint32_t i, const SliceMeta& m)
    {
#if TSP_TUNE_SLOT16
        if (i < d.nF && (i & (kSlice - 1)) == 0 && m.format != kFmtDelta16)
        {
            SliceMeta& w = d.meta[REL_FF][i / kSlice];
            w.format = kFmtDelta16;
            w.base = 0;
        }
#else
        (void)d; (void)i; (void)m;
#endif
    }
    constexpr uint32_t kSortKeys = kHaloMax <= 1024 ? 1024 : 2048;    // power of two >= kHaloMax

    __device__ __forceinline__ uint32_t haloHash(uint32_t j) { return (j * 2654435761u) >> (32 - 11); }

    // slot of j in the hash table (j must be present)
    __device__ __forceinline__ uint32_t haloFind(const uint32_t* table, uint32_t j)
    {
        uint32_t h = haloHash(j);
        for (uint32_t probe = 0; probe < kHashSlots && table[h] != j; probe++) { h = (h + 1) & (kHashSlots - 1); }
        return h;
    }

#if TSP_TUNE_HALO_PAGES
#error "TSP_TUNE_HALO_PAGES=1 stopped responding in a timing run after the lists were ordered by length within each block; it stays disabled until the cause is found"
    // The block's FF neighbors by pages of 4096 indices: one LDS bitmap per page (at most
    // kHaloPages pages, taken in order of arrival), the pages ranked by their number, so the set
    // bits in page order are the halo in ascending order and an entry's slot is the count of set
    // bits before it. No hash table and no sort; a block whose neighbors span more pages keeps
    // particle indices.
    constexpr uint32_t kPageBits = 12;
    constexpr uint32_t kPageWords = (1u << kPageBits) / 32;
    constexpr uint32_t kHaloPages = 32;
    static_assert(kHaloPages <= kBlock && (kHaloPages * kPageWords) % kBlock == 0, "page table sizes");

    // the slot of page pg in the table (pages arrive in any order; slots fill from 0), or
    // kHaloPages when the table is full
    __device__ __forceinline__ uint32_t haloPageSlot(uint32_t* pageId, uint32_t pg)
    {
        for (uint32_t s = 0; s < kHaloPages; s++)
        {
            uint32_t v = *static_cast<volatile uint32_t*>(&pageId[s]);
            if (v == kEmpty32) { v = atomicCAS(&pageId[s], kEmpty32, pg); if (v == kEmpty32) { return s; } }
            if (v == pg) { return s; }
        }
        return kHaloPages;
    }

    __global__ void __launch_bounds__(kBlock) TSP_WAVES_ATTR(TSP_TUNE_QUERY_WAVES) kHaloBuild(Dev d)
    {
        if (stopped(d)) { return; }
        __shared__ uint32_t bits[kHaloPages * kPageWords];
        __shared__ uint16_t wrank[kHaloPages * kPageWords];
        __shared__ uint32_t pageId[kHaloPages], slotOfRank[kHaloPages];
        __shared__ uint32_t full;
        __shared__ uint32_t waveSum[kBlock / 64];
        const uint32_t blk = blockIdx.x;
        const uint32_t t = threadIdx.x;
        const uint32_t i = blk * kBlock + t;
        for (uint32_t q = t; q < kHaloPages * kPageWords; q += kBlock) { bits[q] = 0u; }
        if (t < kHaloPages) { pageId[t] = kEmpty32; }
        if (t == 0) { full = 0u; }
        __syncthreads();
        uint32_t n = 0;
        SliceMeta m{};
        if (i < d.nF)
        {
            // thread t walks the lists at list position i
            n = d.cnt[REL_FF][particleAt(d, i)];
            m = d.meta[REL_FF][i / kSlice];
        }
        const uint32_t lane = i & (kSlice - 1);
        // the entries into their pages' bitmaps (consecutive entries mostly share a page)
        uint32_t lastPage = kEmpty32, lastSlot = 0;
        for (uint32_t k = 0; k < n; k++)
        {
            const uint32_t j = sellRead(d.pool[REL_FF], m, lane, k);
            const uint32_t pg = j >> kPageBits;
            if (pg != lastPage)
            {
                lastSlot = haloPageSlot(pageId, pg);
                if (lastSlot == kHaloPages) { atomicOr(&full, 1u); break; }
                lastPage = pg;
            }
            atomicOr(&bits[lastSlot * kPageWords + ((j >> 5) & (kPageWords - 1u))], 1u << (j & 31u));
        }
        __syncthreads();
        if (full)
        {
            if (t == 0) { d.haloCount[blk] = kNoHalo; }
            return;
        }
        // the pages in ascending order: slotOfRank[r] is the slot of the r-th smallest page
        if (t < kHaloPages && pageId[t] != kEmpty32)
        {
            const uint32_t pg = pageId[t];
            uint32_t r = 0;
            for (uint32_t s = 0; s < kHaloPages; s++) { r += (pageId[s] != kEmpty32 && pageId[s] < pg) ? 1u : 0u; }
            slotOfRank[r] = t;
        }
        __syncthreads();
        uint32_t np = 0;
        for (uint32_t s = 0; s < kHaloPages; s++) { np += pageId[s] != kEmpty32 ? 1u : 0u; }
        // set bits before each word in page order: thread t takes kPer consecutive words of that order
        constexpr uint32_t kPer = kHaloPages * kPageWords / kBlock;
        uint32_t pc[kPer], sum = 0;
#pragma unroll
        for (uint32_t u = 0; u < kPer; u++)
        {
            const uint32_t v = kPer * t + u, r = v / kPageWords;
            pc[u] = r < np ? static_cast<uint32_t>(__popc(bits[slotOfRank[r] * kPageWords + v % kPageWords])) : 0u;
            sum += pc[u];
        }
        uint32_t incl = sum;
        for (int o = 1; o < 64; o <<= 1)
        {
            const uint32_t v = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), o, 64));
            if ((t & 63u) >= static_cast<uint32_t>(o)) { incl += v; }
        }
        if ((t & 63u) == 63u) { waveSum[t >> 6] = incl; }
        __syncthreads();
        uint32_t count = 0, before = 0;
        for (uint32_t w = 0; w < kBlock / 64; w++) { before += w < (t >> 6) ? waveSum[w] : 0u; count += waveSum[w]; }
        if (count > kHaloMax)
        {
            if (t == 0) { d.haloCount[blk] = kNoHalo; }
            return;
        }
        uint32_t* out = d.haloIdx + static_cast<size_t>(blk) * kHaloMax;
        uint32_t r = before + incl - sum;
#pragma unroll
        for (uint32_t u = 0; u < kPer; u++)
        {
            const uint32_t v = kPer * t + u, pr = v / kPageWords;
            if (pr >= np) { break; }
            const uint32_t slot = slotOfRank[pr], w = v % kPageWords;
            wrank[slot * kPageWords + w] = static_cast<uint16_t>(r);
            uint32_t b = bits[slot * kPageWords + w];
            const uint32_t base = (pageId[slot] << kPageBits) + 32u * w;
            while (b)
            {
                out[r++] = base + static_cast<uint32_t>(__ffs(static_cast<int>(b))) - 1u;
                b &= b - 1u;
            }
        }
        if (t == 0) { d.haloCount[blk] = count; }
        __syncthreads();
        // entries become the byte offsets of their slots, in their storage format
        lastPage = kEmpty32;
        for (uint32_t k = 0; k < n; k++)
        {
            const uint32_t j = sellRead(d.pool[REL_FF], m, lane, k);
            const uint32_t pg = j >> kPageBits;
            if (pg != lastPage)
            {
                lastSlot = 0;
                while (pageId[lastSlot] != pg) { lastSlot++; }
                lastPage = pg;
            }
            const uint32_t q = lastSlot * kPageWords + ((j >> 5) & (kPageWords - 1u));
            const uint32_t slot = (wrank[q] + static_cast<uint32_t>(__popc(bits[q] & ((1u << (j & 31u)) - 1u)))) << kSlotShift;
            if (m.format == kFmtDelta16 || TSP_TUNE_SLOT16) { d.pool[REL_FF][m.offset + k * kSlice + lane] = static_cast<uint16_t>(slot); }
            else { reinterpret_cast<uint32_t*>(d.pool[REL_FF] + m.offset)[k * kSlice + lane] = slot; }
        }
        slotsToU16(d, i, m);
    }
#else
    __global__ void __launch_bounds__(kBlock) TSP_WAVES_ATTR(TSP_TUNE_QUERY_WAVES) kHaloBuild(Dev d)
    {
        if (stopped(d)) { return; }
        __shared__ uint32_t table[kHashSlots];
        __shared__ uint16_t rank[kHashSlots];
        __shared__ uint32_t keys[kSortKeys];
        __shared__ uint32_t unique, filled, bmin, notDelta, outOfSpan;
        __shared__ uint32_t waveSum[kBlock / 64];
        const uint32_t blk = blockIdx.x;
        const uint32_t t = threadIdx.x;
        const uint32_t i = blk * kBlock + t;
        if (t == 0) { unique = 0; filled = 0; bmin = kEmpty32; notDelta = 0; outOfSpan = 0; }
        __syncthreads();
        uint32_t n = 0;
        SliceMeta m{};
        if (i < d.nF)
        {
            // thread t walks the lists at list position i
            n = d.cnt[REL_FF][particleAt(d, i)];
            m = d.meta[REL_FF][i / kSlice];
        }
        const uint32_t lane = i & (kSlice - 1);
        uint32_t* out = d.haloIdx + static_cast<size_t>(blk) * kHaloMax;
        // When every FF entry of the block lies within 0xffff of the block's smallest (u16 delta
        // slices whose bases are close), the halo is a bitmap over that window: the set bits in
        // ascending order are the halo, and an entry's slot is the count of set bits below it.
        if (n > 0)
        {
            if (m.format != kFmtDelta16) { atomicOr(&notDelta, 1u); }
            else { atomicMin(&bmin, m.base); }
        }
        __syncthreads();
        if (!notDelta)
        {
            constexpr uint32_t kWords = kHashSlots;   // 2048 x 32 bits: a window of 0x10000 indices
            for (uint32_t q = t; q < kWords; q += kBlock) { table[q] = 0u; }
            __syncthreads();
            const uint32_t b0 = bmin;
            for (uint32_t k = 0; k < n; k++)
            {
                const uint32_t off = sellRead(d.pool[REL_FF], m, lane, k) - b0;
                if (off > 0xffffu) { atomicOr(&outOfSpan, 1u); break; }
                atomicOr(&table[off >> 5], 1u << (off & 31u));
            }
            __syncthreads();
            if (!outOfSpan)
            {
                // ranks of the words (eight consecutive words per thread), as u16 in rank[]
                constexpr uint32_t kPer = kWords / kBlock;
                uint32_t pc[kPer], sum = 0;
#pragma unroll
                for (uint32_t u = 0; u < kPer; u++) { pc[u] = static_cast<uint32_t>(__popc(table[kPer * t + u])); sum += pc[u]; }
                uint32_t incl = sum;
                for (int o = 1; o < 64; o <<= 1)
                {
                    const uint32_t v = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), o, 64));
                    if ((t & 63u) >= static_cast<uint32_t>(o)) { incl += v; }
                }
                if ((t & 63u) == 63u) { waveSum[t >> 6] = incl; }
                __syncthreads();
                uint32_t count = 0, before = 0;
                for (uint32_t w = 0; w < kBlock / 64; w++) { before += w < (t >> 6) ? waveSum[w] : 0u; count += waveSum[w]; }
                if (count > kHaloMax)
                {
                    if (t == 0) { d.haloCount[blk] = kNoHalo; }
                    return;
                }
                uint32_t r = before + incl - sum;
#pragma unroll
                for (uint32_t u = 0; u < kPer; u++)
                {
                    const uint32_t w = kPer * t + u;
                    rank[w] = static_cast<uint16_t>(r);
                    uint32_t bits = table[w];
                    while (bits)
                    {
                        const uint32_t b = static_cast<uint32_t>(__ffs(static_cast<int>(bits))) - 1u;
                        out[r++] = b0 + 32u * w + b;
                        bits &= bits - 1u;
                    }
                }
                if (t == 0) { d.haloCount[blk] = count; }
                __syncthreads();
                for (uint32_t k = 0; k < n; k++)
                {
                    const uint32_t off = sellRead(d.pool[REL_FF], m, lane, k) - b0;
                    const uint32_t w = off >> 5;
                    const uint32_t slot = rank[w] + static_cast<uint32_t>(__popc(table[w] & ((1u << (off & 31u)) - 1u)));
                    d.pool[REL_FF][m.offset + k * kSlice + lane] = static_cast<uint16_t>(slot << kSlotShift);
                }
                return;
            }
        }
        // otherwise a hash set of the block's entries, sorted
        for (uint32_t q = t; q < kHashSlots; q += kBlock) { table[q] = kEmpty32; }
        __syncthreads();
        // insert: open addressing, a relaxed LDS read before the compare-and-swap (most entries
        // are already present); a thread stops once the halo is known to be too large, so at most
        // kHaloMax + kBlock values enter the kHashSlots slots and a probe always ends
        for (uint32_t k = 0; k < n; k++)
        {
            if (__hip_atomic_load(&unique, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) > kHaloMax) { break; }
            const uint32_t j = sellRead(d.pool[REL_FF], m, lane, k);
            uint32_t h = haloHash(j);
            for (uint32_t probe = 0; probe < kHashSlots; probe++)
            {
                uint32_t v = __hip_atomic_load(&table[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (v == kEmpty32)
                {
                    v = atomicCAS(&table[h], kEmpty32, j);
                    if (v == kEmpty32) { atomicAdd(&unique, 1u); break; }
                }
                if (v == j) { break; }
                h = (h + 1) & (kHashSlots - 1);
            }
        }
        __syncthreads();
        static_assert(kHaloMax + kBlock < kHashSlots && kHaloMax <= kSortKeys, "halo table sizes");
        const uint32_t count = unique;
        if (count > kHaloMax)
        {
            if (t == 0) { d.haloCount[blk] = kNoHalo; }
            return;
        }
        // compact, pad, sort ascending (bitonic over the power of two at or above the count)
        const uint32_t nk = 1u << (32 - __clz(static_cast<int>(max(count, 2u) - 1u)));
        for (uint32_t q = t; q < nk; q += kBlock) { keys[q] = kEmpty32; }
        __syncthreads();
        for (uint32_t q = t; q < kHashSlots; q += kBlock)
        {
            const uint32_t v = table[q];
            if (v != kEmpty32) { keys[atomicAdd(&filled, 1u)] = v; }
        }
        __syncthreads();
        for (uint32_t size = 2; size <= nk; size <<= 1)
        {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1)
            {
                for (uint32_t q = t; q < nk; q += kBlock)
                {
                    const uint32_t r = q ^ stride;
                    if (r > q)
                    {
                        const uint32_t a = keys[q], b = keys[r];
                        const bool up = (q & size) == 0;
                        if ((a > b) == up) { keys[q] = b; keys[r] = a; }
                    }
                }
                __syncthreads();
            }
        }
        for (uint32_t q = t; q < count; q += kBlock)
        {
            out[q] = keys[q];
            rank[haloFind(table, keys[q])] = static_cast<uint16_t>(q);
        }
        if (t == 0) { d.haloCount[blk] = count; }
        __syncthreads();
        // entries become their rank in the sorted halo, in their storage format
        for (uint32_t k = 0; k < n; k++)
        {
            const uint32_t j = sellRead(d.pool[REL_FF], m, lane, k);
            const uint32_t slot = rank[haloFind(table, j)] << kSlotShift;
            if (m.format == kFmtDelta16 || TSP_TUNE_SLOT16) { d.pool[REL_FF][m.offset + k * kSlice + lane] = static_cast<uint16_t>(slot); }
            else { reinterpret_cast<uint32_t*>(d.pool[REL_FF] + m.offset)[k * kSlice + lane] = slot; }
        }
        slotsToU16(d, i, m);
    }
#endif

    // ----------------------------------------------------------------------------------------
    // rigid-source lists, one wave per rigid particle: mode 0 counts RF, RB and RR; mode 1 fills
    // the RF CSR; mode 2 fills RF and the RB and RR lists of the contact solver. The candidates of
    // a relation -- cell by cell in z-y-x order, by index inside a cell -- are tested 64 at a
    // time, and ballots keep that order in the lists.

    __global__ void __launch_bounds__(kBlock) kQueryRigid(Dev d, int mode)
    {
        if (stopped(d)) { return; }
        const uint32_t i = (blockIdx.x * kBlock + threadIdx.x) >> 6;
        const uint32_t lane = threadIdx.x & 63u;
        if (i >= d.nR) { return; }
        const bool fill = mode != 0;
        const bool contact = mode == 2;
        const uint32_t sc = d.bodyScene[d.rbody[i]];
        const Frame& f = d.frames[sc];
        const float r2 = d.consts[sc].r2;
        const int reach = f.reach;
        const float4 pi = d.rpos[i];
        const int cx = floorToInt(pi.x * f.cellSizeInv) - f.orig[0];
        const int cy = floorToInt(pi.y * f.cellSizeInv) - f.orig[1];
        const int cz = floorToInt(pi.z * f.cellSizeInv) - f.orig[2];
        uint32_t cnt[3] = { 0, 0, 0 };
        uint32_t w[3] = { fill ? d.rfOff[i] : 0u, contact ? d.rbOff[i] : 0u, contact ? d.rrOff[i] : 0u };
        const bool inside = cx >= 0 && cx < static_cast<int>(f.dim[0]) && cy >= 0 && cy < static_cast<int>(f.dim[1])
                         && cz >= 0 && cz < static_cast<int>(f.dim[2]);
        if (inside)
        {
            const int lo[3] = { max(cx - reach, 0), max(cy - reach, 0), max(cz - reach, 0) };
            const int hi[3] = { min(cx + reach, static_cast<int>(f.dim[0]) - 1), min(cy + reach, static_cast<int>(f.dim[1]) - 1),
                                min(cz + reach, static_cast<int>(f.dim[2]) - 1) };
            const uint32_t nx = static_cast<uint32_t>(hi[0] - lo[0] + 1), ny = static_cast<uint32_t>(hi[1] - lo[1] + 1);
            const uint32_t nCells = nx * ny * static_cast<uint32_t>(hi[2] - lo[2] + 1);
            const uint32_t body = d.rbody[i];
            const uint64_t below = (1ull << lane) - 1ull;
            const int nRel = (fill && !contact) ? 1 : 3;
            for (uint32_t c0 = 0; c0 < nCells; c0 += 64)
            {
                // this lane's cell of the batch and its three ranges
                const uint32_t c = c0 + lane;
                uint2 rng[3] = { make_uint2(0u, 0u), make_uint2(0u, 0u), make_uint2(0u, 0u) };
                if (c < nCells)
                {
                    const uint32_t lin = linearCell(f, lo[0] + static_cast<int>(c % nx), lo[1] + static_cast<int>((c / nx) % ny),
                                                    lo[2] + static_cast<int>(c / (nx * ny)));
                    rng[0] = d.cellF[lin];
                    rng[1] = d.cellB[lin];
                    if (d.nR) { rng[2] = d.cellR[lin]; }
                }
                for (int rel = 0; rel < nRel; rel++)
                {
                    // inclusive prefix of the cells' candidate counts over the lanes
                    const uint32_t size = rng[rel].y - rng[rel].x;
                    uint32_t incl = size;
                    for (int o = 1; o < 64; o <<= 1)
                    {
                        const uint32_t v = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), o, 64));
                        if (lane >= static_cast<uint32_t>(o)) { incl += v; }
                    }
                    const uint32_t total = static_cast<uint32_t>(__shfl(static_cast<int>(incl), 63, 64));
                    for (uint32_t f0 = 0; f0 < total; f0 += 64)
                    {
                        const uint32_t fi = f0 + lane;
                        bool hit = false;
                        uint32_t j = 0;
                        // the lane (cell) whose prefix range holds fi: the first with incl > fi
                        // (six halvings of [0, 63], taken by every lane together)
                        uint32_t a = 0, b = 63;
                        for (int step = 0; step < 6; step++)
                        {
                            const uint32_t m = (a + b) >> 1;
                            const uint32_t im = static_cast<uint32_t>(__shfl(static_cast<int>(incl), static_cast<int>(m), 64));
                            if (im > fi) { b = m; } else { a = m + 1; }
                        }
                        const uint32_t inclA = static_cast<uint32_t>(__shfl(static_cast<int>(incl), static_cast<int>(a), 64));
                        const uint32_t sizeA = static_cast<uint32_t>(__shfl(static_cast<int>(size), static_cast<int>(a), 64));
                        const uint32_t startA = static_cast<uint32_t>(__shfl(static_cast<int>(rng[rel].x), static_cast<int>(a), 64));
                        if (fi < total)
                        {
                            const uint32_t t = startA + (fi - (inclA - sizeA));
                            if (rel == 0) { j = t; hit = within(pi.x, pi.y, pi.z, d.pos[j], r2); }
                            else if (rel == 1) { j = t; hit = within(pi.x, pi.y, pi.z, d.bpos[j], r2); }
                            else
                            {
                                j = d.idxR[t];
                                hit = j != i && d.rbody[j] != body && within(pi.x, pi.y, pi.z, d.rpos[j], r2);
                            }
                        }
                        const uint64_t bal = __ballot(hit);
                        const bool store = rel == 0 ? fill : contact;
                        if (store && hit)
                        {
                            uint32_t* dst = rel == 0 ? d.rfIdx : (rel == 1 ? d.rbIdx : d.rrIdx);
                            dst[w[rel] + static_cast<uint32_t>(__popcll(bal & below))] = j;
                        }
                        const uint32_t n = static_cast<uint32_t>(__popcll(bal));
                        w[rel] += n;
                        cnt[rel] += n;
                    }
                }
            }
        }
        if (!fill && lane == 0)
        {
            if (i == 0) { d.rfOff[d.nR] = 0u; }
            d.rfCnt[i] = cnt[0];
            d.rfOff[i] = cnt[0];
            d.rbCnt[i] = cnt[1];
            d.rrCnt[i] = cnt[2];
        }
    }

    // totals of the rigid-source relations and the RF capacity check (after the RF offset scan);
    // a scene whose rigid particles have RB or RR entries runs the contact solve
    // (BuildRigidBoundaryNeighbors, rigid_solver.cpp:90-149)
    __global__ void __launch_bounds__(kBlock) kRigidTotals(Dev d)
    {
        __shared__ uint32_t red[2][kBlock];
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i < d.nR && (d.rbCnt[i] | d.rrCnt[i]) != 0u)
        {
            d.solve[d.bodyScene[d.rbody[i]]].contactActive = 1u;
        }
        red[0][threadIdx.x] = i < d.nR ? d.rbCnt[i] : 0u;
        red[1][threadIdx.x] = i < d.nR ? d.rrCnt[i] : 0u;
        __syncthreads();
        for (uint32_t o = kBlock / 2; o > 0; o >>= 1)
        {
            if (threadIdx.x < o) { red[0][threadIdx.x] += red[0][threadIdx.x + o]; red[1][threadIdx.x] += red[1][threadIdx.x + o]; }
            __syncthreads();
        }
        if (threadIdx.x == 0)
        {
            atomicAdd(reinterpret_cast<unsigned long long*>(&d.status->rbTotal), static_cast<unsigned long long>(red[0][0]));
            atomicAdd(reinterpret_cast<unsigned long long*>(&d.status->rrTotal), static_cast<unsigned long long>(red[1][0]));
        }
        if (i == 0)
        {
            d.status->rfTotal = d.rfOff[d.nR];
            if (d.rfOff[d.nR] > d.rfCap) { atomicOr(&d.status->overflow, 4u); }
        }
    }

    // the decision point of a captured step, after the search's counts and scans: what the host
    // would otherwise read to grow storage, rebuild a frame or take the contact solve calls off
    // the rest of the captured step (status->stop bit 1); the host then runs the step anew
    __global__ void kStepGate(Dev d)
    {
        if (threadIdx.x != 0) { return; }
        const Status& st = *d.status;
        const bool refuse = st.abort != 0u || st.overflow != 0u || st.rbTotal != 0ull || st.rrTotal != 0ull
                            || (st.pairTiny != 0u && d.fastPairs != 0u);
        if (refuse) { d.status->stop |= 1u; }
    }

    // after the captured sweeps: a scene with fluid that has not converged calls off the captured
    // tail (bit 2); the host adds sweeps and runs the tail itself
    __global__ void __launch_bounds__(kBlock) kSolveGate(Dev d)
    {
        if (stopped(d)) { return; }
        bool open = false;
        for (uint32_t sc = threadIdx.x; sc < d.nScenes; sc += kBlock)
        {
            open = open || (d.scenes[sc].nF != 0u && d.solve[sc].converged == 0u);
        }
        if (open) { atomicOr(&d.status->stop, 2u); }
    }

    __global__ void kClearStop(Dev d)
    {
        if (threadIdx.x == 0) { d.status->stop = 0u; }
    }

    // scene AABB of the positions of one particle class (ComputeSceneAABB_ParticleClass,
    // neighbor_search.cpp:227-258); min/max are exact in any order
    __device__ __forceinline__ uint32_t orderedBits(float v)
    {
        const uint32_t u = __float_as_uint(v);
        return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    }

    // per-scene bounds of one population: out[6 s + a] holds the ordered bits of the min x, y, z
    // and then the max; a wave whose valid lanes share one scene reduces before its atomics
    // per-scene bounds: each wave walks a contiguous range of 64-particle chunks, keeps per-lane
    // running bounds while the chunks belong to one scene, and reduces them into that scene's
    // bounds with one set of atomics when the scene changes or the range ends (chunks that mix
    // scenes update lane by lane)
    __global__ void __launch_bounds__(kBlock) kAabb(Dev d, int pop, const float4* __restrict__ p, uint32_t n,
        uint32_t* __restrict__ out)
    {
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t nWaves = gridDim.x * (kBlock / 64);
        const uint32_t gw = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
        const uint32_t nChunks = (n + 63u) / 64u;
        const uint32_t c0 = static_cast<uint32_t>(static_cast<uint64_t>(nChunks) * gw / nWaves);
        const uint32_t c1 = static_cast<uint32_t>(static_cast<uint64_t>(nChunks) * (gw + 1) / nWaves);
        uint32_t rs = 0xffffffffu;
        uint32_t v[6] = { 0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u };
        const auto flush = [&]()
        {
            for (int o = 32; o > 0; o >>= 1)
            {
                for (int a = 0; a < 3; a++)
                {
                    v[a] = min(v[a], static_cast<uint32_t>(__shfl_xor(static_cast<int>(v[a]), o, 64)));
                    v[3 + a] = max(v[3 + a], static_cast<uint32_t>(__shfl_xor(static_cast<int>(v[3 + a]), o, 64)));
                }
            }
            if (lane == 0 && rs != 0xffffffffu)
            {
                for (int a = 0; a < 3; a++) { atomicMin(&out[6 * rs + a], v[a]); atomicMax(&out[6 * rs + 3 + a], v[3 + a]); }
            }
            for (int a = 0; a < 3; a++) { v[a] = 0xffffffffu; v[3 + a] = 0u; }
        };
        for (uint32_t c = c0; c < c1; c++)
        {
            const uint32_t i = c * 64u + lane;
            bool valid = i < n;
            uint32_t sc = 0;
            if (valid)
            {
                sc = sceneOfParticle(d, pop, i);
                if (pop == POP_F && i - d.scenes[sc].offF >= d.scenes[sc].nF) { valid = false; }
            }
            uint32_t b[3] = { 0, 0, 0 };
            if (valid)
            {
                const float4 q = p[i];
                b[0] = orderedBits(q.x); b[1] = orderedBits(q.y); b[2] = orderedBits(q.z);
            }
            const uint64_t vm = __ballot(valid);
            if (vm == 0) { continue; }
            const int first = __ffsll(static_cast<unsigned long long>(vm)) - 1;
            const uint32_t s0 = static_cast<uint32_t>(__shfl(static_cast<int>(sc), first, 64));
            if (__ballot(valid && sc != s0) == 0)
            {
                if (s0 != rs) { flush(); rs = s0; }
                if (valid)
                {
                    for (int a = 0; a < 3; a++) { v[a] = min(v[a], b[a]); v[3 + a] = max(v[3 + a], b[a]); }
                }
            }
            else if (valid)
            {
                for (int a = 0; a < 3; a++) { atomicMin(&out[6 * sc + a], b[a]); atomicMax(&out[6 * sc + 3 + a], b[a]); }
            }
        }
        flush();
    }

    // ----------------------------------------------------------------------------------------
    // host launchers

    static inline uint32_t blocksFor(uint64_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

    void launchResetSearch(hipStream_t s, const Dev& d) { kResetSearch<<<1, kBlock, 0, s>>>(d); }

    // key arrays hold u32 keys, or u64 keys (two words each) when d.key64
    void launchKeys(hipStream_t s, const Dev& d, int pop, const float4* p, uint32_t n, bool check, uint32_t* key,
                    uint32_t* idx)
    {
        if (!n) { return; }
        if (d.key64) { kKeys<<<blocksFor(n), kBlock, 0, s>>>(d, pop, p, n, check, reinterpret_cast<uint64_t*>(key), idx); }
        else { kKeys<<<blocksFor(n), kBlock, 0, s>>>(d, pop, p, n, check, key, idx); }
    }

    void launchKeysSorted(hipStream_t s, const Dev& d, const float4* p, uint32_t n, uint32_t* key, uint32_t* idx)
    {
        if (!n) { return; }
        if (d.key64) { kKeysSorted<<<blocksFor(n), kBlock, 0, s>>>(d, p, n, reinterpret_cast<uint64_t*>(key), idx); }
        else { kKeysSorted<<<blocksFor(n), kBlock, 0, s>>>(d, p, n, key, idx); }
    }

    template <typename K>
    static bool radixSort(hipStream_t s, K* key, uint32_t* idx, K* key2, uint32_t* idx2, uint32_t n, uint32_t bits,
                          uint32_t* hist, uint32_t* tmp, uint32_t* fail, bool copyBack)
    {
        if (n <= kSmallSort)
        {
            kSortSmall<K><<<1, kSmallThreads, 0, s>>>(key, idx, key2, idx2, n, bits);
            return false;
        }
        const uint32_t nTiles = (n + kTile - 1) / kTile;
        K* kIn = key; uint32_t* iIn = idx; K* kOut = key2; uint32_t* iOut = idx2;
        for (uint32_t shift = 0; shift < bits; shift += kRadixBits)
        {
            kRadixHist<K><<<nTiles, kBlock, 0, s>>>(kIn, n, shift, nTiles, hist);
            launchScan(s, hist, kRadixBins * nTiles, tmp, fail);
            kRadixScatter<K><<<nTiles, kBlock, 0, s>>>(kIn, iIn, kOut, iOut, n, shift, nTiles, hist);
            K* t = kIn; kIn = kOut; kOut = t;
            uint32_t* ti = iIn; iIn = iOut; iOut = ti;
        }
        if (kIn == key) { return false; }
        if (!copyBack) { return true; }
        (void)hipMemcpyAsync(key, kIn, sizeof(K) * n, hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(idx, iIn, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s);
        return false;
    }

    // a sort of n keys runs as one block and uses neither the digit counts nor the scan state
    bool radixSortSingleBlock(uint32_t n) { return n <= kSmallSort; }

    // sorts (key, idx) stably by the low `bits` bits; tmp is a scan state for 256 x ceil(n / 2048)
    // values, fail its give-up word. The result ends in (key, idx), or, without copyBack, in
    // (key2, idx2) after an odd number of passes: the return value says which.
    bool launchRadixSort(hipStream_t s, bool key64, uint32_t* key, uint32_t* idx, uint32_t* key2, uint32_t* idx2,
                         uint32_t n, uint32_t bits, uint32_t* hist, uint32_t* tmp, uint32_t* fail, bool copyBack)
    {
        if (n < 2) { return false; }
        if (key64)
        {
            return radixSort(s, reinterpret_cast<uint64_t*>(key), idx, reinterpret_cast<uint64_t*>(key2), idx2, n, bits, hist,
                             tmp, fail, copyBack);
        }
        return radixSort(s, key, idx, key2, idx2, n, bits, hist, tmp, fail, copyBack);
    }

    void launchPermute(hipStream_t s, const Gather& g, const uint32_t* idx)
    {
        if (g.n) { kPermute<<<blocksFor(g.n), kBlock, 0, s>>>(g, idx); }
    }

    void launchClearCells(hipStream_t s, const Dev& d, const uint32_t* key, uint32_t n, uint2* cell)
    {
        if (n) { kClearCells<<<blocksFor(n), kBlock, 0, s>>>(d, key, n, cell); }
    }

    void launchClearAll(hipStream_t s, uint2* cell, uint64_t n)
    {
        if (n) { kClearAll<<<blocksFor(n), kBlock, 0, s>>>(cell, n); }
    }

    void launchBuildCells(hipStream_t s, const Dev& d, const uint32_t* key, uint32_t n, uint2* cell)
    {
        if (n) { kBuildCells<<<blocksFor(n), kBlock, 0, s>>>(d, key, n, cell); }
    }

    void launchQueryCount(hipStream_t s, const Dev& d)
    {
        const uint32_t nSlices = (d.nF + kSlice - 1) / kSlice;
        if (d.nF) { kQueryCount<<<blocksFor(d.nF), kBlock, 0, s>>>(d, nSlices); }
    }

    // units is 3 x (nSlices + 1) u32
    void launchScanSlices(hipStream_t s, const Dev& d, uint32_t* units, uint32_t* tmp)
    {
        const uint32_t nSlices = (d.nF + kSlice - 1) / kSlice;
        const dim3 g(blocksFor(nSlices + 1ull), 3);
        kSliceUnits<<<g, kBlock, 0, s>>>(d, nSlices, units);
        for (uint32_t r = 0; r < 3; r++) { launchScan(s, units + r * (nSlices + 1), nSlices + 1, tmp, &d.status->scanFail); }
        kSliceApply<<<g, kBlock, 0, s>>>(d, nSlices, units);
    }

    void launchQueryFill(hipStream_t s, const Dev& d)
    {
        if (d.nF)
        {
            kQueryFill<<<blocksFor(d.nF), kBlock, 0, s>>>(d);
            kQueryFillWide<<<blocksFor(d.nF), kBlock, 0, s>>>(d);
        }
    }

    void launchHaloBuild(hipStream_t s, const Dev& d)
    {
        if (d.nF) { kHaloBuild<<<blocksFor(d.nF), kBlock, 0, s>>>(d); }
    }

    void launchQueryRigid(hipStream_t s, const Dev& d, int mode, uint32_t* tmp)
    {
        if (!d.nR) { return; }
        kQueryRigid<<<blocksFor(64ull * d.nR), kBlock, 0, s>>>(d, mode);
        if (mode == 0)
        {
            launchScan(s, d.rfOff, d.nR + 1, tmp, &d.status->scanFail);
            kRigidTotals<<<blocksFor(d.nR), kBlock, 0, s>>>(d);
        }
    }

    void launchStepGate(hipStream_t s, const Dev& d) { kStepGate<<<1, 64, 0, s>>>(d); }
    void launchSolveGate(hipStream_t s, const Dev& d) { kSolveGate<<<1, kBlock, 0, s>>>(d); }
    void launchClearStop(hipStream_t s, const Dev& d) { kClearStop<<<1, 64, 0, s>>>(d); }

    void launchAabb(hipStream_t s, const Dev& d, int pop, const float4* p, uint32_t n, uint32_t* out)
    {
        // about 8 chunks of 64 particles per wave, at most 256 blocks
        const uint32_t chunks = (n + 63u) / 64u;
        const uint32_t blocks = std::max(1u, std::min(256u, (chunks + 31u) / 32u));
        if (n) { kAabb<<<blocks, kBlock, 0, s>>>(d, pop, p, n, out); }
    }
}
