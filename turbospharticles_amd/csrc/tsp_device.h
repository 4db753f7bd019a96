// tsp_device.h -- device data layout and shared helpers of the gfx950 IISPH step.
//
// Layout in HBM (indexed in the current Morton order; one neighbor of a sum costs one or two
// 16-byte gathers):
//   fluid      pos (x, y, z, pressure) | vel (vx, vy, vz, density) | vp (v*, 0) | ap (a^p, 0)
//              as float4; source, diag (f32); id (u64). Every array has a second buffer, so the
//              re-order is one gather + pointer swap
//   boundary   bpos (x, y, z, psi) float4, V (f32), id (u64): sorted by cell id when the frame
//              is built, static afterwards
//   rigid      world (x, y, z, psi), velocity (v, V), predicted velocity (v*, 0) as float4;
//              body-frame offsets, body (u32), id (u64), reaction force (f32); per-body state
//              field-major (f32) with FP64 force/torque sums
//   cell grid  dense over the cell frame, one uint2 {begin, end} per cell and particle class
//   SELL-32    per slice of 32 fluid list positions and relation (FF, FB, FR): one
//              SliceMeta {offset in u16 units, base, pitch, format}; entry k of lane l at
//              offset + k*32 + l, as u16 (j - base) or u32 j. FF entries of a block with a halo
//              hold the slot of j in the block's halo instead (same u16/u32 storage, no base).
//              List position p holds the lists of one particle of p's 256-particle block: p
//              itself, or, with lane sorting, the block's particles by list length (laneRank,
//              rankWho)
//   halo       per 256-particle block: up to kHaloMax (1064 tuned) distinct FF neighbor indices,
//              ascending
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// build-time variants of the solver sweeps (all give the same bits; the build picks one):
//   TSP_TUNE_PREFETCH      the next four list entries load while the current four are summed
//   TSP_TUNE_GATHER4       the sweeps read the LDS halo records of four entries before summing them
//   TSP_TUNE_SWEEP_AHEAD   the first sweep group of a step: the previous step's sweeps plus this
//   TSP_TUNE_BOX_CELLS     the fill stages the fluid cell ranges of its block's stencil box in LDS
//                          when the box has at most this many cells (0: off)
//   TSP_TUNE_RANGE_LOADS   the stencil walk loads a row's nine cell ranges in every lane and selects
//                          (1), or loads only the ranges a lane walks (0)
//   TSP_TUNE_QUERY_WAVES   register budget of the neighbor query and halo kernels: waves per SIMD
//                          (0: the compiler's choice)
//   TSP_TUNE_HALO_PAGES    the FF halo of a block from bitmaps over pages of 4096 indices (1), or
//                          from one 64 Ki-index window bitmap, else a sorted hash set (0)
//   TSP_TUNE_STEP_GRAPH    a step runs as one captured graph launch with one status read (1), or
//                          launch by launch with the host reading the status between phases (0)
//   TSP_TUNE_JACOBI_REC24  kJacobi's halo records take 24 bytes instead of 32 (six blocks per CU)
//   TSP_TUNE_KG_STORE      kSourceDiag stores each FF pair's gradient factor; the sweeps read it
//   TSP_TUNE_HALO_MAX      FF halo records per 256-particle block (LDS per sweep block)
//   TSP_TUNE_GRAPHS        each group of solver sweeps runs as one instantiated HIP graph
//   TSP_TUNE_PA_WAVES, TSP_TUNE_JACOBI_WAVES   waves per SIMD the compiler must leave room for
//                          (register budget of the sweep kernels); 0 leaves the choice to it
//   TSP_TUNE_PERSISTENT    the solver sweeps run as persistent kernels: a resident grid whose
//                          blocks walk the 256-particle tiles, instead of one block per tile
//   TSP_TUNE_SLICE         SELL slice height: 32 (a wave covers two slices) or 64 (one slice per wave)
//   TSP_TUNE_RADIX_BITS    digit width of the LSD radix sorts (8: 256 buckets, 6, 4)
//   TSP_TUNE_SWEEP_THREADS threads per 256-particle tile of the solver sweeps (256, or 128: two
//                          particles per thread)
//   TSP_TUNE_SLOT16        the halo build stores the halo slots of u32 slices as u16 (1), or keeps
//                          their u32 storage (0)
//   TSP_TUNE_LANE_SORT     the lists of each 256-particle block take its list positions longest
//                          first, so a wave's lanes walk lists of similar length (1: by FF count,
//                          2: by FB count, then FF count; 0: positions in particle order)
//   TSP_TUNE_SEARCH_FLAT   the count pass folds its hits in through selects, its row streams loading
//                          past their end (1) or branching around those loads (2); or it branches
//                          on each candidate (0)
//   TSP_TUNE_SEARCH_ROW    the neighbor query walks the stencil by rows of three cells, loading
//                          the rows' ranges together and TSP_TUNE_SEARCH_BATCH candidates at a time
//                          (1: the row's candidates as one stream, 2: cell by cell; 0: the cell walk)
// the settings chosen by turbospharticles_amd/tune.py, when present, come first
#if __has_include("tsp_tuned.h")
#include "tsp_tuned.h"
#endif
#ifndef TSP_TUNE_BOX_CELLS
#define TSP_TUNE_BOX_CELLS 0
#endif
#ifndef TSP_TUNE_RANGE_LOADS
#define TSP_TUNE_RANGE_LOADS 0
#endif
#ifndef TSP_TUNE_QUERY_WAVES
#define TSP_TUNE_QUERY_WAVES 0
#endif
#ifndef TSP_TUNE_HALO_PAGES
#define TSP_TUNE_HALO_PAGES 0
#endif
#ifndef TSP_TUNE_STEP_GRAPH
#define TSP_TUNE_STEP_GRAPH 1
#endif
#ifndef TSP_TUNE_SWEEP_AHEAD
#define TSP_TUNE_SWEEP_AHEAD -1
#endif
#ifndef TSP_TUNE_GATHER4
#define TSP_TUNE_GATHER4 0
#endif
#ifndef TSP_TUNE_PREFETCH
#define TSP_TUNE_PREFETCH 1
#endif
#ifndef TSP_TUNE_JACOBI_REC24
#define TSP_TUNE_JACOBI_REC24 1
#endif
#ifndef TSP_TUNE_KG_STORE
#define TSP_TUNE_KG_STORE 0
#endif
#ifndef TSP_TUNE_HALO_MAX
#define TSP_TUNE_HALO_MAX 1000
#endif
#ifndef TSP_TUNE_GRAPHS
#define TSP_TUNE_GRAPHS 1
#endif
#ifndef TSP_TUNE_PA_WAVES
#define TSP_TUNE_PA_WAVES 0
#endif
#ifndef TSP_TUNE_JACOBI_WAVES
#define TSP_TUNE_JACOBI_WAVES 0
#endif
#ifndef TSP_TUNE_SEARCH_ROW
#define TSP_TUNE_SEARCH_ROW 1
#endif
#ifndef TSP_TUNE_PERSISTENT
#define TSP_TUNE_PERSISTENT 0
#endif
#ifndef TSP_TUNE_SEARCH_FLAT
#define TSP_TUNE_SEARCH_FLAT 0
#endif
#ifndef TSP_TUNE_SLOT16
#define TSP_TUNE_SLOT16 1
#endif
#ifndef TSP_TUNE_LANE_SORT
#define TSP_TUNE_LANE_SORT 1
#endif
#ifndef TSP_TUNE_SEARCH_BATCH
#define TSP_TUNE_SEARCH_BATCH 4
#endif
#ifndef TSP_TUNE_SLICE
#define TSP_TUNE_SLICE 32
#endif
#ifndef TSP_TUNE_RADIX_BITS
#define TSP_TUNE_RADIX_BITS 8
#endif
#ifndef TSP_TUNE_SWEEP_THREADS
#define TSP_TUNE_SWEEP_THREADS 256
#endif
#define TSP_WAVES_ATTR_0
#define TSP_WAVES_ATTR_1 __attribute__((amdgpu_waves_per_eu(1)))
#define TSP_WAVES_ATTR_2 __attribute__((amdgpu_waves_per_eu(2)))
#define TSP_WAVES_ATTR_3 __attribute__((amdgpu_waves_per_eu(3)))
#define TSP_WAVES_ATTR_4 __attribute__((amdgpu_waves_per_eu(4)))
#define TSP_WAVES_ATTR_5 __attribute__((amdgpu_waves_per_eu(5)))
#define TSP_WAVES_ATTR_6 __attribute__((amdgpu_waves_per_eu(6)))
#define TSP_WAVES_ATTR_7 __attribute__((amdgpu_waves_per_eu(7)))
#define TSP_WAVES_ATTR_8 __attribute__((amdgpu_waves_per_eu(8)))
#define TSP_WAVES_CAT(n) TSP_WAVES_ATTR_##n
#define TSP_WAVES_ATTR(n) TSP_WAVES_CAT(n)

namespace tsp
{
    constexpr int kBlock = 256;          // threads per block: 4 waves of 64
    constexpr uint32_t kSlice = TSP_TUNE_SLICE;   // SELL slice height (32: a wave covers two slices)
    static_assert(kSlice == 32 || kSlice == 64, "slice height");
    constexpr uint32_t kRadixBits = TSP_TUNE_RADIX_BITS;
    constexpr uint32_t kRadixBins = 1u << kRadixBits;
    static_assert(kRadixBits >= 1 && kRadixBits <= 8, "radix digit width");
    constexpr uint32_t kEmpty32 = 0xffffffffu;
    constexpr uint32_t kFmtDelta16 = 0;
    constexpr uint32_t kFmtAbs32 = 1;
    // FF lists of more than kDictMax entries take the reference's normal (per-neighbor scaled)
    // formulation: its per-particle chunks become dict chunks only up to 256 distinct neighbors
    // (N_SELL_DICT_SIZE_MAX, neighbor_lists.cpp:326-332)
    constexpr uint32_t kDictMax = 256;

    // state of a one-pass scan of m values: counter, epoch, two spare words, then one u64 status
    // word per 2048-value tile; zeroed when allocated
    constexpr uint32_t kScanHeaderWords = 4;
    inline constexpr uint64_t scanTmpWords(uint64_t m) { return kScanHeaderWords + 2ull * ((m + 2047ull) / 2048ull); }

    // FF hit masks of the count pass (one word per stencil row of three cells, reach 1): the fill
    // takes the hits from them and loads no candidate position. A row with more than 32 fluid
    // candidates is marked in the last word and tested again by the fill.
    constexpr uint32_t kMaskRows = 9;
    constexpr uint32_t kMaskWords = kMaskRows + 1;
    // FF neighbor halo of a 256-particle block: the distinct FF neighbors of its particles,
    // ascending; the FF entries of the block then hold halo slots instead of particle indices
    constexpr uint32_t kHaloMax = TSP_TUNE_HALO_MAX; // 24 B x 1064 records in kJacobi: six blocks per CU
    constexpr uint32_t kNoHalo = 0xffffffffu;   // halo count of a block that keeps particle indices

    enum Rel { REL_FF = 0, REL_FB = 1, REL_FR = 2 };

    constexpr uint32_t kWideBit = 0x80000000u;   // in ffMin: the particle's FF list spans more than 0xffff

    // a compile-time flag passed to a generic lambda (one loop body per value)
    template <bool B>
    struct Flag
    {
        static constexpr bool value = B;
    };

    struct SliceMeta
    {
        uint32_t offset;  // in u16 units from the start of the pool (storage units before the scan)
        uint32_t base;
        uint16_t pitch;
        uint16_t format;
        uint32_t entries; // list entries of the slice without padding
    };

    // flags and totals of one neighbor search over all scenes of a context, read back by the host
    struct Status
    {
        uint32_t abort;          // some scene's particle left its inner cell frame: restart the search
        uint32_t overflow;       // a SELL pool is too small
        uint32_t bUnsorted;      // boundary ids are not non-decreasing under the frames
        uint32_t pairTiny;       // some FF or FB pair has 0 < d2 < 2^-96: the kernels take the general sqrt
        uint32_t scanFail;       // sticky: a one-pass scan gave up waiting (kResetSearch keeps it)
        uint32_t wideCount;      // particles in wideList (FF lists spanning more than 0xffff)
        uint32_t stop;           // a captured step stops early: 1 = its search needs the host, 2 = its
                                 // solve needs more sweeps (kernels launched with stopMask skip)
        uint64_t poolUnits[3];   // u16 units used by FF, FB, FR
        uint64_t entries[3];     // entries (without padding) of FF, FB, FR
        uint64_t rfTotal, rbTotal, rrTotal;
    };

    // pressure solve state of one scene (SimBackend_CPU members N_solverIters_,
    // avgRelDensityErrorF_*), decided on the device and read back after groups of sweeps
    struct Solve
    {
        uint32_t iters;
        uint32_t compute;        // the current sweep reduces errors
        uint32_t converged;
        uint32_t sweeps;
        float avgRaw, avgPos, avgAbs;
        uint32_t pending;        // a Jacobi sweep ran whose errors are not yet summed
        double sumRaw, sumPos, sumAbs;
        uint32_t abort;          // a particle of this scene left the inner frame
        uint32_t rigidDegenerate;
        // rigid contact solve of the step (RigidContactReport, rigid_solver.h)
        uint32_t contactActive;  // the scene has RB or RR entries: its contact solve runs
        uint32_t contactDone;
        uint32_t contactIters;
        uint32_t contactConverged;
        float contactAvg;
        uint32_t contactBoundary;   // boundary particles in contact
        uint32_t contactBad;        // kContactBad* flags
        uint32_t pad1;
        uint64_t contacts;          // N_contacts of the relaxation
    };

    // Solve::contactBad
    constexpr uint32_t kContactBadDiag = 1u, kContactBadResidual = 2u, kContactBadPressure = 4u,
                       kContactManyBodies = 8u;

    // where a scene's particles live in the arrays of a context. Fluid segments start at a
    // multiple of kBlock and are padded to one, so a thread block never spans two scenes;
    // the padding slots are dead (they sort last, own no cell, and no kernel updates them).
    struct Scene
    {
        uint32_t offF, nF;       // fluid: first slot, active particles
        uint32_t blk0, nBlk;     // fluid thread blocks
        uint32_t offB, nB;
        uint32_t offR, nR;
        uint32_t offRB, nRB;
    };

    struct Frame
    {
        int32_t orig[3];
        uint32_t dim[3];
        float cellSizeInv;
        int32_t reach;       // stencil half width in cells
        uint64_t keyBase;    // the scene's Morton ids are keyBase + morton(local cell)
        uint64_t emptyKey;   // keyBase + morton(dim - 1) + 1: particles outside the frame sort last
        uint64_t deadKey;    // emptyKey + 1: the padding slots of the fluid segment
        uint32_t cellOff;    // the scene's first cell in the cell tables
        int32_t innerLo[3], innerHi[3];
        int32_t domainSet;
        float domMin[3], domMax[3];
    };

    struct Consts
    {
        float dt, hinv, hsq, alpha, rho0, rho0SqInv, r2, mass;
        float ksize, ktwo, kinSlope, kinOff, koutScale;
        float weightScale;     // mass * alpha * hinv^3 (dict FF density)
        float hsqEps;          // 0.01 h^2
        float visF, visB;      // 10 nu
        float gx, gy, gz;
        float cFacF, cFacB, dFacB, omega, warm;
        float presDis, presOuterB, presOuterFF;
        float dtsq;
        float densityDevTarget;
        uint32_t minIters, maxIters, fixedIters, termCond;
        uint32_t nF;           // active fluid particles of the scene (the error averages divide by it)
        float contactDevTarget;     // rigidContactDensityDevTarget
        uint32_t contactMaxIters;   // rigidContactMaxSolverIters
    };

    // records gathered by one permutation launch
    struct Gather
    {
        const float4* s4[4];
        float4* d4[4];
        uint32_t n4;
        const float* s1[4];
        float* d1[4];
        uint32_t n1;
        const uint64_t* idS;
        uint64_t* idD;
        uint32_t n;
    };

    // --- device scene builder (kernels_build.hip): the parts of a scene plan its kernels read ---
    struct BuildVec
    {
        float x, y, z;
    };

    // the fluid block: nu x nv x nw lattice, index = (kw nv + jv) nu + iu
    struct BuildFluid
    {
        BuildVec origin, u, v, w;
        float uLen, vLen, wLen;
        float uOff, vOff, wOff, radius, jitAmp, size, rho0;
        uint64_t seed;
        uint32_t nu, nv, nw;
    };

    // one container face: layers x nu x nv particles from index `first`
    struct BuildFace
    {
        BuildVec origin, u, v, n;
        float uLen, vLen;
        uint32_t nu, nv;
        uint64_t first;
    };

    // the boundary layers and their psi / V sums (cells of size 2h; keys pack the cell offsets
    // from cellMin as z | y | x with bits[] bits each)
    struct BuildBoundary
    {
        BuildFace faces[6];
        uint32_t nFaces, layers;
        float size, sizeInv, radiusSqu, alpha, rho0, massB, cellSizeInv;
        int32_t cellMin[3], cellMax[3];
        uint32_t bits[3];
    };

    // fields of the body state Dev::rb (field-major: rb[field * nRB + body], TSP_RB_* order)
    enum
    {
        RB_XCM = 0, RB_VX = 3, RB_WX = 6, RB_LX = 9, RB_A00 = 12, RB_MASS = 21, RB_MASS_INV = 22, RB_I00 = 23
    };

    // fields of Dev::cR (per rigid particle) and Dev::cB (per contact boundary particle)
    enum { CR_RHO, CR_DIAG, CR_P, CR_SRC, CR_ERR, CR_OM, CR_VX, CR_VY, CR_VZ, CR_FX, CR_FY, CR_FZ, kCR };
    enum { CB_RHO, CB_DIAG, CB_P, CB_SRC, CB_ERR, CB_OM, kCB };

    // device pointers of one context (passed by value to every kernel)
    struct Dev
    {
        uint32_t nF, nB, nR, nRB;    // array lengths over all scenes (nF counts padding slots)
        uint32_t nScenes;
        uint32_t fastDiv;            // every scene's kernel constants lie in divCR's range
        uint32_t fastPairs;          // this search's FF and FB pairs may take sqrtCR / divCR (kgFactor<true>)
        Scene *scenes;
        Frame *frames;
        Consts *consts;
        Solve *solve;
        uint32_t key64;              // the keys are 64-bit (some frame needs more than 32 bits of key space)
        uint32_t stopMask;           // nonzero in a captured step: a kernel returns at once when status->stop
                                     // has one of these bits (0 everywhere else)
        uint64_t *sceneKey0;         // nScenes + 1 ascending key bases (the last one ends the key space)
        uint32_t *blockScene;        // scene of each fluid thread block
        uint16_t *sceneB;            // scene of each boundary particle
        uint32_t *bodyScene;         // scene of each rigid body
        // fluid, packed so that one neighbor costs one 16-byte gather per quantity:
        //   pos (x, y, z, pressure), vel (vx, vy, vz, density), vp (v*, 0), ap (a^p, 0)
        float4 *pos, *vel, *vp, *ap;
        float *source, *diag;
        uint64_t *id;
        // boundary: (x, y, z, psi) and V
        float4 *bpos;
        float *bvol;
        // rigid: world (x, y, z, psi), velocity (v, V), predicted velocity (v*, 0); body-frame offsets
        float4 *rpos, *rvel, *rvp;
        float *rxrel, *ryrel, *rzrel;
        uint32_t *rbody;
        float *rfx, *rfy, *rfz;
        float *rb;               // nRB x 32 state (TSP_RB_* field order, field-major)
        float *rbPred;           // nRB x 9: v, L, omega predicted
        float *rbForce;          // nRB x 6: force, torque (float, like the reference members)
        uint32_t *byBodyOff, *byBodyIdx;
        // neighbor search
        uint32_t *keyF, *keyF2, *idxF, *idxF2;   // keys: u32, or u64 (two words each) when key64
        uint32_t *keyR, *keyR2, *idxR, *idxR2;
        uint2 *cellF, *cellB, *cellR;
        // SELL-32
        uint16_t *cnt[3];
        SliceMeta *meta[3];
        uint16_t *pool[3];
        uint64_t poolCap[3];     // u16 units
        float *kgFF;             // per FF entry (by its u16 position): gradient factor of the step (build variant)
        uint8_t *laneRank;       // per particle: its list position within its block (lane sorting)
        uint8_t *rankWho;        // per list position: the particle's index within the block (lane sorting)
        uint32_t *ffMin;         // per particle: smallest FF neighbor index (0 without FF neighbors), with
                                 // kWideBit set when the list spans more than 0xffff
        uint32_t *wideList;      // the particles with kWideBit, in no particular order (status->wideCount)
        uint32_t *ffMask;        // kMaskWords x nF: per stencil row, bit q = the row's q-th fluid candidate
                                 // is an FF neighbor; then the rows the masks do not cover
        uint32_t *haloIdx;       // per block: kHaloMax particle indices, ascending
        uint32_t *haloCount;     // per block: halo size, or kNoHalo
        // rigid-source CSR (RF), counts of RB and RR
        uint32_t *rfCnt, *rfOff, *rfIdx;
        uint32_t *rbCnt, *rrCnt;
        uint64_t rfCap;
        // rigid contacts: RB and RR lists (filled when some scene has contacts), the boundary
        // particles in contact, and the state of the contact solve
        uint32_t *rbOff, *rrOff;     // nR + 1
        uint32_t *rbIdx, *rrIdx;     // boundary / rigid indices
        uint64_t rbCap, rrCap;
        uint32_t *cbScan;            // nB + 1: contact flags of the boundary particles, then their exclusive scan
        uint32_t *cbLocal;           // per contact boundary particle: its boundary index (rbCap)
        uint32_t *brOff, *brPos, *brIdx;   // per contact boundary particle: rigid neighbors, ascending
        float *cRw;                  // 3 x nR: world offsets A r_rel
        float *cR;                   // kCR x nR: contact state of the rigid particles
        float *cB;                   // kCB x rbCap: contact state of the contact boundary particles
        double *cErrR, *cErrB;       // FP64 error terms of the last evaluation
        float *cIw;                  // 9 x nRB: world inverse inertia
        float *cLoad;                // 6 x nRB: contact force and torque
        uint32_t *cBodyB;            // nRB: boundary contacts of each movable body
        // reductions and scans
        double *errPartial;      // Jacobi error partials, [3][blocks of the fluid grid]
        uint32_t *scanTmp;       // block sums of the multi-block scans
        Status *status;
    };

    __device__ __forceinline__ float rbv(const Dev& d, int f, uint32_t b) { return d.rb[f * d.nRB + b]; }

    // a kernel of a captured step whose part of the step was called off (uniform over the grid)
    __device__ __forceinline__ bool stopped(const Dev& d)
    {
        return d.stopMask != 0u && (d.status->stop & d.stopMask) != 0u;
    }

    // key i of a key array (u32, or u64 when the context's keys are 64-bit)
    __device__ __forceinline__ uint64_t keyAt(const Dev& d, const uint32_t* keys, uint32_t i)
    {
        return d.key64 ? reinterpret_cast<const uint64_t*>(keys)[i] : static_cast<uint64_t>(keys[i]);
    }

    // scene of a key: the last scene whose key base is not above it
    __device__ __forceinline__ uint32_t sceneOfKey(const Dev& d, uint64_t key)
    {
        uint32_t lo = 0, hi = d.nScenes;
        while (hi - lo > 1)
        {
            const uint32_t mid = (lo + hi) >> 1;
            if (d.sceneKey0[mid] <= key) { lo = mid; } else { hi = mid; }
        }
        return lo;
    }

    // --- FP32 helpers matching the reference's operation order (no contraction) ---

    __device__ __forceinline__ float vmax(float a, float b) { return a > b ? a : b; }

    // FloorToInt, CPU/backend/neighbor_search.cpp:157-162
    __device__ __forceinline__ int floorToInt(float v)
    {
        const int i = static_cast<int>(v);
        return (v < static_cast<float>(i)) ? i - 1 : i;
    }

    // SpreadBitsWithGaps_21to63 (neighbor_search.cpp:2057-2070): bit b of v to bit 3b
    __device__ __forceinline__ uint64_t spread21(uint32_t v)
    {
        uint64_t x = v & 0x1fffffu;
        x = (x | (x << 32)) & 0x1f00000000ffffull;
        x = (x | (x << 16)) & 0x1f0000ff0000ffull;
        x = (x | (x << 8)) & 0x100f00f00f00f00full;
        x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
        x = (x | (x << 2)) & 0x1249249249249249ull;
        return x;
    }

    __device__ __forceinline__ uint32_t compact21(uint64_t x)
    {
        x &= 0x1249249249249249ull;
        x = (x | (x >> 2)) & 0x10c30c30c30c30c3ull;
        x = (x | (x >> 4)) & 0x100f00f00f00f00full;
        x = (x | (x >> 8)) & 0x1f0000ff0000ffull;
        x = (x | (x >> 16)) & 0x1f00000000ffffull;
        x = (x | (x >> 32)) & 0x1fffffull;
        return static_cast<uint32_t>(x);
    }

    __device__ __forceinline__ uint32_t spread10(uint32_t v)
    {
        uint32_t x = v & 0x3ffu;
        x = (x | (x << 16)) & 0x030000ffu;
        x = (x | (x << 8)) & 0x0300f00fu;
        x = (x | (x << 4)) & 0x030c30c3u;
        x = (x | (x << 2)) & 0x09249249u;
        return x;
    }

    __device__ __forceinline__ uint32_t compact10(uint32_t x)
    {
        x &= 0x09249249u;
        x = (x | (x >> 2)) & 0x030c30c3u;
        x = (x | (x >> 4)) & 0x0300f00fu;
        x = (x | (x >> 8)) & 0x030000ffu;
        x = (x | (x >> 16)) & 0x3ffu;
        return x;
    }

    // Morton id of local cell coordinates (x bits at 0,3,6..., y at 1,4,..., z at 2,5,...),
    // MapToCellId_Morton, neighbor_search.cpp:2032-2055: 32 bits for frames of up to 1024 cells
    // per axis, 63 bits for frames of up to 2^21
    __device__ __forceinline__ uint32_t morton(uint32_t lx, uint32_t ly, uint32_t lz)
    {
        return spread10(lx) | (spread10(ly) << 1) | (spread10(lz) << 2);
    }

    __device__ __forceinline__ uint64_t morton64(uint32_t lx, uint32_t ly, uint32_t lz)
    {
        return spread21(lx) | (spread21(ly) << 1) | (spread21(lz) << 2);
    }

    // index of a local cell of a scene's frame in the cell tables
    __device__ __forceinline__ uint32_t linearCell(const Frame& f, int lx, int ly, int lz)
    {
        return f.cellOff + static_cast<uint32_t>(lx) + f.dim[0] * (static_cast<uint32_t>(ly) + f.dim[1] * static_cast<uint32_t>(lz));
    }

    // --- correctly rounded f32 square root and quotient, in fewer instructions than the
    //     general lowering: the pair kernels must round exactly like the CPU restatement ---

    // sqrt(x) for x = 0 or x >= 2^-96: the hardware estimate (within 1 ulp) moved to the
    // neighbour whose residual has the right sign -- the correction of the general lowering,
    // without its input scaling (needed only below 2^-96) and its +-0 / inf cases (x = 0 gives
    // +0 here). Checked bit for bit against the general lowering over every such float.
    __device__ __forceinline__ float sqrtCR(float x)
    {
        const float s = __builtin_amdgcn_sqrtf(x);
        const float sDn = __int_as_float(__float_as_int(s) - 1);
        const float sUp = __int_as_float(__float_as_int(s) + 1);
        const float rDn = __builtin_fmaf(-sDn, s, x);
        const float rUp = __builtin_fmaf(-sUp, s, x);
        const float t = rDn <= 0.0f ? sDn : s;
        return rUp > 0.0f ? sUp : t;
    }

    // squared distance in the reference's operation order
    __device__ __forceinline__ float dist2(float dx, float dy, float dz)
    {
        return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    }

    // squared distances that sqrtCR does not cover (0 < d2 < 2^-96), as one unsigned compare
    __device__ __forceinline__ bool tinyDist2(float d2)
    {
        return static_cast<uint32_t>(__float_as_int(d2)) - 1u < 0x0f800000u - 1u;
    }

    // a / b: the reciprocal estimate refined by one Newton step, then two residual corrections
    // of the quotient -- the general lowering's core, which is the whole computation when it
    // needs no operand scaling. Callers guarantee 2^-60 <= |a| <= 2^60 and
    // 2^-30 <= |b| <= 2^31 (the exponents then differ by less than the 96 that trigger
    // scaling, and every intermediate is a normal number); checked bit for bit against the
    // general lowering over such operands.
    __device__ __forceinline__ float divCR(float a, float b)
    {
        float y = __builtin_amdgcn_rcpf(b);
        const float e = __builtin_fmaf(-b, y, 1.0f);
        y = __builtin_fmaf(e, y, y);
        float q = a * y;
        float r = __builtin_fmaf(-b, q, a);
        q = __builtin_fmaf(r, y, q);
        r = __builtin_fmaf(-b, q, a);
        return __builtin_fmaf(r, y, q);
    }

    // KernelWeight_CubicSpline_SIMD, CPU/simd.h:279-309 (with kFast the root takes sqrtCR)
    template <bool kFast = false>
    __device__ __forceinline__ float kwNormal(float dx, float dy, float dz, const Consts& k)
    {
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        const float dist = kFast ? sqrtCR(d2) : __builtin_sqrtf(d2);
        const float q = dist * k.hinv;
        const float t1 = vmax(1.0f - q, 0.0f);
        const float t2 = vmax(2.0f - q, 0.0f);
        const float t1_2 = t1 * t1, t2_2 = t2 * t2;
        const float t1_3 = t1_2 * t1, t2_3 = t2_2 * t2;
        const float w = k.alpha * __builtin_fmaf(-4.0f, t1_3, t2_3);
        return d2 < k.r2 ? w : 0.0f;
    }

    // KernelWeightUnscaled_CubicSpline_SIMD, CPU/simd.h:392-404
    template <bool kFast = false>
    __device__ __forceinline__ float kwUnscaled(float d2, const Consts& k)
    {
        const float dist = kFast ? sqrtCR(d2) : __builtin_sqrtf(d2);
        const float t1 = vmax(k.ksize - dist, 0.0f);
        const float t2 = vmax(k.ktwo - dist, 0.0f);
        return __builtin_fmaf(-4.0f, t1 * t1 * t1, t2 * t2 * t2);
    }

    // KernelGradientFactorNoMinDist_CubicSpline_SIMD, CPU/simd.h:374-388. With kFast the root
    // and the quotient take sqrtCR and divCR. The host allows it for a launch when no FF or FB
    // pair of the search has 0 < d2 < 2^-96 and every scene has |koutScale| and h in divCR's
    // range (the quotient's operands are koutScale and dist, and it only matters where
    // dist >= h). FR pairs keep the general lowering: rigid particles move within a step, after
    // the search has seen their distances.
    template <bool kFast = false>
    __device__ __forceinline__ float kgFactor(float d2, const Consts& k)
    {
        const float dist = kFast ? sqrtCR(d2) : __builtin_sqrtf(d2);
        const float inner = __builtin_fmaf(k.kinSlope, dist, k.kinOff);
        const float t2 = vmax(k.ktwo - dist, 0.0f);
        const float outer = (t2 * t2) * (kFast ? divCR(k.koutScale, dist) : k.koutScale / dist);
        return dist < k.ksize ? inner : outer;
    }

    // KernelGradient_CubicSpline_Scalar, CPU/simd.h:311-347 (rigid reaction forces)
    __device__ __forceinline__ void kgScalar(float dx, float dy, float dz, const Consts& k, float& gx, float& gy, float& gz)
    {
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        if (d2 >= k.r2) { gx = gy = gz = 0.0f; return; }
        const float dist = __builtin_sqrtf(d2);
        const float q = dist * k.hinv;
        float f;
        if (q < 1.0f)
        {
            f = k.alpha * k.hinv * k.hinv * __builtin_fmaf(9.0f, q, -12.0f);
        }
        else
        {
            const float t2v = 2.0f - q;
            const float t2 = t2v < 0.0f ? 0.0f : t2v;
            f = k.alpha * (-3.0f * t2 * t2) * (k.hinv / dist);
        }
        gx = f * dx; gy = f * dy; gz = f * dz;
    }

    // entry k of lane `lane` of a SELL-32 slice
    __device__ __forceinline__ uint32_t sellRead(const uint16_t* pool, const SliceMeta& m, uint32_t lane, uint32_t k)
    {
        if (m.format == kFmtDelta16) { return m.base + static_cast<uint32_t>(pool[m.offset + k * kSlice + lane]); }
        return reinterpret_cast<const uint32_t*>(pool + m.offset)[k * kSlice + lane];
    }

    __device__ __forceinline__ void sellWrite(uint16_t* pool, const SliceMeta& m, uint32_t lane, uint32_t k, uint32_t j)
    {
        if (m.format == kFmtDelta16) { pool[m.offset + k * kSlice + lane] = static_cast<uint16_t>(j - m.base); }
        else { reinterpret_cast<uint32_t*>(pool + m.offset)[k * kSlice + lane] = j; }
    }

    // blocks are dispatched round-robin over the 8 XCDs; this maps them so that each XCD walks
    // one contiguous range of particles, which keeps neighbor gathers inside its own L2
    __device__ __forceinline__ uint32_t xcdBlock()
    {
        const uint32_t nb = gridDim.x, b = blockIdx.x;
        const uint32_t xcd = b & 7u, k = b >> 3;
        const uint32_t per = nb >> 3, rem = nb & 7u;
        return xcd < rem ? xcd * (per + 1u) + k : rem * (per + 1u) + (xcd - rem) * per + k;
    }
    // FF entries of a block with a halo hold the byte offset of the neighbor's 16-byte record in
    // the block's LDS halo (slot << kSlotShift), so a sweep addresses the record without a shift
    constexpr uint32_t kSlotShift = 4;
    static_assert((TSP_TUNE_HALO_MAX << 4) <= 0x10000u, "halo byte offsets fit the u16 entries");

    // the record of a halo entry in an LDS array: 16-byte records at the entry's byte offset,
    // 8-byte records at half of it
    template <typename T>
    __device__ __forceinline__ const T& recAt(const T* base, uint32_t off)
    {
        static_assert(sizeof(T) == 16 || sizeof(T) == 8, "16- or 8-byte records");
        return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (sizeof(T) == 16 ? off : off >> 1));
    }

    // --- SELL-32 list access (shared by the SPH, statistics and export kernels) ---

    struct List
    {
        const uint16_t* pool;
        SliceMeta m;
        uint32_t lane, n;
        __device__ __forceinline__ uint32_t operator[](uint32_t k) const { return sellRead(pool, m, lane, k); }
    };

    // the list position of particle i, and the particle at list position p (both in its block)
    __device__ __forceinline__ uint32_t sellPos(const Dev& d, uint32_t i)
    {
#if TSP_TUNE_LANE_SORT
        return (i & ~(kBlock - 1u)) + d.laneRank[i];
#else
        (void)d;
        return i;
#endif
    }

    __device__ __forceinline__ uint32_t particleAt(const Dev& d, uint32_t p)
    {
#if TSP_TUNE_LANE_SORT
        return (p & ~(kBlock - 1u)) + d.rankWho[p];
#else
        (void)d;
        return p;
#endif
    }

    // the list of relation rel of particle i, stored at list position p
    __device__ __forceinline__ List listAt(const Dev& d, int rel, uint32_t p, uint32_t i)
    {
        List l;
        l.pool = d.pool[rel];
        l.m = d.meta[rel][p / kSlice];
        l.lane = p & (kSlice - 1);
        l.n = d.cnt[rel][i];
        return l;
    }

    __device__ __forceinline__ List listOf(const Dev& d, int rel, uint32_t i) { return listAt(d, rel, sellPos(d, i), i); }

    // visits the entries of one list in storage order, as use(load(j), e): j the neighbor (or its
    // halo slot when kSlots: the stored value without the slice base) and e the entry's position
    // in u16 units of the pool, which also indexes the per-entry gradient factors. The index loads
    // of four entries are issued together, then the four gathers (load), then the four sums (use)
    // in entry order.
    template <bool kSlots, typename Load, typename Use>
    __device__ __forceinline__ void visitSplit(const List& l, Load&& load, Use&& use)
    {
        const auto fn4 = [&](uint32_t j0, uint32_t j1, uint32_t j2, uint32_t j3, uint32_t e, uint32_t step)
        {
            const auto r0 = load(j0);
            const auto r1 = load(j1);
            const auto r2 = load(j2);
            const auto r3 = load(j3);
            use(r0, e); use(r1, e + step); use(r2, e + 2u * step); use(r3, e + 3u * step);
        };
        const auto fn = [&](uint32_t j, uint32_t e) { use(load(j), e); };
        uint32_t q = 0;
        if (l.m.format == kFmtDelta16)
        {
            const uint32_t e0 = l.m.offset + l.lane;
            const uint16_t* p = l.pool + e0;
            const uint32_t base = kSlots ? 0u : l.m.base;
#if TSP_TUNE_PREFETCH
            // the next four indices load while the current four are summed
            uint32_t j0 = 0, j1 = 0, j2 = 0, j3 = 0;
            if (l.n >= 4) { j0 = p[0]; j1 = p[kSlice]; j2 = p[2 * kSlice]; j3 = p[3 * kSlice]; }
            for (; q + 4 <= l.n; q += 4)
            {
                uint32_t n0 = 0, n1 = 0, n2 = 0, n3 = 0;
                if (q + 8 <= l.n)
                {
                    n0 = p[(q + 4) * kSlice]; n1 = p[(q + 5) * kSlice];
                    n2 = p[(q + 6) * kSlice]; n3 = p[(q + 7) * kSlice];
                }
                fn4(base + j0, base + j1, base + j2, base + j3, e0 + q * kSlice, kSlice);
                j0 = n0; j1 = n1; j2 = n2; j3 = n3;
            }
#else
            for (; q + 4 <= l.n; q += 4)
            {
                const uint32_t j0 = p[(q + 0) * kSlice], j1 = p[(q + 1) * kSlice];
                const uint32_t j2 = p[(q + 2) * kSlice], j3 = p[(q + 3) * kSlice];
                fn4(base + j0, base + j1, base + j2, base + j3, e0 + q * kSlice, kSlice);
            }
#endif
            for (; q < l.n; q++) { fn(base + p[q * kSlice], e0 + q * kSlice); }
        }
        else
        {
            const uint32_t e0 = l.m.offset + 2u * l.lane;
            const uint32_t* p = reinterpret_cast<const uint32_t*>(l.pool + l.m.offset) + l.lane;
            for (; q + 4 <= l.n; q += 4)
            {
                const uint32_t j0 = p[(q + 0) * kSlice], j1 = p[(q + 1) * kSlice];
                const uint32_t j2 = p[(q + 2) * kSlice], j3 = p[(q + 3) * kSlice];
                fn4(j0, j1, j2, j3, e0 + 2u * q * kSlice, 2u * kSlice);
            }
            for (; q < l.n; q++) { fn(p[q * kSlice], e0 + 2u * q * kSlice); }
        }
    }

    // the same with the neighbor passed through: fn(j, e)
    template <bool kSlots, typename Fn>
    __device__ __forceinline__ void visitEntries(const List& l, Fn&& fn)
    {
        visitSplit<kSlots>(l, [](uint32_t j) { return j; }, fn);
    }

    template <typename Fn>
    __device__ __forceinline__ void visit(const List& l, Fn&& fn) { visitEntries<false>(l, fn); }

    // FF neighbors of a particle: from the block's LDS halo when it has one, else from HBM
    template <typename FromHalo, typename FromGlobal>
    __device__ __forceinline__ void visitFF(const List& l, bool halo, FromHalo&& fh, FromGlobal&& fg)
    {
        if (halo) { visitEntries<true>(l, fh); }
        else { visitEntries<false>(l, fg); }
    }

    // the same with the halo records read as load(slot) and summed as use(record, e): with
    // TSP_TUNE_GATHER4 the four LDS reads of a group issue before its four sums
    template <typename HLoad, typename HUse, typename FromGlobal>
    __device__ __forceinline__ void visitFFSplit(const List& l, bool halo, HLoad&& hl, HUse&& hu, FromGlobal&& fg)
    {
#if TSP_TUNE_GATHER4
        if (halo) { visitSplit<true>(l, hl, hu); }
#else
        if (halo) { visitEntries<true>(l, [&](uint32_t q, uint32_t e) { hu(hl(q), e); }); }
#endif
        else { visitEntries<false>(l, fg); }
    }

    // --- rigid bodies, one body or particle (shared by the rigid kernels and the contact solve) ---

    // UpdateRigidParticles(predicted), cpu_backend.cpp:2009-2063
    __device__ __forceinline__ void rigidUpdateOne(const Dev& d, uint32_t i, bool predicted)
    {
        const uint32_t b = d.rbody[i];
        const float lx = d.rxrel[i], ly = d.ryrel[i], lz = d.rzrel[i];
        const float rx = __builtin_fmaf(rbv(d, RB_A00 + 0, b), lx, __builtin_fmaf(rbv(d, RB_A00 + 1, b), ly, rbv(d, RB_A00 + 2, b) * lz));
        const float ry = __builtin_fmaf(rbv(d, RB_A00 + 3, b), lx, __builtin_fmaf(rbv(d, RB_A00 + 4, b), ly, rbv(d, RB_A00 + 5, b) * lz));
        const float rz = __builtin_fmaf(rbv(d, RB_A00 + 6, b), lx, __builtin_fmaf(rbv(d, RB_A00 + 7, b), ly, rbv(d, RB_A00 + 8, b) * lz));
        if (!predicted)
        {
            const float psi = d.rpos[i].w;
            d.rpos[i] = make_float4(rbv(d, RB_XCM + 0, b) + rx, rbv(d, RB_XCM + 1, b) + ry, rbv(d, RB_XCM + 2, b) + rz, psi);
        }
        const uint32_t nb = d.nRB;
        const float wx = predicted ? d.rbPred[6 * nb + b] : rbv(d, RB_WX + 0, b);
        const float wy = predicted ? d.rbPred[7 * nb + b] : rbv(d, RB_WX + 1, b);
        const float wz = predicted ? d.rbPred[8 * nb + b] : rbv(d, RB_WX + 2, b);
        const float vbx = predicted ? d.rbPred[0 * nb + b] : rbv(d, RB_VX + 0, b);
        const float vby = predicted ? d.rbPred[1 * nb + b] : rbv(d, RB_VX + 1, b);
        const float vbz = predicted ? d.rbPred[2 * nb + b] : rbv(d, RB_VX + 2, b);
        float4* out = predicted ? d.rvp : d.rvel;
        out[i] = make_float4(vbx + wy * rz - wz * ry, vby + wz * rx - wx * rz, vbz + wx * ry - wy * rx, out[i].w);
    }

    // PredictRigidBodies, cpu_backend.cpp:2174-2235
    __device__ __forceinline__ void rigidPredictOne(const Dev& d, uint32_t b)
    {
        const Consts& k = d.consts[d.bodyScene[b]];
        const uint32_t nb = d.nRB;
        const float mi = rbv(d, RB_MASS_INV, b);
        float lp[3];
        for (int a = 0; a < 3; a++)
        {
            d.rbPred[a * nb + b] = rbv(d, RB_VX + a, b) + k.dt * d.rbForce[a * nb + b] * mi;
            lp[a] = rbv(d, RB_LX + a, b) + k.dt * d.rbForce[(3 + a) * nb + b];
            d.rbPred[(3 + a) * nb + b] = lp[a];
        }
        float A[9], B[9], T[9], I[9];
        for (int q = 0; q < 9; q++) { A[q] = rbv(d, RB_A00 + q, b); B[q] = rbv(d, RB_I00 + q, b); }
        for (int r = 0; r < 3; r++)
        {
            for (int c = 0; c < 3; c++) { T[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c]; }
        }
        for (int r = 0; r < 3; r++)
        {
            for (int c = 0; c < 3; c++) { I[3 * r + c] = T[3 * r] * A[3 * c] + T[3 * r + 1] * A[3 * c + 1] + T[3 * r + 2] * A[3 * c + 2]; }
        }
        for (int r = 0; r < 3; r++)
        {
            d.rbPred[(6 + r) * nb + b] = __builtin_fmaf(I[3 * r], lp[0], __builtin_fmaf(I[3 * r + 1], lp[1], I[3 * r + 2] * lp[2]));
        }
    }

}
