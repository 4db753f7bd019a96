// tsp_capi.hip -- host orchestration of the gfx950 IISPH step and its C ABI (include/tsp_hip.h).
//
// One context = one device, one HIP stream, and any number of independent scenes whose state
// stays resident in HBM. The scenes share every array: each owns a segment of the fluid,
// boundary, rigid and body arrays (fluid segments aligned to the 256-thread block), a range of the
// 32-bit sort-key space and a range of the cell tables, so one launch of each kernel advances all
// scenes. A context with one scene is the plain backend.
//
// The step sequence follows SimBackend_CPU::Step_IISPH (CPU/backend/sph_solver.cpp:34-319). The
// differences from the reference's CUDA host (GPU/backend/gpu_backend.cpp:1274-1594) are that list
// sizes and solver convergence live on the device: the host reads one small status block after the
// neighbor count (pool growth and frame hysteresis) and the scenes' solver states after groups of
// Jacobi sweeps; sweeps queued past a scene's convergence exit at once for that scene.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <unordered_map>
#include <string>
#include <vector>

#include "scene_plan.h"
#include "tsp_device.h"
#include "tsp_hip.h"

namespace tsp
{
    void launchResetSearch(hipStream_t, const Dev&);
    void launchKeys(hipStream_t, const Dev&, int, const float4*, uint32_t, bool, uint32_t*, uint32_t*);
    void launchKeysSorted(hipStream_t, const Dev&, const float4*, uint32_t, uint32_t*, uint32_t*);
    bool launchRadixSort(hipStream_t, bool, uint32_t*, uint32_t*, uint32_t*, uint32_t*, uint32_t, uint32_t, uint32_t*, uint32_t*,
                         uint32_t*, bool);
    bool radixSortSingleBlock(uint32_t n);
    void launchPermute(hipStream_t, const Gather&, const uint32_t*);
    void launchClearCells(hipStream_t, const Dev&, const uint32_t*, uint32_t, uint2*);
    void launchClearAll(hipStream_t, uint2*, uint64_t);
    void launchBuildCells(hipStream_t, const Dev&, const uint32_t*, uint32_t, uint2*);
    void launchQueryCount(hipStream_t, const Dev&);
    void launchScanSlices(hipStream_t, const Dev&, uint32_t*, uint32_t*);
    void launchQueryFill(hipStream_t, const Dev&);
    void launchQueryRigid(hipStream_t, const Dev&, int, uint32_t*);
    void launchStepGate(hipStream_t, const Dev&);
    void launchSolveGate(hipStream_t, const Dev&);
    void launchClearStop(hipStream_t, const Dev&);
    void launchContactOffsets(hipStream_t, const Dev&, uint32_t*);
    void launchContactSolve(hipStream_t, const Dev&, uint32_t, uint32_t*);
    void launchExportSurface(hipStream_t, const Dev&, uint32_t, uint32_t, uint32_t, uint32_t*, uint32_t*, float*, float*);
    void launchBuildFluid(hipStream_t, const Dev&, uint32_t, uint32_t, const BuildFluid&);
    void launchBuildBoundary(hipStream_t, const BuildBoundary&, uint32_t, float4*, float*, uint64_t*, float*, uint32_t*,
                             uint32_t*, uint32_t*, uint32_t*, uint32_t*, uint32_t*);
    void launchHaloBuild(hipStream_t, const Dev&);
    void launchAabb(hipStream_t, const Dev&, int, const float4*, uint32_t, uint32_t*);
    void launchDensity(hipStream_t, const Dev&);
    void launchNonPressure(hipStream_t, const Dev&);
    void launchSourceDiag(hipStream_t, const Dev&);
    void launchPressureAcc(hipStream_t, const Dev&, bool);
    void launchJacobi(hipStream_t, const Dev&);
    void launchFinalize(hipStream_t, const Dev&);
    void launchIntegration(hipStream_t, const Dev&);
    void launchRigidUpdate(hipStream_t, const Dev&, bool);
    void launchRigidClearGravity(hipStream_t, const Dev&);
    void launchRigidForces(hipStream_t, const Dev&, bool);
    void launchRigidReduce(hipStream_t, const Dev&);
    void launchRigidPredict(hipStream_t, const Dev&);
    void launchRigidIntegrate(hipStream_t, const Dev&);
    void launchStatsCounts(hipStream_t, const uint16_t*, const uint32_t*, uint32_t, uint32_t, uint32_t*, uint64_t*);
    void launchStatsFluid(hipStream_t, const float4*, uint32_t, uint32_t, float, double*, double*);
    hipError_t statsScratchInit(int);
    bool statsScratch(int, uint32_t**, uint64_t**, double**, double**, uint32_t*);
}

using namespace tsp;

namespace
{
    constexpr int kPopF = 0, kPopB = 1, kPopR = 2;

    struct Error
    {
        int code;
        std::string msg;
    };

    void hipCheck(hipError_t e, const char* what)
    {
        if (e != hipSuccess) { throw Error{ TSP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e) }; }
    }

    uint32_t roundUp(uint32_t n, uint32_t m) { return (n + m - 1) / m * m; }

    // events around each named launch group while profiling is enabled
    struct Profiler
    {
        bool on = false;
        struct Rec
        {
            std::string name;
            hipEvent_t a, b;
        };
        std::vector<Rec> pending;
        std::vector<hipEvent_t> pool;
        std::map<std::string, std::pair<uint64_t, double>> acc;
        std::vector<std::string> order;

        hipEvent_t get()
        {
            if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
            hipEvent_t e;
            hipCheck(hipEventCreate(&e), "hipEventCreate");
            return e;
        }
        void resolve()
        {
            for (Rec& r : pending)
            {
                hipCheck(hipEventSynchronize(r.b), "hipEventSynchronize");
                float ms = 0.0f;
                hipCheck(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
                if (!acc.count(r.name)) { order.push_back(r.name); }
                auto& v = acc[r.name];
                v.first++;
                v.second += ms;
                pool.push_back(r.a);
                pool.push_back(r.b);
            }
            pending.clear();
        }
        ~Profiler()
        {
            for (Rec& r : pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
            for (hipEvent_t e : pool) { (void)hipEventDestroy(e); }
        }
    };

    struct Scope
    {
        Profiler& p;
        hipStream_t s;
        const char* name;
        hipEvent_t a = nullptr;
        Scope(Profiler& p_, hipStream_t s_, const char* n) : p(p_), s(s_), name(n)
        {
            if (p.on) { a = p.get(); (void)hipEventRecord(a, s); }
        }
        ~Scope()
        {
            if (!p.on || !a) { return; }
            hipEvent_t b = p.get();
            (void)hipEventRecord(b, s);
            p.pending.push_back({ name, a, b });
        }
    };

    // constants of the kernels (solver_constants.h:20-36, simd.h:359-370, hoisted kernel constants)
    Consts makeConsts(const tsp_params& p, uint32_t nF)
    {
        Consts k{};
        const double h = static_cast<double>(p.particle_size);
        const double hinv = 1.0 / h;
        const double rho0 = static_cast<double>(p.rest_density);
        k.hinv = static_cast<float>(hinv);
        k.hsq = static_cast<float>(h * h);
        k.alpha = static_cast<float>(hinv * hinv * hinv / (4.0 * 3.14159265358979323846));
        k.rho0SqInv = static_cast<float>(1.0 / (rho0 * rho0));
        k.dt = p.time_step;
        k.r2 = p.kernel_radius_sq;
        k.mass = p.mass;
        k.rho0 = p.rest_density;
        k.ksize = 1.0f / k.hinv;
        k.ktwo = 2.0f / k.hinv;
        k.kinSlope = 9.0f * k.alpha * k.hinv * k.hinv * k.hinv;
        k.kinOff = -12.0f * k.alpha * k.hinv * k.hinv;
        k.koutScale = -3.0f * k.alpha * k.hinv * k.hinv * k.hinv;
        k.weightScale = k.mass * k.alpha * k.hinv * k.hinv * k.hinv;
        k.hsqEps = 0.01f * k.hsq;
        k.visF = 10.0f * p.viscosity_fluid;
        k.visB = 10.0f * p.viscosity_boundary;
        k.gx = p.gravity[0]; k.gy = p.gravity[1]; k.gz = p.gravity[2];
        k.cFacF = -k.mass * k.rho0SqInv;
        k.cFacB = -2.0f * p.press_gamma * k.rho0SqInv;
        k.dFacB = -k.mass * k.rho0SqInv;
        k.omega = p.jacobi_omega;
        k.warm = p.jacobi_warm_start;
        k.presDis = 1.0f / (k.rho0 * k.rho0);
        k.presOuterB = 2.0f * p.press_gamma * k.presDis;
        k.presOuterFF = k.mass * k.presDis;
        k.dtsq = k.dt * k.dt;
        k.densityDevTarget = p.density_dev_target;
        k.minIters = p.min_solver_iters;
        k.maxIters = p.max_solver_iters;
        k.fixedIters = p.fixed_iters;
        k.termCond = p.termination_cond;
        k.nF = nF;
        k.contactDevTarget = p.rigid_contact_density_dev_target;
        k.contactMaxIters = p.rigid_contact_max_solver_iters;
        return k;
    }

    void validateParams(const tsp_params& p)
    {
        if (!(p.particle_size > 0.0f) || !(p.kernel_radius > 0.0f) || !(p.kernel_radius_sq > 0.0f)
            || !(p.rest_density > 0.0f) || !(p.time_step > 0.0f) || !(p.grid_cell_size_ratio > 0.0f)
            || p.max_solver_iters < 1)
        {
            throw Error{ TSP_ERR_INVALID, "invalid solver parameters" };
        }
    }

    // SolveRigidContacts' check (rigid_solver.cpp:604-610), made before every step of a scene
    // with bodies as in the reference
    void validateContactParams(const tsp_params& p)
    {
        if (!std::isfinite(p.rigid_contact_density_dev_target) || !(p.rigid_contact_density_dev_target > 0.0f)
            || p.rigid_contact_max_solver_iters == 0)
        {
            throw Error{ TSP_ERR_INVALID, "Invalid rigid contact solver tolerance or iteration limit." };
        }
    }

    // one scene as the host sees it
    struct SceneHost
    {
        uint32_t id = 0;
        tsp_params params{};
        float domMin[3] = { 0, 0, 0 }, domMax[3] = { 0, 0, 0 };
        bool domainSet = false;
        uint32_t nF = 0, nB = 0, nR = 0, nRB = 0;
        // place in the arena
        uint32_t offF = 0, segF = 0, blk0 = 0, nBlk = 0, offB = 0, offR = 0, offRB = 0;
        // static rigid topology (scene-local body of each rigid particle) and body records
        std::vector<uint32_t> rbodyLocal, bodyId, bodyCount;
        // cell frame
        Frame frame{};
        bool frameValid = false;
        uint32_t frameBuilds = 0;
        // solver bookkeeping
        uint64_t step = 0;
        double time = 0.0;
        uint64_t sweepsTotal = 0;
        uint32_t lastSweeps = 0;
        uint32_t flags = 0;
        Solve last{};
    };

    // where a scene of a new layout comes from: a resident scene, or a host state
    struct Source
    {
        int keep = -1;
        const tsp_state_view* view = nullptr;
        const tsp::scene::Plan* plan = nullptr;   // generated on the device (the view holds its rigid body)
        SceneHost host;
    };

    BuildVec vec(const tsp::scene::V3& v) { return BuildVec{ v.x, v.y, v.z }; }

    // the view of a plan: counts, the rigid particles and bodies; fluid and boundary come from the
    // device builder
    tsp_state_view planView(const tsp::scene::Plan& plan)
    {
        tsp_state_view v{};
        v.n_fluid = static_cast<uint32_t>(plan.nF);
        v.n_boundary = static_cast<uint32_t>(plan.nB);
        v.n_rigid = static_cast<uint32_t>(plan.idR.size());
        v.n_bodies = static_cast<uint32_t>(plan.idRB.size());
        auto& P = const_cast<tsp::scene::Plan&>(plan);
        v.rigid_id = P.idR.empty() ? nullptr : P.idR.data();
        v.rigid_body = P.bodyR.empty() ? nullptr : P.bodyR.data();
        for (int f = 0; f < TSP_R_COUNT; f++) { v.rigid[f] = P.R[f].empty() ? nullptr : P.R[f].data(); }
        v.body_id = P.idRB.empty() ? nullptr : P.idRB.data();
        v.body_count = P.countRB.empty() ? nullptr : P.countRB.data();
        for (int f = 0; f < TSP_RB_COUNT; f++) { v.body[f] = P.RB[f].empty() ? nullptr : P.RB[f].data(); }
        return v;
    }
}

struct tsp_ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    // the rigid particles' keys and single-block sort run beside the fluid sort on a second stream
    // of the same priority, joined before the rigid cell table is built
    hipStream_t side = nullptr;
    hipEvent_t forkEv = nullptr, joinEv = nullptr;
    std::string err;
    tsp_params defaults{};
    bool paramsSet = false;
    float domMin[3] = { 0, 0, 0 }, domMax[3] = { 0, 0, 0 };
    bool domainSet = false;
    uint32_t nextId = 1;

    std::vector<SceneHost> scenes;
    Dev d{};
    // second buffers of the re-order (swapped with the ones in d after each gather)
    float4 *pos2 = nullptr, *vel2 = nullptr, *vp2 = nullptr, *ap2 = nullptr;
    float *source2 = nullptr, *diag2 = nullptr;
    uint64_t* idF2 = nullptr;
    float4* bpos2 = nullptr;
    float* bvol2 = nullptr;
    uint64_t *idB = nullptr, *idB2 = nullptr;
    uint64_t* idR = nullptr;
    uint32_t* radixHist = nullptr;
    uint32_t* sliceUnits = nullptr; // 3 x (slices + 1) scan buffer
    uint32_t* aabb = nullptr;       // 6 per scene
    std::vector<void*> arena;       // particle and body arrays (replaced by a relayout)
    std::vector<void*> work;        // search and scene-table buffers sized by the layout
    std::vector<void*> frameAllocs; // cell tables
    std::vector<void*> poolAllocs;  // SELL-32 pools and the RF list
    std::vector<void*> contactAllocs; // RB/RR lists and the contact boundary arrays (grown on demand)
    uint64_t cellCap = 0;
    uint64_t cellLimit = 0xffffffffull;   // most cells of all scenes' frames (tsp_set_cell_limit)

    // rigid contacts of the last search
    bool contactLists = false;      // RB or RR entries exist: the lists are stored
    uint64_t contactRB = 0;         // RB entries (bounds the contact boundary particles)
    uint32_t* contactScanTmp = nullptr;

    bool layoutDirty = true;        // key bases and cell offsets must be assigned again
    bool listsValid = false;        // a neighbor search ran since the last layout
    bool cellsClean = false;        // fluid and rigid cell tables hold no runs
    uint32_t keyBits = 1;

    // the search status and the scenes' solver states share one device block and one pinned
    // host block, so one copy reads them all back
    static constexpr size_t kSolveOff = 256;
    unsigned char* hostBlock = nullptr;
    Status* hostStatus = nullptr;   // pinned, at the start of hostBlock
    Solve* hostSolve = nullptr;     // pinned, solveCap entries from kSolveOff
    uint32_t solveCap = 0;
    Status last{};

    Profiler prof;

    // instantiated graphs of sweep groups, keyed by the kernels' arguments and the sweep count
    std::map<std::string, hipGraphExec_t> graphs;

    void dropGraphs()
    {
        for (auto& g : graphs) { (void)hipGraphExecDestroy(g.second); }
        graphs.clear();
    }

    ~tsp_ctx()
    {
        dropGraphs();
        dropStepGraphs();
        releaseList(statAllocs);
        releaseList(poolAllocs);
        releaseList(contactAllocs);
        releaseList(expAllocs);
        releaseList(frameAllocs);
        releaseList(work);
        releaseList(arena);
        if (hostBlock) { (void)hipHostFree(hostBlock); }
        if (forkEv) { (void)hipEventDestroy(forkEv); }
        if (joinEv) { (void)hipEventDestroy(joinEv); }
        if (side) { (void)hipStreamDestroy(side); }
        if (stream) { (void)hipStreamDestroy(stream); }
    }

    // device memory the context holds, by what it stores (tsp_memory)
    enum MemCat : uint8_t { MEM_PF, MEM_PB, MEM_PR, MEM_SF, MEM_SB, MEM_SR, MEM_LF, MEM_LR, MEM_MISC, kMemCats };
    std::unordered_map<void*, std::pair<uint64_t, uint8_t>> allocated;
    uint64_t memBytes[kMemCats] = {};
    uint64_t memTotal = 0, memPeak = 0;

    template <typename T>
    T* alloc(size_t n, std::vector<void*>& list, MemCat cat = MEM_MISC)
    {
        void* p = nullptr;
        const uint64_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        hipCheck(hipMalloc(&p, bytes), "hipMalloc");
        list.push_back(p);
        allocated[p] = { bytes, static_cast<uint8_t>(cat) };
        memBytes[cat] += bytes;
        memTotal += bytes;
        memPeak = std::max(memPeak, memTotal);
        return static_cast<T*>(p);
    }

    void releaseList(std::vector<void*>& list)
    {
        for (void* p : list)
        {
            const auto it = allocated.find(p);
            if (it != allocated.end())
            {
                memBytes[it->second.second] -= it->second.first;
                memTotal -= it->second.first;
                allocated.erase(it);
            }
            (void)hipFree(p);
        }
        list.clear();
    }

    void sync() { hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize"); }

    // a synchronous device-to-host copy on the context stream
    void d2h(void* dst, const void* src, size_t bytes)
    {
        if (bytes) { hipCheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream), "D2H"); sync(); }
    }

    void h2d(void* dst, const void* src, size_t bytes)
    {
        if (bytes) { hipCheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream), "H2D"); sync(); }
    }

    void d2d(void* dst, const void* src, size_t bytes)
    {
        if (bytes) { hipCheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream), "D2D"); }
    }

    void zero(void* dst, size_t bytes)
    {
        if (bytes) { hipCheck(hipMemsetAsync(dst, 0, bytes, stream), "hipMemsetAsync"); }
    }

    uint32_t nScenes() const { return static_cast<uint32_t>(scenes.size()); }

    int indexOf(uint32_t id) const
    {
        for (uint32_t s = 0; s < nScenes(); s++) { if (scenes[s].id == id) { return static_cast<int>(s); } }
        return -1;
    }

    // the search status and every scene's solver state
    void readStatus()
    {
        hipCheck(hipGetLastError(), "kernel launch");
        hipCheck(hipMemcpyAsync(hostBlock, d.status, kSolveOff + sizeof(Solve) * nScenes(), hipMemcpyDeviceToHost, stream),
                 "status D2H");
        sync();
        prof.resolve();
        last = *hostStatus;
        if (last.scanFail) { throw Error{ TSP_ERR_STATE, "a one-pass scan gave up waiting for a predecessor tile" }; }
    }

    void uploadConsts()
    {
        std::vector<Consts> k(nScenes());
        bool fast = true;
        for (uint32_t s = 0; s < nScenes(); s++)
        {
            k[s] = makeConsts(scenes[s].params, scenes[s].nF);
            // the short exact quotient of the gradient factor needs its operands in range
            const float a = std::fabs(k[s].koutScale);
            fast = fast && a >= 0x1p-60f && a <= 0x1p60f && k[s].ksize >= 0x1p-30f && k[s].ktwo <= 0x1p31f;
        }
        d.fastDiv = fast ? 1u : 0u;
        h2d(d.consts, k.data(), sizeof(Consts) * k.size());
    }

    // --- pools of the SELL-32 lists ---
    void allocPools(const uint64_t cap[3], uint64_t rfCap)
    {
        releaseList(poolAllocs);
        for (int r = 0; r < 3; r++)
        {
            d.poolCap[r] = cap[r];
            d.pool[r] = alloc<uint16_t>(cap[r], poolAllocs, MEM_LF);
        }
        d.rfCap = rfCap;
        d.rfIdx = alloc<uint32_t>(rfCap, poolAllocs, MEM_LR);
#if TSP_TUNE_KG_STORE
        d.kgFF = alloc<float>(cap[REL_FF], poolAllocs, MEM_LF);
#else
        d.kgFF = nullptr;
#endif
    }

    // --- a planned scene's fluid and boundary generated in the new arena (n) at h's offsets ---
    void buildOnDevice(const Dev& n, const SceneHost& h, const tsp::scene::Plan& plan, uint64_t* nidB)
    {
        const tsp::scene::Params& p = plan.p;
        const tsp::scene::FluidBlock& fb = plan.fluid;
        BuildFluid bf{};
        bf.origin = vec(fb.blk.origin); bf.u = vec(fb.blk.u); bf.v = vec(fb.blk.v); bf.w = vec(fb.blk.w);
        bf.uLen = fb.blk.uLen; bf.vLen = fb.blk.vLen; bf.wLen = fb.blk.wLen;
        bf.uOff = fb.uOff; bf.vOff = fb.vOff; bf.wOff = fb.wOff;
        bf.radius = fb.radius; bf.jitAmp = fb.jitAmp; bf.size = p.size; bf.rho0 = p.rho0;
        bf.seed = fb.seed;
        bf.nu = fb.nu; bf.nv = fb.nv; bf.nw = fb.nw;
        launchBuildFluid(stream, n, h.offF, h.nF, bf);
        if (!h.nB) { return; }
        BuildBoundary bb{};
        bb.nFaces = plan.nFaces;
        bb.layers = p.layers;
        bb.size = p.size; bb.sizeInv = p.sizeInv; bb.radiusSqu = p.radiusSqu; bb.alpha = p.alpha;
        bb.rho0 = p.rho0; bb.massB = p.massB;
        bb.cellSizeInv = 1.0f / p.radius;
        // cell bounds of the layers: each face's positions are monotone in i, j and the layer, so
        // its corners bound it
        float lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
        for (uint32_t q = 0; q < plan.nFaces; q++)
        {
            const tsp::scene::BoundaryFace& f = plan.faces[q];
            BuildFace& g = bb.faces[q];
            g.origin = vec(f.f.origin); g.u = vec(f.f.u); g.v = vec(f.f.v); g.n = vec(f.f.n);
            g.uLen = f.f.uLen; g.vLen = f.f.vLen;
            g.nu = f.nu; g.nv = f.nv; g.first = f.first;
            for (uint32_t layer : { 0u, p.layers - 1u })
            {
                for (uint32_t i : { 0u, f.nu - 1u })
                {
                    for (uint32_t j : { 0u, f.nv - 1u })
                    {
                        const tsp::scene::V3 x = tsp::scene::BoundaryPosition(f.f, p.size, layer, i, j);
                        const float c[3] = { x.x, x.y, x.z };
                        for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], c[a]); hi[a] = std::max(hi[a], c[a]); }
                    }
                }
            }
        }
        uint32_t totalBits = 0;
        for (int a = 0; a < 3; a++)
        {
            bb.cellMin[a] = static_cast<int32_t>(std::floor(lo[a] * bb.cellSizeInv)) - 1;
            bb.cellMax[a] = static_cast<int32_t>(std::floor(hi[a] * bb.cellSizeInv)) + 1;
            const uint64_t range = static_cast<uint64_t>(static_cast<int64_t>(bb.cellMax[a]) - bb.cellMin[a]) + 1;
            uint32_t bits = 1;
            while ((1ull << bits) < range) { bits++; }
            bb.bits[a] = bits;
            totalBits += bits;
        }
        if (totalBits > 32) { throw Error{ TSP_ERR_CAPACITY, "container too large for the device builder's 32-bit cell keys" }; }
        const uint32_t nB = h.nB;
        const uint64_t nTiles = (nB + 2047ull) / 2048ull;
        std::vector<void*> tmp;
        try
        {
            float* xyz = alloc<float>(3ull * nB, tmp);
            uint32_t* key = alloc<uint32_t>(nB, tmp);
            uint32_t* idx = alloc<uint32_t>(nB, tmp);
            uint32_t* key2 = alloc<uint32_t>(nB, tmp);
            uint32_t* idx2 = alloc<uint32_t>(nB, tmp);
            uint32_t* hist = alloc<uint32_t>(256ull * nTiles, tmp);
            uint32_t* sums = alloc<uint32_t>(scanTmpWords(256ull * nTiles), tmp);
            zero(sums, sizeof(uint32_t) * scanTmpWords(256ull * nTiles));
            launchBuildBoundary(stream, bb, nB, n.bpos + h.offB, n.bvol + h.offB, nidB + h.offB, xyz, key, idx, key2, idx2,
                                hist, sums);
            hipCheck(hipGetLastError(), "device builder launch");
            uint32_t scanFailed = 0;
            d2h(&scanFailed, sums + 2, sizeof(uint32_t));
            if (scanFailed) { throw Error{ TSP_ERR_STATE, "a one-pass scan gave up waiting for a predecessor tile" }; }
        }
        catch (...)
        {
            (void)hipStreamSynchronize(stream);
            releaseList(tmp);
            throw;
        }
        releaseList(tmp);
    }

    // --- storage of the contact solve sized by the RB and RR entries of a search ---
    void ensureContactCapacity(uint64_t rbTotal, uint64_t rrTotal)
    {
        if (rbTotal <= d.rbCap && rrTotal <= d.rrCap && !contactAllocs.empty()) { return; }
        if (rbTotal > 0x7fffffffull || rrTotal > 0x7fffffffull)
        {
            throw Error{ TSP_ERR_CAPACITY, "rigid contact lists exceed 2^31 entries" };
        }
        const uint64_t rbCap = std::max<uint64_t>(d.rbCap, rbTotal + rbTotal / 4 + 256);
        const uint64_t rrCap = std::max<uint64_t>(d.rrCap, rrTotal + rrTotal / 4 + 256);
        sync();
        releaseList(contactAllocs);
        d.rbCap = rbCap;
        d.rrCap = rrCap;
        d.rbIdx = alloc<uint32_t>(rbCap, contactAllocs, MEM_LR);
        d.rrIdx = alloc<uint32_t>(rrCap, contactAllocs, MEM_LR);
        d.cbLocal = alloc<uint32_t>(rbCap, contactAllocs, MEM_LR);
        d.brOff = alloc<uint32_t>(rbCap + 1, contactAllocs, MEM_LR);
        d.brPos = alloc<uint32_t>(rbCap + 1, contactAllocs, MEM_LR);
        d.brIdx = alloc<uint32_t>(rbCap, contactAllocs, MEM_LR);
        d.cB = alloc<float>(static_cast<uint64_t>(kCB) * rbCap, contactAllocs);
        d.cErrB = alloc<double>(rbCap, contactAllocs);
        // block sums of the contact scans (boundary flags, BR offsets)
        const uint64_t m = std::max<uint64_t>(d.nB + 1ull, rbCap + 1);
        contactScanTmp = alloc<uint32_t>(scanTmpWords(m), contactAllocs);
        zero(contactScanTmp, sizeof(uint32_t) * scanTmpWords(m));
    }

    // SolveRigidContacts (CPU/backend/rigid_solver.cpp:602-799) for every scene whose rigid
    // particles have RB or RR entries: the contact lists of all scenes, then one block per scene
    // runs its solve to the end and adds the loads to its bodies. Nothing is read back here; the
    // reports and any failure come with the status read at the end of the step call.
    void solveContacts()
    {
        Scope sc(prof, stream, "contact");
        launchContactSolve(stream, d, static_cast<uint32_t>(contactRB), contactScanTmp);
    }

    // --- the arena: the particles and bodies of every scene, laid out anew. Resident scenes
    //     move on the device; new scenes are packed from their host state ---
    void relayout(std::vector<Source>& src)
    {
        sync();
        std::vector<SceneHost> ns(src.size());
        uint64_t oF = 0, oB = 0, oR = 0, oRB = 0;
        for (size_t q = 0; q < src.size(); q++)
        {
            SceneHost& h = ns[q];
            h = src[q].keep >= 0 ? scenes[static_cast<size_t>(src[q].keep)] : src[q].host;
            h.offF = static_cast<uint32_t>(oF);
            h.segF = roundUp(h.nF, kBlock);
            h.blk0 = static_cast<uint32_t>(oF / kBlock);
            h.nBlk = h.segF / kBlock;
            h.offB = static_cast<uint32_t>(oB);
            h.offR = static_cast<uint32_t>(oR);
            h.offRB = static_cast<uint32_t>(oRB);
            oF += h.segF; oB += h.nB; oR += h.nR; oRB += h.nRB;
        }
        if (oF > 0x7fffffffull || oB > 0x7fffffffull || oR > 0x7fffffffull || oRB > 0x7fffffffull)
        {
            throw Error{ TSP_ERR_CAPACITY, "particle count exceeds the local particle limit" };
        }
        if (ns.size() > 0xffff) { throw Error{ TSP_ERR_CAPACITY, "more than 65535 scenes in one context" }; }
        const uint32_t nF = static_cast<uint32_t>(oF), nB = static_cast<uint32_t>(oB);
        const uint32_t nR = static_cast<uint32_t>(oR), nRB = static_cast<uint32_t>(oRB);
        const uint32_t S = static_cast<uint32_t>(ns.size());
        const size_t nRB1 = std::max<uint32_t>(nRB, 1), oldRB1 = std::max<uint32_t>(d.nRB, 1);

        std::vector<void*> A;
        Dev n = d;
        float4 *npos2, *nvel2, *nvp2, *nap2, *nbpos2;
        float *nsource2, *ndiag2, *nbvol2;
        uint64_t *nidF2, *nidB, *nidB2, *nidR;
        try
        {
            n.nF = nF; n.nB = nB; n.nR = nR; n.nRB = nRB; n.nScenes = S;
            n.pos = alloc<float4>(nF, A, MEM_PF); npos2 = alloc<float4>(nF, A, MEM_PF);
            n.vel = alloc<float4>(nF, A, MEM_PF); nvel2 = alloc<float4>(nF, A, MEM_PF);
            n.vp = alloc<float4>(nF, A, MEM_PF); nvp2 = alloc<float4>(nF, A, MEM_PF);
            n.ap = alloc<float4>(nF, A, MEM_PF); nap2 = alloc<float4>(nF, A, MEM_PF);
            n.source = alloc<float>(nF, A, MEM_PF); nsource2 = alloc<float>(nF, A, MEM_PF);
            n.diag = alloc<float>(nF, A, MEM_PF); ndiag2 = alloc<float>(nF, A, MEM_PF);
            n.id = alloc<uint64_t>(nF, A, MEM_PF); nidF2 = alloc<uint64_t>(nF, A, MEM_PF);
            // the padding slots of the fluid segments hold zeros
            zero(n.pos, sizeof(float4) * nF); zero(n.vel, sizeof(float4) * nF);
            zero(n.vp, sizeof(float4) * nF); zero(n.ap, sizeof(float4) * nF);
            zero(n.source, sizeof(float) * nF); zero(n.diag, sizeof(float) * nF);
            zero(n.id, sizeof(uint64_t) * nF);
            n.bpos = alloc<float4>(nB, A, MEM_PB); nbpos2 = alloc<float4>(nB, A, MEM_PB);
            n.bvol = alloc<float>(nB, A, MEM_PB); nbvol2 = alloc<float>(nB, A, MEM_PB);
            nidB = alloc<uint64_t>(nB, A, MEM_PB); nidB2 = alloc<uint64_t>(nB, A, MEM_PB);
            n.rxrel = alloc<float>(nR, A, MEM_PR); n.ryrel = alloc<float>(nR, A, MEM_PR); n.rzrel = alloc<float>(nR, A, MEM_PR);
            n.rpos = alloc<float4>(nR, A, MEM_PR); n.rvel = alloc<float4>(nR, A, MEM_PR); n.rvp = alloc<float4>(nR, A, MEM_PR);
            n.rfx = alloc<float>(nR, A, MEM_PR); n.rfy = alloc<float>(nR, A, MEM_PR); n.rfz = alloc<float>(nR, A, MEM_PR);
            n.rbody = alloc<uint32_t>(nR, A, MEM_PR);
            nidR = alloc<uint64_t>(nR, A, MEM_PR);
            n.rb = alloc<float>(static_cast<size_t>(TSP_RB_COUNT) * nRB1, A, MEM_PR);
            n.rbPred = alloc<float>(9 * nRB1, A, MEM_PR);
            n.rbForce = alloc<float>(6 * nRB1, A, MEM_PR);
            zero(n.rbPred, sizeof(float) * 9 * nRB1);
            zero(n.rbForce, sizeof(float) * 6 * nRB1);
            n.byBodyOff = alloc<uint32_t>(nRB + 1ull, A, MEM_PR);
            n.byBodyIdx = alloc<uint32_t>(nR, A, MEM_PR);
            n.blockScene = alloc<uint32_t>(nF / kBlock, A, MEM_PF);
            n.bodyScene = alloc<uint32_t>(nRB, A, MEM_PR);
            n.sceneB = alloc<uint16_t>(nB, A, MEM_PB);

            const auto field = [](const float* f, uint32_t i, float dflt) { return f ? f[i] : dflt; };
            std::vector<uint64_t> seq;
            const auto ids = [&](const uint64_t* given, uint32_t cnt) -> const uint64_t* {
                if (given) { return given; }
                seq.resize(cnt);
                for (uint32_t i = 0; i < cnt; i++) { seq[i] = i; }
                return seq.data();
            };
            for (size_t q = 0; q < src.size(); q++)
            {
                const SceneHost& h = ns[q];
                if (src[q].keep >= 0)
                {
                    const SceneHost& o = scenes[static_cast<size_t>(src[q].keep)];
                    d2d(n.pos + h.offF, d.pos + o.offF, sizeof(float4) * h.nF);
                    d2d(n.vel + h.offF, d.vel + o.offF, sizeof(float4) * h.nF);
                    d2d(n.vp + h.offF, d.vp + o.offF, sizeof(float4) * h.nF);
                    d2d(n.ap + h.offF, d.ap + o.offF, sizeof(float4) * h.nF);
                    d2d(n.source + h.offF, d.source + o.offF, sizeof(float) * h.nF);
                    d2d(n.diag + h.offF, d.diag + o.offF, sizeof(float) * h.nF);
                    d2d(n.id + h.offF, d.id + o.offF, sizeof(uint64_t) * h.nF);
                    d2d(n.bpos + h.offB, d.bpos + o.offB, sizeof(float4) * h.nB);
                    d2d(n.bvol + h.offB, d.bvol + o.offB, sizeof(float) * h.nB);
                    d2d(nidB + h.offB, idB + o.offB, sizeof(uint64_t) * h.nB);
                    d2d(n.rxrel + h.offR, d.rxrel + o.offR, sizeof(float) * h.nR);
                    d2d(n.ryrel + h.offR, d.ryrel + o.offR, sizeof(float) * h.nR);
                    d2d(n.rzrel + h.offR, d.rzrel + o.offR, sizeof(float) * h.nR);
                    d2d(n.rpos + h.offR, d.rpos + o.offR, sizeof(float4) * h.nR);
                    d2d(n.rvel + h.offR, d.rvel + o.offR, sizeof(float4) * h.nR);
                    d2d(n.rvp + h.offR, d.rvp + o.offR, sizeof(float4) * h.nR);
                    d2d(nidR + h.offR, idR + o.offR, sizeof(uint64_t) * h.nR);
                    for (size_t f = 0; f < TSP_RB_COUNT; f++)
                    {
                        d2d(n.rb + f * nRB1 + h.offRB, d.rb + f * oldRB1 + o.offRB, sizeof(float) * h.nRB);
                    }
                    for (size_t f = 0; f < 9; f++)
                    {
                        d2d(n.rbPred + f * nRB1 + h.offRB, d.rbPred + f * oldRB1 + o.offRB, sizeof(float) * h.nRB);
                    }
                    for (size_t f = 0; f < 6; f++)
                    {
                        d2d(n.rbForce + f * nRB1 + h.offRB, d.rbForce + f * oldRB1 + o.offRB, sizeof(float) * h.nRB);
                    }
                    continue;
                }
                const tsp_state_view* s = src[q].view;
                if (src[q].plan) { buildOnDevice(n, h, *src[q].plan, nidB); }
                // fluid: the SoA fields of the view packed into the float4 records
                if (!src[q].plan)
                {
                    std::vector<float4> r[4];
                    for (auto& v : r) { v.resize(h.nF); }
                    std::vector<float> src1(h.nF), diag1(h.nF);
                    float* const* f = s->fluid;
                    for (uint32_t i = 0; i < h.nF; i++)
                    {
                        r[0][i] = make_float4(field(f[TSP_F_X], i, 0), field(f[TSP_F_Y], i, 0), field(f[TSP_F_Z], i, 0),
                                              field(f[TSP_F_PRESSURE], i, 0));
                        r[1][i] = make_float4(field(f[TSP_F_VX], i, 0), field(f[TSP_F_VY], i, 0), field(f[TSP_F_VZ], i, 0),
                                              field(f[TSP_F_DENSITY], i, 0));
                        r[2][i] = make_float4(field(f[TSP_F_VPX], i, 0), field(f[TSP_F_VPY], i, 0), field(f[TSP_F_VPZ], i, 0), 0.0f);
                        r[3][i] = make_float4(field(f[TSP_F_APX], i, 0), field(f[TSP_F_APY], i, 0), field(f[TSP_F_APZ], i, 0), 0.0f);
                        src1[i] = field(f[TSP_F_SOURCE], i, 0);
                        diag1[i] = field(f[TSP_F_DIAG], i, 0);
                    }
                    float4* dst[4] = { n.pos, n.vel, n.vp, n.ap };
                    for (int k = 0; k < 4; k++) { h2d(dst[k] + h.offF, r[k].data(), sizeof(float4) * h.nF); }
                    h2d(n.source + h.offF, src1.data(), sizeof(float) * h.nF);
                    h2d(n.diag + h.offF, diag1.data(), sizeof(float) * h.nF);
                    h2d(n.id + h.offF, ids(s->fluid_id, h.nF), sizeof(uint64_t) * h.nF);
                }
                // boundary (ensureBoundaryCoefficientsPresent defaults: psi = massB, V = psi / rho0)
                if (!src[q].plan)
                {
                    const float hh = h.params.particle_size, rho0 = h.params.rest_density;
                    const float massB = rho0 * hh * hh * hh;
                    const float volB = rho0 > 0.0f ? massB / rho0 : 0.0f;
                    std::vector<float4> r(h.nB);
                    std::vector<float> vol(h.nB);
                    float* const* f = s->boundary;
                    for (uint32_t i = 0; i < h.nB; i++)
                    {
                        r[i] = make_float4(field(f[TSP_B_X], i, 0), field(f[TSP_B_Y], i, 0), field(f[TSP_B_Z], i, 0),
                                           field(f[TSP_B_PSI], i, massB));
                        vol[i] = field(f[TSP_B_VOL], i, volB);
                    }
                    h2d(n.bpos + h.offB, r.data(), sizeof(float4) * h.nB);
                    h2d(n.bvol + h.offB, vol.data(), sizeof(float) * h.nB);
                    h2d(nidB + h.offB, ids(s->boundary_id, h.nB), sizeof(uint64_t) * h.nB);
                }
                // rigid particles: body-frame offsets; the world records follow from the body state
                {
                    float* const* f = s->rigid;
                    std::vector<float> rel[3];
                    std::vector<float4> wp(h.nR), wv(h.nR), wvp(h.nR);
                    for (int a = 0; a < 3; a++) { rel[a].resize(h.nR); }
                    for (uint32_t i = 0; i < h.nR; i++)
                    {
                        rel[0][i] = field(f[TSP_R_XREL], i, 0);
                        rel[1][i] = field(f[TSP_R_YREL], i, 0);
                        rel[2][i] = field(f[TSP_R_ZREL], i, 0);
                        wp[i] = make_float4(0.0f, 0.0f, 0.0f, field(f[TSP_R_PSI], i, 0));
                        wv[i] = make_float4(0.0f, 0.0f, 0.0f, field(f[TSP_R_VOL], i, 0));
                        wvp[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    }
                    h2d(n.rxrel + h.offR, rel[0].data(), sizeof(float) * h.nR);
                    h2d(n.ryrel + h.offR, rel[1].data(), sizeof(float) * h.nR);
                    h2d(n.rzrel + h.offR, rel[2].data(), sizeof(float) * h.nR);
                    h2d(n.rpos + h.offR, wp.data(), sizeof(float4) * h.nR);
                    h2d(n.rvel + h.offR, wv.data(), sizeof(float4) * h.nR);
                    h2d(n.rvp + h.offR, wvp.data(), sizeof(float4) * h.nR);
                    h2d(nidR + h.offR, ids(s->rigid_id, h.nR), sizeof(uint64_t) * h.nR);
                }
                // bodies (a missing orientation field is the identity)
                for (size_t f = 0; f < TSP_RB_COUNT; f++)
                {
                    std::vector<float> v(h.nRB);
                    const bool diagonal = f == TSP_RB_A00 || f == TSP_RB_A11 || f == TSP_RB_A22;
                    for (uint32_t b = 0; b < h.nRB; b++) { v[b] = s->body[f] ? s->body[f][b] : (diagonal ? 1.0f : 0.0f); }
                    h2d(n.rb + f * nRB1 + h.offRB, v.data(), sizeof(float) * h.nRB);
                }
            }
            // tables rebuilt from the host records: body of each rigid particle, the particles of
            // each body (BuildRigidParticleIndicesByBody, sync/state_sim.h:609-631), scene maps
            std::vector<uint32_t> rbody(nR), off(nRB + 1ull, 0), idx(nR);
            std::vector<uint32_t> blockScene(nF / kBlock), bodyScene(nRB);
            std::vector<uint16_t> sceneB(nB);
            for (uint32_t s = 0; s < S; s++)
            {
                const SceneHost& h = ns[s];
                for (uint32_t i = 0; i < h.nR; i++) { rbody[h.offR + i] = h.offRB + h.rbodyLocal[i]; }
                for (uint32_t b = 0; b < h.nBlk; b++) { blockScene[h.blk0 + b] = s; }
                for (uint32_t b = 0; b < h.nRB; b++) { bodyScene[h.offRB + b] = s; }
                for (uint32_t i = 0; i < h.nB; i++) { sceneB[h.offB + i] = static_cast<uint16_t>(s); }
            }
            for (uint32_t i = 0; i < nR; i++) { off[rbody[i] + 1]++; }
            for (uint32_t b = 0; b < nRB; b++) { off[b + 1] += off[b]; }
            std::vector<uint32_t> w(off);
            for (uint32_t i = 0; i < nR; i++) { idx[w[rbody[i]]++] = i; }
            h2d(n.rbody, rbody.data(), sizeof(uint32_t) * nR);
            h2d(n.byBodyOff, off.data(), sizeof(uint32_t) * (nRB + 1ull));
            h2d(n.byBodyIdx, idx.data(), sizeof(uint32_t) * nR);
            h2d(n.blockScene, blockScene.data(), sizeof(uint32_t) * blockScene.size());
            h2d(n.bodyScene, bodyScene.data(), sizeof(uint32_t) * nRB);
            h2d(n.sceneB, sceneB.data(), sizeof(uint16_t) * nB);
            sync();
        }
        catch (...)
        {
            (void)hipStreamSynchronize(stream);
            releaseList(A);
            throw;
        }

        // install the new arena
        releaseList(arena);
        arena = std::move(A);
        d = n;
        pos2 = npos2; vel2 = nvel2; vp2 = nvp2; ap2 = nap2; source2 = nsource2; diag2 = ndiag2; idF2 = nidF2;
        bpos2 = nbpos2; bvol2 = nbvol2; idB = nidB; idB2 = nidB2; idR = nidR;
        scenes = std::move(ns);
        layoutDirty = true;
        listsValid = false;
        cellsClean = false;

        // buffers sized by the layout
        releaseList(work);
        const uint32_t nSort = std::max(nF, std::max(nB, nR));
        // key arrays: room for u64 keys (two words each), which frames beyond 1024 cells per axis need
        d.keyF = alloc<uint32_t>(2ull * nSort, work, MEM_SF);
        d.keyF2 = alloc<uint32_t>(2ull * nSort, work, MEM_SF);
        d.idxF = alloc<uint32_t>(nSort, work, MEM_SF);
        d.idxF2 = alloc<uint32_t>(nSort, work, MEM_SF);
        d.keyR = alloc<uint32_t>(2ull * nR, work, MEM_SR);
        d.keyR2 = alloc<uint32_t>(2ull * nR, work, MEM_SR);
        d.idxR = alloc<uint32_t>(nR, work, MEM_SR);
        d.idxR2 = alloc<uint32_t>(nR, work, MEM_SR);
        const uint64_t nTiles = (nSort + 2047ull) / 2048ull;
        radixHist = alloc<uint32_t>(256ull * nTiles, work);
        aabb = alloc<uint32_t>(6ull * S, work);
        const uint32_t nSlices = (nF + kSlice - 1) / kSlice;
        for (int r = 0; r < 3; r++)
        {
            d.cnt[r] = alloc<uint16_t>(nF, work, MEM_LF);
            d.meta[r] = alloc<SliceMeta>(nSlices, work, MEM_LF);
            zero(d.cnt[r], sizeof(uint16_t) * nF);
        }
        sliceUnits = alloc<uint32_t>(3ull * (nSlices + 1), work);
        {
            // block sums of the scans: radix digit offsets, slice units, RF offsets
            const uint64_t m = std::max<uint64_t>(256ull * nTiles, std::max<uint64_t>(nSlices + 1ull, nR + 1ull));
            d.scanTmp = alloc<uint32_t>(scanTmpWords(m), work);
            zero(d.scanTmp, sizeof(uint32_t) * scanTmpWords(m));
        }
        d.ffMin = alloc<uint32_t>(nF, work, MEM_LF);
        d.wideList = alloc<uint32_t>(nF, work, MEM_LF);
        d.laneRank = alloc<uint8_t>(nF, work, MEM_LF);
        d.rankWho = alloc<uint8_t>(nF, work, MEM_LF);
        d.ffMask = alloc<uint32_t>(static_cast<size_t>(kMaskWords) * nF, work, MEM_SF);
        {
            const uint64_t nBlocks = nF / kBlock;
            d.haloIdx = alloc<uint32_t>(nBlocks * kHaloMax, work, MEM_LF);
            d.haloCount = alloc<uint32_t>(nBlocks, work, MEM_LF);
            // no block has a halo until the first search builds one
            hipCheck(hipMemsetAsync(d.haloCount, 0xff, sizeof(uint32_t) * std::max<uint64_t>(nBlocks, 1), stream), "memset");
        }
        d.rfCnt = alloc<uint32_t>(nR, work, MEM_LR);
        d.rbCnt = alloc<uint32_t>(nR, work, MEM_LR);
        d.rrCnt = alloc<uint32_t>(nR, work, MEM_LR);
        zero(d.rfCnt, sizeof(uint32_t) * nR);
        zero(d.rbCnt, sizeof(uint32_t) * nR);
        zero(d.rrCnt, sizeof(uint32_t) * nR);
        d.rfOff = alloc<uint32_t>(nR + 1ull, work, MEM_LR);
        d.rbOff = alloc<uint32_t>(nR + 1ull, work, MEM_LR);
        d.rrOff = alloc<uint32_t>(nR + 1ull, work, MEM_LR);
        d.cbScan = alloc<uint32_t>(nB + 1ull, work);
        zero(d.cbScan, sizeof(uint32_t) * (nB + 1ull));
        d.cRw = alloc<float>(3ull * nR, work);
        d.cR = alloc<float>(static_cast<uint64_t>(kCR) * nR, work);
        d.cErrR = alloc<double>(nR, work);
        d.cIw = alloc<float>(9ull * std::max<uint32_t>(nRB, 1), work);
        d.cLoad = alloc<float>(6ull * std::max<uint32_t>(nRB, 1), work);
        d.cBodyB = alloc<uint32_t>(std::max<uint32_t>(nRB, 1), work);
        // the contact boundary arrays and their scan buffer depend on the boundary count
        releaseList(contactAllocs);
        d.rbCap = d.rrCap = 0;
        contactLists = false;
        contactRB = 0;
        d.errPartial = alloc<double>(3ull * (nF / kBlock), work);
        {
            static_assert(sizeof(Status) <= kSolveOff, "status block layout");
            unsigned char* blk = alloc<unsigned char>(kSolveOff + sizeof(Solve) * S, work);
            zero(blk, kSolveOff + sizeof(Solve) * S);
            d.status = reinterpret_cast<Status*>(blk);
            d.solve = reinterpret_cast<Solve*>(blk + kSolveOff);
        }
        d.scenes = alloc<Scene>(S, work);
        d.frames = alloc<Frame>(S, work);
        d.consts = alloc<Consts>(S, work);
        d.sceneKey0 = alloc<uint64_t>(S + 1ull, work);
        {
            std::vector<Scene> t(S);
            for (uint32_t s = 0; s < S; s++)
            {
                const SceneHost& h = scenes[s];
                t[s] = Scene{ h.offF, h.nF, h.blk0, h.nBlk, h.offB, h.nB, h.offR, h.nR, h.offRB, h.nRB };
            }
            h2d(d.scenes, t.data(), sizeof(Scene) * S);
        }
        uploadConsts();
        if (solveCap < S)
        {
            if (hostBlock) { (void)hipHostFree(hostBlock); hostBlock = nullptr; hostStatus = nullptr; hostSolve = nullptr; solveCap = 0; }
            hipCheck(hipHostMalloc(reinterpret_cast<void**>(&hostBlock), kSolveOff + sizeof(Solve) * S), "hipHostMalloc");
            std::memset(hostBlock, 0, kSolveOff);
            hostStatus = reinterpret_cast<Status*>(hostBlock);
            hostSolve = reinterpret_cast<Solve*>(hostBlock + kSolveOff);
            solveCap = S;
        }
        if (S) { std::memset(hostSolve, 0, sizeof(Solve) * S); }
        // pools: kept when large enough (about 56 FF, 16 FB and 8 FR entries per particle to
        // start, grown on demand)
        {
            uint64_t cap[3] = { 56ull * nF + 1024, 16ull * nF + 1024, 8ull * nF + 1024 };
            uint64_t rfCap = 64ull * nR + 1024;
            const bool have = !poolAllocs.empty();
            if (!have || cap[0] > d.poolCap[0] || cap[1] > d.poolCap[1] || cap[2] > d.poolCap[2] || rfCap > d.rfCap)
            {
                if (have)
                {
                    for (int r = 0; r < 3; r++) { cap[r] = std::max(cap[r], d.poolCap[r]); }
                    rfCap = std::max(rfCap, d.rfCap);
                }
                allocPools(cap, rfCap);
            }
        }
        // world records of the rigid particles from the body state
        launchRigidUpdate(stream, d, false);
        sync();
    }

    // a scene record for a new scene from a host state (validates it)
    SceneHost newScene(const tsp_params& p, const tsp_state_view* s, const float* mn, const float* mx)
    {
        validateParams(p);
        if (s->n_fluid > 0x7fffffffu || s->n_boundary > 0x7fffffffu || s->n_rigid > 0x7fffffffu)
        {
            throw Error{ TSP_ERR_CAPACITY, "particle count exceeds the local particle limit" };
        }
        if (s->n_rigid && !s->rigid_body) { throw Error{ TSP_ERR_INVALID, "rigid particles need body ids" }; }
        SceneHost h;
        h.params = p;
        h.domainSet = mn && mx;
        if (h.domainSet) { for (int a = 0; a < 3; a++) { h.domMin[a] = mn[a]; h.domMax[a] = mx[a]; } }
        h.nF = s->n_fluid; h.nB = s->n_boundary; h.nR = s->n_rigid; h.nRB = s->n_bodies;
        h.rbodyLocal.resize(h.nR);
        for (uint32_t i = 0; i < h.nR; i++)
        {
            if (s->rigid_body[i] >= h.nRB) { throw Error{ TSP_ERR_INVALID, "Invalid rigid body index." }; }
            h.rbodyLocal[i] = s->rigid_body[i];
        }
        h.bodyId.resize(h.nRB);
        h.bodyCount.resize(h.nRB);
        for (uint32_t b = 0; b < h.nRB; b++)
        {
            h.bodyId[b] = s->body_id ? s->body_id[b] : b;
            h.bodyCount[b] = s->body_count ? s->body_count[b] : 0;
        }
        h.step = s->step;
        h.time = s->time;
        return h;
    }

    // --- cell frames: UpdateCellFrame (CPU/backend/neighbor_search.cpp:1902-2006) for every scene
    //     whose frame is invalid; then every scene's key range and cell range, the cleared cell
    //     tables and the boundary order ---
    void layoutFrames()
    {
        const uint32_t S = nScenes();
        bool rebuild = false;
        for (const SceneHost& h : scenes) { rebuild = rebuild || !h.frameValid; }
        if (rebuild)
        {
            // per-scene bounds of F, B and R on the device (order-free min/max)
            std::vector<uint32_t> bits(6ull * S);
            for (uint32_t s = 0; s < S; s++)
            {
                for (int a = 0; a < 3; a++) { bits[6 * s + a] = 0xffffffffu; bits[6 * s + 3 + a] = 0u; }
            }
            h2d(aabb, bits.data(), sizeof(uint32_t) * bits.size());
            launchAabb(stream, d, kPopF, d.pos, d.nF, aabb);
            launchAabb(stream, d, kPopB, d.bpos, d.nB, aabb);
            launchAabb(stream, d, kPopR, d.rpos, d.nR, aabb);
            hipCheck(hipMemcpyAsync(bits.data(), aabb, sizeof(uint32_t) * bits.size(), hipMemcpyDeviceToHost, stream), "aabb D2H");
            sync();
            const auto toFloat = [](uint32_t u)
            {
                const uint32_t v = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
                float f;
                std::memcpy(&f, &v, 4);
                return f;
            };
            const auto floorToIntHost = [](float v)
            {
                const int i = static_cast<int>(v);
                return (v < static_cast<float>(i)) ? i - 1 : i;
            };
            for (uint32_t s = 0; s < S; s++)
            {
                SceneHost& h = scenes[s];
                if (h.frameValid) { continue; }
                const tsp_params& p = h.params;
                const float cs = p.kernel_radius * p.grid_cell_size_ratio;
                const float span = p.kernel_radius / cs;
                const float sm = std::max(0.0f, span - 1e-6f);
                const int32_t reach = std::max<int32_t>(1, static_cast<int32_t>(std::ceil(sm)));
                const float inv = 1.0f / cs;
                const bool empty = bits[6 * s] == 0xffffffffu && bits[6 * s + 3] == 0u;
                float mn[3], mx[3];
                for (int a = 0; a < 3; a++)
                {
                    mn[a] = empty ? 0.0f : toFloat(bits[6 * s + a]);
                    mx[a] = empty ? 0.0f : toFloat(bits[6 * s + 3 + a]);
                }
                const bool domOk = h.domainSet && h.domMax[0] > h.domMin[0] && h.domMax[1] > h.domMin[1] && h.domMax[2] > h.domMin[2];
                if (domOk)
                {
                    for (int a = 0; a < 3; a++)
                    {
                        mn[a] = std::clamp(mn[a], h.domMin[a], h.domMax[a]);
                        mx[a] = std::clamp(mx[a], h.domMin[a], h.domMax[a]);
                    }
                }
                const int64_t margin = static_cast<int64_t>(reach) + 8;
                Frame f{};
                for (int a = 0; a < 3; a++)
                {
                    const int64_t o = static_cast<int64_t>(floorToIntHost(mn[a] * inv)) - margin;
                    const int64_t e = static_cast<int64_t>(floorToIntHost(mx[a] * inv)) + margin;
                    // TryBuildCellCoordFrame (core/cell_frame.h:32-66): 21 Morton bits per axis
                    if (o > e || o < INT32_MIN || e > INT32_MAX || e - o + 1 > (1ll << 21))
                    {
                        throw Error{ TSP_ERR_CAPACITY, "cell frame exceeds the supported range (2^21 cells per axis)" };
                    }
                    f.orig[a] = static_cast<int32_t>(o);
                    f.dim[a] = static_cast<uint32_t>(e - o + 1);
                    f.innerLo[a] = f.orig[a] + reach;
                    f.innerHi[a] = f.orig[a] + static_cast<int32_t>(f.dim[a]) - 1 - reach;
                }
                f.cellSizeInv = inv;
                f.reach = reach;
                f.domainSet = domOk ? 1 : 0;
                for (int a = 0; a < 3; a++) { f.domMin[a] = h.domMin[a]; f.domMax[a] = h.domMax[a]; }
                h.frame = f;
                h.frameValid = true;
                h.frameBuilds++;
                h.flags |= TSP_FLAG_FRAME_REBUILT;
            }
        }
        // key range of each scene: its Morton ids, then its empty key (outside the frame) and its
        // dead key (padding slots), so a sort of all scenes keeps every scene's segment in place
        // The keys are 32-bit while the whole key space fits (every frame within 1024 cells per
        // axis, and scenes together), else 64-bit: the sort then moves 8-byte keys.
        uint64_t key = 0, cells = 0;
        std::vector<uint64_t> key0(S + 1ull);
        std::vector<Frame> frames(S);
        for (uint32_t s = 0; s < S; s++)
        {
            Frame& f = scenes[s].frame;
            // the scene's Morton ids lie below morton(dim - 1) + 1: the interleave is monotone in
            // each coordinate (a range no wider than the frame needs, so the sort runs fewer passes)
            const auto spread = [](uint64_t v)
            {
                uint64_t r = 0;
                for (uint32_t b = 0; b < 21; b++) { r |= ((v >> b) & 1ull) << (3 * b); }
                return r;
            };
            const uint64_t range = (spread(f.dim[0] - 1) | (spread(f.dim[1] - 1) << 1) | (spread(f.dim[2] - 1) << 2)) + 1;
            if (range + 2 > ~0ull - key)
            {
                throw Error{ TSP_ERR_CAPACITY, "the scenes' cell frames exceed the 64-bit key space of one context" };
            }
            f.keyBase = key;
            f.emptyKey = key + range;
            f.deadKey = f.emptyKey + 1;
            f.cellOff = static_cast<uint32_t>(cells);
            key0[s] = f.keyBase;
            key += range + 2;
            cells += static_cast<uint64_t>(f.dim[0]) * f.dim[1] * f.dim[2];
            if (cells > cellLimit)
            {
                throw Error{ TSP_ERR_CAPACITY, "the scenes' cell frames exceed the context's cell limit" };
            }
            frames[s] = f;
        }
        key0[S] = key;
        keyBits = 1;
        for (uint64_t v = key > 0 ? key - 1 : 0; v >> 1; v >>= 1) { keyBits++; }
        d.key64 = keyBits > 32 ? 1u : 0u;
        h2d(d.frames, frames.data(), sizeof(Frame) * S);
        h2d(d.sceneKey0, key0.data(), sizeof(uint64_t) * key0.size());
        if (cells > cellCap || frameAllocs.empty())
        {
            releaseList(frameAllocs);
            cellCap = std::max<uint64_t>(cells, 1);
            d.cellF = alloc<uint2>(cellCap, frameAllocs, MEM_SF);
            d.cellB = alloc<uint2>(cellCap, frameAllocs, MEM_SB);
            d.cellR = alloc<uint2>(cellCap, frameAllocs, MEM_SR);
        }
        launchClearAll(stream, d.cellF, cells);
        launchClearAll(stream, d.cellB, cells);
        launchClearAll(stream, d.cellR, cells);
        cellsClean = true;

        // boundary: re-sorted only if its ids are not already non-decreasing (neighbor_search.cpp:1755-1811)
        if (d.nB)
        {
            launchResetSearch(stream, d);
            launchKeysSorted(stream, d, d.bpos, d.nB, d.keyF, d.idxF);
            uint32_t unsorted = 0;
            hipCheck(hipMemcpyAsync(&unsorted, &d.status->bUnsorted, 4, hipMemcpyDeviceToHost, stream), "flag D2H");
            sync();
            if (unsorted)
            {
                if (launchRadixSort(stream, d.key64, d.keyF, d.idxF, d.keyF2, d.idxF2, d.nB, keyBits, radixHist, d.scanTmp,
                                    &d.status->scanFail, false))
                {
                    std::swap(d.keyF, d.keyF2);
                    std::swap(d.idxF, d.idxF2);
                }
                Gather g{};
                g.s4[0] = d.bpos; g.d4[0] = bpos2; g.n4 = 1;
                g.s1[0] = d.bvol; g.d1[0] = bvol2; g.n1 = 1;
                g.idS = idB; g.idD = idB2;
                g.n = d.nB;
                launchPermute(stream, g, d.idxF);
                std::swap(d.bpos, bpos2);
                std::swap(d.bvol, bvol2);
                std::swap(idB, idB2);
            }
            launchBuildCells(stream, d, d.keyF, d.nB, d.cellB);
            sync();
        }
        layoutDirty = false;
    }

    // --- re-order, construct and query (DataReOrder + Construct + Query_Count/Fill), all scenes ---
    // the fluid fields a search re-orders trade places with their second buffers: positions,
    // velocities (with density) and pressure (pos.w, the warm start) in a step, every field in
    // the standalone search
    void swapFields(bool allFields)
    {
        std::swap(d.pos, pos2);
        std::swap(d.vel, vel2);
        std::swap(d.id, idF2);
        if (allFields)
        {
            std::swap(d.vp, vp2);
            std::swap(d.ap, ap2);
            std::swap(d.source, source2);
            std::swap(d.diag, diag2);
        }
    }

    // the search up to its status: keys, sorts, re-order, cell tables, counts, list offsets
    void searchFront(bool allFields)
    {
        {
            Scope sc(prof, stream, "reset");
            launchResetSearch(stream, d);
            // the runs of the previous sorted keys are the only non-empty cells
            if (!cellsClean)
            {
                launchClearCells(stream, d, d.keyF, d.nF, d.cellF);
                launchClearCells(stream, d, d.keyR, d.nR, d.cellR);
            }
        }
        // a single-block rigid sort shares no scratch with the fluid sort: it runs beside it
        const bool fork = d.nR && radixSortSingleBlock(d.nR) && !prof.on;
        if (d.nR)
        {
            Scope sc(prof, stream, "keys_sort_rigid");
            hipStream_t rs = stream;
            if (fork)
            {
                hipCheck(hipEventRecord(forkEv, stream), "hipEventRecord");
                hipCheck(hipStreamWaitEvent(side, forkEv, 0), "hipStreamWaitEvent");
                rs = side;
            }
            launchKeys(rs, d, kPopR, d.rpos, d.nR, true, d.keyR, d.idxR);
            if (launchRadixSort(rs, d.key64, d.keyR, d.idxR, d.keyR2, d.idxR2, d.nR, keyBits, radixHist, d.scanTmp,
                                &d.status->scanFail, false))
            {
                std::swap(d.keyR, d.keyR2);
                std::swap(d.idxR, d.idxR2);
            }
            if (fork) { hipCheck(hipEventRecord(joinEv, side), "hipEventRecord"); }
        }
        {
            Scope sc(prof, stream, "keys_sort");
            launchKeys(stream, d, kPopF, d.pos, d.nF, true, d.keyF, d.idxF);
            if (launchRadixSort(stream, d.key64, d.keyF, d.idxF, d.keyF2, d.idxF2, d.nF, keyBits, radixHist, d.scanTmp,
                                &d.status->scanFail, false))
            {
                std::swap(d.keyF, d.keyF2);
                std::swap(d.idxF, d.idxF2);
            }
        }
        {
            Scope sc(prof, stream, "permute");
            Gather g{};
            g.s4[0] = d.pos; g.d4[0] = pos2;
            g.s4[1] = d.vel; g.d4[1] = vel2;
            g.n4 = 2;
            if (allFields)
            {
                g.s4[2] = d.vp; g.d4[2] = vp2;
                g.s4[3] = d.ap; g.d4[3] = ap2;
                g.n4 = 4;
                g.s1[0] = d.source; g.d1[0] = source2;
                g.s1[1] = d.diag; g.d1[1] = diag2;
                g.n1 = 2;
            }
            g.idS = d.id; g.idD = idF2;
            g.n = d.nF;
            launchPermute(stream, g, d.idxF);
        }
        swapFields(allFields);
        if (fork) { hipCheck(hipStreamWaitEvent(stream, joinEv, 0), "hipStreamWaitEvent"); }
        {
            Scope sc(prof, stream, "cells");
            launchBuildCells(stream, d, d.keyF, d.nF, d.cellF);
            launchBuildCells(stream, d, d.keyR, d.nR, d.cellR);
            cellsClean = false;
        }
        {
            Scope sc(prof, stream, "query_count");
            launchQueryCount(stream, d);
            launchQueryRigid(stream, d, 0, d.scanTmp);
        }
        {
            Scope sc(prof, stream, "scan");
            launchScanSlices(stream, d, sliceUnits, d.scanTmp);
        }
    }

    void neighborSearch(bool allFields)
    {
        bool rebuild = layoutDirty;
        for (const SceneHost& h : scenes) { rebuild = rebuild || !h.frameValid; }
        if (rebuild) { layoutFrames(); }
        for (int attempt = 0; attempt < 4; attempt++)
        {
            searchFront(allFields);
            readStatus();
            if (last.abort)
            {
                // a particle left its scene's inner frame: undo the permutation, rebuild that
                // scene's frame first
                swapFields(allFields);
                for (uint32_t s = 0; s < nScenes(); s++)
                {
                    if (hostSolve[s].abort) { scenes[s].frameValid = false; }
                }
                layoutFrames();
                continue;
            }
            // the short exact root and quotient cover this search's FF and FB pairs unless one is tiny
            d.fastPairs = (d.fastDiv && !last.pairTiny) ? 1u : 0u;
            // capacity errors leave the particles in their order before the search (the lists
            // of this search are not written)
            const auto refuse = [&](const char* why)
            {
                swapFields(allFields);
                listsValid = false;
                throw Error{ TSP_ERR_CAPACITY, why };
            };
            if (last.overflow & 2u) { refuse("neighbor count exceeds the u16 count storage"); }
            for (int r = 0; r < 3; r++)
            {
                // slice offsets are u32 in u16 units
                if (last.poolUnits[r] > 0xffffffffull) { refuse("SELL pool exceeds 2^32 u16 units"); }
            }
            if (last.overflow & 5u)
            {
                uint64_t cap[3];
                for (int r = 0; r < 3; r++)
                {
                    cap[r] = std::max<uint64_t>(d.poolCap[r], last.poolUnits[r] + last.poolUnits[r] / 8 + 64);
                    cap[r] = std::min<uint64_t>(cap[r], 0xffffffffull);
                }
                const uint64_t rfCap = std::max<uint64_t>(d.rfCap, last.rfTotal + last.rfTotal / 8 + 64);
                allocPools(cap, rfCap);
                for (SceneHost& h : scenes) { h.flags |= TSP_FLAG_LIST_REGROWN; }
            }
            // the contact solve's RB and RR lists are stored when some scene has entries
            contactLists = last.rbTotal + last.rrTotal > 0;
            contactRB = last.rbTotal;
            if (contactLists)
            {
                ensureContactCapacity(last.rbTotal, last.rrTotal);
                launchContactOffsets(stream, d, d.scanTmp);
            }
            {
                Scope sc(prof, stream, "query_fill");
                launchQueryFill(stream, d);
                launchQueryRigid(stream, d, contactLists ? 2 : 1, d.scanTmp);
            }
            listsValid = true;
            {
                Scope sc(prof, stream, "halo");
                launchHaloBuild(stream, d);
            }
            return;
        }
        throw Error{ TSP_ERR_STATE, "cell frame rebuild did not converge" };
    }

    // --- n relaxed Jacobi sweeps of every scene (a^p, then p) and the bookkeeping of the last ---
    void sweepLaunches(uint32_t n)
    {
        for (uint32_t s = 0; s < n; s++)
        {
            {
                Scope sc(prof, stream, "pressure_acc");
                launchPressureAcc(stream, d, true);
            }
            Scope sc(prof, stream, "jacobi");
            launchJacobi(stream, d);
        }
        Scope sc(prof, stream, "finalize");
        launchFinalize(stream, d);
    }

    // the group runs as one instantiated HIP graph (captured on first use for these kernel
    // arguments and this count), unless launches are being timed one by one
    void launchSweeps(uint32_t n)
    {
        if (!TSP_TUNE_GRAPHS || prof.on)
        {
            sweepLaunches(n);
            return;
        }
        // keyed by the kernel arguments the sweeps read: the sort buffers, which trade places
        // from search to search, are left out
        Dev k = d;
        k.keyF = k.keyF2 = k.idxF = k.idxF2 = nullptr;
        k.keyR = k.keyR2 = k.idxR = k.idxR2 = nullptr;
        std::string key(reinterpret_cast<const char*>(&k), sizeof(Dev));
        key.append(reinterpret_cast<const char*>(&n), sizeof(n));
        auto it = graphs.find(key);
        if (it == graphs.end())
        {
            if (graphs.size() >= 64) { dropGraphs(); }
            hipCheck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            hipGraph_t g = nullptr;
            try
            {
                sweepLaunches(n);
            }
            catch (...)
            {
                (void)hipStreamEndCapture(stream, &g);
                if (g) { (void)hipGraphDestroy(g); }
                throw;
            }
            hipCheck(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
            hipGraphExec_t e = nullptr;
            const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hipCheck(rc, "hipGraphInstantiate");
            it = graphs.emplace(std::move(key), e).first;
        }
        hipCheck(hipGraphLaunch(it->second, stream), "hipGraphLaunch");
    }

    // --- one IISPH step of every scene (Step_IISPH, sph_solver.cpp:34-319), in phases ---

    // density, non-pressure forces, rigid prediction, source terms and diagonals
    void predictLaunches()
    {
        {
            Scope sc(prof, stream, "density");
            launchDensity(stream, d);
        }
        if (d.nRB) { launchRigidClearGravity(stream, d); }
        {
            Scope sc(prof, stream, "nonpressure");
            launchNonPressure(stream, d);
        }
        if (d.nRB)
        {
            Scope sc(prof, stream, "rigid");
            launchRigidForces(stream, d, false);
            launchRigidReduce(stream, d);
            launchRigidPredict(stream, d);
            launchRigidUpdate(stream, d, true);
        }
        {
            Scope sc(prof, stream, "source_diag");
            launchSourceDiag(stream, d);
        }
    }

    // the first sweep group: the previous step's sweeps of the slowest scene plus
    // TSP_TUNE_SWEEP_AHEAD (sweeps queued past convergence return at once; one short of it costs
    // a status read); maxSweeps bounds the solve, anyFluid says whether there is one
    uint32_t firstSweepGroup(bool& anyFluid, uint32_t& maxSweeps) const
    {
        uint32_t maxIters = 1, lastMax = 0;
        anyFluid = false;
        for (const SceneHost& h : scenes)
        {
            if (!h.nF) { continue; }
            anyFluid = true;
            maxIters = std::max(maxIters, h.params.max_solver_iters);
            lastMax = std::max(lastMax, h.lastSweeps);
        }
        maxSweeps = maxIters > 1 ? maxIters - 1 : 1;
        const int64_t guess = static_cast<int64_t>(lastMax) + TSP_TUNE_SWEEP_AHEAD;
        return static_cast<uint32_t>(std::max<int64_t>(1, std::min<int64_t>(guess, maxSweeps)));
    }

    // relaxed Jacobi sweeps of all scenes until every scene has converged; each group is followed
    // by one read of the solver states. queued sweeps have run already (their states read).
    void solveLoop(uint32_t queued, uint32_t group, uint32_t maxSweeps)
    {
        for (;;)
        {
            if (queued > 0)
            {
                bool all = true;
                for (uint32_t s = 0; s < nScenes(); s++) { all = all && (scenes[s].nF == 0 || hostSolve[s].converged); }
                if (all) { return; }
            }
            const uint32_t n = std::min(group, maxSweeps - queued);
            if (n == 0) { throw Error{ TSP_ERR_STATE, "pressure solve did not stop at max_solver_iters" }; }
            launchSweeps(n);
            queued += n;
            readStatus();
            group = 1;
        }
    }

    // final pressure accelerations, rigid coupling, integration and the contact solve
    void tailLaunches()
    {
        {
            Scope sc(prof, stream, "pressure_acc_final");
            launchPressureAcc(stream, d, false);
        }
        if (d.nRB)
        {
            Scope sc(prof, stream, "rigid");
            launchRigidForces(stream, d, true);
            launchRigidClearGravity(stream, d);
            launchRigidReduce(stream, d);
        }
        {
            Scope sc(prof, stream, "integration");
            launchIntegration(stream, d);
        }
        if (d.nRB && contactLists) { solveContacts(); }
        if (d.nRB)
        {
            Scope sc(prof, stream, "rigid");
            launchRigidPredict(stream, d);
            launchRigidIntegrate(stream, d);
            launchRigidUpdate(stream, d, false);
        }
    }

    // the scenes' counters after a step, from the solver states read last
    void stepBookkeeping()
    {
        for (uint32_t s = 0; s < nScenes(); s++)
        {
            SceneHost& h = scenes[s];
            const Solve& so = hostSolve[s];
            h.last = so;
            if (so.contactActive) { h.flags |= TSP_FLAG_RIGID_CONTACT; }
            h.lastSweeps = h.nF ? so.sweeps : 0;
            h.sweepsTotal += h.lastSweeps;
            const tsp_params& p = h.params;
            if (h.nF && p.termination_cond != 0 && so.iters >= p.max_solver_iters)
            {
                const float e = p.termination_cond == 1 ? so.avgRaw : p.termination_cond == 2 ? so.avgPos : so.avgAbs;
                if (!(e <= p.density_dev_target)) { h.flags |= TSP_FLAG_MAX_ITERS; }
            }
            h.step++;
            h.time += static_cast<double>(p.time_step);
        }
    }

    // --- the step as captured graphs (TSP_TUNE_STEP_GRAPH, tsp_set_step_graph) ---
    // The launches of a step are captured once per argument set and replayed: the search with
    // the lists and the prediction as one graph, the first sweep group as the sweep-group graph,
    // the tail as one graph, with one status read at the end. What the host decides from the
    // search's status in the launch-by-launch form is decided on the device instead: the gate
    // after the scans calls the rest of the step off (stop bit 1) when a frame must be rebuilt,
    // storage must grow, a contact solve or the general square root is needed; the gate after the
    // sweeps calls the tail off (bit 2) when a scene needs more sweeps. A called-off step is
    // finished launch by launch: anew after bit 1 (the permutation undone), from the sweeps after
    // bit 2. The order of every operation is the launch-by-launch form's, so the bits are too.
    struct StepGraph
    {
        hipGraphExec_t exec = nullptr;
        Dev after{};                       // the device arguments after the piece (buffers traded)
        float4 *pos2 = nullptr, *vel2 = nullptr;
        uint64_t* idF2 = nullptr;
    };
    std::map<std::string, StepGraph> stepGraphs;
    bool stepGraphOn = TSP_TUNE_STEP_GRAPH != 0;

    void dropStepGraphs()
    {
        for (auto& g : stepGraphs) { (void)hipGraphExecDestroy(g.second.exec); }
        stepGraphs.clear();
    }

    // the captured form covers steps that need no host decision in the common case
    bool stepGraphAllowed() const
    {
        if (!stepGraphOn || prof.on || layoutDirty || !listsValid || cellsClean || contactLists || !d.nF) { return false; }
        if (d.fastPairs != d.fastDiv) { return false; }
        for (const SceneHost& h : scenes) { if (!h.frameValid) { return false; } }
        return true;
    }

    // launches piece `id` of the step as a graph (captured on first use for these arguments) and
    // takes over the host state the launches leave (the traded buffers)
    template <typename Fn>
    void launchPiece(int id, Fn&& launches)
    {
        std::string key(reinterpret_cast<const char*>(&id), sizeof(id));
        key.append(reinterpret_cast<const char*>(&d), sizeof(Dev));
        const void* extra[5] = { pos2, vel2, idF2, sliceUnits, radixHist };
        key.append(reinterpret_cast<const char*>(extra), sizeof(extra));
        key.append(reinterpret_cast<const char*>(&keyBits), sizeof(keyBits));
        auto it = stepGraphs.find(key);
        if (it == stepGraphs.end())
        {
            if (stepGraphs.size() >= 32) { dropStepGraphs(); }
            const Dev before = d;
            float4* const p2 = pos2;
            float4* const v2 = vel2;
            uint64_t* const i2 = idF2;
            hipCheck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            hipGraph_t g = nullptr;
            try
            {
                launches();
            }
            catch (...)
            {
                (void)hipStreamEndCapture(stream, &g);
                if (g) { (void)hipGraphDestroy(g); }
                d = before; pos2 = p2; vel2 = v2; idF2 = i2;
                throw;
            }
            hipCheck(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
            StepGraph sg;
            const hipError_t rc = hipGraphInstantiate(&sg.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            sg.after = d; sg.pos2 = pos2; sg.vel2 = vel2; sg.idF2 = idF2;
            d = before; pos2 = p2; vel2 = v2; idF2 = i2;
            hipCheck(rc, "hipGraphInstantiate");
            it = stepGraphs.emplace(std::move(key), sg).first;
        }
        hipCheck(hipGraphLaunch(it->second.exec, stream), "hipGraphLaunch");
        d = it->second.after;
        pos2 = it->second.pos2; vel2 = it->second.vel2; idF2 = it->second.idF2;
    }

    // 0: the step is done; 1: called off at the search (run the step launch by launch);
    // 2: the first group did not finish the solve (continue with more sweeps and the tail)
    int stepGraph(uint32_t group)
    {
        launchPiece(0, [&]()
        {
            searchFront(false);
            launchStepGate(stream, d);
            d.stopMask = 1u;
            launchQueryFill(stream, d);
            launchQueryRigid(stream, d, 1, d.scanTmp);
            launchHaloBuild(stream, d);
            predictLaunches();
        });
        // d.stopMask is 1 here (the state the piece leaves)
        launchSweeps(group);
        launchSolveGate(stream, d);
        d.stopMask = 3u;
        launchPiece(1, [&]() { tailLaunches(); });
        d.stopMask = 0u;
        cellsClean = false;
        listsValid = true;
        readStatus();
        if (last.stop & 1u)
        {
            // nothing after the scans ran: back to the order before the search, lists unwritten
            swapFields(false);
            listsValid = false;
            return 1;
        }
        return (last.stop & 2u) ? 2 : 0;
    }

    void stepOnce()
    {
        bool anyFluid = false;
        uint32_t maxSweeps = 1;
        const uint32_t group = firstSweepGroup(anyFluid, maxSweeps);
        if (anyFluid && stepGraphAllowed())
        {
            const int r = stepGraph(group);
            if (r != 1) { for (SceneHost& h : scenes) { h.flags |= TSP_FLAG_STEP_GRAPH; } }
            if (r == 0)
            {
                stepBookkeeping();
                return;
            }
            if (r == 2)
            {
                launchClearStop(stream, d);
                solveLoop(group, 1, maxSweeps);
                tailLaunches();
                stepBookkeeping();
                return;
            }
        }
        neighborSearch(false);
        predictLaunches();
        if (anyFluid) { solveLoop(0, group, maxSweeps); }
        else { readStatus(); }
        tailLaunches();
        stepBookkeeping();
    }

    void steps(uint32_t n)
    {
        for (SceneHost& h : scenes) { h.flags = 0; }
        for (const SceneHost& h : scenes) { if (h.nRB) { validateContactParams(h.params); } }
        bool work = false;
        for (const SceneHost& h : scenes) { work = work || h.nF > 0 || h.nRB > 0; }
        for (uint32_t s = 0; s < n; s++)
        {
            if (!work)
            {
                for (SceneHost& h : scenes) { h.step++; h.time += static_cast<double>(h.params.time_step); }
                continue;
            }
            stepOnce();
        }
        readStatus();
        for (uint32_t s = 0; s < nScenes(); s++)
        {
            // the last step's contact report (the contact solve runs after the step's other reads)
            Solve& l = scenes[s].last;
            const Solve& so = hostSolve[s];
            l.contactActive = so.contactActive; l.contactDone = so.contactDone; l.contactIters = so.contactIters;
            l.contactConverged = so.contactConverged; l.contactAvg = so.contactAvg; l.contactBoundary = so.contactBoundary;
            l.contactBad = so.contactBad; l.contacts = so.contacts;
            if (hostSolve[s].rigidDegenerate) { throw Error{ TSP_ERR_NUMERIC, "Degenerate rigid body orientation." }; }
            const uint32_t bad = hostSolve[s].contactBad;
            if (bad & kContactManyBodies)
            {
                throw Error{ TSP_ERR_CAPACITY, "more than 8 bodies around one rigid contact particle" };
            }
            if (bad & kContactBadDiag) { throw Error{ TSP_ERR_NUMERIC, "Invalid rigid contact diagonal." }; }
            if (bad & kContactBadResidual) { throw Error{ TSP_ERR_NUMERIC, "Non-finite rigid contact density error." }; }
            if (bad & kContactBadPressure) { throw Error{ TSP_ERR_NUMERIC, "Non-finite rigid contact pressure." }; }
        }
    }

    // --- host copies of one scene ---
    void pushScene(uint32_t si, tsp_state_view* s)
    {
        const SceneHost& h = scenes[si];
        if (s->n_fluid != h.nF || s->n_boundary != h.nB || s->n_rigid != h.nR || s->n_bodies != h.nRB)
        {
            throw Error{ TSP_ERR_INVALID, "state view counts differ from the backend" };
        }
        const size_t nRB1 = std::max<uint32_t>(d.nRB, 1);
        sync();
        const auto d2h = [&](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) { hipCheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream), "D2H"); }
        };
        // fluid records unpacked into the SoA fields of the view
        {
            const uint32_t n = h.nF;
            std::vector<float4> q[4];
            const float4* srcs[4] = { d.pos, d.vel, d.vp, d.ap };
            for (int r = 0; r < 4; r++)
            {
                q[r].resize(n);
                d2h(q[r].data(), srcs[r] + h.offF, sizeof(float4) * n);
            }
            sync();
            float* const* f = s->fluid;
            const int comp[4][4] = { { TSP_F_X, TSP_F_Y, TSP_F_Z, TSP_F_PRESSURE }, { TSP_F_VX, TSP_F_VY, TSP_F_VZ, TSP_F_DENSITY },
                                     { TSP_F_VPX, TSP_F_VPY, TSP_F_VPZ, -1 }, { TSP_F_APX, TSP_F_APY, TSP_F_APZ, -1 } };
            for (int r = 0; r < 4; r++)
            {
                for (int c = 0; c < 4; c++)
                {
                    float* out = comp[r][c] >= 0 ? f[comp[r][c]] : nullptr;
                    if (!out) { continue; }
                    for (uint32_t i = 0; i < n; i++) { out[i] = reinterpret_cast<const float*>(&q[r][i])[c]; }
                }
            }
            d2h(f[TSP_F_SOURCE], d.source + h.offF, sizeof(float) * n);
            d2h(f[TSP_F_DIAG], d.diag + h.offF, sizeof(float) * n);
        }
        d2h(s->fluid_id, d.id + h.offF, sizeof(uint64_t) * h.nF);
        {
            const uint32_t n = h.nB;
            std::vector<float4> q(n);
            d2h(q.data(), d.bpos + h.offB, sizeof(float4) * n);
            sync();
            float* const* f = s->boundary;
            const int comp[4] = { TSP_B_X, TSP_B_Y, TSP_B_Z, TSP_B_PSI };
            for (int c = 0; c < 4; c++)
            {
                if (!f[comp[c]]) { continue; }
                for (uint32_t i = 0; i < n; i++) { f[comp[c]][i] = reinterpret_cast<const float*>(&q[i])[c]; }
            }
            d2h(f[TSP_B_VOL], d.bvol + h.offB, sizeof(float) * n);
        }
        d2h(s->boundary_id, idB + h.offB, sizeof(uint64_t) * h.nB);
        {
            const uint32_t n = h.nR;
            float* const* f = s->rigid;
            d2h(f[TSP_R_XREL], d.rxrel + h.offR, sizeof(float) * n);
            d2h(f[TSP_R_YREL], d.ryrel + h.offR, sizeof(float) * n);
            d2h(f[TSP_R_ZREL], d.rzrel + h.offR, sizeof(float) * n);
            std::vector<float4> wp(n), wv(n);
            d2h(wp.data(), d.rpos + h.offR, sizeof(float4) * n);
            d2h(wv.data(), d.rvel + h.offR, sizeof(float4) * n);
            sync();
            for (uint32_t i = 0; i < n; i++)
            {
                if (f[TSP_R_PSI]) { f[TSP_R_PSI][i] = wp[i].w; }
                if (f[TSP_R_VOL]) { f[TSP_R_VOL][i] = wv[i].w; }
                if (s->rigid_body) { s->rigid_body[i] = h.rbodyLocal[i]; }
            }
        }
        d2h(s->rigid_id, idR + h.offR, sizeof(uint64_t) * h.nR);
        for (size_t f = 0; f < TSP_RB_COUNT; f++) { d2h(s->body[f], d.rb + f * nRB1 + h.offRB, sizeof(float) * h.nRB); }
        sync();
        for (uint32_t b = 0; b < h.nRB; b++)
        {
            if (s->body_id) { s->body_id[b] = h.bodyId[b]; }
            if (s->body_count) { s->body_count[b] = h.bodyCount[b]; }
        }
        s->step = h.step;
        s->time = h.time;
    }

    void sceneInfo(uint32_t si, tsp_step_info* info)
    {
        std::memset(info, 0, sizeof(*info));
        const SceneHost& h = scenes[si];
        const Solve& s = h.last;
        info->step = h.step;
        info->time = h.time;
        info->solver_iters = h.nF ? s.iters : 0;
        info->flags = h.flags;
        info->avg_err_raw = s.avgRaw; info->avg_err_pos = s.avgPos; info->avg_err_abs = s.avgAbs;
        info->err_sum_raw = s.sumRaw; info->err_sum_pos = s.sumPos; info->err_sum_abs = s.sumAbs;
        info->sweeps_total = h.sweepsTotal;
        info->frame_builds = h.frameBuilds;
        for (int a = 0; a < 3; a++) { info->frame_orig[a] = h.frame.orig[a]; info->frame_dim[a] = h.frame.dim[a]; }
        // list totals count every scene of the context
        info->n_ff = last.entries[0]; info->n_fb = last.entries[1]; info->n_fr = last.entries[2];
        info->n_rf = last.rfTotal; info->n_rb = last.rbTotal; info->n_rr = last.rrTotal;
        info->list_bytes = 2ull * (last.poolUnits[0] + last.poolUnits[1] + last.poolUnits[2]);
        info->contact_iters = s.contactIters;
        info->contact_avg_err = s.contactAvg;
        info->contact_converged = s.contactConverged;
        info->contact_boundary = s.contactBoundary;
        info->contacts = s.contacts;
        if (h.nRB)
        {
            sync();
            const size_t nRB1 = std::max<uint32_t>(d.nRB, 1);
            for (size_t f = 0; f < 6; f++)
            {
                hipCheck(hipMemcpy(&info->rigid_force_torque[f], d.rbForce + f * nRB1 + h.offRB, sizeof(float), hipMemcpyDeviceToHost),
                         "D2H");
            }
        }
    }

    // --- statistics of one scene on the device (FillStatisticsRecord, gpu_backend.cpp:1067-1208) ---
    std::vector<void*> statAllocs;
    uint32_t* statHist = nullptr;
    uint64_t* statCounts = nullptr;   // 6 relations x (min, max, p99, sum)
    double* statFluid = nullptr;      // sums of raw, positive, absolute error and |v|, min and max |v|
    double* statPartial = nullptr;
    uint64_t statPartialCap = 0;

    // with noAlloc the kernels take the device's static statistics scratch instead of the
    // context's buffers, and nothing is allocated
    void statistics(uint32_t si, tsp_statistics* out, bool noAlloc = false)
    {
        const SceneHost& h = scenes[si];
        const uint64_t nb = (static_cast<uint64_t>(h.nF) + kBlock - 1) / kBlock;
        uint32_t* statHist = this->statHist;
        uint64_t* statCounts = this->statCounts;
        double* statFluid = this->statFluid;
        double* statPartial = this->statPartial;
        if (noAlloc)
        {
            uint32_t maxFluid = 0;
            if (!statsScratch(device, &statHist, &statCounts, &statFluid, &statPartial, &maxFluid))
            {
                throw Error{ TSP_ERR_STATE, "the statistics scratch of this device is not set up" };
            }
            if (h.nF > maxFluid)
            {
                throw Error{ TSP_ERR_CAPACITY, "the allocation-free statistics take scenes of at most 2^24 fluid particles" };
            }
        }
        else if (!statHist || statPartialCap < nb)
        {
            sync();
            releaseList(statAllocs);
            statHist = this->statHist = alloc<uint32_t>(65536, statAllocs);
            statCounts = this->statCounts = alloc<uint64_t>(24, statAllocs);
            statFluid = this->statFluid = alloc<double>(6, statAllocs);
            statPartialCap = std::max<uint64_t>(nb, 1024);
            statPartial = this->statPartial = alloc<double>(6 * statPartialCap, statAllocs);
        }
        const uint16_t* c16[3] = { d.cnt[REL_FF] + h.offF, d.cnt[REL_FB] + h.offF, d.cnt[REL_FR] + h.offF };
        for (int r = 0; r < 3; r++)
        {
            launchStatsCounts(stream, c16[r], nullptr, 0xffffu, h.nF, statHist, statCounts + 4 * r);
        }
        const uint32_t* c32[3] = { d.rfCnt + h.offR, d.rbCnt + h.offR, d.rrCnt + h.offR };
        for (int r = 0; r < 3; r++) { launchStatsCounts(stream, nullptr, c32[r], 0u, h.nR, statHist, statCounts + 4 * (3 + r)); }
        launchStatsFluid(stream, d.vel, h.offF, h.nF, h.params.rest_density, statPartial, statFluid);
        uint64_t cnt[24];
        double fl[6];
        hipCheck(hipMemcpyAsync(cnt, statCounts, sizeof(cnt), hipMemcpyDeviceToHost, stream), "D2H");
        hipCheck(hipMemcpyAsync(fl, statFluid, sizeof(fl), hipMemcpyDeviceToHost, stream), "D2H");
        sync();
        std::memset(out, 0, sizeof(*out));
        out->n_fluid = h.nF; out->n_boundary = h.nB; out->n_rigid = h.nR;
        out->solver_iters = h.nF ? h.last.iters : 0;
        for (int r = 0; r < 6; r++)
        {
            const uint32_t n = r < 3 ? h.nF : h.nR;
            out->nbr_min[r] = static_cast<uint32_t>(cnt[4 * r]);
            out->nbr_max[r] = static_cast<uint32_t>(cnt[4 * r + 1]);
            out->nbr_p99[r] = static_cast<uint32_t>(cnt[4 * r + 2]);
            out->nbr_avg[r] = n ? static_cast<double>(cnt[4 * r + 3]) / n : 0.0;
        }
        const double inv = h.nF ? 1.0 / h.nF : 0.0;
        for (int q = 0; q < 3; q++) { out->dens_rel_err_avg[q] = fl[q] * inv; }
        out->vel_avg = fl[3] * inv;
        out->vel_min = h.nF ? fl[4] : 0.0;
        out->vel_max = h.nF ? fl[5] : 0.0;
        out->cfl = out->vel_max * h.params.time_step / h.params.particle_size;
    }

    // --- export snapshot of one scene (SimBackend_GPU::PackExportSnapshot, gpu_backend.cpp:346-551) ---
    std::vector<void*> expAllocs;
    uint32_t* expFlags = nullptr;     // per fluid particle of the scene, then their scan
    uint32_t* expTmp = nullptr;
    float* expXyz = nullptr;
    float* expSpeed = nullptr;
    uint64_t expCap = 0;

    void exportSnapshot(uint32_t si, uint32_t threshold, uint64_t meshIdBegin, tsp_export_view* v)
    {
        if (!listsValid) { throw Error{ TSP_ERR_STATE, "export needs a neighbor search since the last layout" }; }
        const SceneHost& h = scenes[si];
        v->step = h.step;
        v->time = h.time;
        v->n_fluid = 0;
        v->n_boundary = v->boundary_xyz ? h.nB : 0;
        v->n_rigid = h.nR;
        if (expCap < h.nF + 1ull)
        {
            sync();
            releaseList(expAllocs);
            expCap = std::max<uint64_t>(h.nF + 1ull, 1024);
            expFlags = alloc<uint32_t>(expCap, expAllocs);
            expTmp = alloc<uint32_t>(scanTmpWords(expCap), expAllocs);
            zero(expTmp, sizeof(uint32_t) * scanTmpWords(expCap));
            expXyz = alloc<float>(3 * expCap, expAllocs);
            expSpeed = alloc<float>(expCap, expAllocs);
        }
        launchExportSurface(stream, d, h.offF, h.nF, threshold, expFlags, expTmp, expXyz, expSpeed);
        uint32_t count = 0, scanFailed = 0;
        hipCheck(hipMemcpyAsync(&count, expFlags + h.nF, sizeof(uint32_t), hipMemcpyDeviceToHost, stream), "D2H");
        hipCheck(hipMemcpyAsync(&scanFailed, &d.status->scanFail, sizeof(uint32_t), hipMemcpyDeviceToHost, stream), "D2H");
        sync();
        if (scanFailed) { throw Error{ TSP_ERR_STATE, "a one-pass scan gave up waiting for a predecessor tile" }; }
        v->n_fluid = count;
        if (count > v->cap_fluid) { throw Error{ TSP_ERR_CAPACITY, "the fluid export arrays hold fewer particles than the surface set" }; }
        if (count && v->fluid_xyz) { d2h(v->fluid_xyz, expXyz, sizeof(float) * 3ull * count); }
        if (count && v->fluid_vel_mag) { d2h(v->fluid_vel_mag, expSpeed, sizeof(float) * count); }
        if (v->boundary_xyz && h.nB)
        {
            std::vector<float4> b(h.nB);
            d2h(b.data(), d.bpos + h.offB, sizeof(float4) * h.nB);
            for (uint32_t i = 0; i < h.nB; i++)
            {
                v->boundary_xyz[3ull * i] = b[i].x; v->boundary_xyz[3ull * i + 1] = b[i].y; v->boundary_xyz[3ull * i + 2] = b[i].z;
            }
            if (v->boundary_vol) { d2h(v->boundary_vol, d.bvol + h.offB, sizeof(float) * h.nB); }
            if (v->boundary_is_mesh)
            {
                std::vector<uint64_t> ids(h.nB);
                d2h(ids.data(), idB + h.offB, sizeof(uint64_t) * h.nB);
                for (uint32_t i = 0; i < h.nB; i++) { v->boundary_is_mesh[i] = ids[i] >= meshIdBegin ? 1u : 0u; }
            }
        }
        if (h.nR)
        {
            std::vector<float4> r(h.nR);
            if (v->rigid_xyz)
            {
                d2h(r.data(), d.rpos + h.offR, sizeof(float4) * h.nR);
                for (uint32_t i = 0; i < h.nR; i++)
                {
                    v->rigid_xyz[3ull * i] = r[i].x; v->rigid_xyz[3ull * i + 1] = r[i].y; v->rigid_xyz[3ull * i + 2] = r[i].z;
                }
            }
            if (v->rigid_vol)
            {
                d2h(r.data(), d.rvel + h.offR, sizeof(float4) * h.nR);
                for (uint32_t i = 0; i < h.nR; i++) { v->rigid_vol[i] = r[i].w; }
            }
            if (v->rigid_body)
            {
                d2h(v->rigid_body, d.rbody + h.offR, sizeof(uint32_t) * h.nR);
                for (uint32_t i = 0; i < h.nR; i++) { v->rigid_body[i] -= h.offRB; }
            }
        }
    }

    // one relayout: the listed scenes leave, new scenes join at the end
    void update(uint32_t nRemove, const uint32_t* removeIds, uint32_t nAdd, const tsp_params* params,
                const tsp_state_view* const* views, const float* domains, uint32_t* outIds,
                const tsp::scene::Plan* const* plans = nullptr)
    {
        if ((nRemove && !removeIds) || (nAdd && (!params || !views))) { throw Error{ TSP_ERR_INVALID, "null scene arrays" }; }
        std::vector<bool> drop(nScenes(), false);
        for (uint32_t q = 0; q < nRemove; q++)
        {
            const int at = indexOf(removeIds[q]);
            if (at < 0) { throw Error{ TSP_ERR_INVALID, "unknown scene id" }; }
            drop[static_cast<size_t>(at)] = true;
        }
        std::vector<Source> src;
        for (uint32_t q = 0; q < nScenes(); q++)
        {
            if (drop[q]) { continue; }
            Source x;
            x.keep = static_cast<int>(q);
            src.push_back(std::move(x));
        }
        for (uint32_t q = 0; q < nAdd; q++)
        {
            if (!views[q]) { throw Error{ TSP_ERR_INVALID, "null state" }; }
            Source x;
            x.view = views[q];
            x.plan = plans ? plans[q] : nullptr;
            x.host = newScene(params[q], views[q], domains ? domains + 6 * q : nullptr, domains ? domains + 6 * q + 3 : nullptr);
            src.push_back(std::move(x));
        }
        std::vector<uint32_t> ids;
        for (Source& x : src)
        {
            if (x.keep < 0) { x.host.id = nextId++; ids.push_back(x.host.id); }
        }
        const uint32_t nKept = static_cast<uint32_t>(src.size()) - nAdd;
        relayout(src);
        // the new scenes' frames and key ranges are laid out now, so a context that cannot hold
        // them refuses them here rather than at the next step
        try
        {
            layoutFrames();
        }
        catch (const Error& e)
        {
            if (e.code != TSP_ERR_CAPACITY || nAdd == 0) { throw; }
            std::vector<Source> back(nKept);
            for (uint32_t q = 0; q < nKept; q++) { back[q].keep = static_cast<int>(q); }
            relayout(back);
            throw;
        }
        if (outIds) { for (uint32_t q = 0; q < nAdd; q++) { outIds[q] = ids[q]; } }
    }

    uint32_t sceneAt(uint32_t id) const
    {
        const int at = indexOf(id);
        if (at < 0) { throw Error{ TSP_ERR_INVALID, "unknown scene id" }; }
        return static_cast<uint32_t>(at);
    }

    void requireScene() const
    {
        if (scenes.empty()) { throw Error{ TSP_ERR_STATE, "no state" }; }
    }
};

namespace
{
    template <typename Fn>
    int guarded(tsp_ctx* ctx, Fn&& fn)
    {
        if (!ctx) { return TSP_ERR_INVALID; }
        try
        {
            hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
            ctx->err.clear();
            return fn();
        }
        catch (const Error& e)
        {
            ctx->err = e.msg;
            return e.code;
        }
        catch (const std::exception& e)
        {
            ctx->err = e.what();
            return TSP_ERR_HIP;
        }
    }
}

extern "C" {

const char* tsp_name(void) { return "HIP (gfx950)"; }

int tsp_create(int device, tsp_ctx** out)
{
    if (!out) { return TSP_ERR_INVALID; }
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) { return TSP_ERR_HIP; }
    tsp_ctx* c = new tsp_ctx();
    c->device = device;
    // statsScratchInit: the device address of this device's statistics scratch, once per device
    if (hipSetDevice(device) != hipSuccess || statsScratchInit(device) != hipSuccess
        || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess
        || hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess
        || hipEventCreateWithFlags(&c->forkEv, hipEventDisableTiming) != hipSuccess
        || hipEventCreateWithFlags(&c->joinEv, hipEventDisableTiming) != hipSuccess
        || hipHostMalloc(reinterpret_cast<void**>(&c->hostBlock), tsp_ctx::kSolveOff) != hipSuccess)
    {
        delete c;
        return TSP_ERR_HIP;
    }
    std::memset(c->hostBlock, 0, tsp_ctx::kSolveOff);
    c->hostStatus = reinterpret_cast<Status*>(c->hostBlock);
    c->hostSolve = reinterpret_cast<Solve*>(c->hostBlock + tsp_ctx::kSolveOff);
    *out = c;
    return TSP_OK;
}

void tsp_destroy(tsp_ctx* ctx)
{
    if (!ctx) { return; }
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) { (void)hipStreamSynchronize(ctx->stream); }
    delete ctx;
}

const char* tsp_last_error(const tsp_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tsp_set_params(tsp_ctx* ctx, const tsp_params* p)
{
    return guarded(ctx, [&]() {
        if (!p) { throw Error{ TSP_ERR_INVALID, "null params" }; }
        validateParams(*p);
        ctx->defaults = *p;
        ctx->paramsSet = true;
        // every resident scene takes the parameters; a new kernel radius or cell ratio needs a
        // new cell frame
        bool changed = false;
        for (SceneHost& h : ctx->scenes)
        {
            if (std::memcmp(&h.params, p, sizeof(tsp_params)) == 0) { continue; }
            if (p->kernel_radius != h.params.kernel_radius || p->grid_cell_size_ratio != h.params.grid_cell_size_ratio)
            {
                h.frameValid = false;
            }
            h.params = *p;
            changed = true;
        }
        if (changed) { ctx->uploadConsts(); }
        return TSP_OK;
    });
}

int tsp_set_step_graph(tsp_ctx* ctx, int enable)
{
    return guarded(ctx, [&]() {
        ctx->stepGraphOn = enable != 0;
        return TSP_OK;
    });
}

int tsp_set_cell_limit(tsp_ctx* ctx, uint64_t max_cells)
{
    return guarded(ctx, [&]() {
        if (max_cells == 0 || max_cells > 0xffffffffull) { throw Error{ TSP_ERR_INVALID, "cell limit outside 1..2^32-1" }; }
        ctx->cellLimit = max_cells;
        return TSP_OK;
    });
}

int tsp_set_domain(tsp_ctx* ctx, const float mn[3], const float mx[3])
{
    return guarded(ctx, [&]() {
        if (!mn || !mx) { throw Error{ TSP_ERR_INVALID, "null domain" }; }
        for (int a = 0; a < 3; a++) { ctx->domMin[a] = mn[a]; ctx->domMax[a] = mx[a]; }
        ctx->domainSet = true;
        // the hysteresis test reads the domain of every step (IsInsideSimDomain); a frame's
        // extent changes only at its next rebuild
        for (SceneHost& h : ctx->scenes)
        {
            for (int a = 0; a < 3; a++) { h.domMin[a] = mn[a]; h.domMax[a] = mx[a]; }
            h.domainSet = true;
            Frame& f = h.frame;
            f.domainSet = (mx[0] > mn[0] && mx[1] > mn[1] && mx[2] > mn[2]) ? 1 : 0;
            for (int a = 0; a < 3; a++) { f.domMin[a] = mn[a]; f.domMax[a] = mx[a]; }
        }
        if (!ctx->scenes.empty() && !ctx->layoutDirty)
        {
            std::vector<Frame> frames;
            for (const SceneHost& h : ctx->scenes) { frames.push_back(h.frame); }
            ctx->h2d(ctx->d.frames, frames.data(), sizeof(Frame) * frames.size());
        }
        return TSP_OK;
    });
}

int tsp_pull(tsp_ctx* ctx, const tsp_state_view* s)
{
    return guarded(ctx, [&]() {
        if (!s) { throw Error{ TSP_ERR_INVALID, "null state" }; }
        if (!ctx->paramsSet) { throw Error{ TSP_ERR_STATE, "pull before params" }; }
        // the context holds exactly this one scene afterwards
        std::vector<uint32_t> all;
        for (const SceneHost& h : ctx->scenes) { all.push_back(h.id); }
        float dom[6];
        for (int a = 0; a < 3; a++) { dom[a] = ctx->domMin[a]; dom[3 + a] = ctx->domMax[a]; }
        ctx->update(static_cast<uint32_t>(all.size()), all.data(), 1, &ctx->defaults, &s, ctx->domainSet ? dom : nullptr, nullptr);
        return TSP_OK;
    });
}

int tsp_pull_spec(tsp_ctx* ctx, const tsp_scene_spec* spec)
{
    return guarded(ctx, [&]() {
        if (!spec) { throw Error{ TSP_ERR_INVALID, "null scene spec" }; }
        if (!ctx->paramsSet) { throw Error{ TSP_ERR_STATE, "pull before params" }; }
        tsp::scene::Plan plan;
        const int rc = tsp::scene::MakePlan(*spec, plan);
        if (rc != TSP_OK) { throw Error{ rc, "invalid scene spec" }; }
        if (plan.nF > 0x7fffffffull || plan.nB > 0x7fffffffull) { throw Error{ TSP_ERR_CAPACITY, "particle count exceeds the local particle limit" }; }
        const tsp_state_view v = planView(plan);
        const tsp_state_view* vp = &v;
        const tsp::scene::Plan* pp = &plan;
        // the spec's domain becomes the context's, as SetSimDomainBounds with the scene's bounds
        for (int a = 0; a < 3; a++) { ctx->domMin[a] = plan.domMin[a]; ctx->domMax[a] = plan.domMax[a]; }
        ctx->domainSet = true;
        std::vector<uint32_t> all;
        for (const SceneHost& h : ctx->scenes) { all.push_back(h.id); }
        float dom[6];
        for (int a = 0; a < 3; a++) { dom[a] = plan.domMin[a]; dom[3 + a] = plan.domMax[a]; }
        ctx->update(static_cast<uint32_t>(all.size()), all.data(), 1, &ctx->defaults, &vp, dom, nullptr, &pp);
        return TSP_OK;
    });
}

int tsp_push(tsp_ctx* ctx, tsp_state_view* s)
{
    return guarded(ctx, [&]() {
        if (!s) { throw Error{ TSP_ERR_INVALID, "null state" }; }
        if (ctx->scenes.empty()) { throw Error{ TSP_ERR_STATE, "push before pull" }; }
        ctx->pushScene(0, s);
        return TSP_OK;
    });
}

int tsp_neighbor_search(tsp_ctx* ctx)
{
    return guarded(ctx, [&]() {
        if (ctx->scenes.empty() || !ctx->paramsSet) { throw Error{ TSP_ERR_STATE, "neighbor search before pull/params" }; }
        for (SceneHost& h : ctx->scenes) { h.flags = 0; }
        if (ctx->d.nF == 0 && ctx->d.nB == 0 && ctx->d.nR == 0) { return TSP_OK; }
        ctx->neighborSearch(true);
        ctx->sync();
        ctx->prof.resolve();
        return TSP_OK;
    });
}

int tsp_step(tsp_ctx* ctx, uint32_t n)
{
    return guarded(ctx, [&]() {
        if (ctx->scenes.empty() || !ctx->paramsSet) { throw Error{ TSP_ERR_STATE, "step before pull/params" }; }
        ctx->steps(n);
        return TSP_OK;
    });
}

int tsp_get_step_info(tsp_ctx* ctx, tsp_step_info* info)
{
    return guarded(ctx, [&]() {
        if (!info) { throw Error{ TSP_ERR_INVALID, "null info" }; }
        ctx->requireScene();
        ctx->sceneInfo(0, info);
        return TSP_OK;
    });
}

int tsp_get_counts(tsp_ctx* ctx, uint32_t counts[4])
{
    return guarded(ctx, [&]() {
        if (!counts) { throw Error{ TSP_ERR_INVALID, "null counts" }; }
        if (ctx->scenes.empty()) { counts[0] = counts[1] = counts[2] = counts[3] = 0; return TSP_OK; }
        const SceneHost& h = ctx->scenes[0];
        counts[0] = h.nF; counts[1] = h.nB; counts[2] = h.nR; counts[3] = h.nRB;
        return TSP_OK;
    });
}

int tsp_get_neighbors(tsp_ctx* ctx, int rel, uint32_t* counts, uint32_t* idx, uint64_t cap, uint64_t* total)
{
    return guarded(ctx, [&]() {
        if (rel < 0 || rel > 5) { throw Error{ TSP_ERR_INVALID, "relation must be 0..5" }; }
        ctx->requireScene();
        const Dev& d = ctx->d;
        const SceneHost& h = ctx->scenes[0];
        ctx->sync();
        if (rel >= 3)
        {
            // RF is stored as CSR; RB and RR are stored when some scene has entries (the contact
            // solve reads them), which is whenever this scene has any
            const uint32_t n = h.nR;
            std::vector<uint32_t> cnt(n);
            const uint32_t* src = rel == 3 ? d.rfCnt : (rel == 4 ? d.rbCnt : d.rrCnt);
            if (n) { hipCheck(hipMemcpy(cnt.data(), src + h.offR, sizeof(uint32_t) * n, hipMemcpyDeviceToHost), "D2H"); }
            uint64_t tot = 0;
            for (uint32_t i = 0; i < n; i++) { tot += cnt[i]; if (counts) { counts[i] = cnt[i]; } }
            if (total) { *total = tot; }
            if (idx && cap >= tot && tot)
            {
                const uint32_t* offs = rel == 3 ? d.rfOff : (rel == 4 ? d.rbOff : d.rrOff);
                const uint32_t* list = rel == 3 ? d.rfIdx : (rel == 4 ? d.rbIdx : d.rrIdx);
                if (rel != 3 && !ctx->contactLists) { throw Error{ TSP_ERR_STATE, "RB/RR lists were not stored" }; }
                uint32_t first = 0;
                hipCheck(hipMemcpy(&first, offs + h.offR, sizeof(uint32_t), hipMemcpyDeviceToHost), "D2H");
                hipCheck(hipMemcpy(idx, list + first, sizeof(uint32_t) * tot, hipMemcpyDeviceToHost), "D2H");
                const uint32_t base = rel == 3 ? h.offF : (rel == 4 ? h.offB : h.offR);
                for (uint64_t q = 0; q < tot; q++) { idx[q] -= base; }
            }
            return TSP_OK;
        }
        // the scene's slices (its fluid segment starts at a multiple of the block size and is
        // padded to one)
        const uint32_t n = h.nF;
        const uint32_t s0 = h.offF / kSlice, nSlices = roundUp(n, kBlock) / kSlice;
        std::vector<uint16_t> cnt(n);
        std::vector<SliceMeta> meta(nSlices);
        // particle i's lists sit at list position (i - i % 256) + laneRank[i] of the scene
        std::vector<uint8_t> laneRank(n);
        if (n)
        {
            hipCheck(hipMemcpy(cnt.data(), d.cnt[rel] + h.offF, sizeof(uint16_t) * n, hipMemcpyDeviceToHost), "D2H");
            hipCheck(hipMemcpy(meta.data(), d.meta[rel] + s0, sizeof(SliceMeta) * nSlices, hipMemcpyDeviceToHost), "D2H");
#if TSP_TUNE_LANE_SORT
            hipCheck(hipMemcpy(laneRank.data(), d.laneRank + h.offF, n, hipMemcpyDeviceToHost), "D2H");
#else
            for (uint32_t i = 0; i < n; i++) { laneRank[i] = static_cast<uint8_t>(i % kBlock); }
#endif
        }
        const uint32_t mask = 0xffffu;
        uint64_t tot = 0;
        for (uint32_t i = 0; i < n; i++) { tot += cnt[i] & mask; if (counts) { counts[i] = cnt[i] & mask; } }
        if (total) { *total = tot; }
        if (!idx || cap < tot || !tot) { return TSP_OK; }
        const uint64_t units = ctx->last.poolUnits[rel];
        std::vector<uint16_t> pool(units);
        if (units) { hipCheck(hipMemcpy(pool.data(), d.pool[rel], sizeof(uint16_t) * units, hipMemcpyDeviceToHost), "D2H"); }
        // FF entries of blocks with a halo are halo slots
        const uint32_t b0 = h.blk0, nBlocks = h.nBlk;
        std::vector<uint32_t> haloCount, haloIdx;
        if (rel == REL_FF && nBlocks)
        {
            haloCount.resize(nBlocks);
            haloIdx.resize(static_cast<size_t>(nBlocks) * kHaloMax);
            hipCheck(hipMemcpy(haloCount.data(), d.haloCount + b0, sizeof(uint32_t) * nBlocks, hipMemcpyDeviceToHost), "D2H");
            hipCheck(hipMemcpy(haloIdx.data(), d.haloIdx + static_cast<size_t>(b0) * kHaloMax, sizeof(uint32_t) * haloIdx.size(),
                               hipMemcpyDeviceToHost), "D2H");
        }
        const uint32_t sub = rel == REL_FF ? h.offF : (rel == REL_FB ? h.offB : h.offR);
        // host decoder of the SELL-32 storage (ReadCompressedSell32Neighbor, gpu_backend.cpp:42-65)
        uint64_t w = 0;
        for (uint32_t i = 0; i < n; i++)
        {
            const uint32_t p = i - i % kBlock + laneRank[i];
            const SliceMeta& m = meta[p / kSlice];
            const uint32_t lane = p % kSlice;
            const uint32_t blk = i / kBlock;
            const bool slots = rel == REL_FF && haloCount[blk] <= kHaloMax;
            for (uint32_t k = 0; k < (cnt[i] & mask); k++)
            {
                uint32_t j;
                if (m.format == kFmtDelta16) { j = (slots ? 0u : m.base) + pool[m.offset + k * kSlice + lane]; }
                else { std::memcpy(&j, &pool[m.offset + 2ull * (k * kSlice + lane)], 4); }
                if (slots) { j = haloIdx[static_cast<size_t>(blk) * kHaloMax + (j >> kSlotShift)]; }
                idx[w++] = j - sub;
            }
        }
        return TSP_OK;
    });
}

int tsp_get_sort(tsp_ctx* ctx, uint32_t* orderF, uint64_t* keysF, uint32_t* orderB, uint64_t* keysB)
{
    return guarded(ctx, [&]() {
        ctx->requireScene();
        const Dev& d = ctx->d;
        const SceneHost& h = ctx->scenes[0];
        ctx->sync();
        // scene-local Morton ids; out-of-frame ids as all ones
        // key i of a key array (u32, or u64 in two words)
        const auto widen = [&](uint64_t* dst, const uint32_t* src, uint32_t off, uint32_t n) {
            std::vector<uint64_t> k(n);
            if (n && d.key64)
            {
                hipCheck(hipMemcpy(k.data(), reinterpret_cast<const uint64_t*>(src) + off, sizeof(uint64_t) * n,
                                   hipMemcpyDeviceToHost), "D2H");
            }
            else if (n)
            {
                std::vector<uint32_t> k32(n);
                hipCheck(hipMemcpy(k32.data(), src + off, sizeof(uint32_t) * n, hipMemcpyDeviceToHost), "D2H");
                for (uint32_t i = 0; i < n; i++) { k[i] = k32[i]; }
            }
            for (uint32_t i = 0; i < n; i++) { dst[i] = k[i] >= h.frame.emptyKey ? ~0ull : k[i] - h.frame.keyBase; }
        };
        if (orderF && h.nF)
        {
            hipCheck(hipMemcpy(orderF, d.idxF + h.offF, sizeof(uint32_t) * h.nF, hipMemcpyDeviceToHost), "D2H");
            for (uint32_t i = 0; i < h.nF; i++) { orderF[i] -= h.offF; }
        }
        if (keysF) { widen(keysF, d.keyF, h.offF, h.nF); }
        if ((orderB || keysB) && h.nB)
        {
            // the boundary order relative to the pulled order, from the ids (sequence ids on pull)
            launchKeys(ctx->stream, d, kPopB, d.bpos, d.nB, false, d.keyF2, d.idxF2);
            ctx->sync();
            if (keysB) { widen(keysB, d.keyF2, h.offB, h.nB); }
            if (orderB)
            {
                std::vector<uint64_t> idb(h.nB);
                hipCheck(hipMemcpy(idb.data(), ctx->idB + h.offB, sizeof(uint64_t) * h.nB, hipMemcpyDeviceToHost), "D2H");
                for (uint32_t i = 0; i < h.nB; i++) { orderB[i] = static_cast<uint32_t>(idb[i]); }
            }
        }
        return TSP_OK;
    });
}

int tsp_get_rigid_world(tsp_ctx* ctx, float* xyz, float* vel)
{
    return guarded(ctx, [&]() {
        ctx->requireScene();
        const Dev& d = ctx->d;
        const SceneHost& h = ctx->scenes[0];
        ctx->sync();
        std::vector<float4> q(h.nR);
        const float4* src[2] = { d.rpos, d.rvel };
        float* dst[2] = { xyz, vel };
        for (int r = 0; r < 2; r++)
        {
            if (!dst[r] || !h.nR) { continue; }
            hipCheck(hipMemcpy(q.data(), src[r] + h.offR, sizeof(float4) * h.nR, hipMemcpyDeviceToHost), "D2H");
            for (uint32_t i = 0; i < h.nR; i++) { dst[r][3 * i] = q[i].x; dst[r][3 * i + 1] = q[i].y; dst[r][3 * i + 2] = q[i].z; }
        }
        return TSP_OK;
    });
}

int tsp_profile_enable(tsp_ctx* ctx, int on)
{
    return guarded(ctx, [&]() {
        ctx->sync();
        ctx->prof.resolve();
        ctx->prof.on = on != 0;
        ctx->prof.acc.clear();
        ctx->prof.order.clear();
        return TSP_OK;
    });
}

int tsp_profile_read(tsp_ctx* ctx, char* names, uint64_t* launches, double* ms, int cap)
{
    if (!ctx) { return TSP_ERR_INVALID; }
    const int rc = guarded(ctx, [&]() {
        ctx->sync();
        ctx->prof.resolve();
        return TSP_OK;
    });
    if (rc != TSP_OK) { return rc; }
    int q = 0;
    for (const std::string& n : ctx->prof.order)
    {
        if (q < cap)
        {
            const auto& v = ctx->prof.acc[n];
            if (names) { std::snprintf(names + 48 * q, 48, "%s", n.c_str()); }
            if (launches) { launches[q] = v.first; }
            if (ms) { ms[q] = v.second; }
        }
        q++;
    }
    return q;
}

int tsp_get_statistics(tsp_ctx* ctx, tsp_statistics* out)
{
    return guarded(ctx, [&]() {
        if (!out) { throw Error{ TSP_ERR_INVALID, "null statistics" }; }
        ctx->requireScene();
        ctx->statistics(0, out);
        return TSP_OK;
    });
}

int tsp_get_memory(tsp_ctx* ctx, tsp_memory* out)
{
    return guarded(ctx, [&]() {
        if (!out) { throw Error{ TSP_ERR_INVALID, "null memory record" }; }
        std::memset(out, 0, sizeof(*out));
        for (int c = 0; c < 3; c++)
        {
            out->particles[c] = ctx->memBytes[tsp_ctx::MEM_PF + c];
            out->search[c] = ctx->memBytes[tsp_ctx::MEM_SF + c];
        }
        out->lists[0] = ctx->memBytes[tsp_ctx::MEM_LF];
        out->lists[2] = ctx->memBytes[tsp_ctx::MEM_LR];
        out->misc = ctx->memBytes[tsp_ctx::MEM_MISC];
        out->alloc_bytes = ctx->memTotal;
        out->peak_alloc_bytes = ctx->memPeak;
        return TSP_OK;
    });
}

int tsp_get_statistics_noalloc(tsp_ctx* ctx, tsp_statistics* out)
{
    return guarded(ctx, [&]() {
        if (!out) { throw Error{ TSP_ERR_INVALID, "null statistics" }; }
        ctx->requireScene();
        ctx->statistics(0, out, true);
        return TSP_OK;
    });
}

int tsp_batch_scene_statistics(tsp_ctx* ctx, uint32_t scene_id, tsp_statistics* out)
{
    return guarded(ctx, [&]() {
        if (!out) { throw Error{ TSP_ERR_INVALID, "null statistics" }; }
        ctx->statistics(ctx->sceneAt(scene_id), out);
        return TSP_OK;
    });
}

int tsp_export_snapshot(tsp_ctx* ctx, uint32_t surface_nbr_count_threshold, uint64_t mesh_id_begin, tsp_export_view* out)
{
    return guarded(ctx, [&]() {
        if (!out) { throw Error{ TSP_ERR_INVALID, "null export view" }; }
        ctx->requireScene();
        ctx->exportSnapshot(0, surface_nbr_count_threshold, mesh_id_begin, out);
        return TSP_OK;
    });
}

int tsp_batch_export_snapshot(tsp_ctx* ctx, uint32_t scene_id, uint32_t surface_nbr_count_threshold, uint64_t mesh_id_begin,
                              tsp_export_view* out)
{
    return guarded(ctx, [&]() {
        if (!out) { throw Error{ TSP_ERR_INVALID, "null export view" }; }
        ctx->exportSnapshot(ctx->sceneAt(scene_id), surface_nbr_count_threshold, mesh_id_begin, out);
        return TSP_OK;
    });
}

int tsp_synchronize(tsp_ctx* ctx)
{
    return guarded(ctx, [&]() {
        ctx->sync();
        return TSP_OK;
    });
}

// --- several scenes in one context ---

int tsp_batch_update(tsp_ctx* ctx, uint32_t n_remove, const uint32_t* remove_ids, uint32_t n_add, const tsp_params* params,
                     const tsp_state_view* const* views, const float* domains, uint32_t* out_ids)
{
    return guarded(ctx, [&]() {
        ctx->update(n_remove, remove_ids, n_add, params, views, domains, out_ids);
        if (!ctx->paramsSet && n_add) { ctx->defaults = params[0]; ctx->paramsSet = true; }
        return TSP_OK;
    });
}

int tsp_batch_add_scene(tsp_ctx* ctx, const tsp_params* params, const tsp_state_view* state, const float domain[6],
                        uint32_t* scene_id)
{
    return tsp_batch_update(ctx, 0, nullptr, 1, params, &state, domain, scene_id);
}

int tsp_batch_add_scene_spec(tsp_ctx* ctx, const tsp_params* params, const tsp_scene_spec* spec, uint32_t* scene_id)
{
    return guarded(ctx, [&]() {
        if (!params || !spec) { throw Error{ TSP_ERR_INVALID, "null params or scene spec" }; }
        tsp::scene::Plan plan;
        const int rc = tsp::scene::MakePlan(*spec, plan);
        if (rc != TSP_OK) { throw Error{ rc, "invalid scene spec" }; }
        if (plan.nF > 0x7fffffffull || plan.nB > 0x7fffffffull) { throw Error{ TSP_ERR_CAPACITY, "particle count exceeds the local particle limit" }; }
        const tsp_state_view v = planView(plan);
        const tsp_state_view* vp = &v;
        const tsp::scene::Plan* pp = &plan;
        float dom[6];
        for (int a = 0; a < 3; a++) { dom[a] = plan.domMin[a]; dom[3 + a] = plan.domMax[a]; }
        if (!ctx->paramsSet) { ctx->defaults = *params; ctx->paramsSet = true; }
        ctx->update(0, nullptr, 1, params, &vp, dom, scene_id, &pp);
        return TSP_OK;
    });
}

int tsp_batch_update_specs(tsp_ctx* ctx, uint32_t n_remove, const uint32_t* remove_ids, uint32_t n_add,
                           const tsp_params* params, const tsp_scene_spec* const* specs, uint32_t* out_ids)
{
    return guarded(ctx, [&]() {
        if (n_add && (!params || !specs)) { throw Error{ TSP_ERR_INVALID, "null params or scene specs" }; }
        std::vector<tsp::scene::Plan> plans(n_add);
        std::vector<tsp_state_view> views(n_add);
        std::vector<const tsp_state_view*> vptr(n_add);
        std::vector<const tsp::scene::Plan*> pptr(n_add);
        std::vector<float> dom(6ull * n_add);
        for (uint32_t q = 0; q < n_add; q++)
        {
            if (!specs[q]) { throw Error{ TSP_ERR_INVALID, "null scene spec" }; }
            const int rc = tsp::scene::MakePlan(*specs[q], plans[q]);
            if (rc != TSP_OK) { throw Error{ rc, "invalid scene spec" }; }
            if (plans[q].nF > 0x7fffffffull || plans[q].nB > 0x7fffffffull)
            {
                throw Error{ TSP_ERR_CAPACITY, "particle count exceeds the local particle limit" };
            }
            views[q] = planView(plans[q]);
            vptr[q] = &views[q];
            pptr[q] = &plans[q];
            for (int a = 0; a < 3; a++) { dom[6 * q + a] = plans[q].domMin[a]; dom[6 * q + 3 + a] = plans[q].domMax[a]; }
        }
        if (!ctx->paramsSet && n_add) { ctx->defaults = params[0]; ctx->paramsSet = true; }
        ctx->update(n_remove, remove_ids, n_add, params, vptr.data(), dom.data(), out_ids, pptr.data());
        return TSP_OK;
    });
}

int tsp_batch_remove_scene(tsp_ctx* ctx, uint32_t scene_id)
{
    return tsp_batch_update(ctx, 1, &scene_id, 0, nullptr, nullptr, nullptr, nullptr);
}

int tsp_batch_scene_ids(tsp_ctx* ctx, uint32_t* ids, uint32_t cap, uint32_t* n)
{
    return guarded(ctx, [&]() {
        if (n) { *n = ctx->nScenes(); }
        for (uint32_t q = 0; ids && q < ctx->nScenes() && q < cap; q++) { ids[q] = ctx->scenes[q].id; }
        return TSP_OK;
    });
}

int tsp_batch_scene_counts(tsp_ctx* ctx, uint32_t scene_id, uint32_t counts[4])
{
    return guarded(ctx, [&]() {
        if (!counts) { throw Error{ TSP_ERR_INVALID, "null counts" }; }
        const SceneHost& h = ctx->scenes[ctx->sceneAt(scene_id)];
        counts[0] = h.nF; counts[1] = h.nB; counts[2] = h.nR; counts[3] = h.nRB;
        return TSP_OK;
    });
}

int tsp_batch_push_scene(tsp_ctx* ctx, uint32_t scene_id, tsp_state_view* state)
{
    return guarded(ctx, [&]() {
        if (!state) { throw Error{ TSP_ERR_INVALID, "null state" }; }
        ctx->pushScene(ctx->sceneAt(scene_id), state);
        return TSP_OK;
    });
}

int tsp_batch_scene_info(tsp_ctx* ctx, uint32_t scene_id, tsp_step_info* info)
{
    return guarded(ctx, [&]() {
        if (!info) { throw Error{ TSP_ERR_INVALID, "null info" }; }
        ctx->sceneInfo(ctx->sceneAt(scene_id), info);
        return TSP_OK;
    });
}

int tsp_batch_step(tsp_ctx* ctx, uint32_t n_steps)
{
    return guarded(ctx, [&]() {
        ctx->requireScene();
        ctx->steps(n_steps);
        return TSP_OK;
    });
}

uint64_t tsp_batch_footprint(uint32_t n_fluid, uint32_t n_boundary, uint32_t n_rigid, uint32_t n_bodies)
{
    // device bytes a scene holds while resident: per fluid slot (segment rounded to 256) about
    // 190 B of records and second buffers, 16 B of sort keys, 160 B of SELL-32 lists at the
    // initial pool sizes, 26 B of halo, counts, list positions and the wide-list index, and up to
    // 24 B of cell tables; per boundary particle 64 B; per rigid particle about 400 B (records,
    // keys, the RF list); per body 256 B
    const uint64_t f = static_cast<uint64_t>(roundUp(n_fluid, kBlock));
    return f * 416ull + static_cast<uint64_t>(n_boundary) * 64ull + static_cast<uint64_t>(n_rigid) * 400ull
         + static_cast<uint64_t>(n_bodies) * 256ull + (64ull << 10);
}

} // extern "C"
