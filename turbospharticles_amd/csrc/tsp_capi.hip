// tsp_capi.hip -- host orchestration of the gfx950 IISPH step and its C ABI (include/tsp_hip.h).
//
// One context = one device, one HIP stream, all particle state resident in HBM. The step
// sequence follows SimBackend_CPU::Step_IISPH (CPU/backend/sph_solver.cpp:34-319); the
// differences from the reference's CUDA host (GPU/backend/gpu_backend.cpp:1274-1594) are that
// list sizes and solver convergence live on the device: the host reads one small status block
// after the neighbor count (pool growth and frame hysteresis) and after groups of Jacobi sweeps,
// and sweeps queued past convergence exit at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "tsp_device.h"
#include "tsp_hip.h"

namespace tsp
{
    void launchResetSearch(hipStream_t, Status*);
    void launchKeys(hipStream_t, const float4*, uint32_t, const Frame*, Status*, bool, uint32_t*, uint32_t*);
    void launchKeysSorted(hipStream_t, const float4*, uint32_t, const Frame*, Status*, uint32_t*, uint32_t*);
    void launchRadixSort(hipStream_t, uint32_t*, uint32_t*, uint32_t*, uint32_t*, uint32_t, uint32_t, uint32_t*, uint32_t*);
    void launchPermute(hipStream_t, const Gather&, const uint32_t*);
    void launchClearCells(hipStream_t, const uint32_t*, uint32_t, const Frame*, uint2*);
    void launchClearAll(hipStream_t, uint2*, uint64_t);
    void launchBuildCells(hipStream_t, const uint32_t*, uint32_t, const Frame*, uint2*);
    void launchQueryCount(hipStream_t, const Dev&, float, int);
    void launchScanSlices(hipStream_t, const Dev&, uint32_t*, uint32_t*);
    void launchQueryFill(hipStream_t, const Dev&, float, int);
    void launchQueryRigid(hipStream_t, const Dev&, float, int, bool, uint32_t*);
    void launchHaloBuild(hipStream_t, const Dev&);
    void launchAabb(hipStream_t, const float4*, uint32_t, uint32_t*);
    void launchDensity(hipStream_t, const Dev&, const Consts&);
    void launchNonPressure(hipStream_t, const Dev&, const Consts&);
    void launchSourceDiag(hipStream_t, const Dev&, const Consts&);
    void launchPressureAcc(hipStream_t, const Dev&, const Consts&, bool);
    void launchJacobi(hipStream_t, const Dev&, const Consts&);
    void launchFinalize(hipStream_t, const Dev&, const Consts&);
    void launchIntegration(hipStream_t, const Dev&, const Consts&);
    void launchRigidUpdate(hipStream_t, const Dev&, bool);
    void launchRigidClearGravity(hipStream_t, const Dev&, const Consts&);
    void launchRigidForces(hipStream_t, const Dev&, const Consts&, bool);
    void launchRigidReduce(hipStream_t, const Dev&);
    void launchRigidPredict(hipStream_t, const Dev&, const Consts&);
    void launchRigidIntegrate(hipStream_t, const Dev&, const Consts&);
}

using namespace tsp;

namespace
{
    struct Error
    {
        int code;
        std::string msg;
    };

    void hipCheck(hipError_t e, const char* what)
    {
        if (e != hipSuccess) { throw Error{ TSP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e) }; }
    }

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
}

struct tsp_ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    tsp_params params{};
    bool paramsSet = false;
    float domMin[3] = { 0, 0, 0 }, domMax[3] = { 0, 0, 0 };
    bool domainSet = false;
    bool pulled = false;

    Dev d{};
    // second buffers of the re-order (swapped with the ones in d after each gather)
    float4 *pos2 = nullptr, *vel2 = nullptr, *vp2 = nullptr, *ap2 = nullptr;
    float *source2 = nullptr, *diag2 = nullptr;
    uint64_t* idF2 = nullptr;
    float4* bpos2 = nullptr;
    float* bvol2 = nullptr;
    uint64_t *idB = nullptr, *idB2 = nullptr;
    uint64_t* idR = nullptr;
    std::vector<uint32_t> hostIdRB, hostCountRB;
    uint32_t* radixHist = nullptr;
    uint32_t* sliceUnits = nullptr; // 3 x (slices + 1) scan buffer
    uint32_t* aabb = nullptr;
    std::vector<void*> allocs;      // particle-count sized
    std::vector<void*> frameAllocs; // frame sized (cell grids)
    std::vector<void*> poolAllocs;

    Frame frame{};
    bool frameValid = false;
    bool cellsClean = false;        // fluid and rigid cell tables hold no runs (frame just built)
    uint32_t keyBits = 0;
    uint32_t reach = 1;
    float cellSize = 0.0f;
    uint32_t frameBuilds = 0;
    bool bSorted = false;

    Status* hostStatus = nullptr; // pinned
    uint64_t step = 0;
    double time = 0.0;
    uint64_t sweepsTotal = 0;
    uint32_t lastSweeps = 0;
    uint32_t flags = 0;
    Status last{};

    Profiler prof;

    ~tsp_ctx()
    {
        releaseAll();
        if (hostStatus) { (void)hipHostFree(hostStatus); }
        if (stream) { (void)hipStreamDestroy(stream); }
    }

    template <typename T>
    T* alloc(size_t n, std::vector<void*>& list)
    {
        void* p = nullptr;
        hipCheck(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
        list.push_back(p);
        return static_cast<T*>(p);
    }

    static void releaseList(std::vector<void*>& list)
    {
        for (void* p : list) { (void)hipFree(p); }
        list.clear();
    }

    void releaseAll()
    {
        releaseList(poolAllocs);
        releaseList(frameAllocs);
        releaseList(allocs);
    }

    void sync() { hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize"); }

    void readStatus()
    {
        hipCheck(hipGetLastError(), "kernel launch");
        hipCheck(hipMemcpyAsync(hostStatus, d.status, sizeof(Status), hipMemcpyDeviceToHost, stream), "status D2H");
        sync();
        prof.resolve();
        last = *hostStatus;
    }

    // --- constants of the kernels (solver_constants.h:20-36, simd.h:359-370, hoisted kernel constants) ---
    Consts consts() const
    {
        const tsp_params& p = params;
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
        k.nF = d.nF;
        return k;
    }

    // --- pools of the SELL-32 lists ---
    void allocPools(const uint64_t cap[3], uint64_t rfCap)
    {
        releaseList(poolAllocs);
        for (int r = 0; r < 3; r++)
        {
            d.poolCap[r] = cap[r];
            d.pool[r] = alloc<uint16_t>(cap[r], poolAllocs);
        }
        d.rfCap = rfCap;
        d.rfIdx = alloc<uint32_t>(rfCap, poolAllocs);
    }

    // --- cell frame: UpdateCellFrame, CPU/backend/neighbor_search.cpp:1902-2006 ---
    void buildFrame()
    {
        const tsp_params& p = params;
        const float cs = p.kernel_radius * p.grid_cell_size_ratio;
        const float span = p.kernel_radius / cs;
        const float sm = std::max(0.0f, span - 1e-6f);
        reach = std::max<uint32_t>(1u, static_cast<uint32_t>(std::ceil(sm)));
        cellSize = cs;
        const float inv = 1.0f / cs;

        // scene AABB of F, B and R on the device (order-free min/max)
        const uint32_t init[6] = { 0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u };
        hipCheck(hipMemcpyAsync(aabb, init, sizeof(init), hipMemcpyHostToDevice, stream), "aabb init");
        launchAabb(stream, d.pos, d.nF, aabb);
        launchAabb(stream, d.bpos, d.nB, aabb);
        launchAabb(stream, d.rpos, d.nR, aabb);
        uint32_t bits[6];
        hipCheck(hipMemcpyAsync(bits, aabb, sizeof(bits), hipMemcpyDeviceToHost, stream), "aabb D2H");
        sync();
        const auto toFloat = [](uint32_t u)
        {
            const uint32_t v = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
            float f;
            std::memcpy(&f, &v, 4);
            return f;
        };
        float mn[3], mx[3];
        for (int a = 0; a < 3; a++) { mn[a] = toFloat(bits[a]); mx[a] = toFloat(bits[3 + a]); }
        const bool domOk = domainSet && domMax[0] > domMin[0] && domMax[1] > domMin[1] && domMax[2] > domMin[2];
        if (domOk)
        {
            for (int a = 0; a < 3; a++)
            {
                mn[a] = std::clamp(mn[a], domMin[a], domMax[a]);
                mx[a] = std::clamp(mx[a], domMin[a], domMax[a]);
            }
        }
        const auto floorToIntHost = [](float v)
        {
            const int i = static_cast<int>(v);
            return (v < static_cast<float>(i)) ? i - 1 : i;
        };
        const int64_t margin = static_cast<int64_t>(reach) + 8;
        Frame f{};
        uint32_t dimMax = 0;
        for (int a = 0; a < 3; a++)
        {
            const int64_t o = static_cast<int64_t>(floorToIntHost(mn[a] * inv)) - margin;
            const int64_t e = static_cast<int64_t>(floorToIntHost(mx[a] * inv)) + margin;
            if (o > e || e - o + 1 > 1024)
            {
                throw Error{ TSP_ERR_CAPACITY, "cell frame exceeds 1024 cells per axis (32-bit Morton keys)" };
            }
            f.orig[a] = static_cast<int32_t>(o);
            f.dim[a] = static_cast<uint32_t>(e - o + 1);
            dimMax = std::max(dimMax, f.dim[a]);
            f.innerLo[a] = f.orig[a] + static_cast<int32_t>(reach);
            f.innerHi[a] = f.orig[a] + static_cast<int32_t>(f.dim[a]) - 1 - static_cast<int32_t>(reach);
        }
        uint32_t axisBits = 1;
        if (dimMax >= 2) { axisBits = 0; for (uint32_t v = dimMax - 1; v; v >>= 1) { axisBits++; } }
        keyBits = 3u * axisBits + 1u;
        f.emptyKey = 1u << (3u * axisBits);
        f.cellSizeInv = inv;
        f.domainSet = domOk ? 1 : 0;
        for (int a = 0; a < 3; a++) { f.domMin[a] = domMin[a]; f.domMax[a] = domMax[a]; }
        frame = f;
        hipCheck(hipMemcpyAsync(d.frame, &frame, sizeof(Frame), hipMemcpyHostToDevice, stream), "frame H2D");

        // dense cell tables over the frame
        releaseList(frameAllocs);
        const uint64_t nCells = static_cast<uint64_t>(f.dim[0]) * f.dim[1] * f.dim[2];
        d.cellF = alloc<uint2>(nCells, frameAllocs);
        d.cellB = alloc<uint2>(nCells, frameAllocs);
        d.cellR = alloc<uint2>(nCells, frameAllocs);
        launchClearAll(stream, d.cellF, nCells);
        launchClearAll(stream, d.cellB, nCells);
        launchClearAll(stream, d.cellR, nCells);
        cellsClean = true;

        // boundary: re-sorted only if its ids are not already non-decreasing (neighbor_search.cpp:1755-1811)
        if (d.nB)
        {
            launchResetSearch(stream, d.status);
            launchKeysSorted(stream, d.bpos, d.nB, d.frame, d.status, d.keyF, d.idxF);
            uint32_t unsorted = 0;
            hipCheck(hipMemcpyAsync(&unsorted, &d.status->bUnsorted, 4, hipMemcpyDeviceToHost, stream), "flag D2H");
            sync();
            if (unsorted)
            {
                launchRadixSort(stream, d.keyF, d.idxF, d.keyF2, d.idxF2, d.nB, keyBits, radixHist, d.scanTmp);
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
            launchBuildCells(stream, d.keyF, d.nB, d.frame, d.cellB);
            sync();
        }
        frameValid = true;
        frameBuilds++;
        flags |= TSP_FLAG_FRAME_REBUILT;
    }

    // --- re-order, construct and query (DataReOrder + Construct + Query_Count/Fill) ---
    void neighborSearch(bool allFields, const Consts& k)
    {
        if (!frameValid) { buildFrame(); }
        for (int attempt = 0; attempt < 4; attempt++)
        {
            {
                Scope sc(prof, stream, "reset");
                launchResetSearch(stream, d.status);
                // the runs of the previous sorted keys are the only non-empty cells
                if (!cellsClean)
                {
                    launchClearCells(stream, d.keyF, d.nF, d.frame, d.cellF);
                    launchClearCells(stream, d.keyR, d.nR, d.frame, d.cellR);
                }
            }
            if (d.nR)
            {
                Scope sc(prof, stream, "keys_sort_rigid");
                launchKeys(stream, d.rpos, d.nR, d.frame, d.status, true, d.keyR, d.idxR);
                launchRadixSort(stream, d.keyR, d.idxR, d.keyR2, d.idxR2, d.nR, keyBits, radixHist, d.scanTmp);
            }
            {
                Scope sc(prof, stream, "keys_sort");
                launchKeys(stream, d.pos, d.nF, d.frame, d.status, true, d.keyF, d.idxF);
                launchRadixSort(stream, d.keyF, d.idxF, d.keyF2, d.idxF2, d.nF, keyBits, radixHist, d.scanTmp);
            }
            // the step re-orders positions, velocities (with density) and pressure (pos.w, the
            // warm start); the standalone search re-orders every field
            const auto swapFields = [&]()
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
            };
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
            swapFields();
            {
                Scope sc(prof, stream, "cells");
                launchBuildCells(stream, d.keyF, d.nF, d.frame, d.cellF);
                launchBuildCells(stream, d.keyR, d.nR, d.frame, d.cellR);
                cellsClean = false;
            }
            {
                Scope sc(prof, stream, "query_count");
                launchQueryCount(stream, d, params.kernel_radius_sq, static_cast<int>(reach));
                launchQueryRigid(stream, d, params.kernel_radius_sq, static_cast<int>(reach), false, d.scanTmp);
            }
            {
                Scope sc(prof, stream, "scan");
                launchScanSlices(stream, d, sliceUnits, d.scanTmp);
            }
            readStatus();
            if (last.abort)
            {
                // a particle left the inner frame: undo the permutation and rebuild the frame first
                swapFields();
                frameValid = false;
                buildFrame();
                continue;
            }
            if (last.overflow & 2u) { throw Error{ TSP_ERR_CAPACITY, "neighbor count exceeds the u16 count storage" }; }
            if (last.overflow & 5u)
            {
                uint64_t cap[3];
                for (int r = 0; r < 3; r++) { cap[r] = std::max<uint64_t>(d.poolCap[r], last.poolUnits[r] + last.poolUnits[r] / 8 + 64); }
                const uint64_t rfCap = std::max<uint64_t>(d.rfCap, last.rfTotal + last.rfTotal / 8 + 64);
                for (int r = 0; r < 3; r++)
                {
                    if (cap[r] > 0xffffffffull) { throw Error{ TSP_ERR_CAPACITY, "SELL pool exceeds 2^32 u16 units" }; }
                }
                allocPools(cap, rfCap);
                flags |= TSP_FLAG_LIST_REGROWN;
            }
            {
                Scope sc(prof, stream, "query_fill");
                launchQueryFill(stream, d, params.kernel_radius_sq, static_cast<int>(reach));
                launchQueryRigid(stream, d, params.kernel_radius_sq, static_cast<int>(reach), true, d.scanTmp);
            }
            {
                Scope sc(prof, stream, "halo");
                launchHaloBuild(stream, d);
            }
            if (last.rbTotal + last.rrTotal > 0) { flags |= TSP_FLAG_RIGID_CONTACT; }
            (void)k;
            return;
        }
        throw Error{ TSP_ERR_STATE, "cell frame rebuild did not converge" };
    }

    // --- one IISPH step (Step_IISPH, sph_solver.cpp:34-319) ---
    void stepOnce()
    {
        const Consts k = consts();
        neighborSearch(false, k);
        {
            Scope sc(prof, stream, "density");
            launchDensity(stream, d, k);
        }
        if (d.nRB) { launchRigidClearGravity(stream, d, k); }
        {
            Scope sc(prof, stream, "nonpressure");
            launchNonPressure(stream, d, k);
        }
        if (d.nRB)
        {
            Scope sc(prof, stream, "rigid");
            launchRigidForces(stream, d, k, false);
            launchRigidReduce(stream, d);
            launchRigidPredict(stream, d, k);
            launchRigidUpdate(stream, d, true);
        }
        {
            Scope sc(prof, stream, "source_diag");
            launchSourceDiag(stream, d, k);
        }
        // relaxed Jacobi sweeps; each group is followed by one status read, sweeps queued past
        // convergence return at once
        uint32_t queued = 0;
        const uint32_t maxSweeps = params.max_solver_iters > 1 ? params.max_solver_iters - 1 : 1;
        uint32_t group = std::max<uint32_t>(1u, lastSweeps > 1 ? lastSweeps - 1 : 1u);
        if (d.nF == 0) { group = 0; }
        while (d.nF > 0)
        {
            const uint32_t n = std::min(group, maxSweeps - queued);
            if (n == 0) { throw Error{ TSP_ERR_STATE, "pressure solve did not stop at max_solver_iters" }; }
            for (uint32_t s = 0; s < n; s++)
            {
                {
                    Scope sc(prof, stream, "pressure_acc");
                    launchPressureAcc(stream, d, k, true);
                }
                Scope sc(prof, stream, "jacobi");
                launchJacobi(stream, d, k);
            }
            {
                Scope sc(prof, stream, "finalize");
                launchFinalize(stream, d, k);
            }
            queued += n;
            readStatus();
            if (last.converged) { break; }
            group = 1;
        }
        lastSweeps = d.nF ? last.sweeps : 0;
        sweepsTotal += lastSweeps;
        {
            Scope sc(prof, stream, "pressure_acc_final");
            launchPressureAcc(stream, d, k, false);
        }
        if (d.nRB)
        {
            Scope sc(prof, stream, "rigid");
            launchRigidForces(stream, d, k, true);
            launchRigidClearGravity(stream, d, k);
            launchRigidReduce(stream, d);
        }
        {
            Scope sc(prof, stream, "integration");
            launchIntegration(stream, d, k);
        }
        if (d.nRB)
        {
            Scope sc(prof, stream, "rigid");
            launchRigidPredict(stream, d, k);
            launchRigidIntegrate(stream, d, k);
            launchRigidUpdate(stream, d, false);
        }
        if (d.nF && params.termination_cond != 0 && last.iters >= params.max_solver_iters)
        {
            const float e = params.termination_cond == 1 ? last.avgRaw : params.termination_cond == 2 ? last.avgPos : last.avgAbs;
            if (!(e <= params.density_dev_target)) { flags |= TSP_FLAG_MAX_ITERS; }
        }
        step++;
        time += static_cast<double>(params.time_step);
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

    void validateParams(const tsp_params& p)
    {
        if (!(p.particle_size > 0.0f) || !(p.kernel_radius > 0.0f) || !(p.kernel_radius_sq > 0.0f)
            || !(p.rest_density > 0.0f) || !(p.time_step > 0.0f) || !(p.grid_cell_size_ratio > 0.0f)
            || p.max_solver_iters < 1)
        {
            throw Error{ TSP_ERR_INVALID, "invalid solver parameters" };
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
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess
        || hipHostMalloc(reinterpret_cast<void**>(&c->hostStatus), sizeof(Status)) != hipSuccess)
    {
        delete c;
        return TSP_ERR_HIP;
    }
    std::memset(c->hostStatus, 0, sizeof(Status));
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
        const bool frameChanged = ctx->paramsSet && (p->kernel_radius != ctx->params.kernel_radius
                                                     || p->grid_cell_size_ratio != ctx->params.grid_cell_size_ratio);
        ctx->params = *p;
        ctx->paramsSet = true;
        if (frameChanged) { ctx->frameValid = false; }
        return TSP_OK;
    });
}

int tsp_set_domain(tsp_ctx* ctx, const float mn[3], const float mx[3])
{
    return guarded(ctx, [&]() {
        if (!mn || !mx) { throw Error{ TSP_ERR_INVALID, "null domain" }; }
        for (int a = 0; a < 3; a++) { ctx->domMin[a] = mn[a]; ctx->domMax[a] = mx[a]; }
        ctx->domainSet = true;
        // the hysteresis test reads the domain of every step (IsInsideSimDomain); the frame
        // extent itself changes only at the next rebuild
        if (ctx->frameValid)
        {
            Frame& f = ctx->frame;
            f.domainSet = (mx[0] > mn[0] && mx[1] > mn[1] && mx[2] > mn[2]) ? 1 : 0;
            for (int a = 0; a < 3; a++) { f.domMin[a] = mn[a]; f.domMax[a] = mx[a]; }
            hipCheck(hipMemcpyAsync(ctx->d.frame, &f, sizeof(Frame), hipMemcpyHostToDevice, ctx->stream), "frame H2D");
            ctx->sync();
        }
        return TSP_OK;
    });
}

int tsp_pull(tsp_ctx* ctx, const tsp_state_view* s)
{
    return guarded(ctx, [&]() {
        if (!s) { throw Error{ TSP_ERR_INVALID, "null state" }; }
        if (s->n_fluid > 0x7fffffffu || s->n_boundary > 0x7fffffffu || s->n_rigid > 0x7fffffffu)
        {
            throw Error{ TSP_ERR_CAPACITY, "particle count exceeds the local particle limit" };
        }
        if (s->n_rigid && !s->rigid_body) { throw Error{ TSP_ERR_INVALID, "rigid particles need body ids" }; }
        for (uint32_t i = 0; i < s->n_rigid; i++)
        {
            if (s->rigid_body[i] >= s->n_bodies) { throw Error{ TSP_ERR_INVALID, "Invalid rigid body index." }; }
        }
        ctx->sync();
        ctx->releaseAll();
        ctx->d = Dev{};
        Dev& d = ctx->d;
        d.nF = s->n_fluid; d.nB = s->n_boundary; d.nR = s->n_rigid; d.nRB = s->n_bodies;
        auto& A = ctx->allocs;
        hipStream_t st = ctx->stream;
        const auto h2d = [&](void* dst, const void* src, size_t bytes) {
            if (bytes) { hipCheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), "H2D"); }
            ctx->sync();
        };
        const auto field = [](const float* src, uint32_t i, float dflt) { return src ? src[i] : dflt; };
        // fluid: the SoA fields of the view packed into the float4 records
        {
            const uint32_t n = d.nF;
            std::vector<float4> q[4];
            for (auto& v : q) { v.resize(n); }
            std::vector<float> src1(n), diag1(n);
            float* const* f = s->fluid;
            for (uint32_t i = 0; i < n; i++)
            {
                q[0][i] = make_float4(field(f[TSP_F_X], i, 0), field(f[TSP_F_Y], i, 0), field(f[TSP_F_Z], i, 0), field(f[TSP_F_PRESSURE], i, 0));
                q[1][i] = make_float4(field(f[TSP_F_VX], i, 0), field(f[TSP_F_VY], i, 0), field(f[TSP_F_VZ], i, 0), field(f[TSP_F_DENSITY], i, 0));
                q[2][i] = make_float4(field(f[TSP_F_VPX], i, 0), field(f[TSP_F_VPY], i, 0), field(f[TSP_F_VPZ], i, 0), 0.0f);
                q[3][i] = make_float4(field(f[TSP_F_APX], i, 0), field(f[TSP_F_APY], i, 0), field(f[TSP_F_APZ], i, 0), 0.0f);
                src1[i] = field(f[TSP_F_SOURCE], i, 0);
                diag1[i] = field(f[TSP_F_DIAG], i, 0);
            }
            float4** dst[4] = { &d.pos, &d.vel, &d.vp, &d.ap };
            float4** dst2[4] = { &ctx->pos2, &ctx->vel2, &ctx->vp2, &ctx->ap2 };
            for (int r = 0; r < 4; r++)
            {
                *dst[r] = ctx->alloc<float4>(n, A);
                *dst2[r] = ctx->alloc<float4>(n, A);
                h2d(*dst[r], q[r].data(), sizeof(float4) * n);
            }
            d.source = ctx->alloc<float>(n, A);
            d.diag = ctx->alloc<float>(n, A);
            ctx->source2 = ctx->alloc<float>(n, A);
            ctx->diag2 = ctx->alloc<float>(n, A);
            h2d(d.source, src1.data(), sizeof(float) * n);
            h2d(d.diag, diag1.data(), sizeof(float) * n);
        }
        std::vector<uint64_t> seq;
        const auto ids = [&](const uint64_t* src, uint32_t n) -> const uint64_t* {
            if (src) { return src; }
            seq.resize(n);
            for (uint32_t i = 0; i < n; i++) { seq[i] = i; }
            return seq.data();
        };
        d.id = ctx->alloc<uint64_t>(d.nF, A);
        ctx->idF2 = ctx->alloc<uint64_t>(d.nF, A);
        h2d(d.id, ids(s->fluid_id, d.nF), sizeof(uint64_t) * d.nF);
        // boundary (ensureBoundaryCoefficientsPresent defaults: psi = massB, V = psi / rho0)
        {
            const uint32_t n = d.nB;
            const float h = ctx->params.particle_size, rho0 = ctx->params.rest_density;
            const float massB = rho0 * h * h * h;
            const float volB = rho0 > 0.0f ? massB / rho0 : 0.0f;
            std::vector<float4> q(n);
            std::vector<float> vol(n);
            float* const* f = s->boundary;
            for (uint32_t i = 0; i < n; i++)
            {
                q[i] = make_float4(field(f[TSP_B_X], i, 0), field(f[TSP_B_Y], i, 0), field(f[TSP_B_Z], i, 0), field(f[TSP_B_PSI], i, massB));
                vol[i] = field(f[TSP_B_VOL], i, volB);
            }
            d.bpos = ctx->alloc<float4>(n, A);
            ctx->bpos2 = ctx->alloc<float4>(n, A);
            d.bvol = ctx->alloc<float>(n, A);
            ctx->bvol2 = ctx->alloc<float>(n, A);
            h2d(d.bpos, q.data(), sizeof(float4) * n);
            h2d(d.bvol, vol.data(), sizeof(float) * n);
        }
        ctx->idB = ctx->alloc<uint64_t>(d.nB, A);
        ctx->idB2 = ctx->alloc<uint64_t>(d.nB, A);
        h2d(ctx->idB, ids(s->boundary_id, d.nB), sizeof(uint64_t) * d.nB);
        // rigid particles: body-frame offsets; world records are written from the body state below
        {
            const uint32_t n = d.nR;
            float* const* f = s->rigid;
            std::vector<float> rel[3];
            std::vector<float4> wp(n), wv(n), wvp(n);
            for (int a2 = 0; a2 < 3; a2++) { rel[a2].resize(n); }
            for (uint32_t i = 0; i < n; i++)
            {
                rel[0][i] = field(f[TSP_R_XREL], i, 0);
                rel[1][i] = field(f[TSP_R_YREL], i, 0);
                rel[2][i] = field(f[TSP_R_ZREL], i, 0);
                wp[i] = make_float4(0.0f, 0.0f, 0.0f, field(f[TSP_R_PSI], i, 0));
                wv[i] = make_float4(0.0f, 0.0f, 0.0f, field(f[TSP_R_VOL], i, 0));
                wvp[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            float** rr[3] = { &d.rxrel, &d.ryrel, &d.rzrel };
            for (int a2 = 0; a2 < 3; a2++)
            {
                *rr[a2] = ctx->alloc<float>(n, A);
                h2d(*rr[a2], rel[a2].data(), sizeof(float) * n);
            }
            d.rpos = ctx->alloc<float4>(n, A);
            d.rvel = ctx->alloc<float4>(n, A);
            d.rvp = ctx->alloc<float4>(n, A);
            h2d(d.rpos, wp.data(), sizeof(float4) * n);
            h2d(d.rvel, wv.data(), sizeof(float4) * n);
            h2d(d.rvp, wvp.data(), sizeof(float4) * n);
        }
        d.rbody = ctx->alloc<uint32_t>(d.nR, A);
        h2d(d.rbody, s->rigid_body, sizeof(uint32_t) * d.nR);
        ctx->idR = ctx->alloc<uint64_t>(d.nR, A);
        h2d(ctx->idR, ids(s->rigid_id, d.nR), sizeof(uint64_t) * d.nR);
        d.rfx = ctx->alloc<float>(d.nR, A);
        d.rfy = ctx->alloc<float>(d.nR, A);
        d.rfz = ctx->alloc<float>(d.nR, A);
        // bodies
        d.rb = ctx->alloc<float>(static_cast<size_t>(TSP_RB_COUNT) * d.nRB, A);
        d.rbPred = ctx->alloc<float>(9ull * std::max<uint32_t>(d.nRB, 1), A);
        d.rbForce = ctx->alloc<float>(6ull * std::max<uint32_t>(d.nRB, 1), A);
        {
            std::vector<float> zeros(9ull * std::max<uint32_t>(d.nRB, 1), 0.0f);
            h2d(d.rbForce, zeros.data(), sizeof(float) * 6 * std::max<uint32_t>(d.nRB, 1));
            h2d(d.rbPred, zeros.data(), sizeof(float) * 9 * std::max<uint32_t>(d.nRB, 1));
        }
        for (int f = 0; f < TSP_RB_COUNT; f++)
        {
            if (s->body[f]) { h2d(d.rb + static_cast<size_t>(f) * d.nRB, s->body[f], sizeof(float) * d.nRB); }
            else
            {
                std::vector<float> v(d.nRB, (f == TSP_RB_A00 || f == TSP_RB_A11 || f == TSP_RB_A22) ? 1.0f : 0.0f);
                h2d(d.rb + static_cast<size_t>(f) * d.nRB, v.data(), sizeof(float) * d.nRB);
            }
        }
        ctx->hostIdRB.assign(d.nRB, 0);
        ctx->hostCountRB.assign(d.nRB, 0);
        for (uint32_t b = 0; b < d.nRB; b++)
        {
            ctx->hostIdRB[b] = s->body_id ? s->body_id[b] : b;
            ctx->hostCountRB[b] = s->body_count ? s->body_count[b] : 0;
        }
        // BuildRigidParticleIndicesByBody (sync/state_sim.h:609-631)
        {
            std::vector<uint32_t> off(d.nRB + 1, 0), idx(d.nR);
            for (uint32_t i = 0; i < d.nR; i++) { off[s->rigid_body[i] + 1]++; }
            for (uint32_t b = 0; b < d.nRB; b++) { off[b + 1] += off[b]; }
            std::vector<uint32_t> w(off);
            for (uint32_t i = 0; i < d.nR; i++) { idx[w[s->rigid_body[i]]++] = i; }
            d.byBodyOff = ctx->alloc<uint32_t>(d.nRB + 1, A);
            d.byBodyIdx = ctx->alloc<uint32_t>(d.nR, A);
            h2d(d.byBodyOff, off.data(), sizeof(uint32_t) * (d.nRB + 1));
            h2d(d.byBodyIdx, idx.data(), sizeof(uint32_t) * d.nR);
        }
        // neighbor search buffers
        const uint32_t nSort = std::max(d.nF, std::max(d.nB, d.nR));
        d.keyF = ctx->alloc<uint32_t>(nSort, A);
        d.keyF2 = ctx->alloc<uint32_t>(nSort, A);
        d.idxF = ctx->alloc<uint32_t>(nSort, A);
        d.idxF2 = ctx->alloc<uint32_t>(nSort, A);
        d.keyR = ctx->alloc<uint32_t>(d.nR, A);
        d.keyR2 = ctx->alloc<uint32_t>(d.nR, A);
        d.idxR = ctx->alloc<uint32_t>(d.nR, A);
        d.idxR2 = ctx->alloc<uint32_t>(d.nR, A);
        const uint64_t nTiles = (nSort + 2047ull) / 2048ull;
        ctx->radixHist = ctx->alloc<uint32_t>(256ull * nTiles, A);
        ctx->aabb = ctx->alloc<uint32_t>(6, A);
        const uint32_t nSlices = (d.nF + kSlice - 1) / kSlice;
        for (int r = 0; r < 3; r++)
        {
            d.cnt[r] = ctx->alloc<uint16_t>(d.nF, A);
            d.meta[r] = ctx->alloc<SliceMeta>(nSlices, A);
        }
        ctx->sliceUnits = ctx->alloc<uint32_t>(3ull * (nSlices + 1), A);
        {
            // block sums of the scans: radix digit offsets, slice units, RF offsets
            const uint64_t m = std::max<uint64_t>(256ull * nTiles, std::max<uint64_t>(nSlices + 1ull, d.nR + 1ull));
            d.scanTmp = ctx->alloc<uint32_t>((m + 2047ull) / 2048ull + 1ull, A);
        }
        d.ffWideMin = ctx->alloc<uint32_t>(d.nF, A);
        {
            const uint64_t nBlocks = (d.nF + kBlock - 1) / kBlock;
            d.haloIdx = ctx->alloc<uint32_t>(nBlocks * kHaloMax, A);
            d.haloCount = ctx->alloc<uint32_t>(nBlocks, A);
            // no block has a halo until the first search builds one
            hipCheck(hipMemsetAsync(d.haloCount, 0xff, sizeof(uint32_t) * std::max<uint64_t>(nBlocks, 1), st), "memset");
        }
        d.rfCnt = ctx->alloc<uint32_t>(d.nR, A);
        d.rbCnt = ctx->alloc<uint32_t>(d.nR, A);
        d.rrCnt = ctx->alloc<uint32_t>(d.nR, A);
        d.rfOff = ctx->alloc<uint32_t>(d.nR + 1ull, A);
        d.errPartial = ctx->alloc<double>(3ull * ((d.nF + kBlock - 1) / kBlock), A);
        d.status = ctx->alloc<Status>(1, A);
        d.frame = ctx->alloc<Frame>(1, A);
        {
            const Status zero{};
            h2d(d.status, &zero, sizeof(Status));
        }
        // initial pool sizes: about 40 FF, 16 FB and 8 FR entries per particle, grown on demand
        const uint64_t cap[3] = { 48ull * d.nF + 1024, 16ull * d.nF + 1024, 8ull * d.nF + 1024 };
        ctx->allocPools(cap, 64ull * d.nR + 1024);
        ctx->sync();
        launchRigidUpdate(st, d, false);
        ctx->sync();
        ctx->frameValid = false;
        ctx->step = s->step;
        ctx->time = s->time;
        ctx->sweepsTotal = 0;
        ctx->lastSweeps = 0;
        ctx->flags = 0;
        ctx->last = Status{};
        ctx->pulled = true;
        return TSP_OK;
    });
}

int tsp_push(tsp_ctx* ctx, tsp_state_view* s)
{
    return guarded(ctx, [&]() {
        if (!s) { throw Error{ TSP_ERR_INVALID, "null state" }; }
        if (!ctx->pulled) { throw Error{ TSP_ERR_STATE, "push before pull" }; }
        const Dev& d = ctx->d;
        if (s->n_fluid != d.nF || s->n_boundary != d.nB || s->n_rigid != d.nR || s->n_bodies != d.nRB)
        {
            throw Error{ TSP_ERR_INVALID, "state view counts differ from the backend" };
        }
        hipStream_t st = ctx->stream;
        const auto d2h = [&](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) { hipCheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st), "D2H"); }
        };
        // fluid records unpacked into the SoA fields of the view
        {
            const uint32_t n = d.nF;
            std::vector<float4> q[4];
            const float4* srcs[4] = { d.pos, d.vel, d.vp, d.ap };
            for (int r = 0; r < 4; r++)
            {
                q[r].resize(n);
                d2h(q[r].data(), srcs[r], sizeof(float4) * n);
            }
            ctx->sync();
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
            d2h(f[TSP_F_SOURCE], d.source, sizeof(float) * n);
            d2h(f[TSP_F_DIAG], d.diag, sizeof(float) * n);
        }
        d2h(s->fluid_id, d.id, sizeof(uint64_t) * d.nF);
        {
            const uint32_t n = d.nB;
            std::vector<float4> q(n);
            d2h(q.data(), d.bpos, sizeof(float4) * n);
            ctx->sync();
            float* const* f = s->boundary;
            const int comp[4] = { TSP_B_X, TSP_B_Y, TSP_B_Z, TSP_B_PSI };
            for (int c = 0; c < 4; c++)
            {
                if (!f[comp[c]]) { continue; }
                for (uint32_t i = 0; i < n; i++) { f[comp[c]][i] = reinterpret_cast<const float*>(&q[i])[c]; }
            }
            d2h(f[TSP_B_VOL], d.bvol, sizeof(float) * n);
        }
        d2h(s->boundary_id, ctx->idB, sizeof(uint64_t) * d.nB);
        {
            const uint32_t n = d.nR;
            float* const* f = s->rigid;
            d2h(f[TSP_R_XREL], d.rxrel, sizeof(float) * n);
            d2h(f[TSP_R_YREL], d.ryrel, sizeof(float) * n);
            d2h(f[TSP_R_ZREL], d.rzrel, sizeof(float) * n);
            std::vector<float4> wp(n), wv(n);
            d2h(wp.data(), d.rpos, sizeof(float4) * n);
            d2h(wv.data(), d.rvel, sizeof(float4) * n);
            ctx->sync();
            for (uint32_t i = 0; i < n; i++)
            {
                if (f[TSP_R_PSI]) { f[TSP_R_PSI][i] = wp[i].w; }
                if (f[TSP_R_VOL]) { f[TSP_R_VOL][i] = wv[i].w; }
            }
        }
        d2h(s->rigid_body, d.rbody, sizeof(uint32_t) * d.nR);
        d2h(s->rigid_id, ctx->idR, sizeof(uint64_t) * d.nR);
        for (int f = 0; f < TSP_RB_COUNT; f++) { d2h(s->body[f], d.rb + static_cast<size_t>(f) * d.nRB, sizeof(float) * d.nRB); }
        ctx->sync();
        for (uint32_t b = 0; b < d.nRB; b++)
        {
            if (s->body_id) { s->body_id[b] = ctx->hostIdRB[b]; }
            if (s->body_count) { s->body_count[b] = ctx->hostCountRB[b]; }
        }
        s->step = ctx->step;
        s->time = ctx->time;
        return TSP_OK;
    });
}

int tsp_neighbor_search(tsp_ctx* ctx)
{
    return guarded(ctx, [&]() {
        if (!ctx->pulled || !ctx->paramsSet) { throw Error{ TSP_ERR_STATE, "neighbor search before pull/params" }; }
        ctx->flags = 0;
        if (ctx->d.nF == 0 && ctx->d.nB == 0 && ctx->d.nR == 0) { return TSP_OK; }
        ctx->neighborSearch(true, ctx->consts());
        ctx->sync();
        ctx->prof.resolve();
        return TSP_OK;
    });
}

int tsp_step(tsp_ctx* ctx, uint32_t n)
{
    return guarded(ctx, [&]() {
        if (!ctx->pulled || !ctx->paramsSet) { throw Error{ TSP_ERR_STATE, "step before pull/params" }; }
        ctx->flags = 0;
        for (uint32_t s = 0; s < n; s++)
        {
            if (ctx->d.nF == 0 && ctx->d.nRB == 0)
            {
                ctx->step++;
                ctx->time += static_cast<double>(ctx->params.time_step);
                continue;
            }
            ctx->stepOnce();
        }
        ctx->readStatus();
        if (ctx->last.rigidDegenerate) { throw Error{ TSP_ERR_NUMERIC, "Degenerate rigid body orientation." }; }
        return TSP_OK;
    });
}

int tsp_get_step_info(tsp_ctx* ctx, tsp_step_info* info)
{
    return guarded(ctx, [&]() {
        if (!info) { throw Error{ TSP_ERR_INVALID, "null info" }; }
        std::memset(info, 0, sizeof(*info));
        const Status& s = ctx->last;
        info->step = ctx->step;
        info->time = ctx->time;
        info->solver_iters = ctx->d.nF ? s.iters : 0;
        info->flags = ctx->flags;
        info->avg_err_raw = s.avgRaw; info->avg_err_pos = s.avgPos; info->avg_err_abs = s.avgAbs;
        info->err_sum_raw = s.sumRaw; info->err_sum_pos = s.sumPos; info->err_sum_abs = s.sumAbs;
        info->sweeps_total = ctx->sweepsTotal;
        info->frame_builds = ctx->frameBuilds;
        for (int a = 0; a < 3; a++) { info->frame_orig[a] = ctx->frame.orig[a]; info->frame_dim[a] = ctx->frame.dim[a]; }
        info->n_ff = s.entries[0]; info->n_fb = s.entries[1]; info->n_fr = s.entries[2];
        info->n_rf = s.rfTotal; info->n_rb = s.rbTotal; info->n_rr = s.rrTotal;
        info->list_bytes = 2ull * (s.poolUnits[0] + s.poolUnits[1] + s.poolUnits[2]);
        if (ctx->d.nRB)
        {
            float ft[6];
            for (int f = 0; f < 6; f++)
            {
                hipCheck(hipMemcpy(&ft[f], ctx->d.rbForce + static_cast<size_t>(f) * ctx->d.nRB, sizeof(float),
                                   hipMemcpyDeviceToHost), "D2H");
                info->rigid_force_torque[f] = ft[f];
            }
        }
        return TSP_OK;
    });
}

int tsp_get_counts(tsp_ctx* ctx, uint32_t counts[4])
{
    return guarded(ctx, [&]() {
        if (!counts) { throw Error{ TSP_ERR_INVALID, "null counts" }; }
        counts[0] = ctx->d.nF; counts[1] = ctx->d.nB; counts[2] = ctx->d.nR; counts[3] = ctx->d.nRB;
        return TSP_OK;
    });
}

int tsp_get_neighbors(tsp_ctx* ctx, int rel, uint32_t* counts, uint32_t* idx, uint64_t cap, uint64_t* total)
{
    return guarded(ctx, [&]() {
        if (rel < 0 || rel > 5) { throw Error{ TSP_ERR_INVALID, "relation must be 0..5" }; }
        if (!ctx->pulled) { throw Error{ TSP_ERR_STATE, "no state" }; }
        const Dev& d = ctx->d;
        ctx->sync();
        if (rel >= 3)
        {
            // RF is stored as CSR; RB and RR are only counted (they drive the contact flag)
            const uint32_t n = d.nR;
            std::vector<uint32_t> cnt(n);
            const uint32_t* src = rel == 3 ? d.rfCnt : (rel == 4 ? d.rbCnt : d.rrCnt);
            if (n) { hipCheck(hipMemcpy(cnt.data(), src, sizeof(uint32_t) * n, hipMemcpyDeviceToHost), "D2H"); }
            uint64_t tot = 0;
            for (uint32_t i = 0; i < n; i++) { tot += cnt[i]; if (counts) { counts[i] = cnt[i]; } }
            if (total) { *total = tot; }
            if (rel == 3 && idx && cap >= tot && tot)
            {
                hipCheck(hipMemcpy(idx, d.rfIdx, sizeof(uint32_t) * tot, hipMemcpyDeviceToHost), "D2H");
            }
            else if (rel != 3 && idx && tot) { throw Error{ TSP_ERR_INVALID, "RB/RR lists are counted, not stored" }; }
            return TSP_OK;
        }
        const uint32_t n = d.nF;
        const uint32_t nSlices = (n + kSlice - 1) / kSlice;
        std::vector<uint16_t> cnt(n);
        std::vector<SliceMeta> meta(nSlices);
        if (n)
        {
            hipCheck(hipMemcpy(cnt.data(), d.cnt[rel], sizeof(uint16_t) * n, hipMemcpyDeviceToHost), "D2H");
            hipCheck(hipMemcpy(meta.data(), d.meta[rel], sizeof(SliceMeta) * nSlices, hipMemcpyDeviceToHost), "D2H");
        }
        const uint32_t mask = rel == 0 ? 0x7fffu : 0xffffu;
        uint64_t tot = 0;
        for (uint32_t i = 0; i < n; i++) { tot += cnt[i] & mask; if (counts) { counts[i] = cnt[i] & mask; } }
        if (total) { *total = tot; }
        if (!idx || cap < tot || !tot) { return TSP_OK; }
        const uint64_t units = ctx->last.poolUnits[rel];
        std::vector<uint16_t> pool(units);
        if (units) { hipCheck(hipMemcpy(pool.data(), d.pool[rel], sizeof(uint16_t) * units, hipMemcpyDeviceToHost), "D2H"); }
        // FF entries of blocks with a halo are halo slots
        const uint32_t nBlocks = (n + kBlock - 1) / kBlock;
        std::vector<uint32_t> haloCount, haloIdx;
        if (rel == REL_FF && n)
        {
            haloCount.resize(nBlocks);
            haloIdx.resize(static_cast<size_t>(nBlocks) * kHaloMax);
            hipCheck(hipMemcpy(haloCount.data(), d.haloCount, sizeof(uint32_t) * nBlocks, hipMemcpyDeviceToHost), "D2H");
            hipCheck(hipMemcpy(haloIdx.data(), d.haloIdx, sizeof(uint32_t) * haloIdx.size(), hipMemcpyDeviceToHost), "D2H");
        }
        // host decoder of the SELL-32 storage (ReadCompressedSell32Neighbor, gpu_backend.cpp:42-65)
        uint64_t w = 0;
        for (uint32_t i = 0; i < n; i++)
        {
            const SliceMeta& m = meta[i / kSlice];
            const uint32_t lane = i % kSlice;
            const uint32_t blk = i / kBlock;
            const bool slots = rel == REL_FF && haloCount[blk] <= kHaloMax;
            for (uint32_t k = 0; k < (cnt[i] & mask); k++)
            {
                uint32_t j;
                if (m.format == kFmtDelta16) { j = (slots ? 0u : m.base) + pool[m.offset + k * kSlice + lane]; }
                else { std::memcpy(&j, &pool[m.offset + 2ull * (k * kSlice + lane)], 4); }
                if (slots) { j = haloIdx[static_cast<size_t>(blk) * kHaloMax + j]; }
                idx[w++] = j;
            }
        }
        return TSP_OK;
    });
}

int tsp_get_sort(tsp_ctx* ctx, uint32_t* orderF, uint64_t* keysF, uint32_t* orderB, uint64_t* keysB)
{
    return guarded(ctx, [&]() {
        const Dev& d = ctx->d;
        ctx->sync();
        const auto widen = [&](uint64_t* dst, const uint32_t* src, uint32_t n) {
            std::vector<uint32_t> k(n);
            if (n) { hipCheck(hipMemcpy(k.data(), src, sizeof(uint32_t) * n, hipMemcpyDeviceToHost), "D2H"); }
            for (uint32_t i = 0; i < n; i++) { dst[i] = k[i] == ctx->frame.emptyKey ? ~0ull : k[i]; }
        };
        if (orderF && d.nF) { hipCheck(hipMemcpy(orderF, d.idxF, sizeof(uint32_t) * d.nF, hipMemcpyDeviceToHost), "D2H"); }
        if (keysF) { widen(keysF, d.keyF, d.nF); }
        if ((orderB || keysB) && d.nB)
        {
            // the boundary order relative to the pulled order, from the ids (stable sequence ids on pull)
            std::vector<uint32_t> kb(d.nB), ib(d.nB);
            launchKeys(ctx->stream, d.bpos, d.nB, d.frame, d.status, false, d.keyF2, d.idxF2);
            ctx->sync();
            if (keysB) { widen(keysB, d.keyF2, d.nB); }
            if (orderB)
            {
                std::vector<uint64_t> idb(d.nB);
                hipCheck(hipMemcpy(idb.data(), ctx->idB, sizeof(uint64_t) * d.nB, hipMemcpyDeviceToHost), "D2H");
                for (uint32_t i = 0; i < d.nB; i++) { orderB[i] = static_cast<uint32_t>(idb[i]); }
            }
        }
        return TSP_OK;
    });
}

int tsp_get_rigid_world(tsp_ctx* ctx, float* xyz, float* vel)
{
    return guarded(ctx, [&]() {
        const Dev& d = ctx->d;
        ctx->sync();
        std::vector<float4> q(d.nR);
        const float4* src[2] = { d.rpos, d.rvel };
        float* dst[2] = { xyz, vel };
        for (int r = 0; r < 2; r++)
        {
            if (!dst[r] || !d.nR) { continue; }
            hipCheck(hipMemcpy(q.data(), src[r], sizeof(float4) * d.nR, hipMemcpyDeviceToHost), "D2H");
            for (uint32_t i = 0; i < d.nR; i++) { dst[r][3 * i] = q[i].x; dst[r][3 * i + 1] = q[i].y; dst[r][3 * i + 2] = q[i].z; }
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

int tsp_synchronize(tsp_ctx* ctx)
{
    return guarded(ctx, [&]() {
        ctx->sync();
        return TSP_OK;
    });
}

} // extern "C"
