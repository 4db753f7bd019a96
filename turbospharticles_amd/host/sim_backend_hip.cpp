// sim_backend_hip.cpp -- SimBackend_HIP over the C ABI of libtsp_hip.so, and MakeSimBackend.
//
// The state crosses the boundary without staging: a tsp_state_view points into the vectors of
// the universal state, and the library copies them to and from the device.
#include "sim_backend_hip.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace tsp
{
    void SimParams::SetParticleSize(float h)
    {
        particleSizeF = h;
        kernelRadius = 2.0f * h;
        kernelRadiusSqu = kernelRadius * kernelRadius;
        massF = restDensity * h * h * h;
    }

    namespace
    {
        tsp_params ToC(const SimParams& p)
        {
            tsp_params c{};
            c.time_step = p.timeStep;
            c.particle_size = p.particleSizeF;
            c.kernel_radius = p.kernelRadius;
            c.kernel_radius_sq = p.kernelRadiusSqu;
            c.rest_density = p.restDensity;
            c.mass = p.massF;
            c.viscosity_fluid = p.viscosityF;
            c.viscosity_boundary = p.viscosityB;
            for (int a = 0; a < 3; a++) { c.gravity[a] = p.gravity[a]; }
            c.press_gamma = p.pressGamma;
            c.jacobi_omega = p.jacobiOmega;
            c.jacobi_warm_start = p.jacobiWarmStart;
            c.density_dev_target = p.densityDevTarget;
            c.grid_cell_size_ratio = p.gridCellSizeRatio;
            c.fixed_iters = p.fixedIters;
            c.min_solver_iters = p.minSolverIters;
            c.max_solver_iters = p.maxSolverIters;
            c.termination_cond = static_cast<uint32_t>(p.solverTerminationCond);
            c.rigid_contact_density_dev_target = p.rigidContactDensityDevTarget;
            c.rigid_contact_max_solver_iters = p.rigidContactMaxSolverIters;
            return c;
        }

        template <typename T>
        T* Ptr(std::vector<T>& v) { return v.empty() ? nullptr : v.data(); }

        // the view over the vectors of a state; a missing (empty) optional field reads as null
        tsp_state_view View(Sync::UniversalSimState& s)
        {
            tsp_state_view v{};
            auto& F = s.particlesF;
            auto& B = s.particlesB;
            auto& R = s.particlesR;
            auto& RB = s.rigidBodies;
            v.n_fluid = F.N_; v.n_boundary = B.N_; v.n_rigid = R.N_; v.n_bodies = RB.N_;
            v.fluid_id = Ptr(F.id);
            std::vector<float>* f[TSP_F_COUNT] = { &F.xPos, &F.yPos, &F.zPos, &F.xVel, &F.yVel, &F.zVel,
                                                   &F.xVelPred, &F.yVelPred, &F.zVelPred, &F.xAccPress, &F.yAccPress,
                                                   &F.zAccPress, &F.sourceTerm, &F.diagElem, &F.pressure, &F.density };
            for (int k = 0; k < TSP_F_COUNT; k++) { v.fluid[k] = Ptr(*f[k]); }
            v.boundary_id = Ptr(B.id);
            std::vector<float>* b[TSP_B_COUNT] = { &B.xPos, &B.yPos, &B.zPos, &B.volB, &B.psiB };
            for (int k = 0; k < TSP_B_COUNT; k++) { v.boundary[k] = Ptr(*b[k]); }
            v.rigid_id = Ptr(R.id);
            v.rigid_body = Ptr(R.bodyId);
            std::vector<float>* r[TSP_R_COUNT] = { &R.xPosRel, &R.yPosRel, &R.zPosRel, &R.volR, &R.psiR };
            for (int k = 0; k < TSP_R_COUNT; k++) { v.rigid[k] = Ptr(*r[k]); }
            v.body_id = Ptr(RB.id);
            v.body_count = Ptr(RB.rigidParticleCount);
            std::vector<float>* q[TSP_RB_COUNT] = {
                &RB.xCM, &RB.yCM, &RB.zCM, &RB.xVel, &RB.yVel, &RB.zVel, &RB.xOmega, &RB.yOmega, &RB.zOmega,
                &RB.xAngMom, &RB.yAngMom, &RB.zAngMom, &RB.a00, &RB.a01, &RB.a02, &RB.a10, &RB.a11, &RB.a12,
                &RB.a20, &RB.a21, &RB.a22, &RB.mass, &RB.massInv, &RB.iBodyInv00, &RB.iBodyInv01, &RB.iBodyInv02,
                &RB.iBodyInv10, &RB.iBodyInv11, &RB.iBodyInv12, &RB.iBodyInv20, &RB.iBodyInv21, &RB.iBodyInv22 };
            for (int k = 0; k < TSP_RB_COUNT; k++) { v.body[k] = Ptr(*q[k]); }
            v.step = s.simStep;
            v.time = s.simTime;
            return v;
        }

        // every vector of a population holds N_ entries or none (an optional field)
        template <typename... V>
        void RequireSizes(uint32_t n, const char* what, const V&... vs)
        {
            const bool ok = ((vs.empty() || vs.size() == n) && ...);
            if (!ok) { throw std::runtime_error(std::string(what) + " arrays do not match the particle count."); }
        }

        template <typename... V>
        void ResizeAll(uint32_t n, V&... vs) { (vs.resize(n), ...); }
    }

    SimBackend_HIP::SimBackend_HIP(int device, Bench::BenchmarksTracker* tracker) : tracker_(tracker)
    {
        const int rc = tsp_create(device, &ctx_);
        if (rc != TSP_OK) { throw std::runtime_error("SimBackend_HIP: no usable gfx950 device (tsp_create " + std::to_string(rc) + ")"); }
    }

    SimBackend_HIP::~SimBackend_HIP() { tsp_destroy(ctx_); }

    const char* SimBackend_HIP::Name() const { return tsp_name(); }

    void SimBackend_HIP::Check(int rc, const char* what) const
    {
        if (rc != TSP_OK) { throw std::runtime_error(std::string(what) + ": " + tsp_last_error(ctx_)); }
    }

    void SimBackend_HIP::ApplyParams(const SimParams& params)
    {
        const tsp_params c = ToC(params);
        if (paramsSet_ && std::memcmp(&c, &last_, sizeof(c)) == 0) { return; }
        Check(tsp_set_params(ctx_, &c), "SimBackend_HIP::SetParams");
        last_ = c;
        paramsSet_ = true;
    }

    void SimBackend_HIP::SetSimDomainBounds(const vec3& min, const vec3& max)
    {
        const float mn[3] = { min.x, min.y, min.z };
        const float mx[3] = { max.x, max.y, max.z };
        Check(tsp_set_domain(ctx_, mn, mx), "SimBackend_HIP::SetSimDomainBounds");
    }

    void SimBackend_HIP::PullFromUniversalSimState(Sync::UniversalSimState& src, const SimParams& params)
    {
        auto& F = src.particlesF;
        auto& B = src.particlesB;
        auto& R = src.particlesR;
        auto& RB = src.rigidBodies;
        RequireSizes(F.N_, "Fluid", F.id, F.xPos, F.yPos, F.zPos, F.xVel, F.yVel, F.zVel, F.xVelPred, F.yVelPred,
                     F.zVelPred, F.xAccPress, F.yAccPress, F.zAccPress, F.sourceTerm, F.diagElem, F.pressure, F.density);
        RequireSizes(B.N_, "Boundary", B.id, B.xPos, B.yPos, B.zPos, B.volB, B.psiB);
        RequireSizes(R.N_, "Rigid", R.id, R.bodyId, R.xPosRel, R.yPosRel, R.zPosRel, R.volR, R.psiR);
        RequireSizes(RB.N_, "Rigid body", RB.id, RB.rigidParticleCount, RB.xCM, RB.a00, RB.mass, RB.massInv);
        if (R.N_ && R.bodyId.size() != R.N_) { throw std::runtime_error("Rigid particles need body ids."); }
        ApplyParams(params);
        tsp_state_view v = View(src);
        Check(tsp_pull(ctx_, &v), "SimBackend_HIP::PullFromUniversalSimState");
        n_[0] = F.N_; n_[1] = B.N_; n_[2] = R.N_; n_[3] = RB.N_;
        cacheValid_ = false;
    }

    void SimBackend_HIP::PushToUniversalSimState(Sync::UniversalSimState& dst, bool backendWillBeRetired)
    {
        (void)backendWillBeRetired;
        auto& F = dst.particlesF;
        auto& B = dst.particlesB;
        auto& R = dst.particlesR;
        auto& RB = dst.rigidBodies;
        F.N_ = n_[0]; B.N_ = n_[1]; R.N_ = n_[2]; RB.N_ = n_[3];
        ResizeAll(n_[0], F.id, F.xPos, F.yPos, F.zPos, F.xVel, F.yVel, F.zVel, F.xVelPred, F.yVelPred, F.zVelPred,
                  F.xAccPress, F.yAccPress, F.zAccPress, F.sourceTerm, F.diagElem, F.pressure, F.density);
        ResizeAll(n_[1], B.id, B.xPos, B.yPos, B.zPos, B.volB, B.psiB);
        ResizeAll(n_[2], R.id, R.bodyId, R.xPosRel, R.yPosRel, R.zPosRel, R.volR, R.psiR);
        ResizeAll(n_[3], RB.id, RB.rigidParticleCount, RB.xCM, RB.yCM, RB.zCM, RB.xVel, RB.yVel, RB.zVel, RB.xOmega,
                  RB.yOmega, RB.zOmega, RB.xAngMom, RB.yAngMom, RB.zAngMom, RB.a00, RB.a01, RB.a02, RB.a10, RB.a11,
                  RB.a12, RB.a20, RB.a21, RB.a22, RB.mass, RB.massInv, RB.iBodyInv00, RB.iBodyInv01, RB.iBodyInv02,
                  RB.iBodyInv10, RB.iBodyInv11, RB.iBodyInv12, RB.iBodyInv20, RB.iBodyInv21, RB.iBodyInv22);
        tsp_state_view v = View(dst);
        Check(tsp_push(ctx_, &v), "SimBackend_HIP::PushToUniversalSimState");
        dst.simStep = v.step;
        dst.simTime = v.time;
    }

    const Sync::UniversalSimState& SimBackend_HIP::Current() const
    {
        tsp_step_info info{};
        Check(tsp_get_step_info(ctx_, &info), "SimBackend_HIP::Current");
        if (!cacheValid_ || cache_.simStep != info.step)
        {
            const_cast<SimBackend_HIP*>(this)->PushToUniversalSimState(cache_, false);
            cacheValid_ = true;
        }
        return cache_;
    }

    void SimBackend_HIP::PackExportSnapshot(Sync::ExportSnapshot& snap, const SimParams& params, bool includeB,
                                            std::optional<uint64_t> meshIdBeginB) const
    {
        const uint64_t mesh = meshIdBeginB ? *meshIdBeginB : UINT64_MAX;
        std::vector<float> fxyz, fvel, bxyz, bvol, rxyz, rvol;
        std::vector<uint8_t> bmesh;
        std::vector<uint32_t> rbody;
        bxyz.resize(includeB ? 3ull * n_[1] : 0);
        bvol.resize(includeB ? n_[1] : 0);
        bmesh.resize(includeB ? n_[1] : 0);
        rxyz.resize(3ull * n_[2]);
        rvol.resize(n_[2]);
        rbody.resize(n_[2]);
        tsp_export_view v{};
        uint32_t cap = n_[0] / 4 + 1024;
        for (int attempt = 0; attempt < 2; attempt++)
        {
            fxyz.resize(3ull * cap);
            fvel.resize(cap);
            v = tsp_export_view{};
            v.cap_fluid = cap;
            v.fluid_xyz = fxyz.data();
            v.fluid_vel_mag = fvel.data();
            v.boundary_xyz = includeB ? bxyz.data() : nullptr;
            v.boundary_vol = includeB ? bvol.data() : nullptr;
            v.boundary_is_mesh = includeB ? bmesh.data() : nullptr;
            v.rigid_xyz = rxyz.data();
            v.rigid_vol = rvol.data();
            v.rigid_body = rbody.data();
            const int rc = tsp_export_snapshot(ctx_, params.surfaceNbrCountThreshold, mesh, &v);
            if (rc == TSP_ERR_CAPACITY && v.n_fluid > cap && attempt == 0) { cap = v.n_fluid; continue; }
            Check(rc, "SimBackend_HIP::PackExportSnapshot");
            break;
        }
        snap = Sync::ExportSnapshot{};
        snap.simStep = v.step;
        snap.simTime = v.time;
        snap.N_F = v.n_fluid;
        snap.xPosF.resize(v.n_fluid); snap.yPosF.resize(v.n_fluid); snap.zPosF.resize(v.n_fluid);
        snap.velMagF.assign(fvel.begin(), fvel.begin() + v.n_fluid);
        for (uint32_t i = 0; i < v.n_fluid; i++)
        {
            snap.xPosF[i] = fxyz[3ull * i]; snap.yPosF[i] = fxyz[3ull * i + 1]; snap.zPosF[i] = fxyz[3ull * i + 2];
        }
        snap.N_B = v.n_boundary;
        if (includeB)
        {
            snap.xPosB.resize(v.n_boundary); snap.yPosB.resize(v.n_boundary); snap.zPosB.resize(v.n_boundary);
            for (uint32_t i = 0; i < v.n_boundary; i++)
            {
                snap.xPosB[i] = bxyz[3ull * i]; snap.yPosB[i] = bxyz[3ull * i + 1]; snap.zPosB[i] = bxyz[3ull * i + 2];
            }
            snap.volB = bvol;
            snap.isMeshB = bmesh;
        }
        snap.N_R = v.n_rigid;
        snap.bodyIdR = rbody;
        snap.volR = rvol;
        snap.xPosR.resize(v.n_rigid); snap.yPosR.resize(v.n_rigid); snap.zPosR.resize(v.n_rigid);
        for (uint32_t i = 0; i < v.n_rigid; i++)
        {
            snap.xPosR[i] = rxyz[3ull * i]; snap.yPosR[i] = rxyz[3ull * i + 1]; snap.zPosR[i] = rxyz[3ull * i + 2];
        }
    }

    void SimBackend_HIP::PackVisualSnapshot(Sync::VisualSnapshot& snap, bool includeF, bool includeR) const
    {
        const Sync::UniversalSimState& s = Current();
        const auto& F = s.particlesF;
        snap.simStep = s.simStep;
        snap.N_F = includeF ? F.N_ : 0;
        snap.posF.resize(snap.N_F);
        for (uint32_t i = 0; i < snap.N_F; i++) { snap.posF[i] = vec3{ F.xPos[i], F.yPos[i], F.zPos[i] }; }
        if (includeF) { FillScalarData(snap, snap.scalarModeF); }
        else { snap.scalarDataF.clear(); }
        snap.N_R = includeR ? n_[2] : 0;
        snap.posR.resize(snap.N_R);
        snap.volR.resize(snap.N_R);
        if (snap.N_R)
        {
            std::vector<float> xyz(3ull * n_[2]), vel(3ull * n_[2]);
            Check(tsp_get_rigid_world(ctx_, xyz.data(), vel.data()), "SimBackend_HIP::PackVisualSnapshot");
            for (uint32_t i = 0; i < snap.N_R; i++)
            {
                snap.posR[i] = vec3{ xyz[3ull * i], xyz[3ull * i + 1], xyz[3ull * i + 2] };
                snap.volR[i] = s.particlesR.volR[i];
            }
        }
        // PackVisualRigidBodyState: centre of mass and the rotation's columns
        const auto& RB = s.rigidBodies;
        snap.N_RB = includeR ? RB.N_ : 0;
        snap.centerOfMassRB.resize(snap.N_RB);
        snap.orientationRB.resize(snap.N_RB);
        for (uint32_t b = 0; b < snap.N_RB; b++)
        {
            snap.centerOfMassRB[b] = vec3{ RB.xCM[b], RB.yCM[b], RB.zCM[b] };
            snap.orientationRB[b].col[0] = vec3{ RB.a00[b], RB.a10[b], RB.a20[b] };
            snap.orientationRB[b].col[1] = vec3{ RB.a01[b], RB.a11[b], RB.a21[b] };
            snap.orientationRB[b].col[2] = vec3{ RB.a02[b], RB.a12[b], RB.a22[b] };
        }
    }

    void SimBackend_HIP::FillScalarData(Sync::VisualSnapshot& snap, Base::ColorCodingMode mode) const
    {
        snap.scalarModeF = mode;
        if (n_[0] == 0 || mode == Base::ColorCodingMode::None)
        {
            snap.scalarDataF.clear();
            return;
        }
        const auto& F = Current().particlesF;
        snap.scalarDataF.resize(F.N_);
        for (uint32_t i = 0; i < F.N_; i++)
        {
            float v = 0.0f;
            switch (mode)
            {
            case Base::ColorCodingMode::VelocityMag:
                v = std::sqrt(F.xVel[i] * F.xVel[i] + F.yVel[i] * F.yVel[i] + F.zVel[i] * F.zVel[i]);
                break;
            case Base::ColorCodingMode::Density: v = F.density[i]; break;
            case Base::ColorCodingMode::Pressure: v = F.pressure[i]; break;
            default: break;
            }
            snap.scalarDataF[i] = v;
        }
    }

    void SimBackend_HIP::FillMemoryRecord(Mem::MemRecordStep& record) const
    {
        using namespace Mem;
        tsp_memory m{};
        Check(tsp_get_memory(ctx_, &m), "SimBackend_HIP::FillMemoryRecord");
        tsp_step_info info{};
        Check(tsp_get_step_info(ctx_, &info), "SimBackend_HIP::FillMemoryRecord");
        record.simStep = info.step;
        record.simTime = info.time;
        const auto set = [&](MemBucket b, uint64_t bytes) { Set(record, MemSpace::VRAM, b, MemUsage{ bytes, bytes }); };
        set(MemBucket::ParticlePropertiesF, m.particles[0]);
        set(MemBucket::ParticlePropertiesB, m.particles[1]);
        set(MemBucket::ParticlePropertiesR, m.particles[2]);
        set(MemBucket::Neighbors_SearchDataStructF, m.search[0]);
        set(MemBucket::Neighbors_SearchDataStructB, m.search[1]);
        set(MemBucket::Neighbors_SearchDataStructR, m.search[2]);
        set(MemBucket::Neighbors_ListStorageF, m.lists[0]);
        set(MemBucket::Neighbors_ListStorageB, m.lists[1]);
        set(MemBucket::Neighbors_ListStorageR, m.lists[2]);
        set(MemBucket::Misc, m.misc);
        record.device.N_allocBytes = m.alloc_bytes;
        record.device.N_peakAllocBytes = m.peak_alloc_bytes;
        // host: the copy of the state this adapter keeps for visual packs
        uint64_t host = 0;
        if (cacheValid_)
        {
            const auto& F = cache_.particlesF;
            host = F.id.capacity() * sizeof(uint64_t) + 16ull * F.xPos.capacity() * sizeof(float);
        }
        Set(record, MemSpace::RAM, MemBucket::ParticlePropertiesF, MemUsage{ host, host });
    }

    void SimBackend_HIP::FillStatisticsRecord(Stats::StatRecordStep& record, const SimParams& params) const
    {
        using namespace Stats;
        (void)params;
        tsp_statistics st{};
        Check(tsp_get_statistics(ctx_, &st), "SimBackend_HIP::FillStatisticsRecord");
        tsp_step_info info{};
        Check(tsp_get_step_info(ctx_, &info), "SimBackend_HIP::FillStatisticsRecord");
        record.simStep = info.step;
        record.simTime = info.time;
        Set(record, StatMetric::NumF, st.n_fluid);
        Set(record, StatMetric::NumB, st.n_boundary);
        Set(record, StatMetric::NumR, st.n_rigid);
        for (int r = 0; r < 6; r++)
        {
            const uint8_t base = static_cast<uint8_t>(GetIdx(StatMetric::NumNbrFF_Avg) + 4 * r);
            record.stats[base + 0] = st.nbr_avg[r];
            record.stats[base + 1] = st.nbr_min[r];
            record.stats[base + 2] = st.nbr_max[r];
            record.stats[base + 3] = st.nbr_p99[r];
        }
        Set(record, StatMetric::DensRelError_Raw_Avg, st.dens_rel_err_avg[0]);
        Set(record, StatMetric::DensRelError_Pos_Avg, st.dens_rel_err_avg[1]);
        Set(record, StatMetric::DensRelError_Abs_Avg, st.dens_rel_err_avg[2]);
        Set(record, StatMetric::NumSolverIters, st.solver_iters);
        Set(record, StatMetric::CFL_Number, st.cfl);
        Set(record, StatMetric::VelMag_Avg, st.vel_avg);
        Set(record, StatMetric::VelMag_Min, st.vel_min);
        Set(record, StatMetric::VelMag_Max, st.vel_max);
    }

    void SimBackend_HIP::ExecuteNeighborSearch(const SimParams& params, bool bench)
    {
        ApplyParams(params);
        const auto t0 = std::chrono::steady_clock::now();
        Check(tsp_neighbor_search(ctx_), "SimBackend_HIP::ExecuteNeighborSearch");
        Check(tsp_synchronize(ctx_), "SimBackend_HIP::ExecuteNeighborSearch");
        if (bench && tracker_)
        {
            const auto ns = std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
            tracker_->Record(Bench::BenchOp::Neighbors_Query, static_cast<uint64_t>(ns));
        }
        cacheValid_ = false;
    }

    void SimBackend_HIP::ExecuteFullSimStep(const SimParams& params, uint32_t N_steps)
    {
        ApplyParams(params);
        const auto t0 = std::chrono::steady_clock::now();
        Check(tsp_step(ctx_, N_steps), "SimBackend_HIP::ExecuteFullSimStep");
        cacheValid_ = false;
        if (tracker_ && N_steps)
        {
            const auto ns = std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
            tracker_->Record(Bench::BenchOp::Total, static_cast<uint64_t>(ns) / N_steps);
        }
    }

    namespace Sim
    {
        std::unique_ptr<I_SimBackend> MakeSimBackend(Base::BackendType backendType,
                                                     Bench::BenchmarksTracker& benchmarksTracker)
        {
            switch (backendType)
            {
            case Base::BackendType::GPU:
                return std::make_unique<SimBackend_HIP>(0, &benchmarksTracker);
            case Base::BackendType::CPU:
                throw std::runtime_error("CPU sim backend requested, but this library holds only the gfx950 backend.");
            }
            throw std::runtime_error("Unsupported sim backend type requested.");
        }
    }
}
