// sim_backend_hip.cpp -- SimBackend_HIP over the C ABI of libtsp_hip.so.
//
// The state crosses the boundary without staging: a tsp_state_view points into the vectors of
// the universal state, and the library copies them to and from the device.
#include "sim_backend_hip.h"

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

    SimBackend_HIP::SimBackend_HIP(int device)
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

    void SimBackend_HIP::SetSimDomainBounds(const float min[3], const float max[3])
    {
        Check(tsp_set_domain(ctx_, min, max), "SimBackend_HIP::SetSimDomainBounds");
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

    void SimBackend_HIP::ExecuteNeighborSearch(const SimParams& params, bool bench)
    {
        (void)bench;
        ApplyParams(params);
        Check(tsp_neighbor_search(ctx_), "SimBackend_HIP::ExecuteNeighborSearch");
    }

    void SimBackend_HIP::ExecuteFullSimStep(const SimParams& params, uint32_t N_steps)
    {
        ApplyParams(params);
        Check(tsp_step(ctx_, N_steps), "SimBackend_HIP::ExecuteFullSimStep");
    }

    bool SimBackend_HIP::FillStatisticsPartial(StatisticsPartial& partial, const SimParams& params) const
    {
        (void)params;
        partial = StatisticsPartial{};
        tsp_step_info info{};
        Check(tsp_get_step_info(ctx_, &info), "SimBackend_HIP::FillStatisticsPartial");
        partial.N_F = n_[0]; partial.N_B = n_[1]; partial.N_R = n_[2];
        partial.N_solverIters = info.solver_iters;
        partial.densRelErrorRawSum = info.err_sum_raw;
        partial.densRelErrorPosSum = info.err_sum_pos;
        partial.densRelErrorAbsSum = info.err_sum_abs;
        DistributionPartial* dist[6] = { &partial.nbrCountFF, &partial.nbrCountFB, &partial.nbrCountFR,
                                         &partial.nbrCountRF, &partial.nbrCountRB, &partial.nbrCountRR };
        for (int rel = 0; rel < 6; rel++)
        {
            const uint32_t n = rel < 3 ? n_[0] : n_[2];
            std::vector<uint32_t> counts(n);
            uint64_t total = 0;
            Check(tsp_get_neighbors(ctx_, rel, counts.data(), nullptr, 0, &total), "SimBackend_HIP::FillStatisticsPartial");
            DistributionPartial& d = *dist[rel];
            d.N = n;
            d.sum = total;
            if (n)
            {
                d.min = counts[0];
                d.max = counts[0];
                for (uint32_t c : counts) { d.min = c < d.min ? c : d.min; d.max = c > d.max ? c : d.max; }
            }
        }
        return true;
    }
}
