// sim_backend_hip.h -- the MI355X step behind the reference's backend interface.
//
// SimBackend_HIP has the hot-path methods of Sim::I_SimBackend
// (source/simulation/core/backend_sim.h:196-266) with the same names, argument meaning and
// error behaviour (std::runtime_error with the reason, like the reference backends), over the C
// ABI of libtsp_hip.so (include/tsp_hip.h). The state and parameter types restate the fields of
// Sync::UniversalSimState (sync/state_sim.h:101-131, 272-285, 382-396, 497-522, 637) and
// Sim::SimParams (core/parameters_sim.h:25-175) that the step reads, with the reference's field
// names, so that the reference's own types map onto them one field at a time.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "../../include/tsp_hip.h"

namespace tsp
{
    namespace Sync
    {
        struct FluidParticles
        {
            uint32_t N_ = 0;
            std::vector<uint64_t> id;
            std::vector<float> xPos, yPos, zPos;
            std::vector<float> xVel, yVel, zVel;
            std::vector<float> xVelPred, yVelPred, zVelPred;
            std::vector<float> xAccPress, yAccPress, zAccPress;
            std::vector<float> sourceTerm, diagElem, pressure, density;
        };

        struct BoundaryParticles
        {
            uint32_t N_ = 0;
            std::vector<uint64_t> id;
            std::vector<float> xPos, yPos, zPos;
            std::vector<float> volB, psiB;
        };

        struct RigidParticles
        {
            uint32_t N_ = 0;
            std::vector<uint64_t> id;
            std::vector<uint32_t> bodyId;
            std::vector<float> xPosRel, yPosRel, zPosRel;
            std::vector<float> volR, psiR;
        };

        struct RigidBodies
        {
            uint32_t N_ = 0;
            std::vector<uint32_t> id, rigidParticleCount;
            std::vector<float> xCM, yCM, zCM, xVel, yVel, zVel, xOmega, yOmega, zOmega, xAngMom, yAngMom, zAngMom;
            std::vector<float> a00, a01, a02, a10, a11, a12, a20, a21, a22;
            std::vector<float> mass, massInv;
            std::vector<float> iBodyInv00, iBodyInv01, iBodyInv02, iBodyInv10, iBodyInv11, iBodyInv12;
            std::vector<float> iBodyInv20, iBodyInv21, iBodyInv22;
        };

        struct UniversalSimState
        {
            FluidParticles particlesF;
            BoundaryParticles particlesB;
            RigidParticles particlesR;
            RigidBodies rigidBodies;
            uint64_t simStep = 0;
            double simTime = 0.0;
        };
    }

    // Base::SolverTerminationCond (base/enum_hub.h:212-218)
    enum class SolverTerminationCond : uint32_t { FixedIters = 0, RelDensityError_Raw = 1, RelDensityError_Pos = 2,
                                                  RelDensityError_Abs = 3 };

    struct SimParams
    {
        float timeStep = 1e-3f;
        float particleSizeF = 0.01f;
        float kernelRadius = 0.02f;
        float kernelRadiusSqu = 0.0004f;
        float restDensity = 1000.0f;
        float massF = 1e-3f;
        float viscosityF = 4e-5f;
        float viscosityB = 8e-5f;
        float gravity[3] = { 0.0f, -9.81f, 0.0f };
        float pressGamma = 0.5f;
        float jacobiOmega = 0.5f;
        float jacobiWarmStart = 0.0f;
        float densityDevTarget = 0.001f;
        float gridCellSizeRatio = 1.0f;
        uint32_t fixedIters = 5;
        uint32_t minSolverIters = 2;
        uint32_t maxSolverIters = 100;
        SolverTerminationCond solverTerminationCond = SolverTerminationCond::RelDensityError_Abs;
        float rigidContactDensityDevTarget = 0.001f;   // parameters_sim.h:100-101
        uint32_t rigidContactMaxSolverIters = 100;

        // the derived values of the reference's setters (core/parameters_sim.cpp:1478-1483)
        void SetParticleSize(float h);
    };

    // Stats::StatisticsPartial (eval/stats/statistics.h:32-52): counts, sums and extremes of the
    // neighbor counts, and the solver figures of the last step
    struct DistributionPartial
    {
        uint64_t N = 0, sum = 0;
        uint32_t min = 0, max = 0;
    };

    struct StatisticsPartial
    {
        uint64_t N_F = 0, N_B = 0, N_R = 0;
        DistributionPartial nbrCountFF, nbrCountFB, nbrCountFR, nbrCountRF, nbrCountRB, nbrCountRR;
        uint32_t N_solverIters = 0;
        double densRelErrorRawSum = 0.0, densRelErrorPosSum = 0.0, densRelErrorAbsSum = 0.0;
    };

    class SimBackend_HIP
    {
    public:
        explicit SimBackend_HIP(int device = 0);
        ~SimBackend_HIP();
        SimBackend_HIP(const SimBackend_HIP&) = delete;
        SimBackend_HIP& operator=(const SimBackend_HIP&) = delete;

        const char* Name() const;

        // bounds the cell frame by the simulation domain (an empty extent leaves it unbounded)
        void SetSimDomainBounds(const float min[3], const float max[3]);

        // copies a validated universal state onto the device (throws on inconsistent sizes or
        // rigid body indices, like Sync::ValidateUniversalSimState)
        void PullFromUniversalSimState(Sync::UniversalSimState& src, const SimParams& params);

        // copies the device state back (every vector resized to the current counts)
        void PushToUniversalSimState(Sync::UniversalSimState& dst, bool backendWillBeRetired);

        void ExecuteNeighborSearch(const SimParams& params, bool bench = false);
        void ExecuteFullSimStep(const SimParams& params, uint32_t N_steps = 1);

        bool FillStatisticsPartial(StatisticsPartial& partial, const SimParams& params) const;

        uint32_t GetNumFluidParticles() const { return n_[0]; }
        uint32_t GetNumBoundaryParticles() const { return n_[1]; }
        uint32_t GetNumRigidParticles() const { return n_[2]; }

        // the C-ABI context, for callers that read lists or profiles directly
        tsp_ctx* Context() const { return ctx_; }

    private:
        void Check(int rc, const char* what) const;
        void ApplyParams(const SimParams& params);

        tsp_ctx* ctx_ = nullptr;
        uint32_t n_[4] = { 0, 0, 0, 0 };
        bool paramsSet_ = false;
        tsp_params last_{};
    };
}
