// backend_sim.h -- the reference's solver-backend interface, restated for the MI355X library.
//
// Sim::I_SimBackend (source/simulation/core/backend_sim.h:194-261) with the same virtual methods,
// argument meaning and error behaviour (std::runtime_error), and MakeSimBackend
// (core/backend_sim.cpp:18-38). The types its methods take are restated with the reference's
// names and the fields a backend fills, so that the reference's own types map onto them one field
// at a time:
//   Base::BackendType, Base::ColorCodingMode          base/enum_hub.h:18-22, :235-241
//   Sync::ExportSnapshot, Sync::VisualSnapshot         sync/snapshots.h:37-95 (glm -> vec3 / mat3)
//   Stats::StatMetric, Stats::StatRecordStep           eval/stats/metrics.h:23-86, :405-413
//   Mem::MemUsage, DeviceMemUsage, MemBucket, MemRecordStep   eval/memory/categories.h:19-80, :274-284
//   Bench::BenchOp, BenchOpList, BenchmarksTracker     eval/bench/operations.h:15-62, :231-252;
//                                                      eval/bench/benchmarks.h:18-86
// and the state and parameter types the step reads: Sync::UniversalSimState and its SoA particle
// structs (sync/state_sim.h:101-131, 272-285, 382-396, 497-522, 637), Sim::SimParams
// (core/parameters_sim.h:25-175) and Base::SolverTerminationCond (base/enum_hub.h:212-218).
#pragma once

#include <array>
#include <chrono>
#include <cstdint>
#include <memory>
#include <optional>
#include <vector>

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
        uint32_t surfaceNbrCountThreshold = 27;        // parameters_sim.h:170 (export snapshots)

        // the derived values of the reference's setters (core/parameters_sim.cpp:1478-1483)
        void SetParticleSize(float h);
    };

    // glm::vec3 / glm::mat3 (column-major) stand-ins
    struct vec3
    {
        float x = 0.0f, y = 0.0f, z = 0.0f;
    };

    struct mat3
    {
        vec3 col[3];
    };

    namespace Base
    {
        enum class BackendType : uint8_t { CPU = 0, GPU = 1 };
        enum class ColorCodingMode : uint8_t { None, VelocityMag, Density, Pressure };
    }

    namespace Sync
    {
        // ExportSnapshot (sync/snapshots.h:37-66); rankF stays empty outside a distributed run
        struct ExportSnapshot
        {
            uint64_t simStep = 0;
            double simTime = 0.0;
            uint32_t N_F = 0;
            std::vector<float> xPosF, yPosF, zPosF, velMagF;
            std::vector<uint16_t> rankF;
            uint32_t N_B = 0;
            std::vector<float> xPosB, yPosB, zPosB, volB;
            std::vector<uint8_t> isMeshB;
            uint32_t N_R = 0;
            std::vector<uint32_t> bodyIdR;
            std::vector<float> xPosR, yPosR, zPosR, volR;
        };

        // VisualSnapshot (sync/snapshots.h:70-95) without the boundary geometry of the visual
        // front end (boxesB, meshPosB, meshVolB), which the scene manager fills, not the backend
        struct VisualSnapshot
        {
            uint64_t simStep = 0;
            uint32_t N_F = 0;
            std::vector<vec3> posF;
            std::vector<float> scalarDataF;
            Base::ColorCodingMode scalarModeF = Base::ColorCodingMode::None;
            uint32_t N_R = 0;
            std::vector<vec3> posR;
            std::vector<float> volR;
            uint32_t N_RB = 0;
            std::vector<vec3> centerOfMassRB;
            std::vector<mat3> orientationRB;
        };
    }

    namespace Stats
    {
        enum class StatMetric : uint8_t
        {
            NumF, NumB, NumR,
            NumNbrFF_Avg, NumNbrFF_Min, NumNbrFF_Max, NumNbrFF_P99,
            NumNbrFB_Avg, NumNbrFB_Min, NumNbrFB_Max, NumNbrFB_P99,
            NumNbrFR_Avg, NumNbrFR_Min, NumNbrFR_Max, NumNbrFR_P99,
            NumNbrRF_Avg, NumNbrRF_Min, NumNbrRF_Max, NumNbrRF_P99,
            NumNbrRB_Avg, NumNbrRB_Min, NumNbrRB_Max, NumNbrRB_P99,
            NumNbrRR_Avg, NumNbrRR_Min, NumNbrRR_Max, NumNbrRR_P99,
            DensRelError_Raw_Avg, DensRelError_Pos_Avg, DensRelError_Abs_Avg,
            NumSolverIters,
            CFL_Number, VelMag_Avg, VelMag_Min, VelMag_Max,
            Count
        };

        constexpr uint8_t GetIdx(StatMetric m) { return static_cast<uint8_t>(m); }

        struct StatRecordStep
        {
            uint64_t simStep = 0;
            double simTime = 0.0;
            std::array<double, GetIdx(StatMetric::Count)> stats{};
        };

        inline double Get(const StatRecordStep& r, StatMetric m) { return r.stats[GetIdx(m)]; }
        inline void Set(StatRecordStep& r, StatMetric m, double x) { r.stats[GetIdx(m)] = x; }
    }

    namespace Mem
    {
        enum class MemSpace : uint8_t { RAM, VRAM, Count };

        struct MemUsage
        {
            uint64_t N_usedBytes = 0;
            uint64_t N_allocBytes = 0;
        };

        struct ProcessMemUsage
        {
            uint64_t N_residentBytes = 0;
            uint64_t N_peakResidentBytes = 0;
        };

        struct DeviceMemUsage
        {
            uint64_t N_residentBytes = 0;
            uint64_t N_peakResidentBytes = 0;
            uint64_t N_allocBytes = 0;
            uint64_t N_peakAllocBytes = 0;
        };

        enum class MemBucket : uint8_t
        {
            ParticlePropertiesF, ParticlePropertiesB, ParticlePropertiesR,
            Neighbors_SearchDataStructF, Neighbors_SearchDataStructB, Neighbors_SearchDataStructR,
            Neighbors_ListStorageF, Neighbors_ListStorageB, Neighbors_ListStorageR,
            Dist_LocalState, Dist_ParticleMetadata, Dist_ExchangeBuffers, Dist_Misc,
            Misc,
            Count
        };

        constexpr uint8_t GetIdx(MemSpace s) { return static_cast<uint8_t>(s); }
        constexpr uint8_t GetIdx(MemBucket b) { return static_cast<uint8_t>(b); }

        using MemBucketRecord = std::array<MemUsage, GetIdx(MemBucket::Count)>;
        using MemSpaceRecord = std::array<MemBucketRecord, GetIdx(MemSpace::Count)>;

        struct MemRecordStep
        {
            uint64_t simStep = 0;
            double simTime = 0.0;
            ProcessMemUsage process{};
            DeviceMemUsage device{};
            MemSpaceRecord mem{};
        };

        inline MemUsage Get(const MemRecordStep& r, MemSpace s, MemBucket b) { return r.mem[GetIdx(s)][GetIdx(b)]; }
        inline void Set(MemRecordStep& r, MemSpace s, MemBucket b, const MemUsage& m) { r.mem[GetIdx(s)][GetIdx(b)] = m; }
    }

    namespace Bench
    {
        enum class BenchOp : uint8_t
        {
            ParticleData_Reordering, Neighbors_Construction, Neighbors_Query, Density_Computation,
            Pressure_SourceAndDiagonal, Pressure_JacobiSolver, Pressure_Acceleration, NonPressure_Acceleration,
            Particle_Integration, Rigid_StateUpdate, Rigid_ForceAccumulation, Statistics_Computation,
            Serialization, Visualization,
            Dist_ParticleRedistribution, Dist_PlanRebuild, Dist_HaloFieldExchange, Dist_LoadBalancing, Dist_Misc,
            Dist_RankBackendStep,
            Total, Test,
            Count
        };

        constexpr uint8_t GetIdx(BenchOp op) { return static_cast<uint8_t>(op); }

        using BenchOpList = std::vector<BenchOp>;

        // the operations a simulation backend reports (MakeDefaultSimBenchOps, operations.h:233-252)
        inline BenchOpList MakeDefaultSimBenchOps()
        {
            return { BenchOp::ParticleData_Reordering, BenchOp::Neighbors_Construction, BenchOp::Neighbors_Query,
                     BenchOp::Density_Computation, BenchOp::NonPressure_Acceleration,
                     BenchOp::Pressure_SourceAndDiagonal, BenchOp::Pressure_JacobiSolver,
                     BenchOp::Pressure_Acceleration, BenchOp::Rigid_ForceAccumulation,
                     BenchOp::Particle_Integration, BenchOp::Rigid_StateUpdate, BenchOp::Statistics_Computation,
                     BenchOp::Serialization, BenchOp::Visualization, BenchOp::Total };
        }

        // per-operation timings of the steps (BenchmarksTracker, benchmarks.h:18-86): the latest
        // value of each operation and the sums over the recorded steps, in nanoseconds
        class BenchmarksTracker
        {
        public:
            void Record(BenchOp op, uint64_t ns)
            {
                latest_[GetIdx(op)] = ns;
                sum_[GetIdx(op)] += ns;
                count_[GetIdx(op)]++;
            }
            uint64_t Latest(BenchOp op) const { return latest_[GetIdx(op)]; }
            uint64_t Sum(BenchOp op) const { return sum_[GetIdx(op)]; }
            uint64_t Count(BenchOp op) const { return count_[GetIdx(op)]; }

        private:
            std::array<uint64_t, GetIdx(BenchOp::Count)> latest_{}, sum_{}, count_{};
        };
    }

    namespace Sim
    {
        class I_DistributedParticleBackend;

        class I_SimBackend
        {
        public:
            virtual ~I_SimBackend() = default;

            virtual const char* Name() const = 0;
            virtual void SetSimDomainBounds(const vec3& min, const vec3& max) = 0;
            virtual void PullFromUniversalSimState(Sync::UniversalSimState& src, const SimParams& params) = 0;
            virtual void PushToUniversalSimState(Sync::UniversalSimState& dst, bool backendWillBeRetired) = 0;
            virtual void PackExportSnapshot(Sync::ExportSnapshot& snap, const SimParams& params, bool includeB,
                                            std::optional<uint64_t> meshIdBeginB) const = 0;
            virtual void PackVisualSnapshot(Sync::VisualSnapshot& snap, bool includeF, bool includeR) const = 0;
            virtual void FillScalarData(Sync::VisualSnapshot& snap, Base::ColorCodingMode mode) const = 0;
            virtual void FillMemoryRecord(Mem::MemRecordStep& record) const = 0;
            virtual void FillStatisticsRecord(Stats::StatRecordStep& record, const SimParams& params) const = 0;
            // the distributed execution capability (none: this backend runs one domain on one GPU)
            virtual I_DistributedParticleBackend* GetDistributedParticleBackend() { return nullptr; }
            virtual void ExecuteNeighborSearch(const SimParams& params, bool bench = false) = 0;
            virtual void ExecuteFullSimStep(const SimParams& params, uint32_t N_steps = 1) = 0;
            virtual void GetBenchOps(Bench::BenchOpList& out) const = 0;
            virtual uint32_t GetNumFluidParticles() const = 0;
            virtual uint32_t GetNumBoundaryParticles() const = 0;
            virtual uint32_t GetNumRigidParticles() const = 0;
        };

        // BackendType::GPU gives the gfx950 backend on device 0; BackendType::CPU is not part of
        // this library (std::runtime_error, as the reference throws for a backend it was built
        // without)
        std::unique_ptr<I_SimBackend> MakeSimBackend(Base::BackendType backendType,
                                                     Bench::BenchmarksTracker& benchmarksTracker);
    }
}
