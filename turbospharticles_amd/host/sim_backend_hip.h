// sim_backend_hip.h -- the MI355X step behind the reference's backend interface.
//
// SimBackend_HIP implements Sim::I_SimBackend (backend_sim.h, restating
// source/simulation/core/backend_sim.h:194-261) over the C ABI of libtsp_hip.so
// (include/tsp_hip.h), with the reference's error behaviour: a failed call throws
// std::runtime_error with the reason. Sim::MakeSimBackend(BackendType::GPU, tracker) returns one
// on device 0.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "../../include/tsp_hip.h"
#include "backend_sim.h"

namespace tsp
{
    class SimBackend_HIP final : public Sim::I_SimBackend
    {
    public:
        explicit SimBackend_HIP(int device = 0, Bench::BenchmarksTracker* tracker = nullptr);
        ~SimBackend_HIP() override;
        SimBackend_HIP(const SimBackend_HIP&) = delete;
        SimBackend_HIP& operator=(const SimBackend_HIP&) = delete;

        const char* Name() const override;

        // bounds the cell frame by the simulation domain (an empty extent leaves it unbounded)
        void SetSimDomainBounds(const vec3& min, const vec3& max) override;

        // copies a validated universal state onto the device (throws on inconsistent sizes or
        // rigid body indices, like Sync::ValidateUniversalSimState)
        void PullFromUniversalSimState(Sync::UniversalSimState& src, const SimParams& params) override;

        // copies the device state back (every vector resized to the current counts)
        void PushToUniversalSimState(Sync::UniversalSimState& dst, bool backendWillBeRetired) override;

        // the surface fluid particles (classified and compacted on the device), the rigid
        // particles and, with includeB, the boundary (PackExportSnapshot, gpu_backend.cpp:346-551)
        void PackExportSnapshot(Sync::ExportSnapshot& snap, const SimParams& params, bool includeB,
                                std::optional<uint64_t> meshIdBeginB) const override;

        // fluid and rigid positions and the bodies' centres and orientations
        void PackVisualSnapshot(Sync::VisualSnapshot& snap, bool includeF, bool includeR) const override;

        // per fluid particle: |v|, density or pressure (nothing for ColorCodingMode::None)
        void FillScalarData(Sync::VisualSnapshot& snap, Base::ColorCodingMode mode) const override;

        // device memory by bucket (VRAM), the device total and high-water mark, and the host
        // staging this adapter holds (RAM)
        void FillMemoryRecord(Mem::MemRecordStep& record) const override;

        // the StatMetric values computed on the device (FillStatisticsRecord, gpu_backend.cpp:1067-1208)
        void FillStatisticsRecord(Stats::StatRecordStep& record, const SimParams& params) const override;

        void ExecuteNeighborSearch(const SimParams& params, bool bench = false) override;
        void ExecuteFullSimStep(const SimParams& params, uint32_t N_steps = 1) override;

        void GetBenchOps(Bench::BenchOpList& out) const override { out = benchOps_; }

        uint32_t GetNumFluidParticles() const override { return n_[0]; }
        uint32_t GetNumBoundaryParticles() const override { return n_[1]; }
        uint32_t GetNumRigidParticles() const override { return n_[2]; }

        // the C-ABI context, for callers that read lists or profiles directly
        tsp_ctx* Context() const { return ctx_; }

    private:
        void Check(int rc, const char* what) const;
        void ApplyParams(const SimParams& params);
        // the current state, copied from the device when the step has moved on since the last copy
        const Sync::UniversalSimState& Current() const;

        tsp_ctx* ctx_ = nullptr;
        Bench::BenchmarksTracker* tracker_ = nullptr;
        uint32_t n_[4] = { 0, 0, 0, 0 };
        bool paramsSet_ = false;
        tsp_params last_{};
        Bench::BenchOpList benchOps_ = Bench::MakeDefaultSimBenchOps();
        mutable Sync::UniversalSimState cache_;
        mutable bool cacheValid_ = false;
    };
}
