// sim_backend_roundtrip.cpp -- drives the gfx950 backend the way the reference's SimulationMgr
// drives a backend (core/simulation.cpp:338-503): through an I_SimBackend* from MakeSimBackend,
// a scene built into a universal state, pull, neighbor search, steps, push, and every reporting
// method of the interface. Writes the pushed fluid state (ids, then the 16 SoA fields, each N_F
// values) to the file named by argv[1], so that a test can compare it with the Python binding of
// the same C ABI, and prints one JSON line with what the reporting methods returned.
//
// usage: sim_backend_roundtrip OUT n_side steps
// exit codes: 0 ok, 2 usage, 3 backend error (message on stderr)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <vector>

#include "../../include/tsp_scene.h"
#include "../../turbospharticles_amd/host/sim_backend_hip.h"

using namespace tsp;

int main(int argc, char** argv)
{
    if (argc != 4) { std::fprintf(stderr, "usage: %s OUT n_side steps\n", argv[0]); return 2; }
    const int nSide = std::atoi(argv[2]);
    const int steps = std::atoi(argv[3]);
    try
    {
        Bench::BenchmarksTracker tracker;
        std::unique_ptr<Sim::I_SimBackend> backend = Sim::MakeSimBackend(Base::BackendType::GPU, tracker);
        tsp_scene_spec spec{};
        spec.particle_size = 0.01f;
        spec.rest_density = 1000.0f;
        spec.boundary_layers = 2;
        const float tank[3] = { 0.4f, 0.3f, 0.3f };
        for (int a = 0; a < 3; a++) { spec.container_size[a] = tank[a]; spec.fluid_counts[a] = static_cast<uint32_t>(nSide); }
        spec.jitter_frac = 0.01f;
        spec.jitter_seed = 67;
        tsp_scene* scene = nullptr;
        if (tsp_scene_build(&spec, &scene) != 0) { std::fprintf(stderr, "scene build failed\n"); return 3; }
        uint32_t n[4];
        tsp_scene_counts(scene, n);
        Sync::UniversalSimState state;
        auto& F = state.particlesF;
        auto& B = state.particlesB;
        F.N_ = n[0];
        B.N_ = n[1];
        std::vector<float>* fluid[TSP_F_COUNT] = { &F.xPos, &F.yPos, &F.zPos, &F.xVel, &F.yVel, &F.zVel, &F.xVelPred,
                                                   &F.yVelPred, &F.zVelPred, &F.xAccPress, &F.yAccPress, &F.zAccPress,
                                                   &F.sourceTerm, &F.diagElem, &F.pressure, &F.density };
        F.id.resize(n[0]);
        for (auto* v : fluid) { v->resize(n[0]); }
        std::vector<float>* bnd[TSP_B_COUNT] = { &B.xPos, &B.yPos, &B.zPos, &B.volB, &B.psiB };
        B.id.resize(n[1]);
        for (auto* v : bnd) { v->resize(n[1]); }
        tsp_state_view view{};
        view.n_fluid = n[0];
        view.n_boundary = n[1];
        view.fluid_id = F.id.data();
        for (int k = 0; k < TSP_F_COUNT; k++) { view.fluid[k] = fluid[k]->data(); }
        view.boundary_id = B.id.data();
        for (int k = 0; k < TSP_B_COUNT; k++) { view.boundary[k] = bnd[k]->data(); }
        if (tsp_scene_export(scene, &view) != 0) { std::fprintf(stderr, "scene export failed\n"); return 3; }
        float mn[3], mx[3];
        tsp_scene_domain(scene, mn, mx);
        tsp_scene_free(scene);

        SimParams params;
        params.SetParticleSize(0.01f);
        params.solverTerminationCond = SolverTerminationCond::RelDensityError_Pos;
        backend->SetSimDomainBounds(vec3{ mn[0], mn[1], mn[2] }, vec3{ mx[0], mx[1], mx[2] });
        backend->PullFromUniversalSimState(state, params);
        backend->ExecuteNeighborSearch(params, true);
        for (int s = 0; s < steps; s++) { backend->ExecuteFullSimStep(params, 1); }
        Sync::UniversalSimState out;
        backend->PushToUniversalSimState(out, false);

        Stats::StatRecordStep stats;
        backend->FillStatisticsRecord(stats, params);
        Mem::MemRecordStep mem;
        backend->FillMemoryRecord(mem);
        Sync::ExportSnapshot exp;
        backend->PackExportSnapshot(exp, params, true, std::nullopt);
        Sync::VisualSnapshot vis;
        vis.scalarModeF = Base::ColorCodingMode::VelocityMag;
        backend->PackVisualSnapshot(vis, true, true);
        Bench::BenchOpList ops;
        backend->GetBenchOps(ops);

        // the visual pack's positions and speeds are the pushed state's
        auto& G = out.particlesF;
        uint32_t visMismatch = vis.N_F == G.N_ && vis.scalarDataF.size() == G.N_ ? 0u : 1u;
        for (uint32_t i = 0; !visMismatch && i < G.N_; i++)
        {
            const float sp = std::sqrt(G.xVel[i] * G.xVel[i] + G.yVel[i] * G.yVel[i] + G.zVel[i] * G.zVel[i]);
            visMismatch += vis.posF[i].x != G.xPos[i] || vis.posF[i].y != G.yPos[i] || vis.posF[i].z != G.zPos[i]
                           || vis.scalarDataF[i] != sp;
        }
        uint64_t vramBuckets = 0;
        for (const auto& u : mem.mem[Mem::GetIdx(Mem::MemSpace::VRAM)]) { vramBuckets += u.N_allocBytes; }

        FILE* f = std::fopen(argv[1], "wb");
        if (!f) { std::fprintf(stderr, "cannot open %s\n", argv[1]); return 3; }
        std::fwrite(G.id.data(), sizeof(uint64_t), G.N_, f);
        std::vector<float>* outF[TSP_F_COUNT] = { &G.xPos, &G.yPos, &G.zPos, &G.xVel, &G.yVel, &G.zVel, &G.xVelPred,
                                                  &G.yVelPred, &G.zVelPred, &G.xAccPress, &G.yAccPress, &G.zAccPress,
                                                  &G.sourceTerm, &G.diagElem, &G.pressure, &G.density };
        for (auto* v : outF) { std::fwrite(v->data(), sizeof(float), v->size(), f); }
        std::fclose(f);
        std::printf("{\"backend\": \"%s\", \"n_fluid\": %u, \"n_boundary\": %u, \"stats\": [", backend->Name(),
                    backend->GetNumFluidParticles(), backend->GetNumBoundaryParticles());
        for (size_t q = 0; q < stats.stats.size(); q++) { std::printf("%s%.17g", q ? ", " : "", stats.stats[q]); }
        std::printf("], \"stats_step\": %llu, \"export_fluid\": %u, \"export_boundary\": %u, \"export_rigid\": %u, "
                    "\"visual_fluid\": %u, \"visual_mismatch\": %u, \"vram_buckets\": %llu, \"vram_alloc\": %llu, "
                    "\"vram_peak\": %llu, \"bench_ops\": %zu, \"tracked_total\": %llu, \"tracked_query\": %llu, "
                    "\"distributed\": %d}\n",
                    static_cast<unsigned long long>(stats.simStep), exp.N_F, exp.N_B, exp.N_R, vis.N_F, visMismatch,
                    static_cast<unsigned long long>(vramBuckets), static_cast<unsigned long long>(mem.device.N_allocBytes),
                    static_cast<unsigned long long>(mem.device.N_peakAllocBytes), ops.size(),
                    static_cast<unsigned long long>(tracker.Count(Bench::BenchOp::Total)),
                    static_cast<unsigned long long>(tracker.Count(Bench::BenchOp::Neighbors_Query)),
                    backend->GetDistributedParticleBackend() != nullptr);
        return 0;
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "%s\n", e.what());
        return 3;
    }
}
