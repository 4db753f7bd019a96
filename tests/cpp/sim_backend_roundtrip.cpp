// sim_backend_roundtrip.cpp -- drives SimBackend_HIP the way the reference's Simulation drives a
// backend: build a scene into a universal state, pull, step, push. Writes the pushed fluid state
// (ids, then the 16 SoA fields, each N_F values) to the file named by argv[1], so that a test can
// compare it with the Python binding of the same C ABI.
//
// usage: sim_backend_roundtrip OUT n_side steps
// exit codes: 0 ok, 2 usage, 3 backend error (message on stderr)
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <vector>

#include "../../include/tsp_scene.h"
#include "../../turbospharticles_amd/host/sim_backend_hip.h"

int main(int argc, char** argv)
{
    if (argc != 4) { std::fprintf(stderr, "usage: %s OUT n_side steps\n", argv[0]); return 2; }
    const int nSide = std::atoi(argv[2]);
    const int steps = std::atoi(argv[3]);
    try
    {
        tsp::SimBackend_HIP backend(0);
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
        tsp::Sync::UniversalSimState state;
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

        tsp::SimParams params;
        params.SetParticleSize(0.01f);
        params.solverTerminationCond = tsp::SolverTerminationCond::RelDensityError_Pos;
        backend.SetSimDomainBounds(mn, mx);
        backend.PullFromUniversalSimState(state, params);
        backend.ExecuteFullSimStep(params, static_cast<uint32_t>(steps));
        tsp::Sync::UniversalSimState out;
        backend.PushToUniversalSimState(out, true);
        tsp::StatisticsPartial stats;
        backend.FillStatisticsPartial(stats, params);

        FILE* f = std::fopen(argv[1], "wb");
        if (!f) { std::fprintf(stderr, "cannot open %s\n", argv[1]); return 3; }
        std::fwrite(out.particlesF.id.data(), sizeof(uint64_t), out.particlesF.N_, f);
        auto& G = out.particlesF;
        std::vector<float>* outF[TSP_F_COUNT] = { &G.xPos, &G.yPos, &G.zPos, &G.xVel, &G.yVel, &G.zVel, &G.xVelPred,
                                                  &G.yVelPred, &G.zVelPred, &G.xAccPress, &G.yAccPress, &G.zAccPress,
                                                  &G.sourceTerm, &G.diagElem, &G.pressure, &G.density };
        for (auto* v : outF) { std::fwrite(v->data(), sizeof(float), v->size(), f); }
        std::fclose(f);
        std::printf("{\"backend\": \"%s\", \"n_fluid\": %u, \"solver_iters\": %u, \"nbr_ff_sum\": %llu, \"nbr_ff_max\": %u}\n",
                    backend.Name(), G.N_, stats.N_solverIters, static_cast<unsigned long long>(stats.nbrCountFF.sum),
                    stats.nbrCountFF.max);
        return 0;
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "%s\n", e.what());
        return 3;
    }
}
