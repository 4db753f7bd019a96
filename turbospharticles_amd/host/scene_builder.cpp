// scene_builder.cpp -- host-side initial states for the IISPH drop-in (tsp_scene.h).
//
// Restates the parts of the reference's scenario builders that the benchmark scenes use:
//   - boundary layers on the faces of an axis-aligned container:
//     ScenarioMgr::BuildBoundaryContainer_OBB / BuildBoundaryLayersOnFace / GetFacesFromOBB
//     (scene/builder.cpp:415-453, :1400-1484, :1354-1372)
//   - the fluid block lattice with SplitMix64 jitter: BuildFluidBlock_OBB, Jitter3D,
//     StartOffset (builder.cpp:455-642, :1486-1509), taking lattice counts per axis
//   - boundary and rigid coefficients psi and V from kernel sums over their own class:
//     ComputeParticleCoefficients (builder.cpp:174-376) with the scalar weight of :49-79
//   - a hollow box rigid body: BuildRigidBody_OBB_Hollow (builder.cpp:925-1087) and its
//     initial angular momentum (scene/scenarios/rigid_cube.h:71-120)
//   - the simulation domain: ComputeSimDomainBounds (scene/scenario_sim.cpp:123-148)
// Built with -ffp-contract=off so every float operation rounds where the reference's does.

#include <cstring>
#include <new>
#include <vector>

#include "scene_plan.h"

using namespace tsp::scene;

struct tsp_scene
{
    std::vector<uint64_t> idF, idB, idR;
    std::vector<float> F[TSP_F_COUNT];
    std::vector<float> B[TSP_B_COUNT];
    std::vector<float> R[TSP_R_COUNT];
    std::vector<uint32_t> bodyR;
    std::vector<uint32_t> idRB, countRB;
    std::vector<float> RB[TSP_RB_COUNT];
    float domMin[3], domMax[3];
};

extern "C" void tsp_scene_random_quat(uint64_t seed, float q[4])
{
    // Shoemake's uniform quaternion from three SplitMix64 uniforms
    const auto u01 = [&](uint64_t k)
    {
        return 0.5 + 0.5 * static_cast<double>(Jitter3D(seed, k, 1.0f));
    };
    const double u1 = u01(0), u2 = u01(1), u3 = u01(2);
    const double twoPi = 6.283185307179586;
    const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
    q[0] = static_cast<float>(b * std::cos(twoPi * u3));
    q[1] = static_cast<float>(a * std::sin(twoPi * u2));
    q[2] = static_cast<float>(a * std::cos(twoPi * u2));
    q[3] = static_cast<float>(b * std::sin(twoPi * u3));
}

extern "C" int tsp_scene_build(const tsp_scene_spec* spec, tsp_scene** out)
{
    if (!spec || !out) { return TSP_ERR_INVALID; }
    Plan plan;
    const int rc = MakePlan(*spec, plan);
    if (rc != TSP_OK) { return rc; }
    tsp_scene* s = new (std::nothrow) tsp_scene();
    if (!s) { return TSP_ERR_INVALID; }
    const Params& p = plan.p;

    // boundary layers of the container faces, then their psi and V
    s->idB.resize(plan.nB);
    for (int f = 0; f < TSP_B_COUNT; f++) { s->B[f].resize(plan.nB); }
    for (uint32_t q = 0; q < plan.nFaces; q++)
    {
        const BoundaryFace& bf = plan.faces[q];
        for (uint32_t layer = 0; layer < p.layers; layer++)
        {
            for (uint32_t i = 0; i < bf.nu; i++)
            {
                for (uint32_t j = 0; j < bf.nv; j++)
                {
                    const uint64_t k = bf.first + (static_cast<uint64_t>(layer) * bf.nu + i) * bf.nv + j;
                    const V3 pos = BoundaryPosition(bf.f, p.size, layer, i, j);
                    s->idB[k] = k;
                    s->B[TSP_B_X][k] = pos.x;
                    s->B[TSP_B_Y][k] = pos.y;
                    s->B[TSP_B_Z][k] = pos.z;
                }
            }
        }
    }
    if (plan.nB)
    {
        ComputeCoefficients(s->B[TSP_B_X], s->B[TSP_B_Y], s->B[TSP_B_Z], s->B[TSP_B_VOL], s->B[TSP_B_PSI],
                            0, static_cast<uint32_t>(plan.nB), p);
    }

    // the fluid block: positions, every other field zero, density rho0
    const uint64_t nF = plan.nF;
    s->idF.resize(nF);
    for (int f = 0; f < TSP_F_COUNT; f++) { s->F[f].assign(nF, 0.0f); }
    s->F[TSP_F_DENSITY].assign(nF, p.rho0);
    for (uint64_t idx = 0; idx < nF; idx++)
    {
        const V3 pos = FluidPosition(plan.fluid, p.size, idx);
        s->idF[idx] = idx;
        s->F[TSP_F_X][idx] = pos.x;
        s->F[TSP_F_Y][idx] = pos.y;
        s->F[TSP_F_Z][idx] = pos.z;
    }

    // the rigid body and the domain as planned
    s->idR = plan.idR;
    s->bodyR = plan.bodyR;
    for (int f = 0; f < TSP_R_COUNT; f++) { s->R[f] = plan.R[f]; }
    s->idRB = plan.idRB;
    s->countRB = plan.countRB;
    for (int f = 0; f < TSP_RB_COUNT; f++) { s->RB[f] = plan.RB[f]; }
    for (int a = 0; a < 3; a++) { s->domMin[a] = plan.domMin[a]; s->domMax[a] = plan.domMax[a]; }
    *out = s;
    return TSP_OK;
}

extern "C" int tsp_scene_spec_counts(const tsp_scene_spec* spec, uint32_t counts[4])
{
    if (!spec || !counts) { return TSP_ERR_INVALID; }
    Plan plan;
    const int rc = MakePlan(*spec, plan);
    if (rc != TSP_OK) { return rc; }
    if (plan.nF > 0x7fffffffull || plan.nB > 0x7fffffffull) { return TSP_ERR_CAPACITY; }
    counts[0] = static_cast<uint32_t>(plan.nF);
    counts[1] = static_cast<uint32_t>(plan.nB);
    counts[2] = static_cast<uint32_t>(plan.idR.size());
    counts[3] = static_cast<uint32_t>(plan.idRB.size());
    return TSP_OK;
}

extern "C" void tsp_scene_counts(const tsp_scene* s, uint32_t counts[4])
{
    counts[0] = static_cast<uint32_t>(s->idF.size());
    counts[1] = static_cast<uint32_t>(s->idB.size());
    counts[2] = static_cast<uint32_t>(s->idR.size());
    counts[3] = static_cast<uint32_t>(s->idRB.size());
}

extern "C" void tsp_scene_domain(const tsp_scene* s, float mn[3], float mx[3])
{
    for (int a = 0; a < 3; a++) { mn[a] = s->domMin[a]; mx[a] = s->domMax[a]; }
}

extern "C" int tsp_scene_export(const tsp_scene* s, const tsp_state_view* d)
{
    if (!s || !d) { return TSP_ERR_INVALID; }
    if (d->n_fluid != s->idF.size() || d->n_boundary != s->idB.size() || d->n_rigid != s->idR.size()
        || d->n_bodies != s->idRB.size())
    {
        return TSP_ERR_INVALID;
    }
    const auto cp = [](auto* dst, const auto& src)
    {
        if (dst && !src.empty()) { std::memcpy(dst, src.data(), sizeof(src[0]) * src.size()); }
    };
    cp(d->fluid_id, s->idF);
    for (int f = 0; f < TSP_F_COUNT; f++) { cp(d->fluid[f], s->F[f]); }
    cp(d->boundary_id, s->idB);
    for (int f = 0; f < TSP_B_COUNT; f++) { cp(d->boundary[f], s->B[f]); }
    cp(d->rigid_id, s->idR);
    cp(d->rigid_body, s->bodyR);
    for (int f = 0; f < TSP_R_COUNT; f++) { cp(d->rigid[f], s->R[f]); }
    cp(d->body_id, s->idRB);
    cp(d->body_count, s->countRB);
    for (int f = 0; f < TSP_RB_COUNT; f++) { cp(d->body[f], s->RB[f]); }
    return TSP_OK;
}

extern "C" void tsp_scene_free(tsp_scene* s) { delete s; }
