// scene_plan.h -- the layout of a generated scene, shared by the host builder (libtsp_host,
// tsp_scene_build) and the device builder (libtsp_hip, tsp_pull_spec / tsp_batch_add_scene_spec).
//
// A plan holds everything but the per-particle loops of the fluid block and the container's
// boundary layers: the six (extended) container faces with their lattice counts, the fluid block
// with its offsets and jitter, the rigid body generated in full (it is small), and the domain.
// The per-particle positions follow from the plan by the formulas below; the device builder
// restates them in its kernels, and the tests require the two builders to agree bit for bit.
//
// Restated from the reference's scenario builders (source/simulation/scene/builder.cpp):
// BuildBoundaryContainer_OBB / BuildBoundaryLayersOnFace / GetFacesFromOBB (:415-453,
// :1400-1484, :1354-1372), BuildFluidBlock_OBB / Jitter3D / StartOffset (:455-642, :1486-1509),
// ComputeParticleCoefficients (:174-376) with the scalar weight of :49-79,
// BuildRigidBody_OBB_Hollow (:925-1087), the rigid cube's initial state
// (scene/scenarios/rigid_cube.h:71-120) and ComputeSimDomainBounds (scene/scenario_sim.cpp:123-148).
#ifndef TSP_SCENE_PLAN_H
#define TSP_SCENE_PLAN_H

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "tsp_scene.h"

namespace tsp::scene
{
    struct V3
    {
        float x, y, z;
    };
    inline V3 operator+(V3 a, V3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
    inline V3 operator-(V3 a, V3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
    inline V3 operator*(V3 a, float s) { return { a.x * s, a.y * s, a.z * s }; }
    inline V3 operator-(V3 a) { return { -a.x, -a.y, -a.z }; }
    inline float Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

    struct Box
    {
        V3 origin, u, v, w;
        float uLen, vLen, wLen;
    };
    struct Face
    {
        V3 origin, u, v, n;
        float uLen, vLen;
    };

    struct Params
    {
        float size, sizeInv, radius, radiusSqu, rho0, alpha, massB;
        uint32_t layers;
    };

    // one face of the container: layers x Nu x Nv particles from index (and id) `first`,
    // index = first + (layer Nu + i) Nv + j
    struct BoundaryFace
    {
        Face f;
        uint32_t nu, nv;
        uint64_t first;
    };

    // the fluid block: Nu x Nv x Nw lattice, index = (kw Nv + jv) Nu + iu
    struct FluidBlock
    {
        Box blk;
        float uOff, vOff, wOff, radius, jitAmp;
        uint64_t seed;
        uint32_t nu, nv, nw;
    };

    struct Plan
    {
        Params p{};
        BoundaryFace faces[6]{};
        uint32_t nFaces = 0;
        uint64_t nB = 0;
        FluidBlock fluid{};
        uint64_t nF = 0;
        // the rigid body in full
        std::vector<uint64_t> idR;
        std::vector<uint32_t> bodyR;
        std::vector<float> R[TSP_R_COUNT];
        std::vector<uint32_t> idRB, countRB;
        std::vector<float> RB[TSP_RB_COUNT];
        float domMin[3] = { 0, 0, 0 }, domMax[3] = { 0, 0, 0 };
    };

    float Jitter3D(uint64_t baseSeed, uint64_t key, float amp);

    // the scalar kernel weight of the coefficient sums (builder.cpp:49-79)
    inline float KernelWeight(float dx, float dy, float dz, float sizeInv, float radiusSqu, float alpha)
    {
        const float distSqu = std::fmaf(dx, dx, std::fmaf(dy, dy, dz * dz));
        if (distSqu >= radiusSqu) { return 0.0f; }
        if (distSqu < 1e-7f) { return 4.0f * alpha; }
        const float dist = std::sqrt(distSqu);
        const float q = dist * sizeInv;
        const float t1 = std::max(1.0f - q, 0.0f);
        const float t2 = std::max(2.0f - q, 0.0f);
        const float t1_3 = (t1 * t1) * t1;
        const float t2_3 = (t2 * t2) * t2;
        return alpha * (t2_3 - 4.0f * t1_3);
    }

    // ComputeParticleCoefficients over one particle class: V = 1 / sum_j W, psi = rho0 V
    void ComputeCoefficients(const std::vector<float>& x, const std::vector<float>& y, const std::vector<float>& z,
                             std::vector<float>& vol, std::vector<float>& psi, uint32_t begin, uint32_t end, const Params& p);

    // the plan of a spec; TSP_OK or a negative tsp_status
    int MakePlan(const tsp_scene_spec& spec, Plan& plan);

    // position of fluid particle idx of the block
    inline V3 FluidPosition(const FluidBlock& b, float size, uint64_t idx)
    {
        const uint32_t iu = static_cast<uint32_t>(idx % b.nu);
        const uint32_t jv = static_cast<uint32_t>((idx / b.nu) % b.nv);
        const uint32_t kw = static_cast<uint32_t>(idx / (static_cast<uint64_t>(b.nu) * b.nv));
        float u = b.uOff + static_cast<float>(iu) * size;
        float v = b.vOff + static_cast<float>(jv) * size;
        float w = b.wOff + static_cast<float>(kw) * size;
        if (b.jitAmp > 0.0f)
        {
            u += Jitter3D(b.seed, idx * 3ull + 0ull, b.jitAmp);
            v += Jitter3D(b.seed, idx * 3ull + 1ull, b.jitAmp);
            w += Jitter3D(b.seed, idx * 3ull + 2ull, b.jitAmp);
        }
        u = std::clamp(u, b.radius, b.blk.uLen - b.radius);
        v = std::clamp(v, b.radius, b.blk.vLen - b.radius);
        w = std::clamp(w, b.radius, b.blk.wLen - b.radius);
        return b.blk.origin + b.blk.u * u + b.blk.v * v + b.blk.w * w;
    }

    // position of boundary particle (layer, i, j) of a face
    inline V3 BoundaryPosition(const Face& f, float size, uint32_t layer, uint32_t i, uint32_t j)
    {
        const float outOffset = size * 0.5f + static_cast<float>(layer) * size;
        const float su = std::min(static_cast<float>(i) * size, f.uLen);
        const float sv = std::min(static_cast<float>(j) * size, f.vLen);
        return f.origin + f.u * su + f.v * sv + f.n * outOffset;
    }
}

#endif
