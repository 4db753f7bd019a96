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

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <numeric>
#include <tuple>
#include <vector>

#include "tsp_scene.h"

namespace
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

    uint32_t TilesCount(float length, float spacing)
    {
        if (spacing <= 0.0f) { return 0; }
        const float intervals = std::floor(length / spacing);
        const double count = static_cast<double>(intervals) + 1.0;
        return static_cast<uint32_t>(std::max(1, static_cast<int>(count)));
    }

    float StartOffset(float length, float spacing, int N)
    {
        if (N <= 1) { return 0.0f; }
        const float covered = static_cast<float>(N - 1) * spacing;
        return 0.5f * std::fmax(0.0f, length - covered);
    }

    float Jitter3D(uint64_t baseSeed, uint64_t key, float amp)
    {
        uint64_t x = baseSeed + key * 0x9E3779B97F4A7C15ull;
        x ^= x >> 30;
        x *= 0xBF58476D1CE4E5B9ull;
        x ^= x >> 27;
        x *= 0x94D049BB133111EBull;
        x ^= x >> 31;
        const double u01 = static_cast<double>(x >> 11) * (1.0 / 9007199254740992.0);
        return static_cast<float>((u01 * 2.0 - 1.0) * amp);
    }

    float KernelWeight(float dx, float dy, float dz, float sizeInv, float radiusSqu, float alpha)
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

    struct Params
    {
        float size, sizeInv, radius, radiusSqu, rho0, alpha, massB;
        uint32_t layers;
    };

    // ComputeParticleCoefficients: V = 1 / sum_j W, psi = rho0 V over the same particle class
    void ComputeCoefficients(const std::vector<float>& x, const std::vector<float>& y,
                             const std::vector<float>& z, std::vector<float>& vol,
                             std::vector<float>& psi, uint32_t begin, uint32_t end, const Params& p)
    {
        struct Cell
        {
            int32_t x, y, z;
        };
        const uint32_t n = end - begin;
        vol.resize(x.size());
        psi.resize(x.size());
        const float defaultPsi = p.massB;
        const float defaultVol = p.rho0 > 0.0f ? defaultPsi / p.rho0 : 0.0f;
        if (n == 0) { return; }
        const float cellSize = p.radius;
        const float cellSizeInv = 1.0f / cellSize;
        std::vector<Cell> cells(n);
        std::vector<uint32_t> sorted(n);
        for (uint32_t k = 0; k < n; k++)
        {
            const uint32_t i = begin + k;
            cells[k] = { static_cast<int32_t>(std::floor(x[i] * cellSizeInv)),
                         static_cast<int32_t>(std::floor(y[i] * cellSizeInv)),
                         static_cast<int32_t>(std::floor(z[i] * cellSizeInv)) };
            sorted[k] = k;
        }
        const auto key = [&](uint32_t k) { return std::make_tuple(cells[k].z, cells[k].y, cells[k].x, k); };
        std::sort(sorted.begin(), sorted.end(), [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
        const auto cellLess = [](const Cell& a, const Cell& b)
        {
            return std::tie(a.z, a.y, a.x) < std::tie(b.z, b.y, b.x);
        };
        const auto rowRange = [&](const Cell& first, const Cell& last)
        {
            const auto lo = std::lower_bound(sorted.begin(), sorted.end(), first,
                [&](uint32_t k, const Cell& t) { return cellLess(cells[k], t); });
            const auto hi = std::upper_bound(lo, sorted.end(), last,
                [&](const Cell& t, uint32_t k) { return cellLess(t, cells[k]); });
            return std::make_pair(static_cast<uint32_t>(lo - sorted.begin()),
                                  static_cast<uint32_t>(hi - sorted.begin()));
        };
        std::vector<std::pair<uint32_t, uint32_t>> rows;
        Cell rowsCell{ 0, 0, 0 };
        bool haveRows = false;
        for (uint32_t s = 0; s < n; s++)
        {
            const uint32_t i = begin + sorted[s];
            const Cell c = cells[sorted[s]];
            if (!haveRows || c.x != rowsCell.x || c.y != rowsCell.y || c.z != rowsCell.z)
            {
                rows.clear();
                for (int dz = -1; dz <= 1; dz++)
                {
                    for (int dy = -1; dy <= 1; dy++)
                    {
                        rows.push_back(rowRange(Cell{ c.x - 1, c.y + dy, c.z + dz },
                                                Cell{ c.x + 1, c.y + dy, c.z + dz }));
                    }
                }
                rowsCell = c;
                haveRows = true;
            }
            float sumW = 0.0f;
            for (const auto& r : rows)
            {
                for (uint32_t t = r.first; t < r.second; t++)
                {
                    const uint32_t j = begin + sorted[t];
                    sumW += KernelWeight(x[i] - x[j], y[i] - y[j], z[i] - z[j], p.sizeInv, p.radiusSqu, p.alpha);
                }
            }
            if (sumW > 0.0f && std::isfinite(sumW))
            {
                vol[i] = 1.0f / sumW;
                psi[i] = p.rho0 * vol[i];
            }
            else
            {
                vol[i] = defaultVol;
                psi[i] = defaultPsi;
            }
        }
    }

    Face ExtendFace(Face f, float padU, float padV)
    {
        f.origin = f.origin - (f.u * padU + f.v * padV);
        f.uLen += 2.0f * padU;
        f.vLen += 2.0f * padV;
        return f;
    }
}

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
    if (!spec || !out || spec->particle_size <= 0.0f) { return TSP_ERR_INVALID; }
    tsp_scene* s = new (std::nothrow) tsp_scene();
    if (!s) { return TSP_ERR_INVALID; }

    Params p{};
    p.size = spec->particle_size;
    const double hd = static_cast<double>(p.size);
    p.sizeInv = static_cast<float>(1.0 / hd);
    p.radius = 2.0f * p.size;
    p.radiusSqu = p.radius * p.radius;
    p.rho0 = spec->rest_density;
    p.alpha = static_cast<float>((1.0 / hd) * (1.0 / hd) * (1.0 / hd) / (4.0 * 3.14159265358979323846));
    p.massB = p.rho0 * p.size * p.size * p.size;
    p.layers = spec->boundary_layers;

    const Box tank{ { spec->container_origin[0], spec->container_origin[1], spec->container_origin[2] },
                    { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 },
                    spec->container_size[0], spec->container_size[1], spec->container_size[2] };

    // --- boundary container (BuildBoundaryContainer_OBB) ---
    if (p.layers > 0)
    {
        const float radiusB = p.size * 0.5f;
        const float pad = radiusB + (static_cast<float>(p.layers) - 1.0f) * p.size;
        Face faces[6] = {
            { tank.origin, tank.u, tank.v, -tank.w, tank.uLen, tank.vLen },
            { tank.origin + tank.w * tank.wLen, tank.u, tank.v, tank.w, tank.uLen, tank.vLen },
            { tank.origin, tank.w, tank.v, -tank.u, tank.wLen, tank.vLen },
            { tank.origin + tank.u * tank.uLen, tank.w, tank.v, tank.u, tank.wLen, tank.vLen },
            { tank.origin, tank.u, tank.w, -tank.v, tank.uLen, tank.wLen },
            { tank.origin + tank.v * tank.vLen, tank.u, tank.w, tank.v, tank.uLen, tank.wLen },
        };
        faces[2] = ExtendFace(faces[2], pad, pad);
        faces[3] = ExtendFace(faces[3], pad, pad);
        faces[4] = ExtendFace(faces[4], 0.0f, pad);
        faces[5] = ExtendFace(faces[5], 0.0f, pad);
        for (const Face& f : faces)
        {
            const uint32_t Nu = TilesCount(f.uLen, p.size);
            const uint32_t Nv = TilesCount(f.vLen, p.size);
            const uint64_t base = s->idB.size();
            for (uint32_t layer = 0; layer < p.layers; layer++)
            {
                const float outOffset = p.size * 0.5f + static_cast<float>(layer) * p.size;
                for (uint32_t i = 0; i < Nu; i++)
                {
                    const float su = std::min(static_cast<float>(i) * p.size, f.uLen);
                    const V3 bu = f.u * su;
                    for (uint32_t j = 0; j < Nv; j++)
                    {
                        const float sv = std::min(static_cast<float>(j) * p.size, f.vLen);
                        const V3 bv = f.v * sv;
                        const V3 pos = f.origin + bu + bv + f.n * outOffset;
                        s->idB.push_back(base + (static_cast<uint64_t>(layer) * Nu + i) * Nv + j);
                        s->B[TSP_B_X].push_back(pos.x);
                        s->B[TSP_B_Y].push_back(pos.y);
                        s->B[TSP_B_Z].push_back(pos.z);
                    }
                }
            }
        }
        ComputeCoefficients(s->B[TSP_B_X], s->B[TSP_B_Y], s->B[TSP_B_Z], s->B[TSP_B_VOL], s->B[TSP_B_PSI],
                            0, static_cast<uint32_t>(s->idB.size()), p);
    }

    // --- fluid block from lattice counts (BuildFluidBlock_OBB) ---
    {
        const uint32_t Nu = spec->fluid_counts[0], Nv = spec->fluid_counts[1], Nw = spec->fluid_counts[2];
        const uint64_t nF64 = static_cast<uint64_t>(Nu) * Nv * Nw;
        if (nF64 > static_cast<uint64_t>(std::numeric_limits<int32_t>::max()))
        {
            delete s;
            return TSP_ERR_CAPACITY;
        }
        const uint32_t nF = static_cast<uint32_t>(nF64);
        const float radius = p.size * 0.5f;
        const Box blk{ { spec->fluid_origin[0], spec->fluid_origin[1], spec->fluid_origin[2] },
                       { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 },
                       static_cast<float>(Nu) * p.size, static_cast<float>(Nv) * p.size,
                       static_cast<float>(Nw) * p.size };
        const float uLenI = blk.uLen - 2.0f * radius;
        const float vLenI = blk.vLen - 2.0f * radius;
        const float wLenI = blk.wLen - 2.0f * radius;
        const float uOff = radius + StartOffset(uLenI, p.size, static_cast<int>(Nu));
        const float vOff = radius + StartOffset(vLenI, p.size, static_cast<int>(Nv));
        const float wOff = radius + StartOffset(wLenI, p.size, static_cast<int>(Nw));
        const float jitAmp = spec->jitter_frac * p.size;
        s->idF.resize(nF);
        for (int f = 0; f < TSP_F_COUNT; f++) { s->F[f].assign(nF, 0.0f); }
        s->F[TSP_F_DENSITY].assign(nF, p.rho0);
        for (uint32_t kw = 0; kw < Nw; kw++)
        {
            for (uint32_t jv = 0; jv < Nv; jv++)
            {
                for (uint32_t iu = 0; iu < Nu; iu++)
                {
                    const uint64_t idx = (static_cast<uint64_t>(kw) * Nv + jv) * Nu + iu;
                    float u = uOff + static_cast<float>(iu) * p.size;
                    float v = vOff + static_cast<float>(jv) * p.size;
                    float w = wOff + static_cast<float>(kw) * p.size;
                    if (jitAmp > 0.0f)
                    {
                        u += Jitter3D(spec->jitter_seed, idx * 3ull + 0ull, jitAmp);
                        v += Jitter3D(spec->jitter_seed, idx * 3ull + 1ull, jitAmp);
                        w += Jitter3D(spec->jitter_seed, idx * 3ull + 2ull, jitAmp);
                    }
                    u = std::clamp(u, radius, blk.uLen - radius);
                    v = std::clamp(v, radius, blk.vLen - radius);
                    w = std::clamp(w, radius, blk.wLen - radius);
                    const V3 pos = blk.origin + blk.u * u + blk.v * v + blk.w * w;
                    s->idF[idx] = idx;
                    s->F[TSP_F_X][idx] = pos.x;
                    s->F[TSP_F_Y][idx] = pos.y;
                    s->F[TSP_F_Z][idx] = pos.z;
                }
            }
        }
    }

    // --- hollow box rigid body (BuildRigidBody_OBB_Hollow + rigid_cube.h initial state) ---
    if (spec->rigid_enable)
    {
        double qw = spec->rigid_quat[0], qx = spec->rigid_quat[1], qy = spec->rigid_quat[2], qz = spec->rigid_quat[3];
        const double qn = std::sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
        if (!(qn > 0.0)) { qw = 1.0; qx = qy = qz = 0.0; }
        else { qw /= qn; qx /= qn; qy /= qn; qz /= qn; }
        // rotation matrix columns = body axes
        const V3 uHat{ static_cast<float>(1.0 - 2.0 * (qy * qy + qz * qz)),
                       static_cast<float>(2.0 * (qx * qy + qw * qz)),
                       static_cast<float>(2.0 * (qx * qz - qw * qy)) };
        const V3 vHat{ static_cast<float>(2.0 * (qx * qy - qw * qz)),
                       static_cast<float>(1.0 - 2.0 * (qx * qx + qz * qz)),
                       static_cast<float>(2.0 * (qy * qz + qw * qx)) };
        const V3 wHat{ static_cast<float>(2.0 * (qx * qz + qw * qy)),
                       static_cast<float>(2.0 * (qy * qz - qw * qx)),
                       static_cast<float>(1.0 - 2.0 * (qx * qx + qy * qy)) };
        const float uLen = spec->rigid_size[0], vLen = spec->rigid_size[1], wLen = spec->rigid_size[2];
        const V3 centerIn{ spec->rigid_center[0], spec->rigid_center[1], spec->rigid_center[2] };
        const V3 origin = centerIn - (uHat * (0.5f * uLen) + vHat * (0.5f * vLen) + wHat * (0.5f * wLen));
        const Box box{ origin, uHat, vHat, wHat, uLen, vLen, wLen };

        const float radius = p.size * 0.5f;
        const float uLenI = box.uLen - 2.0f * radius;
        const float vLenI = box.vLen - 2.0f * radius;
        const float wLenI = box.wLen - 2.0f * radius;
        if (uLenI <= 0.0f || vLenI <= 0.0f || wLenI <= 0.0f)
        {
            delete s;
            return TSP_ERR_INVALID;
        }
        const V3 originI = box.origin + box.u * radius + box.v * radius + box.w * radius;
        const uint32_t Nu = TilesCount(uLenI, p.size), Nv = TilesCount(vLenI, p.size), Nw = TilesCount(wLenI, p.size);
        const V3 center = box.origin + box.u * (0.5f * box.uLen) + box.v * (0.5f * box.vLen) + box.w * (0.5f * box.wLen);
        const uint32_t L = p.layers;
        for (uint32_t kw = 0; kw < Nw; kw++)
        {
            const bool wIn = (kw >= L) && (kw <= Nw - 1 - L);
            const float w = std::min(wLenI, static_cast<float>(kw) * p.size);
            for (uint32_t jv = 0; jv < Nv; jv++)
            {
                const bool vIn = (jv >= L) && (jv <= Nv - 1 - L);
                const float v = std::min(vLenI, static_cast<float>(jv) * p.size);
                for (uint32_t iu = 0; iu < Nu; iu++)
                {
                    const bool uIn = (iu >= L) && (iu <= Nu - 1 - L);
                    if (uIn && vIn && wIn) { continue; }
                    const float u = std::min(uLenI, static_cast<float>(iu) * p.size);
                    const V3 pos = originI + box.u * std::clamp(u, 0.0f, uLenI) + box.v * std::clamp(v, 0.0f, vLenI)
                                 + box.w * std::clamp(w, 0.0f, wLenI);
                    const V3 rel = pos - center;
                    s->idR.push_back(s->idR.size());
                    s->bodyR.push_back(0);
                    s->R[TSP_R_XREL].push_back(Dot(rel, box.u));
                    s->R[TSP_R_YREL].push_back(Dot(rel, box.v));
                    s->R[TSP_R_ZREL].push_back(Dot(rel, box.w));
                }
            }
        }
        ComputeCoefficients(s->R[TSP_R_XREL], s->R[TSP_R_YREL], s->R[TSP_R_ZREL], s->R[TSP_R_VOL], s->R[TSP_R_PSI],
                            0, static_cast<uint32_t>(s->idR.size()), p);

        const float volume = box.uLen * box.vLen * box.wLen;
        const float mass = spec->rigid_density > 0.0f ? spec->rigid_density * volume : 0.0f;
        const float massInv = mass > 0.0f ? 1.0f / mass : 0.0f;
        for (int f = 0; f < TSP_RB_COUNT; f++) { s->RB[f].assign(1, 0.0f); }
        s->idRB.assign(1, 0);
        s->countRB.assign(1, static_cast<uint32_t>(s->idR.size()));
        s->RB[TSP_RB_XCM][0] = center.x;
        s->RB[TSP_RB_YCM][0] = center.y;
        s->RB[TSP_RB_ZCM][0] = center.z;
        s->RB[TSP_RB_MASS][0] = mass;
        s->RB[TSP_RB_MASS_INV][0] = massInv;
        s->RB[TSP_RB_A00][0] = box.u.x; s->RB[TSP_RB_A01][0] = box.v.x; s->RB[TSP_RB_A02][0] = box.w.x;
        s->RB[TSP_RB_A10][0] = box.u.y; s->RB[TSP_RB_A11][0] = box.v.y; s->RB[TSP_RB_A12][0] = box.w.y;
        s->RB[TSP_RB_A20][0] = box.u.z; s->RB[TSP_RB_A21][0] = box.v.z; s->RB[TSP_RB_A22][0] = box.w.z;
        s->RB[TSP_RB_I00][0] = mass > 0.0f ? 12.0f / (mass * (box.vLen * box.vLen + box.wLen * box.wLen)) : 0.0f;
        s->RB[TSP_RB_I11][0] = mass > 0.0f ? 12.0f / (mass * (box.uLen * box.uLen + box.wLen * box.wLen)) : 0.0f;
        s->RB[TSP_RB_I22][0] = mass > 0.0f ? 12.0f / (mass * (box.uLen * box.uLen + box.vLen * box.vLen)) : 0.0f;

        // ApplyInitialRigidBodyState (rigid_cube.h:71-120)
        const float* w0 = spec->rigid_omega;
        s->RB[TSP_RB_VX][0] = spec->rigid_vel[0];
        s->RB[TSP_RB_VY][0] = spec->rigid_vel[1];
        s->RB[TSP_RB_VZ][0] = spec->rigid_vel[2];
        s->RB[TSP_RB_WX][0] = w0[0];
        s->RB[TSP_RB_WY][0] = w0[1];
        s->RB[TSP_RB_WZ][0] = w0[2];
        const float a00 = box.u.x, a01 = box.v.x, a02 = box.w.x;
        const float a10 = box.u.y, a11 = box.v.y, a12 = box.w.y;
        const float a20 = box.u.z, a21 = box.v.z, a22 = box.w.z;
        const float obx = a00 * w0[0] + a10 * w0[1] + a20 * w0[2];
        const float oby = a01 * w0[0] + a11 * w0[1] + a21 * w0[2];
        const float obz = a02 * w0[0] + a12 * w0[1] + a22 * w0[2];
        const float ibx = s->RB[TSP_RB_I00][0] > 0.0f ? 1.0f / s->RB[TSP_RB_I00][0] : 0.0f;
        const float iby = s->RB[TSP_RB_I11][0] > 0.0f ? 1.0f / s->RB[TSP_RB_I11][0] : 0.0f;
        const float ibz = s->RB[TSP_RB_I22][0] > 0.0f ? 1.0f / s->RB[TSP_RB_I22][0] : 0.0f;
        const float lbx = ibx * obx, lby = iby * oby, lbz = ibz * obz;
        s->RB[TSP_RB_LX][0] = a00 * lbx + a01 * lby + a02 * lbz;
        s->RB[TSP_RB_LY][0] = a10 * lbx + a11 * lby + a12 * lbz;
        s->RB[TSP_RB_LZ][0] = a20 * lbx + a21 * lby + a22 * lbz;
    }

    // --- domain (ComputeSimDomainBounds over the container) ---
    {
        const V3 corners[8] = {
            tank.origin, tank.origin + tank.u * tank.uLen, tank.origin + tank.v * tank.vLen,
            tank.origin + tank.w * tank.wLen, tank.origin + tank.u * tank.uLen + tank.v * tank.vLen,
            tank.origin + tank.u * tank.uLen + tank.w * tank.wLen,
            tank.origin + tank.v * tank.vLen + tank.w * tank.wLen,
            tank.origin + tank.u * tank.uLen + tank.v * tank.vLen + tank.w * tank.wLen };
        float mn[3] = { tank.origin.x, tank.origin.y, tank.origin.z };
        float mx[3] = { tank.origin.x, tank.origin.y, tank.origin.z };
        for (const V3& c : corners)
        {
            const float v[3] = { c.x, c.y, c.z };
            for (int a = 0; a < 3; a++) { mn[a] = std::min(mn[a], v[a]); mx[a] = std::max(mx[a], v[a]); }
        }
        const float pad = p.layers > 0 ? (static_cast<float>(p.layers) - 0.5f) * p.size : 0.0f;
        for (int a = 0; a < 3; a++) { s->domMin[a] = mn[a] - pad; s->domMax[a] = mx[a] + pad; }
    }
    *out = s;
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
