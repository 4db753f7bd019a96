// scene_plan.cpp -- the plan of a generated scene (scene_plan.h): the container faces, the fluid
// block, the rigid body in full and the domain; ComputeParticleCoefficients. Compiled into both
// libraries: the host builder and the device builder work from the same plan.
// Built with -ffp-contract=off so every float operation rounds where the reference's does.

#include <limits>
#include <numeric>
#include <tuple>

#include "scene_plan.h"

namespace tsp::scene
{
    namespace
    {
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

        Face ExtendFace(Face f, float padU, float padV)
        {
            f.origin = f.origin - (f.u * padU + f.v * padV);
            f.uLen += 2.0f * padU;
            f.vLen += 2.0f * padV;
            return f;
        }
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


    int MakePlan(const tsp_scene_spec& spec, Plan& plan)
    {
        if (!(spec.particle_size > 0.0f)) { return TSP_ERR_INVALID; }
        Params p{};
        p.size = spec.particle_size;
        const double hd = static_cast<double>(p.size);
        p.sizeInv = static_cast<float>(1.0 / hd);
        p.radius = 2.0f * p.size;
        p.radiusSqu = p.radius * p.radius;
        p.rho0 = spec.rest_density;
        p.alpha = static_cast<float>((1.0 / hd) * (1.0 / hd) * (1.0 / hd) / (4.0 * 3.14159265358979323846));
        p.massB = p.rho0 * p.size * p.size * p.size;
        p.layers = spec.boundary_layers;

        const Box tank{ { spec.container_origin[0], spec.container_origin[1], spec.container_origin[2] },
                        { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 },
                        spec.container_size[0], spec.container_size[1], spec.container_size[2] };

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
            uint64_t first = 0;
            for (const Face& f : faces)
            {
                BoundaryFace& bf = plan.faces[plan.nFaces++];
                bf.f = f;
                bf.nu = TilesCount(f.uLen, p.size);
                bf.nv = TilesCount(f.vLen, p.size);
                bf.first = first;
                first += static_cast<uint64_t>(p.layers) * bf.nu * bf.nv;
            }
            plan.nB = first;
        }

        // --- fluid block from lattice counts (BuildFluidBlock_OBB) ---
        {
            const uint32_t Nu = spec.fluid_counts[0], Nv = spec.fluid_counts[1], Nw = spec.fluid_counts[2];
            const uint64_t nF64 = static_cast<uint64_t>(Nu) * Nv * Nw;
            if (nF64 > static_cast<uint64_t>(std::numeric_limits<int32_t>::max()))
            {
                return TSP_ERR_CAPACITY;
            }
            const uint64_t nF = nF64;
            const float radius = p.size * 0.5f;
            const Box blk{ { spec.fluid_origin[0], spec.fluid_origin[1], spec.fluid_origin[2] },
                           { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 },
                           static_cast<float>(Nu) * p.size, static_cast<float>(Nv) * p.size,
                           static_cast<float>(Nw) * p.size };
            const float uLenI = blk.uLen - 2.0f * radius;
            const float vLenI = blk.vLen - 2.0f * radius;
            const float wLenI = blk.wLen - 2.0f * radius;
            const float uOff = radius + StartOffset(uLenI, p.size, static_cast<int>(Nu));
            const float vOff = radius + StartOffset(vLenI, p.size, static_cast<int>(Nv));
            const float wOff = radius + StartOffset(wLenI, p.size, static_cast<int>(Nw));
            FluidBlock& fb = plan.fluid;
            fb.blk = blk;
            fb.uOff = uOff; fb.vOff = vOff; fb.wOff = wOff;
            fb.radius = radius;
            fb.jitAmp = spec.jitter_frac * p.size;
            fb.seed = spec.jitter_seed;
            fb.nu = Nu; fb.nv = Nv; fb.nw = Nw;
            plan.nF = nF;
        }

        // --- hollow box rigid body (BuildRigidBody_OBB_Hollow + rigid_cube.h initial state) ---
        if (spec.rigid_enable)
        {
            double qw = spec.rigid_quat[0], qx = spec.rigid_quat[1], qy = spec.rigid_quat[2], qz = spec.rigid_quat[3];
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
            const float uLen = spec.rigid_size[0], vLen = spec.rigid_size[1], wLen = spec.rigid_size[2];
            const V3 centerIn{ spec.rigid_center[0], spec.rigid_center[1], spec.rigid_center[2] };
            const V3 origin = centerIn - (uHat * (0.5f * uLen) + vHat * (0.5f * vLen) + wHat * (0.5f * wLen));
            const Box box{ origin, uHat, vHat, wHat, uLen, vLen, wLen };

            const float radius = p.size * 0.5f;
            const float uLenI = box.uLen - 2.0f * radius;
            const float vLenI = box.vLen - 2.0f * radius;
            const float wLenI = box.wLen - 2.0f * radius;
            if (uLenI <= 0.0f || vLenI <= 0.0f || wLenI <= 0.0f)
            {
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
                        plan.idR.push_back(plan.idR.size());
                        plan.bodyR.push_back(0);
                        plan.R[TSP_R_XREL].push_back(Dot(rel, box.u));
                        plan.R[TSP_R_YREL].push_back(Dot(rel, box.v));
                        plan.R[TSP_R_ZREL].push_back(Dot(rel, box.w));
                    }
                }
            }
            ComputeCoefficients(plan.R[TSP_R_XREL], plan.R[TSP_R_YREL], plan.R[TSP_R_ZREL], plan.R[TSP_R_VOL], plan.R[TSP_R_PSI],
                                0, static_cast<uint32_t>(plan.idR.size()), p);

            const float volume = box.uLen * box.vLen * box.wLen;
            const float mass = spec.rigid_density > 0.0f ? spec.rigid_density * volume : 0.0f;
            const float massInv = mass > 0.0f ? 1.0f / mass : 0.0f;
            for (int f = 0; f < TSP_RB_COUNT; f++) { plan.RB[f].assign(1, 0.0f); }
            plan.idRB.assign(1, 0);
            plan.countRB.assign(1, static_cast<uint32_t>(plan.idR.size()));
            plan.RB[TSP_RB_XCM][0] = center.x;
            plan.RB[TSP_RB_YCM][0] = center.y;
            plan.RB[TSP_RB_ZCM][0] = center.z;
            plan.RB[TSP_RB_MASS][0] = mass;
            plan.RB[TSP_RB_MASS_INV][0] = massInv;
            plan.RB[TSP_RB_A00][0] = box.u.x; plan.RB[TSP_RB_A01][0] = box.v.x; plan.RB[TSP_RB_A02][0] = box.w.x;
            plan.RB[TSP_RB_A10][0] = box.u.y; plan.RB[TSP_RB_A11][0] = box.v.y; plan.RB[TSP_RB_A12][0] = box.w.y;
            plan.RB[TSP_RB_A20][0] = box.u.z; plan.RB[TSP_RB_A21][0] = box.v.z; plan.RB[TSP_RB_A22][0] = box.w.z;
            plan.RB[TSP_RB_I00][0] = mass > 0.0f ? 12.0f / (mass * (box.vLen * box.vLen + box.wLen * box.wLen)) : 0.0f;
            plan.RB[TSP_RB_I11][0] = mass > 0.0f ? 12.0f / (mass * (box.uLen * box.uLen + box.wLen * box.wLen)) : 0.0f;
            plan.RB[TSP_RB_I22][0] = mass > 0.0f ? 12.0f / (mass * (box.uLen * box.uLen + box.vLen * box.vLen)) : 0.0f;

            // ApplyInitialRigidBodyState (rigid_cube.h:71-120)
            const float* w0 = spec.rigid_omega;
            plan.RB[TSP_RB_VX][0] = spec.rigid_vel[0];
            plan.RB[TSP_RB_VY][0] = spec.rigid_vel[1];
            plan.RB[TSP_RB_VZ][0] = spec.rigid_vel[2];
            plan.RB[TSP_RB_WX][0] = w0[0];
            plan.RB[TSP_RB_WY][0] = w0[1];
            plan.RB[TSP_RB_WZ][0] = w0[2];
            const float a00 = box.u.x, a01 = box.v.x, a02 = box.w.x;
            const float a10 = box.u.y, a11 = box.v.y, a12 = box.w.y;
            const float a20 = box.u.z, a21 = box.v.z, a22 = box.w.z;
            const float obx = a00 * w0[0] + a10 * w0[1] + a20 * w0[2];
            const float oby = a01 * w0[0] + a11 * w0[1] + a21 * w0[2];
            const float obz = a02 * w0[0] + a12 * w0[1] + a22 * w0[2];
            const float ibx = plan.RB[TSP_RB_I00][0] > 0.0f ? 1.0f / plan.RB[TSP_RB_I00][0] : 0.0f;
            const float iby = plan.RB[TSP_RB_I11][0] > 0.0f ? 1.0f / plan.RB[TSP_RB_I11][0] : 0.0f;
            const float ibz = plan.RB[TSP_RB_I22][0] > 0.0f ? 1.0f / plan.RB[TSP_RB_I22][0] : 0.0f;
            const float lbx = ibx * obx, lby = iby * oby, lbz = ibz * obz;
            plan.RB[TSP_RB_LX][0] = a00 * lbx + a01 * lby + a02 * lbz;
            plan.RB[TSP_RB_LY][0] = a10 * lbx + a11 * lby + a12 * lbz;
            plan.RB[TSP_RB_LZ][0] = a20 * lbx + a21 * lby + a22 * lbz;
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
            for (int a = 0; a < 3; a++) { plan.domMin[a] = mn[a] - pad; plan.domMax[a] = mx[a] + pad; }
        }
        plan.p = p;
        return TSP_OK;
    }
}
