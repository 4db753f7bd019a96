// kernels_contact.hip -- the rigid contact solve of the reference CPU backend
// (CPU/backend/rigid_solver.cpp:23-799, SolveRigidContacts) on the device, for every scene of a
// context at once.
//
// Rigid particles whose neighbor search found boundary particles (RB) or particles of other
// bodies (RR) form a small IISPH problem per scene: contact densities rho = 0.7 + sum 0.7 V_j W,
// diagonals from the body-level gradient sums, relaxation factors shared by each body's
// contacts, and relaxed Jacobi sweeps on non-negative contact pressures whose forces move the
// bodies (inverse mass, world inverse inertia) until the average density error meets the
// scene's target. The contact lists are built for all scenes by a few kernels; then one block
// per scene runs the whole solve -- set-up, every sweep and the final loads -- with block
// barriers between its phases: contact problems are small, and one launch with no host round
// trip beats a launch per phase and sweep. Every FP64 sum runs in a fixed order inside its
// scene's block, so a scene's result does not depend on the other scenes of the batch.
//
// Layout: the RB lists hold global boundary indices; the boundary particles in contact are
// numbered by the exclusive scan of their flags (cbScan), in ascending boundary order, which
// keeps each scene's contact particles contiguous. (The reference numbers them in the order of
// first reference; the numbering only orders the FP64 error sum.) The reverse lists BR hold
// each contact boundary particle's rigid neighbors in ascending order, as the reference builds
// them.
#include "tsp_device.h"

namespace tsp
{
    constexpr float kContactVolume = 0.7f;   // RIGID_CONTACT_VOLUME_FACTOR, rigid_solver.h:23
    constexpr int kMaxBodies = 8;            // distinct bodies around one contact particle

    static inline uint32_t blocksN(uint64_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

    __device__ __forceinline__ uint32_t contactCount(const Dev& d) { return d.cbScan[d.nB]; }

    __device__ __forceinline__ float& cr(const Dev& d, int f, uint32_t i) { return d.cR[static_cast<size_t>(f) * d.nR + i]; }
    __device__ __forceinline__ float& cb(const Dev& d, int f, uint32_t c) { return d.cB[static_cast<size_t>(f) * d.rbCap + c]; }

    // ContactKernel, rigid_solver.cpp:51-63: the weight (no near-zero special case) and the
    // scalar gradient
    __device__ __forceinline__ float contactKernel(float dx, float dy, float dz, const Consts& k, float& gx, float& gy, float& gz)
    {
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        if (d2 >= k.r2) { gx = gy = gz = 0.0f; return 0.0f; }
        const float q = __builtin_sqrtf(d2) * k.hinv;
        const float a1 = 1.0f - q, a2 = 2.0f - q;
        const float t1 = a1 < 0.0f ? 0.0f : a1;
        const float t2 = a2 < 0.0f ? 0.0f : a2;
        kgScalar(dx, dy, dz, k, gx, gy, gz);
        return k.alpha * (t2 * t2 * t2 - 4.0f * t1 * t1 * t1);
    }

    // ----------------------------------------------------------------------------------------
    // contact lists (BuildRigidBoundaryNeighbors, rigid_solver.cpp:90-149)

    // RB / RR offsets from the counts of the search (scanned afterwards)
    __global__ void __launch_bounds__(kBlock) kContactOffsets(Dev d)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i > d.nR) { return; }
        d.rbOff[i] = i < d.nR ? d.rbCnt[i] : 0u;
        d.rrOff[i] = i < d.nR ? d.rrCnt[i] : 0u;
    }

    // flags of the boundary particles named by RB lists (cbScan cleared before)
    __global__ void __launch_bounds__(kBlock) kContactMark(Dev d)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= d.nR) { return; }
        for (uint32_t q = d.rbOff[i]; q < d.rbOff[i + 1]; q++) { d.cbScan[d.rbIdx[q]] = 1u; }
    }

    // contact index -> boundary index, and the BR counts (brOff cleared before)
    __global__ void __launch_bounds__(kBlock) kContactLocal(Dev d)
    {
        const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
        if (b >= d.nB) { return; }
        const uint32_t c = d.cbScan[b];
        if (d.cbScan[b + 1] != c) { d.cbLocal[c] = b; }
    }

    __global__ void __launch_bounds__(kBlock) kContactBrCount(Dev d)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= d.nR) { return; }
        for (uint32_t q = d.rbOff[i]; q < d.rbOff[i + 1]; q++) { atomicAdd(&d.brOff[d.cbScan[d.rbIdx[q]]], 1u); }
    }

    __global__ void __launch_bounds__(kBlock) kContactBrFill(Dev d)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= d.nR) { return; }
        for (uint32_t q = d.rbOff[i]; q < d.rbOff[i + 1]; q++)
        {
            const uint32_t slot = atomicAdd(&d.brPos[d.cbScan[d.rbIdx[q]]], 1u);
            d.brIdx[slot] = i;
        }
    }

    // each BR list in ascending rigid order (insertion sort; the lists are short)
    __global__ void __launch_bounds__(kBlock) kContactBrSort(Dev d)
    {
        const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
        if (c >= contactCount(d)) { return; }
        const uint32_t b0 = d.brOff[c], b1 = d.brOff[c + 1];
        for (uint32_t q = b0 + 1; q < b1; q++)
        {
            const uint32_t v = d.brIdx[q];
            uint32_t t = q;
            while (t > b0 && d.brIdx[t - 1] > v) { d.brIdx[t] = d.brIdx[t - 1]; t--; }
            d.brIdx[t] = v;
        }
    }

    // ----------------------------------------------------------------------------------------
    // one contact neighbor set (ForEachContactNeighbor, rigid_solver.cpp:23-49): of a rigid
    // particle its RR entries, then its RB entries; of a contact boundary particle its BR entries.
    // fn(j, jBoundary, xj, yj, zj, Vj) with j a rigid index or a contact index.

    template <typename Fn>
    __device__ __forceinline__ void forEachContact(const Dev& d, bool boundary, uint32_t i, Fn&& fn)
    {
        if (boundary)
        {
            for (uint32_t q = d.brOff[i]; q < d.brOff[i + 1]; q++)
            {
                const uint32_t j = d.brIdx[q];
                const float4 pj = d.rpos[j];
                fn(j, false, pj.x, pj.y, pj.z, d.rvel[j].w);
            }
            return;
        }
        for (uint32_t q = d.rrOff[i]; q < d.rrOff[i + 1]; q++)
        {
            const uint32_t j = d.rrIdx[q];
            const float4 pj = d.rpos[j];
            fn(j, false, pj.x, pj.y, pj.z, d.rvel[j].w);
        }
        for (uint32_t q = d.rbOff[i]; q < d.rbOff[i + 1]; q++)
        {
            const uint32_t lb = d.rbIdx[q];
            const float4 pj = d.bpos[lb];
            fn(d.cbScan[lb], true, pj.x, pj.y, pj.z, d.bvol[lb]);
        }
    }

    // the position and volume of a rigid particle (boundary = false) or contact boundary particle
    struct ContactSelf
    {
        bool boundary;
        uint32_t i;
        float x, y, z, vol;
    };

    __device__ __forceinline__ ContactSelf contactSelf(const Dev& d, bool boundary, uint32_t i)
    {
        ContactSelf s;
        s.boundary = boundary;
        s.i = i;
        const float4 p = boundary ? d.bpos[d.cbLocal[i]] : d.rpos[i];
        s.x = p.x; s.y = p.y; s.z = p.z;
        s.vol = boundary ? d.bvol[d.cbLocal[i]] : d.rvel[i].w;
        return s;
    }

    // BodyGradient / AddBodyGradient, rigid_solver.cpp:65-86
    struct BodyGradients
    {
        uint32_t body[kMaxBodies];
        float x[kMaxBodies], y[kMaxBodies], z[kMaxBodies], tx[kMaxBodies], ty[kMaxBodies], tz[kMaxBodies];
        int n = 0;
        bool overflow = false;

        __device__ __forceinline__ void add(uint32_t b, float gx, float gy, float gz, float rx, float ry, float rz)
        {
            int e = 0;
            while (e < n && body[e] != b) { e++; }
            if (e == n)
            {
                if (n == kMaxBodies) { overflow = true; return; }
                body[e] = b; x[e] = y[e] = z[e] = tx[e] = ty[e] = tz[e] = 0.0f;
                n++;
            }
            x[e] += gx; y[e] += gy; z[e] += gz;
            tx[e] += ry * gz - rz * gy;
            ty[e] += rz * gx - rx * gz;
            tz[e] += rx * gy - ry * gx;
        }
    };

    // RigidContactGeometry, rigid_solver.cpp:151-227: the world inverse inertia of a body
    __device__ __forceinline__ void geometryBody(const Dev& d, uint32_t b)
    {
        float A[9], B[9], T[9];
        for (int q = 0; q < 9; q++) { A[q] = rbv(d, RB_A00 + q, b); B[q] = rbv(d, RB_I00 + q, b); }
        for (int r = 0; r < 3; r++)
        {
            for (int c = 0; c < 3; c++) { T[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c]; }
        }
        for (int r = 0; r < 3; r++)
        {
            for (int c = 0; c < 3; c++)
            {
                d.cIw[(3 * r + c) * d.nRB + b] = T[3 * r] * A[3 * c] + T[3 * r + 1] * A[3 * c + 1] + T[3 * r + 2] * A[3 * c + 2];
            }
        }
        d.cBodyB[b] = 0u;
    }

    // ... and the world offset A r_rel of a rigid particle
    __device__ __forceinline__ void geometryParticle(const Dev& d, uint32_t i)
    {
        const uint32_t b = d.rbody[i];
        const float lx = d.rxrel[i], ly = d.ryrel[i], lz = d.rzrel[i];
        d.cRw[i] = __builtin_fmaf(rbv(d, RB_A00 + 0, b), lx, __builtin_fmaf(rbv(d, RB_A00 + 1, b), ly, rbv(d, RB_A00 + 2, b) * lz));
        d.cRw[d.nR + i] = __builtin_fmaf(rbv(d, RB_A00 + 3, b), lx, __builtin_fmaf(rbv(d, RB_A00 + 4, b), ly, rbv(d, RB_A00 + 5, b) * lz));
        d.cRw[2 * d.nR + i] = __builtin_fmaf(rbv(d, RB_A00 + 6, b), lx, __builtin_fmaf(rbv(d, RB_A00 + 7, b), ly, rbv(d, RB_A00 + 8, b) * lz));
    }

    // RigidContactDensityAndDiagonal, rigid_solver.cpp:229-344, for one particle; returns the
    // kContactBad* flags it raises
    __device__ uint32_t densityOne(const Dev& d, const ContactSelf& s, const Consts& k)
    {
        const float volRest = kContactVolume * s.vol;
        float rho = kContactVolume;
        bool has = false;
        BodyGradients g;
        const uint32_t ownBody = s.boundary ? 0u : d.rbody[s.i];
        const float rxi = s.boundary ? 0.0f : d.cRw[s.i];
        const float ryi = s.boundary ? 0.0f : d.cRw[d.nR + s.i];
        const float rzi = s.boundary ? 0.0f : d.cRw[2 * d.nR + s.i];
        forEachContact(d, s.boundary, s.i, [&](uint32_t j, bool jb, float xj, float yj, float zj, float vj) {
            float gx, gy, gz;
            const float w = contactKernel(s.x - xj, s.y - yj, s.z - zj, k, gx, gy, gz);
            if (!(w > 0.0f)) { return; }
            has = true;
            const float vn = kContactVolume * vj;
            rho += vn * w;
            gx *= vn; gy *= vn; gz *= vn;
            if (!s.boundary) { g.add(ownBody, gx, gy, gz, rxi, ryi, rzi); }
            if (!jb) { g.add(d.rbody[j], -gx, -gy, -gz, d.cRw[j], d.cRw[d.nR + j], d.cRw[2 * d.nR + j]); }
        });
        if (g.overflow) { return kContactManyBodies; }
        double diag = 0.0;
        const float* iw = d.cIw;
        const uint32_t nb = d.nRB;
        for (int e = 0; e < g.n; e++)
        {
            const uint32_t b = g.body[e];
            const float wx = iw[0 * nb + b] * g.tx[e] + iw[1 * nb + b] * g.ty[e] + iw[2 * nb + b] * g.tz[e];
            const float wy = iw[3 * nb + b] * g.tx[e] + iw[4 * nb + b] * g.ty[e] + iw[5 * nb + b] * g.tz[e];
            const float wz = iw[6 * nb + b] * g.tx[e] + iw[7 * nb + b] * g.ty[e] + iw[8 * nb + b] * g.tz[e];
            const float term = rbv(d, RB_MASS_INV, b) * (g.x[e] * g.x[e] + g.y[e] * g.y[e] + g.z[e] * g.z[e])
                             + g.tx[e] * wx + g.ty[e] * wy + g.tz[e] * wz;
            diag += static_cast<double>(term);
        }
        const float dg = static_cast<float>(static_cast<double>(k.dt) * static_cast<double>(k.dt) * static_cast<double>(volRest)
                                            / static_cast<double>(rho * rho) * diag);
        const float om = has ? 0.5f : 0.0f;
        if (s.boundary)
        {
            cb(d, CB_RHO, s.i) = rho; cb(d, CB_DIAG, s.i) = dg; cb(d, CB_P, s.i) = 0.0f;
            cb(d, CB_SRC, s.i) = rho - 1.0f; cb(d, CB_OM, s.i) = om;
        }
        else
        {
            cr(d, CR_RHO, s.i) = rho; cr(d, CR_DIAG, s.i) = dg; cr(d, CR_P, s.i) = 0.0f;
            cr(d, CR_SRC, s.i) = rho - 1.0f; cr(d, CR_OM, s.i) = om;
        }
        return (!__builtin_isfinite(dg) || dg < 0.0f) ? kContactBadDiag : 0u;
    }

    // the distinct movable bodies among a contact boundary particle's rigid neighbors
    __device__ __forceinline__ int movableBodies(const Dev& d, uint32_t c, uint32_t* bodies, bool& overflow)
    {
        int n = 0;
        for (uint32_t q = d.brOff[c]; q < d.brOff[c + 1]; q++)
        {
            const uint32_t b = d.rbody[d.brIdx[q]];
            if (rbv(d, RB_MASS_INV, b) <= 0.0f) { continue; }
            int e = 0;
            while (e < n && bodies[e] != b) { e++; }
            if (e < n) { continue; }
            if (n == kMaxBodies) { overflow = true; return n; }
            bodies[n++] = b;
        }
        return n;
    }

    // RigidContactPressureForces, rigid_solver.cpp:409-460, for one rigid particle
    __device__ void forcesOne(const Dev& d, uint32_t i, const Consts& k)
    {
        const float4 pi4 = d.rpos[i];
        const float vi = d.rvel[i].w;
        const float rhoi = cr(d, CR_RHO, i);
        const float pi = cr(d, CR_P, i) / (rhoi * rhoi);
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        forEachContact(d, false, i, [&](uint32_t j, bool jb, float xj, float yj, float zj, float vj) {
            float gx, gy, gz;
            const float w = contactKernel(pi4.x - xj, pi4.y - yj, pi4.z - zj, k, gx, gy, gz);
            if (!(w > 0.0f)) { return; }
            const float rho = jb ? cb(d, CB_RHO, j) : cr(d, CR_RHO, j);
            const float pj = (jb ? cb(d, CB_P, j) : cr(d, CR_P, j)) / (rho * rho);
            const float coeff = -kContactVolume * vi * kContactVolume * vj * (pi + pj);
            fx += coeff * gx; fy += coeff * gy; fz += coeff * gz;
        });
        cr(d, CR_FX, i) = fx; cr(d, CR_FY, i) = fy; cr(d, CR_FZ, i) = fz;
    }

    // RigidContactParticleVelocities, rigid_solver.cpp:462-506, for one rigid particle
    __device__ __forceinline__ void velocityOne(const Dev& d, uint32_t i, const Consts& k)
    {
        const uint32_t b = d.rbody[i], nb = d.nRB;
        const float* L = d.cLoad;
        const float* iw = d.cIw;
        const float tx = L[3 * nb + b], ty = L[4 * nb + b], tz = L[5 * nb + b];
        const float wx = iw[0 * nb + b] * tx + iw[1 * nb + b] * ty + iw[2 * nb + b] * tz;
        const float wy = iw[3 * nb + b] * tx + iw[4 * nb + b] * ty + iw[5 * nb + b] * tz;
        const float wz = iw[6 * nb + b] * tx + iw[7 * nb + b] * ty + iw[8 * nb + b] * tz;
        const float rx = d.cRw[i], ry = d.cRw[d.nR + i], rz = d.cRw[2 * d.nR + i];
        const float mi = rbv(d, RB_MASS_INV, b);
        cr(d, CR_VX, i) = k.dt * (mi * L[0 * nb + b] + wy * rz - wz * ry);
        cr(d, CR_VY, i) = k.dt * (mi * L[1 * nb + b] + wz * rx - wx * rz);
        cr(d, CR_VZ, i) = k.dt * (mi * L[2 * nb + b] + wx * ry - wy * rx);
    }

    // RigidContactDensityError, rigid_solver.cpp:508-579, for one particle. initial: with the
    // predicted rigid velocities, into the source term (the swap of SolveRigidContacts);
    // otherwise with the contact velocities, into the residual. Returns the FP64 error term
    // (|residual| where the pressure is positive, max(residual, 0) elsewhere), or NaN when the
    // residual is not finite.
    __device__ double errorOne(const Dev& d, const ContactSelf& s, const Consts& k, bool initial)
    {
        const auto velOf = [&](uint32_t j, float& vx, float& vy, float& vz) {
            if (initial) { const float4 v = d.rvp[j]; vx = v.x; vy = v.y; vz = v.z; }
            else { vx = cr(d, CR_VX, j); vy = cr(d, CR_VY, j); vz = cr(d, CR_VZ, j); }
        };
        float vx = 0.0f, vy = 0.0f, vz = 0.0f;
        if (!s.boundary) { velOf(s.i, vx, vy, vz); }
        float res = s.boundary ? cb(d, CB_SRC, s.i) : cr(d, CR_SRC, s.i);
        forEachContact(d, s.boundary, s.i, [&](uint32_t j, bool jb, float xj, float yj, float zj, float vj) {
            float gx, gy, gz;
            const float w = contactKernel(s.x - xj, s.y - yj, s.z - zj, k, gx, gy, gz);
            if (!(w > 0.0f)) { return; }
            float ux = 0.0f, uy = 0.0f, uz = 0.0f;
            if (!jb) { velOf(j, ux, uy, uz); }
            const float dvx = ux - vx, dvy = uy - vy, dvz = uz - vz;
            const float div = k.dt * (dvx * gx + dvy * gy + dvz * gz);
            res -= kContactVolume * vj * div;
        });
        if (!__builtin_isfinite(res)) { return __builtin_nan(""); }
        if (initial)
        {
            if (s.boundary) { cb(d, CB_SRC, s.i) = res; } else { cr(d, CR_SRC, s.i) = res; }
            return 0.0;
        }
        const float p = s.boundary ? cb(d, CB_P, s.i) : cr(d, CR_P, s.i);
        if (s.boundary) { cb(d, CB_ERR, s.i) = res; } else { cr(d, CR_ERR, s.i) = res; }
        return p > 0.0f ? static_cast<double>(__builtin_fabsf(res)) : static_cast<double>(res < 0.0f ? 0.0f : res);
    }

    // RigidContactPressureRelaxedJacobi, rigid_solver.cpp:581-600, for one particle; false when
    // the new pressure is not finite
    __device__ __forceinline__ bool jacobiOne(const Dev& d, bool boundary, uint32_t i)
    {
        float* P = boundary ? &cb(d, CB_P, i) : &cr(d, CR_P, i);
        const float dg = boundary ? cb(d, CB_DIAG, i) : cr(d, CR_DIAG, i);
        if (!(dg > 0.0f)) { return true; }
        const float om = boundary ? cb(d, CB_OM, i) : cr(d, CR_OM, i);
        const float err = boundary ? cb(d, CB_ERR, i) : cr(d, CR_ERR, i);
        const float p = *P + om * err / dg;
        if (!__builtin_isfinite(p)) { return false; }
        *P = 0.0f < p ? p : 0.0f;
        return true;
    }

    // ----------------------------------------------------------------------------------------
    // the solve of one scene (SolveRigidContacts, rigid_solver.cpp:602-799), one block

    constexpr uint32_t kSolveThreads = 1024;

    // sum over the block in a fixed order (per-wave xor trees, then the waves in order);
    // every thread gets the result
    __device__ double blockSum(double v, double* red)
    {
        for (int o = 32; o > 0; o >>= 1) { v += __shfl_xor(v, o, 64); }
        const uint32_t w = threadIdx.x >> 6;
        __syncthreads();
        if ((threadIdx.x & 63u) == 0) { red[w] = v; }
        __syncthreads();
        double a = red[0];
        for (uint32_t q = 1; q < kSolveThreads / 64; q++) { a += red[q]; }
        return a;
    }

    // blockSum of N values at once, each in blockSum's order (bit-identical), with two barriers
    // instead of 2N
    template <int N>
    __device__ void blockSumN(double (&v)[N], double (*red)[kSolveThreads / 64])
    {
        for (int n = 0; n < N; n++)
        {
            for (int o = 32; o > 0; o >>= 1) { v[n] += __shfl_xor(v[n], o, 64); }
        }
        const uint32_t w = threadIdx.x >> 6;
        __syncthreads();
        if ((threadIdx.x & 63u) == 0) { for (int n = 0; n < N; n++) { red[n][w] = v[n]; } }
        __syncthreads();
        for (int n = 0; n < N; n++)
        {
            double a = red[n][0];
            for (uint32_t q = 1; q < kSolveThreads / 64; q++) { a += red[n][q]; }
            v[n] = a;
        }
    }

    __global__ void __launch_bounds__(kSolveThreads) kContactSolve(Dev d)
    {
        __shared__ double red[kSolveThreads / 64];
        __shared__ double redN[6][kSolveThreads / 64];
        __shared__ uint32_t bad;
        const uint32_t sc = blockIdx.x;
        Solve& so = d.solve[sc];
        if (!so.contactActive) { return; }
        const Scene& S = d.scenes[sc];
        const Consts& k = d.consts[sc];
        const uint32_t t = threadIdx.x, T = kSolveThreads;
        const uint32_t r0 = S.offR, r1 = S.offR + S.nR;
        const uint32_t c0 = d.cbScan[S.offB], c1 = d.cbScan[S.offB + S.nB];
        const uint32_t b0 = S.offRB, b1 = S.offRB + S.nRB;
        if (t == 0) { bad = 0u; }
        // geometry, densities and diagonals
        for (uint32_t b = b0 + t; b < b1; b += T) { geometryBody(d, b); }
        for (uint32_t i = r0 + t; i < r1; i += T) { geometryParticle(d, i); }
        __syncthreads();
        uint32_t myBad = 0;
        for (uint32_t i = r0 + t; i < r1; i += T) { myBad |= densityOne(d, contactSelf(d, false, i), k); }
        for (uint32_t c = c0 + t; c < c1; c += T) { myBad |= densityOne(d, contactSelf(d, true, c), k); }
        __syncthreads();
        // relaxation (RigidContactRelaxation, rigid_solver.cpp:346-407): a body's contacting
        // particles share the factor; boundary particles take the largest count of their bodies
        uint64_t contacts = 0;
        for (uint32_t b = b0; b < b1; b++)
        {
            double n = 0.0;
            for (uint32_t q = d.byBodyOff[b] + t; q < d.byBodyOff[b + 1]; q += T) { n += cr(d, CR_OM, d.byBodyIdx[q]) > 0.0f ? 1.0 : 0.0; }
            const uint32_t count = static_cast<uint32_t>(blockSum(n, red));
            if (count > 0u)
            {
                const float fc = static_cast<float>(count);
                for (uint32_t q = d.byBodyOff[b] + t; q < d.byBodyOff[b + 1]; q += T)
                {
                    float& om = cr(d, CR_OM, d.byBodyIdx[q]);
                    om = om / fc;
                }
                contacts += count;
            }
        }
        double nb = 0.0;
        for (uint32_t c = c0 + t; c < c1; c += T)
        {
            if (cb(d, CB_OM, c) <= 0.0f) { continue; }
            nb += 1.0;
            if (cb(d, CB_DIAG, c) <= 0.0f) { continue; }
            uint32_t bodies[kMaxBodies];
            bool overflow = false;
            const int n = movableBodies(d, c, bodies, overflow);
            if (overflow) { myBad |= kContactManyBodies; continue; }
            for (int e = 0; e < n; e++) { atomicAdd(&d.cBodyB[bodies[e]], 1u); }
        }
        contacts += static_cast<uint64_t>(blockSum(nb, red));
        __syncthreads();
        for (uint32_t c = c0 + t; c < c1; c += T)
        {
            if (cb(d, CB_OM, c) <= 0.0f || cb(d, CB_DIAG, c) <= 0.0f) { continue; }
            uint32_t mx = 0;
            for (uint32_t q = d.brOff[c]; q < d.brOff[c + 1]; q++)
            {
                const uint32_t n = d.cBodyB[d.rbody[d.brIdx[q]]];
                mx = n > mx ? n : mx;
            }
            if (mx > 0u) { cb(d, CB_OM, c) = cb(d, CB_OM, c) / static_cast<float>(mx); }
        }
        // the source motion: the body loads so far (external and fluid), no contact loads
        for (uint32_t b = b0 + t; b < b1; b += T) { rigidPredictOne(d, b); }
        __syncthreads();
        for (uint32_t i = r0 + t; i < r1; i += T) { rigidUpdateOne(d, i, true); }
        __syncthreads();
        for (uint32_t i = r0 + t; i < r1; i += T) { if (errorOne(d, contactSelf(d, false, i), k, true) != 0.0) { myBad |= kContactBadResidual; } }
        for (uint32_t c = c0 + t; c < c1; c += T) { if (errorOne(d, contactSelf(d, true, c), k, true) != 0.0) { myBad |= kContactBadResidual; } }
        if (myBad) { atomicOr(&bad, myBad); }
        __syncthreads();
        uint32_t iters = 0;
        float avg = 0.0f;
        bool converged = false;
        while (!bad)
        {
            for (uint32_t i = r0 + t; i < r1; i += T) { forcesOne(d, i, k); }
            __syncthreads();
            // ReduceRigidParticleForcesToBody(Contact) into the cleared contact loads, FP64 sums
            for (uint32_t b = b0; b < b1; b++)
            {
                const float a[9] = { rbv(d, RB_A00 + 0, b), rbv(d, RB_A00 + 1, b), rbv(d, RB_A00 + 2, b),
                                     rbv(d, RB_A00 + 3, b), rbv(d, RB_A00 + 4, b), rbv(d, RB_A00 + 5, b),
                                     rbv(d, RB_A00 + 6, b), rbv(d, RB_A00 + 7, b), rbv(d, RB_A00 + 8, b) };
                double acc[6] = { 0, 0, 0, 0, 0, 0 };
                for (uint32_t q = d.byBodyOff[b] + t; q < d.byBodyOff[b + 1]; q += T)
                {
                    const uint32_t i = d.byBodyIdx[q];
                    const float lx = d.rxrel[i], ly = d.ryrel[i], lz = d.rzrel[i];
                    const float rx = __builtin_fmaf(a[0], lx, __builtin_fmaf(a[1], ly, a[2] * lz));
                    const float ry = __builtin_fmaf(a[3], lx, __builtin_fmaf(a[4], ly, a[5] * lz));
                    const float rz = __builtin_fmaf(a[6], lx, __builtin_fmaf(a[7], ly, a[8] * lz));
                    const double fx = cr(d, CR_FX, i), fy = cr(d, CR_FY, i), fz = cr(d, CR_FZ, i);
                    acc[0] += fx; acc[1] += fy; acc[2] += fz;
                    acc[3] += static_cast<double>(ry) * fz - static_cast<double>(rz) * fy;
                    acc[4] += static_cast<double>(rz) * fx - static_cast<double>(rx) * fz;
                    acc[5] += static_cast<double>(rx) * fy - static_cast<double>(ry) * fx;
                }
                blockSumN(acc, redN);
                if (t == 0)
                {
                    for (int f = 0; f < 6; f++) { d.cLoad[f * d.nRB + b] = 0.0f + static_cast<float>(acc[f]); }
                }
            }
            __syncthreads();
            for (uint32_t i = r0 + t; i < r1; i += T) { velocityOne(d, i, k); }
            __syncthreads();
            double eR = 0.0, eB = 0.0;
            uint32_t flags = 0;
            for (uint32_t i = r0 + t; i < r1; i += T)
            {
                const double e = errorOne(d, contactSelf(d, false, i), k, false);
                if (e != e) { flags |= kContactBadResidual; } else { eR += e; }
            }
            for (uint32_t c = c0 + t; c < c1; c += T)
            {
                const double e = errorOne(d, contactSelf(d, true, c), k, false);
                if (e != e) { flags |= kContactBadResidual; } else { eB += e; }
            }
            if (flags) { atomicOr(&bad, flags); }
            double e2[2] = { eR, eB };
            blockSumN(e2, redN);
            const double sum = e2[0] + e2[1];
            avg = contacts > 0ull ? static_cast<float>(sum / static_cast<double>(contacts)) : 0.0f;
            converged = avg <= k.contactDevTarget;
            __syncthreads();
            // every wave reads the flags before any wave of the sweep below can raise them, so all
            // waves leave the loop at the same iteration
            const bool stop = converged || iters == k.contactMaxIters || bad != 0u;
            __syncthreads();
            if (stop) { break; }
            flags = 0;
            for (uint32_t i = r0 + t; i < r1; i += T) { if (!jacobiOne(d, false, i)) { flags |= kContactBadPressure; } }
            for (uint32_t c = c0 + t; c < c1; c += T) { if (!jacobiOne(d, true, c)) { flags |= kContactBadPressure; } }
            if (flags) { atomicOr(&bad, flags); }
            __syncthreads();
            iters++;
        }
        __syncthreads();
        const uint32_t fl = bad;
        // the final contact loads added to every body of the scene
        if (!fl)
        {
            for (uint32_t b = b0 + t; b < b1; b += T)
            {
                for (int f = 0; f < 6; f++) { d.rbForce[f * d.nRB + b] += d.cLoad[f * d.nRB + b]; }
            }
        }
        if (t == 0)
        {
            so.contactBoundary = c1 - c0;
            so.contactIters = iters;
            so.contactAvg = avg;
            so.contactConverged = converged ? 1u : 0u;
            so.contacts = contacts;
            so.contactBad = fl;
            so.contactDone = 1u;
        }
    }

    // ----------------------------------------------------------------------------------------
    // launchers

    void launchScan(hipStream_t s, uint32_t* v, uint32_t m, uint32_t* tmp, uint32_t* fail);

    // RB/RR offsets from the search counts (before the contact fill of the rigid query)
    void launchContactOffsets(hipStream_t s, const Dev& d, uint32_t* tmp)
    {
        kContactOffsets<<<blocksN(d.nR + 1ull), kBlock, 0, s>>>(d);
        launchScan(s, d.rbOff, d.nR + 1, tmp, &d.status->scanFail);
        launchScan(s, d.rrOff, d.nR + 1, tmp, &d.status->scanFail);
    }

    // the contact boundary particles and their BR lists for every scene, then one block per scene
    // for the solve; nCBMax bounds the contact boundary particles (the RB entries)
    void launchContactSolve(hipStream_t s, const Dev& d, uint32_t nCBMax, uint32_t* tmp)
    {
        (void)hipMemsetAsync(d.cbScan, 0, sizeof(uint32_t) * (d.nB + 1ull), s);
        (void)hipMemsetAsync(d.brOff, 0, sizeof(uint32_t) * (nCBMax + 1ull), s);
        kContactMark<<<blocksN(d.nR), kBlock, 0, s>>>(d);
        launchScan(s, d.cbScan, d.nB + 1, tmp, &d.status->scanFail);
        if (d.nB) { kContactLocal<<<blocksN(d.nB), kBlock, 0, s>>>(d); }
        kContactBrCount<<<blocksN(d.nR), kBlock, 0, s>>>(d);
        launchScan(s, d.brOff, nCBMax + 1, tmp, &d.status->scanFail);
        (void)hipMemcpyAsync(d.brPos, d.brOff, sizeof(uint32_t) * (nCBMax + 1ull), hipMemcpyDeviceToDevice, s);
        kContactBrFill<<<blocksN(d.nR), kBlock, 0, s>>>(d);
        if (nCBMax) { kContactBrSort<<<blocksN(nCBMax), kBlock, 0, s>>>(d); }
        kContactSolve<<<d.nScenes, kSolveThreads, 0, s>>>(d);
    }
}
