// kernels_contact.hip -- the rigid contact solve of the reference CPU backend
// (CPU/backend/rigid_solver.cpp:23-799, SolveRigidContacts) on the device, for every scene of a
// context at once.
//
// Rigid particles whose neighbor search found boundary particles (RB) or particles of other
// bodies (RR) form a small IISPH problem per scene: contact densities rho = 0.7 + sum 0.7 V_j W,
// diagonals from the body-level gradient sums, relaxation factors shared by each body's
// contacts, and relaxed Jacobi sweeps on non-negative contact pressures whose forces move the
// bodies (inverse mass, world inverse inertia) until the average density error meets the
// scene's target. Every scene decides on its own; the per-scene FP64 sums run in one block per
// scene in a fixed order, so a scene's result does not depend on the other scenes of the batch.
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

    __device__ __forceinline__ uint32_t sceneOfRigid(const Dev& d, uint32_t i) { return d.bodyScene[d.rbody[i]]; }

    __device__ __forceinline__ bool contactRuns(const Dev& d, uint32_t sc)
    {
        const Solve& so = d.solve[sc];
        return so.contactActive != 0u && so.contactDone == 0u;
    }

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
    // RigidContactGeometry, rigid_solver.cpp:151-227, and the per-scene set-up

    __global__ void __launch_bounds__(kBlock) kContactGeometry(Dev d)
    {
        const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
        if (t < d.nRB)
        {
            const uint32_t b = t;
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
        if (t < d.nR)
        {
            const uint32_t i = t;
            const uint32_t b = d.rbody[i];
            const float lx = d.rxrel[i], ly = d.ryrel[i], lz = d.rzrel[i];
            d.cRw[i] = __builtin_fmaf(rbv(d, RB_A00 + 0, b), lx, __builtin_fmaf(rbv(d, RB_A00 + 1, b), ly, rbv(d, RB_A00 + 2, b) * lz));
            d.cRw[d.nR + i] = __builtin_fmaf(rbv(d, RB_A00 + 3, b), lx, __builtin_fmaf(rbv(d, RB_A00 + 4, b), ly, rbv(d, RB_A00 + 5, b) * lz));
            d.cRw[2 * d.nR + i] = __builtin_fmaf(rbv(d, RB_A00 + 6, b), lx, __builtin_fmaf(rbv(d, RB_A00 + 7, b), ly, rbv(d, RB_A00 + 8, b) * lz));
        }
        if (t < d.nScenes)
        {
            const Scene& s = d.scenes[t];
            d.solve[t].contactBoundary = d.cbScan[s.offB + s.nB] - d.cbScan[s.offB];
        }
    }

    // ----------------------------------------------------------------------------------------
    // the contact neighbors of one particle (ForEachContactNeighbor, rigid_solver.cpp:23-49):
    // of a rigid particle its RR entries, then its RB entries; of a contact boundary particle its
    // BR entries. fn(j, jBoundary, xj, yj, zj, Vj) with j a rigid index or a contact index.

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

    // one thread per rigid particle (t < nR) or contact boundary particle (t - nR); the scene
    // and position of the particle; false when the thread has no particle or its scene's solve
    // does not run
    struct ContactSelf
    {
        bool boundary;
        uint32_t i, sc;
        float x, y, z, vol;
    };

    __device__ __forceinline__ bool contactSelf(const Dev& d, uint32_t t, ContactSelf& s)
    {
        if (t < d.nR)
        {
            s.boundary = false;
            s.i = t;
            s.sc = sceneOfRigid(d, t);
            const float4 p = d.rpos[t];
            s.x = p.x; s.y = p.y; s.z = p.z;
            s.vol = d.rvel[t].w;
        }
        else
        {
            const uint32_t c = t - d.nR;
            if (c >= contactCount(d)) { return false; }
            const uint32_t lb = d.cbLocal[c];
            s.boundary = true;
            s.i = c;
            s.sc = d.sceneB[lb];
            const float4 p = d.bpos[lb];
            s.x = p.x; s.y = p.y; s.z = p.z;
            s.vol = d.bvol[lb];
        }
        return contactRuns(d, s.sc);
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

    // RigidContactDensityAndDiagonal, rigid_solver.cpp:229-344
    __global__ void __launch_bounds__(kBlock) kContactDensity(Dev d)
    {
        ContactSelf s;
        if (!contactSelf(d, blockIdx.x * kBlock + threadIdx.x, s)) { return; }
        const Consts& k = d.consts[s.sc];
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
        if (g.overflow) { atomicOr(&d.solve[s.sc].contactBad, kContactManyBodies); return; }
        double diag = 0.0;
        for (int e = 0; e < g.n; e++)
        {
            const uint32_t b = g.body[e];
            const float* iw = d.cIw;
            const uint32_t nb = d.nRB;
            const float wx = iw[0 * nb + b] * g.tx[e] + iw[1 * nb + b] * g.ty[e] + iw[2 * nb + b] * g.tz[e];
            const float wy = iw[3 * nb + b] * g.tx[e] + iw[4 * nb + b] * g.ty[e] + iw[5 * nb + b] * g.tz[e];
            const float wz = iw[6 * nb + b] * g.tx[e] + iw[7 * nb + b] * g.ty[e] + iw[8 * nb + b] * g.tz[e];
            const float term = rbv(d, RB_MASS_INV, b) * (g.x[e] * g.x[e] + g.y[e] * g.y[e] + g.z[e] * g.z[e])
                             + g.tx[e] * wx + g.ty[e] * wy + g.tz[e] * wz;
            diag += static_cast<double>(term);
        }
        const float dg = static_cast<float>(static_cast<double>(k.dt) * static_cast<double>(k.dt) * static_cast<double>(volRest)
                                            / static_cast<double>(rho * rho) * diag);
        if (!__builtin_isfinite(dg) || dg < 0.0f) { atomicOr(&d.solve[s.sc].contactBad, kContactBadDiag); }
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
    }

    // ----------------------------------------------------------------------------------------
    // RigidContactRelaxation, rigid_solver.cpp:346-407

    // one block per body: the body's contacting particles share the factor
    __global__ void __launch_bounds__(kBlock) kContactRelaxRigid(Dev d)
    {
        __shared__ uint32_t red[kBlock];
        const uint32_t b = blockIdx.x;
        const uint32_t sc = d.bodyScene[b];
        if (!contactRuns(d, sc)) { return; }
        uint32_t n = 0;
        for (uint32_t q = d.byBodyOff[b] + threadIdx.x; q < d.byBodyOff[b + 1]; q += kBlock)
        {
            n += cr(d, CR_OM, d.byBodyIdx[q]) > 0.0f ? 1u : 0u;
        }
        red[threadIdx.x] = n;
        __syncthreads();
        for (uint32_t o = kBlock / 2; o > 0; o >>= 1)
        {
            if (threadIdx.x < o) { red[threadIdx.x] += red[threadIdx.x + o]; }
            __syncthreads();
        }
        const uint32_t count = red[0];
        if (count == 0u) { return; }
        const float fc = static_cast<float>(count);
        for (uint32_t q = d.byBodyOff[b] + threadIdx.x; q < d.byBodyOff[b + 1]; q += kBlock)
        {
            float& om = cr(d, CR_OM, d.byBodyIdx[q]);
            om = om / fc;
        }
        if (threadIdx.x == 0) { atomicAdd(reinterpret_cast<unsigned long long*>(&d.solve[sc].contacts), static_cast<unsigned long long>(count)); }
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

    // boundary contacts per movable body
    __global__ void __launch_bounds__(kBlock) kContactRelaxBoundaryCount(Dev d)
    {
        const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
        if (c >= contactCount(d)) { return; }
        const uint32_t sc = d.sceneB[d.cbLocal[c]];
        if (!contactRuns(d, sc) || cb(d, CB_OM, c) <= 0.0f) { return; }
        atomicAdd(reinterpret_cast<unsigned long long*>(&d.solve[sc].contacts), 1ull);
        if (cb(d, CB_DIAG, c) <= 0.0f) { return; }
        uint32_t bodies[kMaxBodies];
        bool overflow = false;
        const int n = movableBodies(d, c, bodies, overflow);
        if (overflow) { atomicOr(&d.solve[sc].contactBad, kContactManyBodies); return; }
        for (int e = 0; e < n; e++) { atomicAdd(&d.cBodyB[bodies[e]], 1u); }
    }

    // a boundary particle shared by several bodies takes the largest of their contact counts
    __global__ void __launch_bounds__(kBlock) kContactRelaxBoundary(Dev d)
    {
        const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
        if (c >= contactCount(d)) { return; }
        const uint32_t sc = d.sceneB[d.cbLocal[c]];
        if (!contactRuns(d, sc) || cb(d, CB_OM, c) <= 0.0f || cb(d, CB_DIAG, c) <= 0.0f) { return; }
        uint32_t mx = 0;
        for (uint32_t q = d.brOff[c]; q < d.brOff[c + 1]; q++)
        {
            const uint32_t n = d.cBodyB[d.rbody[d.brIdx[q]]];
            mx = n > mx ? n : mx;
        }
        if (mx > 0u) { cb(d, CB_OM, c) = cb(d, CB_OM, c) / static_cast<float>(mx); }
    }

    // ----------------------------------------------------------------------------------------
    // RigidContactPressureForces, rigid_solver.cpp:409-460
    __global__ void __launch_bounds__(kBlock) kContactForces(Dev d)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= d.nR) { return; }
        const uint32_t sc = sceneOfRigid(d, i);
        if (!contactRuns(d, sc)) { return; }
        const Consts& k = d.consts[sc];
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

    // ReduceRigidParticleForcesToBody(Contact), cpu_backend.cpp:2089-2172 with FP64 sums, into
    // the cleared contact loads: one block per body
    __global__ void __launch_bounds__(kBlock) kContactLoads(Dev d)
    {
        __shared__ double red[6][kBlock];
        const uint32_t b = blockIdx.x;
        if (!contactRuns(d, d.bodyScene[b])) { return; }
        const uint32_t nb = d.nRB;
        const float a[9] = { rbv(d, RB_A00 + 0, b), rbv(d, RB_A00 + 1, b), rbv(d, RB_A00 + 2, b),
                             rbv(d, RB_A00 + 3, b), rbv(d, RB_A00 + 4, b), rbv(d, RB_A00 + 5, b),
                             rbv(d, RB_A00 + 6, b), rbv(d, RB_A00 + 7, b), rbv(d, RB_A00 + 8, b) };
        double acc[6] = { 0, 0, 0, 0, 0, 0 };
        for (uint32_t q = d.byBodyOff[b] + threadIdx.x; q < d.byBodyOff[b + 1]; q += kBlock)
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
        for (int f = 0; f < 6; f++) { red[f][threadIdx.x] = acc[f]; }
        __syncthreads();
        for (uint32_t o = kBlock / 2; o > 0; o >>= 1)
        {
            if (threadIdx.x < o) { for (int f = 0; f < 6; f++) { red[f][threadIdx.x] += red[f][threadIdx.x + o]; } }
            __syncthreads();
        }
        if (threadIdx.x == 0)
        {
            for (int f = 0; f < 6; f++) { d.cLoad[f * nb + b] = 0.0f + static_cast<float>(red[f][0]); }
        }
    }

    // RigidContactParticleVelocities, rigid_solver.cpp:462-506
    __global__ void __launch_bounds__(kBlock) kContactVelocities(Dev d)
    {
        const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= d.nR) { return; }
        const uint32_t sc = sceneOfRigid(d, i);
        if (!contactRuns(d, sc)) { return; }
        const Consts& k = d.consts[sc];
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

    // RigidContactDensityError, rigid_solver.cpp:508-579. initial: with the predicted rigid
    // velocities, into the source terms (the swap of SolveRigidContacts); otherwise with the
    // contact velocities, into the residuals and the FP64 error terms
    __global__ void __launch_bounds__(kBlock) kContactError(Dev d, bool initial)
    {
        ContactSelf s;
        if (!contactSelf(d, blockIdx.x * kBlock + threadIdx.x, s)) { return; }
        const Consts& k = d.consts[s.sc];
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
        if (!__builtin_isfinite(res)) { atomicOr(&d.solve[s.sc].contactBad, kContactBadResidual); res = 0.0f; }
        if (initial)
        {
            if (s.boundary) { cb(d, CB_SRC, s.i) = res; } else { cr(d, CR_SRC, s.i) = res; }
            return;
        }
        const float p = s.boundary ? cb(d, CB_P, s.i) : cr(d, CR_P, s.i);
        const float term = p > 0.0f ? __builtin_fabsf(res) : (res < 0.0f ? 0.0f : res);
        if (s.boundary) { cb(d, CB_ERR, s.i) = res; d.cErrB[s.i] = term; }
        else { cr(d, CR_ERR, s.i) = res; d.cErrR[s.i] = term; }
    }

    // one block per scene: the FP64 error sum (rigid particles, then contact boundary particles),
    // the average over the contacts and the decision of SolveRigidContacts
    __global__ void __launch_bounds__(kBlock) kContactFinalize(Dev d)
    {
        __shared__ double red[kBlock];
        const uint32_t sc = blockIdx.x;
        if (!contactRuns(d, sc)) { return; }
        const Scene& s = d.scenes[sc];
        double part[2];
        for (int h = 0; h < 2; h++)
        {
            const uint32_t b0 = h ? d.cbScan[s.offB] : s.offR;
            const uint32_t b1 = h ? d.cbScan[s.offB + s.nB] : s.offR + s.nR;
            const double* src = h ? d.cErrB : d.cErrR;
            double a = 0.0;
            for (uint32_t q = b0 + threadIdx.x; q < b1; q += kBlock) { a += src[q]; }
            red[threadIdx.x] = a;
            __syncthreads();
            for (uint32_t o = kBlock / 2; o > 0; o >>= 1)
            {
                if (threadIdx.x < o) { red[threadIdx.x] += red[threadIdx.x + o]; }
                __syncthreads();
            }
            part[h] = red[0];
            __syncthreads();
        }
        if (threadIdx.x != 0) { return; }
        Solve& so = d.solve[sc];
        const Consts& k = d.consts[sc];
        const double sum = part[0] + part[1];
        so.contactAvg = so.contacts > 0ull ? static_cast<float>(sum / static_cast<double>(so.contacts)) : 0.0f;
        so.contactConverged = so.contactAvg <= k.contactDevTarget ? 1u : 0u;
        if (so.contactConverged || so.contactIters == k.contactMaxIters || so.contactBad) { so.contactDone = 1u; }
        else { so.contactIters++; }
    }

    // RigidContactPressureRelaxedJacobi, rigid_solver.cpp:581-600, for the scenes not done
    __global__ void __launch_bounds__(kBlock) kContactJacobi(Dev d)
    {
        ContactSelf s;
        if (!contactSelf(d, blockIdx.x * kBlock + threadIdx.x, s)) { return; }
        float* P = s.boundary ? &cb(d, CB_P, s.i) : &cr(d, CR_P, s.i);
        const float dg = s.boundary ? cb(d, CB_DIAG, s.i) : cr(d, CR_DIAG, s.i);
        if (!(dg > 0.0f)) { return; }
        const float om = s.boundary ? cb(d, CB_OM, s.i) : cr(d, CR_OM, s.i);
        const float err = s.boundary ? cb(d, CB_ERR, s.i) : cr(d, CR_ERR, s.i);
        const float p = *P + om * err / dg;
        if (!__builtin_isfinite(p)) { atomicOr(&d.solve[s.sc].contactBad, kContactBadPressure); return; }
        *P = 0.0f < p ? p : 0.0f;
    }

    // the final contact loads added to the body loads of every body of a scene whose solve ran
    __global__ void __launch_bounds__(kBlock) kContactApply(Dev d)
    {
        const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
        if (b >= d.nRB) { return; }
        const Solve& so = d.solve[d.bodyScene[b]];
        if (!so.contactActive || so.contactBad) { return; }
        const uint32_t nb = d.nRB;
        for (int f = 0; f < 6; f++) { d.rbForce[f * nb + b] += d.cLoad[f * nb + b]; }
    }

    // ----------------------------------------------------------------------------------------
    // launchers

    void launchScan(hipStream_t s, uint32_t* v, uint32_t m, uint32_t* tmp);

    // RB/RR offsets from the search counts (before the contact fill of the rigid query)
    void launchContactOffsets(hipStream_t s, const Dev& d, uint32_t* tmp)
    {
        kContactOffsets<<<blocksN(d.nR + 1ull), kBlock, 0, s>>>(d);
        launchScan(s, d.rbOff, d.nR + 1, tmp);
        launchScan(s, d.rrOff, d.nR + 1, tmp);
    }

    // lists, geometry, densities, diagonals, relaxation factors (the source terms follow after
    // the rigid prediction); nCBMax bounds the contact boundary particles (the RB entries)
    void launchContactSetup(hipStream_t s, const Dev& d, uint32_t nCBMax, uint32_t* tmp)
    {
        (void)hipMemsetAsync(d.cbScan, 0, sizeof(uint32_t) * (d.nB + 1ull), s);
        (void)hipMemsetAsync(d.brOff, 0, sizeof(uint32_t) * (nCBMax + 1ull), s);
        kContactMark<<<blocksN(d.nR), kBlock, 0, s>>>(d);
        launchScan(s, d.cbScan, d.nB + 1, tmp);
        if (d.nB) { kContactLocal<<<blocksN(d.nB), kBlock, 0, s>>>(d); }
        kContactBrCount<<<blocksN(d.nR), kBlock, 0, s>>>(d);
        launchScan(s, d.brOff, nCBMax + 1, tmp);
        (void)hipMemcpyAsync(d.brPos, d.brOff, sizeof(uint32_t) * (nCBMax + 1ull), hipMemcpyDeviceToDevice, s);
        kContactBrFill<<<blocksN(d.nR), kBlock, 0, s>>>(d);
        if (nCBMax) { kContactBrSort<<<blocksN(nCBMax), kBlock, 0, s>>>(d); }
        const uint64_t nGeo = std::max<uint64_t>(std::max<uint64_t>(d.nR, d.nRB), d.nScenes);
        kContactGeometry<<<blocksN(nGeo), kBlock, 0, s>>>(d);
        kContactDensity<<<blocksN(d.nR + static_cast<uint64_t>(nCBMax)), kBlock, 0, s>>>(d);
        if (d.nRB) { kContactRelaxRigid<<<d.nRB, kBlock, 0, s>>>(d); }
        if (nCBMax)
        {
            kContactRelaxBoundaryCount<<<blocksN(nCBMax), kBlock, 0, s>>>(d);
            kContactRelaxBoundary<<<blocksN(nCBMax), kBlock, 0, s>>>(d);
        }
    }

    void launchContactSource(hipStream_t s, const Dev& d, uint32_t nCBMax)
    {
        kContactError<<<blocksN(d.nR + static_cast<uint64_t>(nCBMax)), kBlock, 0, s>>>(d, true);
    }

    // one evaluation of the contact loads and residuals, the decision, and the Jacobi update of
    // the scenes that continue
    void launchContactIteration(hipStream_t s, const Dev& d, uint32_t nCBMax)
    {
        const uint32_t nAll = blocksN(d.nR + static_cast<uint64_t>(nCBMax));
        kContactForces<<<blocksN(d.nR), kBlock, 0, s>>>(d);
        if (d.nRB) { kContactLoads<<<d.nRB, kBlock, 0, s>>>(d); }
        kContactVelocities<<<blocksN(d.nR), kBlock, 0, s>>>(d);
        kContactError<<<nAll, kBlock, 0, s>>>(d, false);
        kContactFinalize<<<d.nScenes, kBlock, 0, s>>>(d);
        kContactJacobi<<<nAll, kBlock, 0, s>>>(d);
    }

    void launchContactApply(hipStream_t s, const Dev& d)
    {
        if (d.nRB) { kContactApply<<<blocksN(d.nRB), kBlock, 0, s>>>(d); }
    }
}
