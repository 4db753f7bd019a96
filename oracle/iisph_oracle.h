/*
 * iisph_oracle.h -- CPU restatement of TurboSPHarticles' TBB/xsimd IISPH step.
 *
 * TEST INFRASTRUCTURE ONLY. This oracle is the checker that the HIP path is compared
 * against; nothing in the product links, loads or calls it. Only tests/, the smoke check
 * and the cpu_baseline leg of bench.py may use it.
 *
 * It restates, in plain C, the algorithm of the reference CPU backend:
 *   - source/simulation/CPU/backend/sph_solver.cpp:34-319 (Step_IISPH), :686-732
 *     (TerminationConditionSatisfied), :734-2089 (the per-chunk SELL kernels)
 *   - source/simulation/CPU/simd.h:279-435 (cubic-spline weight / gradient factors)
 *   - source/simulation/CPU/backend/neighbor_search.cpp:157-564 (floor, cell tables, stable
 *     LSD radix sort), :613-1612 (Construct, chunk query ranges, count/fill queries),
 *     :1624-1848 (particle re-order), :1902-2070 (cell frame with hysteresis, Morton ids)
 *   - source/simulation/CPU/backend/neighbor_lists.{h,cpp} (dict chunks :311-363, the
 *     DictDelta16x2 low/high split :838-858 and its reversed high group, neighbor_lists.h:171-197)
 *   - source/simulation/CPU/backend/cpu_backend.cpp:2001-2303 (rigid prediction, update,
 *     reduction to bodies, integration with Gram-Schmidt)
 *   - source/simulation/core/solver_constants.h:20-36, core/cell_frame.h:32-66
 *
 * The SIMD chunk width of the reference (8 lanes on AVX2, 16 on AVX-512) changes which
 * fluid chunks become "dict" chunks and how a chunk gathers its candidate cells, and with
 * that the FP32 summation formula and order. The oracle takes the chunk width as a
 * parameter (1, 8 or 16). Width 1 is the per-particle form: every chunk is a single
 * particle, so every FF list is a dict list in z-y-x stencil raster order.
 *
 * Two builds of the same source exist: *_f32 (FP32 arithmetic exactly as the reference
 * writes it, explicit fma where the reference calls fma, no contraction elsewhere) and
 * *_f64 (the same algorithm with every value in double: the FP64 oracle).
 *
 * It also restates:
 *   - source/simulation/CPU/backend/rigid_solver.cpp:23-799 (rigid-boundary and rigid-rigid
 *     contact IISPH: contact lists, geometry, density and diagonal, relaxation, pressure
 *     forces, velocities, density error, relaxed Jacobi, SolveRigidContacts)
 *   - source/simulation/core/backend_sim.h:109-117 and CPU/backend/cpu_backend.cpp:829-981
 *     (IsNearSurfaceFluidParticle and the export snapshot's near-surface particles with their
 *     FF neighbors)
 *
 * Parity unpinned (the reference cannot be built in this container, so no test can pin these
 * against it):
 *   - TBB's parallel_reduce join order for the FP64 Jacobi error sums and for the rigid force
 *     sums is run-dependent in the reference; the oracle sums serially in particle order (the
 *     rigid force sums in FP64 unless set_rigid_sums_f64(0)). The same holds for the rigid
 *     contact solver's FP64 error sums and its contact force sums.
 *
 * The ordering parts are pinned by tests/test_oracle_order.py against the reference's rules
 * restated there independently (cell_frame.cpp known answers, frame origin and hysteresis,
 * stable sort of interleaved Morton ids, z-y-x raster list order, the DictDelta16x2 split).
 */
#ifndef TSP_IISPH_ORACLE_H
#define TSP_IISPH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* the fields of Sim::SimParams (core/parameters_sim.h:25-175) that the step reads */
typedef struct orc_params {
    float time_step;
    float particle_size;       /* h, particleSizeF */
    float kernel_radius;       /* 2h */
    float kernel_radius_sq;    /* kernelRadius * kernelRadius */
    float rest_density;
    float mass;                /* massF = rho0 h^3 */
    float viscosity_fluid;
    float viscosity_boundary;
    float gravity[3];
    float press_gamma;
    float jacobi_omega;
    float jacobi_warm_start;
    float density_dev_target;
    float grid_cell_size_ratio;
    uint32_t fixed_iters;
    uint32_t min_solver_iters;
    uint32_t max_solver_iters;
    uint32_t termination_cond;  /* 0 FixedIters, 1 Raw, 2 Pos, 3 Abs (base/enum_hub.h:212-218) */
    float rigid_contact_density_dev_target;   /* rigidContactDensityDevTarget (0.001) */
    uint32_t rigid_contact_max_solver_iters;  /* rigidContactMaxSolverIters (100) */
} orc_params;

/* field order of the SoA structs in sync/state_sim.h:101-131, :272-285, :382-396, :497-522 */
enum {
    ORC_F_X, ORC_F_Y, ORC_F_Z, ORC_F_VX, ORC_F_VY, ORC_F_VZ,
    ORC_F_VPX, ORC_F_VPY, ORC_F_VPZ, ORC_F_APX, ORC_F_APY, ORC_F_APZ,
    ORC_F_SOURCE, ORC_F_DIAG, ORC_F_PRESSURE, ORC_F_DENSITY, ORC_F_COUNT
};
enum { ORC_B_X, ORC_B_Y, ORC_B_Z, ORC_B_VOL, ORC_B_PSI, ORC_B_COUNT };
enum { ORC_R_XREL, ORC_R_YREL, ORC_R_ZREL, ORC_R_VOL, ORC_R_PSI, ORC_R_COUNT };
enum {
    ORC_RB_XCM, ORC_RB_YCM, ORC_RB_ZCM, ORC_RB_VX, ORC_RB_VY, ORC_RB_VZ,
    ORC_RB_WX, ORC_RB_WY, ORC_RB_WZ, ORC_RB_LX, ORC_RB_LY, ORC_RB_LZ,
    ORC_RB_A00, ORC_RB_A01, ORC_RB_A02, ORC_RB_A10, ORC_RB_A11, ORC_RB_A12,
    ORC_RB_A20, ORC_RB_A21, ORC_RB_A22, ORC_RB_MASS, ORC_RB_MASS_INV,
    ORC_RB_I00, ORC_RB_I01, ORC_RB_I02, ORC_RB_I10, ORC_RB_I11, ORC_RB_I12,
    ORC_RB_I20, ORC_RB_I21, ORC_RB_I22, ORC_RB_COUNT
};

/* host SoA views, owned by the caller (pull copies in, push copies out) */
typedef struct orc_state {
    uint32_t n_fluid, n_boundary, n_rigid, n_bodies;
    uint64_t *fluid_id;    float *fluid[ORC_F_COUNT];
    uint64_t *boundary_id; float *boundary[ORC_B_COUNT];
    uint64_t *rigid_id;    uint32_t *rigid_body; float *rigid[ORC_R_COUNT];
    uint32_t *body_id;     uint32_t *body_count; float *body[ORC_RB_COUNT];
} orc_state;

typedef struct orc_info {
    uint32_t solver_iters;          /* N_solverIters_ (the init counts as the first) */
    float avg_err_raw, avg_err_pos, avg_err_abs;
    double err_sum_raw, err_sum_pos, err_sum_abs;
    int32_t frame_orig[3];
    uint32_t frame_dim[3];
    uint32_t frame_builds;
    uint32_t n_dict_chunks, n_chunks;
    uint64_t n_ff, n_fb, n_fr, n_rf, n_rb, n_rr;  /* total list entries per relation */
    uint32_t contact_pairs;         /* RB + RR entries of the last search */
    /* RigidContactReport of the last step (rigid_solver.h) */
    uint32_t contact_iters;         /* N_solverIters */
    float contact_avg_err;          /* avgDensityError */
    uint32_t contact_converged;
    uint32_t contact_boundary;      /* boundary particles in contact (localIdxsB.size()) */
    uint64_t contacts;              /* N_contacts of the relaxation */
} orc_info;

/* relations for orc*_get_list */
enum { ORC_FF, ORC_FB, ORC_FR, ORC_RF, ORC_RB, ORC_RR };

#define ORC_DECLARE(P)                                                                   \
    typedef struct P##_ctx P##_ctx;                                                      \
    P##_ctx *P##_create(void);                                                           \
    void P##_destroy(P##_ctx *c);                                                        \
    void P##_set_params(P##_ctx *c, const orc_params *p);                                \
    void P##_set_domain(P##_ctx *c, const float mn[3], const float mx[3]);               \
    void P##_set_chunk_width(P##_ctx *c, uint32_t w);                                    \
    void P##_set_rigid_sums_f64(P##_ctx *c, int on);                                     \
    void P##_set_threads(P##_ctx *c, int n);                                             \
    /* the cubic-spline alpha the kernels take (0: h^-3/(4 pi) like the solver); tests     \
       of the single kernels pass the reference tests' alpha */                          \
    void P##_set_kernel_alpha(P##_ctx *c, float alpha);                                  \
    int P##_pull(P##_ctx *c, const orc_state *s);                                        \
    /* moves the fluid particles (current order) and keeps the cell frame (tests) */      \
    void P##_set_fluid_positions(P##_ctx *c, const float *x, const float *y,             \
                                 const float *z);                                        \
    int P##_push(P##_ctx *c, orc_state *s);                                              \
    int P##_neighbor_search(P##_ctx *c);                                                 \
    int P##_step(P##_ctx *c);                                                            \
    void P##_get_info(P##_ctx *c, orc_info *info);                                       \
    uint64_t P##_get_list(P##_ctx *c, int rel, uint32_t *counts, uint32_t *idx,          \
                          uint8_t *dict, uint64_t cap);                                  \
    void P##_get_order(P##_ctx *c, uint32_t *order_f, uint64_t *keys_f,                  \
                       uint32_t *order_b, uint64_t *keys_b);                             \
    void P##_get_rigid_world(P##_ctx *c, float *xyz, float *vel);                        \
    void P##_get_body_force(P##_ctx *c, float *ft);                                      \
    /* the contact solver's pieces (tests): after neighbor_search, sets up the contact      \
       problem and, with the given contact pressures (null: zero), evaluates the body loads \
       and the residuals once. Returns the count of contact boundary particles, or < 0.     \
       Every output may be null; rigid arrays have n_rigid entries, boundary arrays the     \
       returned count, loads 6 per body. */                                                 \
    /* the fluid particles the export snapshot carries (FillFluidParticleNearSurfaceFlags and \
       ClassifyExportSurfaceParticles, cpu_backend.cpp:910-981): flags[i] = 1 in the current   \
       particle order, from the lists of the last search */                                   \
    void P##_export_surface(P##_ctx *c, uint32_t threshold, uint8_t *flags);               \
    int64_t P##_contact_probe(P##_ctx *c, const float *p_rigid, const float *p_boundary,   \
                              float *rho_r, float *diag_r, float *omega_r, float *err_r,    \
                              float *rho_b, float *diag_b, float *omega_b, float *err_b,    \
                              uint32_t *local_b, float *loads);                             \
    /* single kernels over explicit per-particle lists (tests of closed forms) */        \
    void P##_kernel_weight(const orc_params *p, const float *dxyz, uint32_t n,           \
                           float *w_out, float *grad_out, float *w_unscaled_out);        \
    int P##_kernels_from_lists(P##_ctx *c, const uint32_t *ff_off, const uint32_t *ff,   \
                               const uint8_t *ff_dict, const uint32_t *fb_off,           \
                               const uint32_t *fb, const uint32_t *fr_off,               \
                               const uint32_t *fr, int stage);

ORC_DECLARE(orc32)
ORC_DECLARE(orc64)

/* stages for orc*_kernels_from_lists (each runs one reference kernel over the lists) */
enum {
    ORC_STAGE_DENSITY = 1, ORC_STAGE_NONPRESSURE = 2, ORC_STAGE_SOURCE_DIAG = 3,
    ORC_STAGE_PRESSURE_ACC = 4, ORC_STAGE_JACOBI = 5, ORC_STAGE_INTEGRATION = 6
};

/* SplitMix64-style jitter of scene/builder.cpp:1486-1501 */
float orc_jitter3d(uint64_t base_seed, uint64_t key, float amp);
/* TryBuildCellCoordFrame of core/cell_frame.h:32-66 (1: built; 0: rejected, frame untouched) */
int orc_try_build_cell_frame(const int64_t cell_min[3], const int64_t cell_max[3], int32_t orig[3],
                             uint32_t dim[3]);
/* Morton spread of neighbor_search.cpp:2057-2070 */
uint64_t orc_spread_bits_21to63(uint32_t v);

#ifdef __cplusplus
}
#endif
#endif
