/*
 * tsp_hip.h -- C ABI of the MI355X (gfx950) IISPH step: libtsp_hip.so.
 *
 * This is the thin layer between the C++ solver backend (SimBackend_HIP in
 * turbospharticles_amd/host/sim_backend_hip.h, with the methods of the reference's
 * Sim::I_SimBackend) and the hand-written HIP kernels. Every entry point
 * takes plain pointers and sizes; the context owns all device memory and one HIP stream,
 * and keeps the particle state resident on the device between calls.
 *
 * Each function names the reference interface it replaces (paths relative to the
 * reference repository). All functions return 0 (TSP_OK) or a negative tsp_status; the
 * message of the last failure is available from tsp_last_error(). A context is not
 * re-entrant: one caller thread at a time, like the reference backend
 * (core/simulation.h:179, one stream per backend at GPU/backend/gpu_backend.h:174).
 */
#ifndef TSP_HIP_H
#define TSP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tsp_status {
    TSP_OK = 0,
    TSP_ERR_INVALID = -1,   /* bad argument, null handle, size mismatch */
    TSP_ERR_HIP = -2,       /* a HIP runtime call failed (GPU::Check, GPU/gpu_utils.h:24-57) */
    TSP_ERR_STATE = -3,     /* call out of order (e.g. step before pull) */
    TSP_ERR_CAPACITY = -4,  /* a count exceeds a limit (u16 neighbor counts, 2^21 cells per frame
                               axis, the context's cell limit) */
    TSP_ERR_NUMERIC = -5    /* degenerate rigid orientation (cpu_backend.cpp:46-57) */
} tsp_status;

/* Fields of Sim::SimParams (core/parameters_sim.h:25-175) that the step reads. Derived values
   follow the reference's setters: mass = rest_density*h*h*h and kernel_radius = 2h, with
   kernel_radius_sq = kernel_radius*kernel_radius (core/parameters_sim.cpp:1478-1483). */
typedef struct tsp_params {
    float time_step;
    float particle_size;
    float kernel_radius;
    float kernel_radius_sq;
    float rest_density;
    float mass;
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
    uint32_t termination_cond; /* Base::SolverTerminationCond: 0 FixedIters, 1 Raw, 2 Pos, 3 Abs */
    float rigid_contact_density_dev_target;   /* rigidContactDensityDevTarget (default 0.001) */
    uint32_t rigid_contact_max_solver_iters;  /* rigidContactMaxSolverIters (default 100) */
} tsp_params;

/* SoA field order of sync/state_sim.h: FluidParticles :101-131, BoundaryParticles :272-285,
   RigidParticles :382-396, RigidBodies :497-522 (orientation and inverse inertia row-major) */
enum {
    TSP_F_X, TSP_F_Y, TSP_F_Z, TSP_F_VX, TSP_F_VY, TSP_F_VZ,
    TSP_F_VPX, TSP_F_VPY, TSP_F_VPZ, TSP_F_APX, TSP_F_APY, TSP_F_APZ,
    TSP_F_SOURCE, TSP_F_DIAG, TSP_F_PRESSURE, TSP_F_DENSITY, TSP_F_COUNT
};
enum { TSP_B_X, TSP_B_Y, TSP_B_Z, TSP_B_VOL, TSP_B_PSI, TSP_B_COUNT };
enum { TSP_R_XREL, TSP_R_YREL, TSP_R_ZREL, TSP_R_VOL, TSP_R_PSI, TSP_R_COUNT };
enum {
    TSP_RB_XCM, TSP_RB_YCM, TSP_RB_ZCM, TSP_RB_VX, TSP_RB_VY, TSP_RB_VZ,
    TSP_RB_WX, TSP_RB_WY, TSP_RB_WZ, TSP_RB_LX, TSP_RB_LY, TSP_RB_LZ,
    TSP_RB_A00, TSP_RB_A01, TSP_RB_A02, TSP_RB_A10, TSP_RB_A11, TSP_RB_A12,
    TSP_RB_A20, TSP_RB_A21, TSP_RB_A22, TSP_RB_MASS, TSP_RB_MASS_INV,
    TSP_RB_I00, TSP_RB_I01, TSP_RB_I02, TSP_RB_I10, TSP_RB_I11, TSP_RB_I12,
    TSP_RB_I20, TSP_RB_I21, TSP_RB_I22, TSP_RB_COUNT
};

/* Host view of a Sync::UniversalSimState (sync/state_sim.h:637): caller-owned arrays.
   tsp_pull reads every non-null array (a null fluid field starts at zero, a null id array
   takes sequential ids like ensureParticleIdsPresent, state_sim.h:219-233); tsp_push writes
   every non-null array. */
typedef struct tsp_state_view {
    uint32_t n_fluid, n_boundary, n_rigid, n_bodies;
    uint64_t *fluid_id;    float *fluid[TSP_F_COUNT];
    uint64_t *boundary_id; float *boundary[TSP_B_COUNT];
    uint64_t *rigid_id;    uint32_t *rigid_body; float *rigid[TSP_R_COUNT];
    uint32_t *body_id;     uint32_t *body_count; float *body[TSP_RB_COUNT];
    uint64_t step;
    double time;
} tsp_state_view;

/* flags of tsp_step_info.flags */
enum {
    TSP_FLAG_MAX_ITERS = 1u << 0,      /* the last solve stopped at max_solver_iters unconverged */
    TSP_FLAG_FRAME_REBUILT = 1u << 1,  /* the cell frame was rebuilt during the call */
    TSP_FLAG_RIGID_CONTACT = 1u << 2,  /* rigid particles had boundary neighbors or neighbors of
                                          other bodies: the rigid contact solve ran */
    TSP_FLAG_LIST_REGROWN = 1u << 3,   /* a neighbor list pool was regrown and the query re-run */
    TSP_FLAG_STEP_GRAPH = 1u << 4      /* a step of the call ran as one captured graph launch (no
                                          reference counterpart: the reference launches each kernel) */
};

typedef struct tsp_step_info {
    uint64_t step;                /* simStep_ */
    double time;                  /* simTime_ */
    uint32_t solver_iters;        /* N_solverIters_ of the last step (the init counts as one) */
    uint32_t flags;
    float avg_err_raw, avg_err_pos, avg_err_abs;   /* avgRelDensityErrorF_* */
    double err_sum_raw, err_sum_pos, err_sum_abs;  /* FP64 error sums of the last sweep */
    uint64_t sweeps_total;        /* Jacobi sweeps over all steps since the last pull */
    uint32_t frame_builds;
    int32_t frame_orig[3];
    uint32_t frame_dim[3];
    uint64_t n_ff, n_fb, n_fr, n_rf, n_rb, n_rr; /* neighbor list entries of the last search */
    uint64_t list_bytes;          /* bytes of the SELL pools as stored */
    float rigid_force_torque[6];  /* body 0: force and torque of the last step */
    /* rigid contact solve of the last step (RigidContactReport, CPU/backend/rigid_solver.h) */
    uint32_t contact_iters;       /* N_solverIters */
    float contact_avg_err;        /* avgDensityError */
    uint32_t contact_converged;
    uint32_t contact_boundary;    /* boundary particles in contact */
    uint64_t contacts;            /* N_contacts of the relaxation */
} tsp_step_info;

typedef struct tsp_ctx tsp_ctx;

/* SimBackend_GPU constructor (GPU/backend/gpu_backend.cpp:69-77): binds a device, creates the
   context's stream. */
int tsp_create(int device, tsp_ctx **out);
void tsp_destroy(tsp_ctx *ctx);
/* message of the last failed call on this context ("" if none) */
const char *tsp_last_error(const tsp_ctx *ctx);
/* I_SimBackend::Name (core/backend_sim.h:200) */
const char *tsp_name(void);

/* the SimParams snapshot passed to ExecuteNeighborSearch/ExecuteFullSimStep
   (core/backend_sim.h:249-252) */
int tsp_set_params(tsp_ctx *ctx, const tsp_params *params);
/* I_SimBackend::SetSimDomainBounds (core/backend_sim.h:204) */
int tsp_set_domain(tsp_ctx *ctx, const float mn[3], const float mx[3]);
/* the most cells the cell frames of the context's scenes may hold together (the dense cell
   tables take 24 bytes per cell); default and maximum 2^32 - 1. A layout beyond it (a pull, an
   admission or a frame rebuild) fails with TSP_ERR_CAPACITY and leaves the resident scenes as
   they were. No reference counterpart: the reference keeps compact tables of the occupied cells
   only (CPU/backend/neighbor_search.h:25-212); callers that budget device memory per context
   use it. */
int tsp_set_cell_limit(tsp_ctx *ctx, uint64_t max_cells);
/* steps as captured graphs (1) or launch by launch (0); the default is the build's
   TSP_TUNE_STEP_GRAPH. Both give the same bits. No reference counterpart (the reference launches
   each kernel from the host): for callers that compare the two forms or replay steps cheaply. */
int tsp_set_step_graph(tsp_ctx *ctx, int enable);
/* I_SimBackend::PullFromUniversalSimState (core/backend_sim.h:207; gpu_backend.cpp:96-240) */
int tsp_pull(tsp_ctx *ctx, const tsp_state_view *src);
/* I_SimBackend::PushToUniversalSimState (core/backend_sim.h:211; gpu_backend.cpp:242-344) */
int tsp_push(tsp_ctx *ctx, tsp_state_view *dst);
/* I_SimBackend::ExecuteNeighborSearch (core/backend_sim.h:249; gpu_backend.cpp:1227-1272):
   re-orders every fluid field, builds the cell tables and the SELL-32 lists */
int tsp_neighbor_search(tsp_ctx *ctx);
/* I_SimBackend::ExecuteFullSimStep (core/backend_sim.h:252; gpu_backend.cpp:1274-1594):
   advances n_steps IISPH steps, neighbor search included */
int tsp_step(tsp_ctx *ctx, uint32_t n_steps);
/* the solver status the reference keeps in SimBackend_CPU/GPU members (N_solverIters_,
   avgRelDensityErrorF_*, simStep_) */
int tsp_get_step_info(tsp_ctx *ctx, tsp_step_info *info);
/* I_SimBackend::GetNumFluidParticles/GetNumBoundaryParticles/GetNumRigidParticles
   (core/backend_sim.h:258-260); counts[3] is the rigid body count */
int tsp_get_counts(tsp_ctx *ctx, uint32_t counts[4]);

/* Neighbor lists of the last search, decoded from the SELL-32 pools (the host decoder of
   GPU/backend/gpu_backend.cpp:42-65). relation: 0 FF, 1 FB, 2 FR, 3 RF, 4 RB, 5 RR. counts has
   one entry per source particle; idx receives the concatenated lists in iteration order when
   cap is large enough; *total is the number of entries. */
int tsp_get_neighbors(tsp_ctx *ctx, int relation, uint32_t *counts, uint32_t *idx,
                      uint64_t cap, uint64_t *total);
/* Sorted cell ids and the permutation of the last fluid re-order (old index of each new
   index), and the boundary order and ids (NeighborSearch::TryGetLastParticleReOrder,
   CPU/backend/neighbor_search.h:52-54). Null pointers are skipped. */
int tsp_get_sort(tsp_ctx *ctx, uint32_t *order_f, uint64_t *keys_f, uint32_t *order_b,
                 uint64_t *keys_b);
/* world positions and velocities of the rigid particles (3 floats each), as updated by the
   last step (K_UpdateRigidParticles, GPU/backend/rigid_kernels.cu:181-219) */
int tsp_get_rigid_world(tsp_ctx *ctx, float *xyz, float *vel);

/* Per-kernel timing with HIP events on the context's stream (the per-op benchmark of
   eval/bench/benchmarks.h:31-38). While enabled, every launch is bracketed by events. */
int tsp_profile_enable(tsp_ctx *ctx, int on);
/* kernel names, launch counts and summed milliseconds since enabling; returns the number of
   kernels, fills up to cap entries. names is cap x 48 chars. */
int tsp_profile_read(tsp_ctx *ctx, char *names, uint64_t *launches, double *ms, int cap);
/* waits for all queued work of the context */
int tsp_synchronize(tsp_ctx *ctx);

/* Statistics of a scene after its last step, computed on the device: the StatMetric values of
   SimBackend_GPU::FillStatisticsRecord (GPU/backend/gpu_backend.cpp:1067-1208; names in
   eval/stats/metrics.h:23-88). Neighbor counts are per source particle, relations FF, FB, FR, RF,
   RB, RR; p99 is the count at sorted position (99 (n - 1)) / 100. */
typedef struct tsp_statistics {
    uint32_t n_fluid, n_boundary, n_rigid;
    uint32_t solver_iters;           /* NumSolverIters */
    double nbr_avg[6];               /* NumNbr*_Avg */
    uint32_t nbr_min[6];             /* NumNbr*_Min */
    uint32_t nbr_max[6];             /* NumNbr*_Max */
    uint32_t nbr_p99[6];             /* NumNbr*_P99 */
    uint32_t pad;
    double dens_rel_err_avg[3];      /* DensRelError_{Raw,Pos,Abs}_Avg: averages of rho/rho0 - 1 */
    double cfl;                      /* CFL_Number: max |v| dt / h */
    double vel_avg, vel_min, vel_max;/* VelMag_{Avg,Min,Max} */
} tsp_statistics;
/* I_SimBackend::FillStatisticsRecord (core/backend_sim.h:243) for the first scene */
int tsp_get_statistics(tsp_ctx *ctx, tsp_statistics *out);
/* tsp_get_statistics without any allocation. No counterpart in the reference: for callers that
   cannot allocate device memory (serving loops, code around HIP graph capture). The kernels take
   a static per-GPU scratch array of 3.3 MB (the library sets up its address at tsp_create) instead
   of the context's buffers; scenes of more than 2^24 fluid particles are refused with
   TSP_ERR_CAPACITY. Every call on one GPU shares that GPU's array: calls that overlap on one GPU
   (contexts on different streams, or threads) race. The results equal tsp_get_statistics'. */
int tsp_get_statistics_noalloc(tsp_ctx *ctx, tsp_statistics *out);

/* Device memory a context holds, by what it stores, in bytes (I_SimBackend::FillMemoryRecord,
   core/backend_sim.h:239, with the buckets of eval/memory/categories.h:59-78). */
typedef struct tsp_memory {
    uint64_t particles[3];      /* fluid, boundary, rigid particle (and body) arrays */
    uint64_t search[3];         /* sort keys and indices, cell tables: fluid, boundary, rigid */
    uint64_t lists[3];          /* fluid-source lists (SELL-32 pools, counts, slice metadata,
                                   halos); none for boundary sources; rigid-source lists */
    uint64_t misc;              /* scans, reductions, solver state, statistics and export scratch */
    uint64_t alloc_bytes;       /* all of the above */
    uint64_t peak_alloc_bytes;  /* the largest alloc_bytes since tsp_create */
} tsp_memory;
int tsp_get_memory(tsp_ctx *ctx, tsp_memory *out);

/* Sync::ExportSnapshot (sync/snapshots.h:37-66) as the caller's arrays. The fluid arrays hold
   the exported particles: those near the surface (IsNearSurfaceFluidParticle, core/backend_sim.h:
   109-117: a boundary or rigid neighbor, or fewer FF neighbors than the threshold) and their FF
   neighbors, in the backend's particle order, with their position triples and speed. */
typedef struct tsp_export_view {
    uint64_t step;                /* written: simStep */
    double time;                  /* written: simTime */
    uint32_t cap_fluid;           /* capacity of the fluid arrays, in particles */
    uint32_t n_fluid;             /* written: exported fluid particles */
    uint32_t n_boundary;          /* written: the scene's boundary count when boundary_xyz is set, else 0 */
    uint32_t n_rigid;             /* written */
    float *fluid_xyz;             /* 3 per exported fluid particle (x, y, z triples) */
    float *fluid_vel_mag;         /* |v| */
    float *boundary_xyz;          /* null: no boundary export (includeB false) */
    float *boundary_vol;
    uint8_t *boundary_is_mesh;    /* id >= mesh_id_begin */
    float *rigid_xyz;             /* world positions, 3 per rigid particle */
    float *rigid_vol;
    uint32_t *rigid_body;         /* the particle's body index */
} tsp_export_view;

/* I_SimBackend::PackExportSnapshot (core/backend_sim.h:215; gpu_backend.cpp:346-551) for the first
   scene: the surface classification and compaction run on the device against the lists of the
   last neighbor search (TSP_ERR_STATE when none ran since the last pull), and only the exported
   particles are copied out. surface_nbr_count_threshold is SimParams::surfaceNbrCountThreshold
   (27 by default); mesh_id_begin the first boundary mesh id (UINT64_MAX: none). Returns
   TSP_ERR_CAPACITY with n_fluid set when cap_fluid is too small. Null arrays are skipped. */
int tsp_export_snapshot(tsp_ctx *ctx, uint32_t surface_nbr_count_threshold, uint64_t mesh_id_begin,
                        tsp_export_view *out);

/*
 * Several independent scenes in one context. The reference runs one scene per backend
 * (core/simulation.h:179, one SimBackend per Simulation); these calls let one context hold many
 * scenes and advance all of them with one launch per kernel. Each scene keeps its own parameters,
 * domain, cell frame, solver state and counters; its results are bit-identical to the same scene
 * run alone in a context of its own. The single-scene calls above act on the first scene (tsp_pull
 * replaces every scene with one; tsp_set_params and tsp_set_domain apply to every scene).
 * Scene ids are assigned by the context and stay valid until the scene is removed.
 */
/* one relayout: removes n_remove scenes, then appends n_add new ones (params[q], *views[q],
   domains[6q..6q+5] = min xyz, max xyz, or domains null for none; a domain with min >= max on
   an axis is no domain, as in SetSimDomainBounds); out_ids receives the new ids. Resident scenes
   keep their device state. */
int tsp_batch_update(tsp_ctx *ctx, uint32_t n_remove, const uint32_t *remove_ids, uint32_t n_add,
                     const tsp_params *params, const tsp_state_view *const *views,
                     const float *domains, uint32_t *out_ids);
/* tsp_batch_update with one scene to add (domain: 6 floats, or null) */
int tsp_batch_add_scene(tsp_ctx *ctx, const tsp_params *params, const tsp_state_view *state,
                        const float domain[6], uint32_t *scene_id);
/* tsp_batch_update with one scene to remove */
int tsp_batch_remove_scene(tsp_ctx *ctx, uint32_t scene_id);
/* ids of the resident scenes in arena order (up to cap), *n = their number */
int tsp_batch_scene_ids(tsp_ctx *ctx, uint32_t *ids, uint32_t cap, uint32_t *n);
/* tsp_get_counts of one scene */
int tsp_batch_scene_counts(tsp_ctx *ctx, uint32_t scene_id, uint32_t counts[4]);
/* tsp_push of one scene */
int tsp_batch_push_scene(tsp_ctx *ctx, uint32_t scene_id, tsp_state_view *dst);
/* tsp_get_step_info of one scene (the neighbor-list totals count every scene) */
int tsp_batch_scene_info(tsp_ctx *ctx, uint32_t scene_id, tsp_step_info *info);
/* A scene described by a tsp_scene_spec (include/tsp_scene.h) generated on the device: the fluid
   block and the container's boundary layers with their psi / V sums are written straight into
   the context's arrays, bit for bit as tsp_scene_build followed by tsp_pull, without host
   arrays (the reference's scenario builders, scene/builder.cpp:174-642, :925-1087; the rigid body
   is generated on the host). tsp_pull_spec replaces the context's scenes with the spec's and sets
   the domain to the spec's (ComputeSimDomainBounds, scene/scenario_sim.cpp:123-148);
   tsp_batch_add_scene_spec adds a scene with its own params. */
typedef struct tsp_scene_spec tsp_scene_spec;
int tsp_pull_spec(tsp_ctx *ctx, const tsp_scene_spec *spec);
int tsp_batch_add_scene_spec(tsp_ctx *ctx, const tsp_params *params, const tsp_scene_spec *spec, uint32_t *scene_id);
/* tsp_batch_update with the new scenes given as specs and generated on the device (one relayout
   for the removals and all the additions) */
int tsp_batch_update_specs(tsp_ctx *ctx, uint32_t n_remove, const uint32_t *remove_ids, uint32_t n_add,
                           const tsp_params *params, const tsp_scene_spec *const *specs, uint32_t *out_ids);

/* tsp_get_statistics of one scene */
int tsp_batch_scene_statistics(tsp_ctx *ctx, uint32_t scene_id, tsp_statistics *out);
/* tsp_export_snapshot of one scene */
int tsp_batch_export_snapshot(tsp_ctx *ctx, uint32_t scene_id, uint32_t surface_nbr_count_threshold,
                              uint64_t mesh_id_begin, tsp_export_view *out);
/* advances every scene n_steps IISPH steps (ExecuteFullSimStep of each scene's backend) */
int tsp_batch_step(tsp_ctx *ctx, uint32_t n_steps);
/* device bytes a scene of these counts holds while resident (an upper estimate at the initial
   list pool sizes), for admission against an HBM budget */
uint64_t tsp_batch_footprint(uint32_t n_fluid, uint32_t n_boundary, uint32_t n_rigid,
                             uint32_t n_bodies);

#ifdef __cplusplus
}
#endif
#endif
