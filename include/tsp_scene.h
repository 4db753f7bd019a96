/*
 * tsp_scene.h -- host-side scene generation for the IISPH drop-in (C ABI).
 *
 * Generates the initial Sync::UniversalSimState SoA arrays that the reference's scenario
 * builders produce (source/simulation/scene/builder.cpp), but from lattice counts per axis
 * rather than from block lengths: the reference's FP32 TilesCount (builder.cpp:1511-1525)
 * turns a 0.4 block at spacing 0.004 into 99 particles per axis, so the configurations are
 * stated as counts. Everything runs on the host; the result is handed to tsp_pull.
 */
#ifndef TSP_SCENE_H
#define TSP_SCENE_H

#include <stdint.h>
#include "tsp_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tsp_scene_spec {
    float particle_size;        /* h: fluid and boundary spacing (particleSizeF = particleSizeB) */
    float rest_density;
    uint32_t boundary_layers;   /* boundaryLayers */
    float container_origin[3];  /* axis-aligned tank (BreakingDam container OBB) */
    float container_size[3];
    float fluid_origin[3];      /* corner of the fluid block */
    uint32_t fluid_counts[3];   /* lattice particles per axis */
    float jitter_frac;          /* jitterFracF: uniform +-frac*h per component */
    uint64_t jitter_seed;       /* jitterSeedF */
    uint32_t rigid_enable;      /* one hollow box body (BuildRigidBody_OBB_Hollow) */
    float rigid_center[3];
    float rigid_size[3];
    float rigid_density;
    float rigid_quat[4];        /* orientation (w, x, y, z), normalised by the builder */
    float rigid_vel[3];
    float rigid_omega[3];
} tsp_scene_spec;

typedef struct tsp_scene tsp_scene;

/* builds a scene; returns 0 or a negative tsp_status */
int tsp_scene_build(const tsp_scene_spec *spec, tsp_scene **out);
/* the counts a spec builds (fluid, boundary, rigid, bodies) without building it */
int tsp_scene_spec_counts(const tsp_scene_spec *spec, uint32_t counts[4]);
/* particle counts of a built scene (fluid, boundary, rigid, bodies) */
void tsp_scene_counts(const tsp_scene *s, uint32_t counts[4]);
/* simulation domain of the scene (ComputeSimDomainBounds, scene/scenario_sim.cpp:123-148) */
void tsp_scene_domain(const tsp_scene *s, float mn[3], float mx[3]);
/* copies the scene into caller-owned arrays; every non-null pointer of the view is written */
int tsp_scene_export(const tsp_scene *s, const tsp_state_view *dst);
void tsp_scene_free(tsp_scene *s);

/* uniform random unit quaternion from a seed (w, x, y, z) */
void tsp_scene_random_quat(uint64_t seed, float q[4]);

/* The reference's simulation state file, TSPSIM2 (eval/serial/state_format.h:10-120,
   serialization.cpp:164-244, 860-935: SaveUniversalSimState / LoadUniversalSimState): a 64-byte
   header, then the SoA arrays of the fluid, boundary, rigid and body populations in the field
   order of tsp_state_view. */
/* counts (fluid, boundary, rigid, bodies), step and time of a state file; checks the header and
   the file length as the reference's reader does */
int tsp_state_file_info(const char *path, uint32_t counts[4], uint64_t *step, double *time);
/* reads a state file into caller-owned arrays sized by tsp_state_file_info (null arrays are
   skipped); sets the view's step and time */
int tsp_state_file_read(const char *path, tsp_state_view *dst);
/* writes a state file from a view (null arrays are written as zeros) through a ".tmp" file
   renamed over the destination; a view with the invalid id (UINT64_MAX), a rigid particle of a
   missing body or body counts that disagree with the memberships is refused and nothing is
   written. tsp_state_file_read refuses the same contents. */
int tsp_state_file_write(const char *path, const tsp_state_view *src);

#ifdef __cplusplus
}
#endif
#endif
