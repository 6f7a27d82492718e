/*
 * cn_hip.h -- C ABI of the MI355X (gfx950) particle-substep library libcn_hip.so.
 *
 * The library replaces the particle/fluid/soft compute passes that CRESSim-Neo issues through
 * PhysicsPassDispatcher (src/physics/physics_pass_dispatcher.h:24-342) and the HLSL kernels they
 * bind (assets/shaders/physics/{particle,fluid,soft,shared}/...).  The substep loop that
 * PhysicsSolver::step runs (src/physics/physics_solver.cpp:336-1401) is cn_step().
 *
 * Conventions
 *   - Every function returns cn_status; CN_OK == 0.  The message of the last failure on a
 *     context is cn_last_error().  No exception crosses this boundary.
 *   - All structs are plain data.  Host pointers passed in are borrowed for the call only.
 *   - One context owns one HIP stream and all device buffers of its scene(s).  A context is not
 *     re-entrant; different contexts may be used from different threads.
 *   - Particle arrays are indexed by the caller's particle index (the reference's pool index).
 *     float4 arrays are 16-byte (x, y, z, w) records.
 */
#ifndef CN_HIP_H
#define CN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CN_ABI_VERSION 1u

typedef enum cn_status {
    CN_OK = 0,
    CN_ERR_INVALID_ARGUMENT = 1,
    CN_ERR_OUT_OF_MEMORY = 2,
    CN_ERR_DEVICE = 3,
    CN_ERR_NOT_READY = 4,
    CN_ERR_CAPACITY = 5,   /* a capacity (neighbor cap, candidate capacity, slot budget) overflowed */
    CN_ERR_UNSUPPORTED = 6
} cn_status;

/* Particle kinds (assets/shaders/include/physics/particle/physics_particle_types.hlsli:9-10). */
#define CN_PARTICLE_KIND_SOLID 0u
#define CN_PARTICLE_KIND_FLUID 1u

/* Broadphase key used to order particles inside a substep.
 *   CN_KEY_HASH    the reference cell hash (physics_particle_grid.hlsli:4-10), 32-bit keys.
 *   CN_KEY_MORTON  Morton code of the wrapped cell coordinate (10 bits per axis).
 * Both produce bit-identical neighbor lists and solver results (a cell's particles are visited in
 * ascending particle index and the 27 cells in dz,dy,dx order either way); Morton keeps
 * spatially adjacent cells adjacent in memory. */
#define CN_KEY_MORTON 0u
#define CN_KEY_HASH 1u

/* GpuFluidMaterial (physics_particle_types.hlsli:115-129), 48 bytes. */
typedef struct cn_fluid_material {
    float rest_density;
    float inv_rest_density;
    float smoothing_radius;
    float density_constraint_scale;
    float viscosity;
    float cohesion;
    float cohesion1;
    float cohesion2;
    float surface_tension;
    float vorticity_confinement;
    float gravity_scale;
    float cfl_radius;
} cn_fluid_material;

/* FluidMaterialDesc (include/physics/physics_types.h:333-343) without the contact block. */
typedef struct cn_fluid_material_desc {
    float viscosity;
    float cohesion;
    float surface_tension;
    float vorticity_confinement;
    float gravity_scale;
    float cfl_coefficient;
} cn_fluid_material_desc;

/* Static axis-aligned box collider.  It is the reference's static rigid body with identity
 * orientation, unit scale and a box collider at local position 0 (GpuColliderGeometryData,
 * physics_fluid_boundary_common.hlsli:10-66 with an identity quaternion).  material is the
 * collider material (friction, restitution, 0, staticFriction) of physics_world.cpp:85-88. */
typedef struct cn_box_collider {
    float center[3];
    float half_extents[3];
    float material[4];
    uint32_t environment;
    uint32_t layer;
    uint32_t mask;
    uint32_t enabled;
    uint32_t reserved[2];
} cn_box_collider;

/* A static collider of any shape: the reference's static rigid body (position, orientation,
 * scale) carrying one collider (shape, local position and orientation, shape parameters), as
 * ComputeFluidBoundaryContactInfo reads them (physics_fluid_boundary_common.hlsli:10-84 with
 * physics_shape_common.hlsli and physics_math.hlsli).  shape_params: sphere (radius), box (half
 * extents), capsule (radius, half height along the local y axis); quaternions are (x, y, z, w) and
 * are normalized like the reference's.  material as for cn_box_collider. */
#define CN_SHAPE_SPHERE 0u
#define CN_SHAPE_BOX 1u
#define CN_SHAPE_CAPSULE 2u
typedef struct cn_collider {
    uint32_t shape;
    uint32_t environment;
    uint32_t layer;
    uint32_t mask;
    float body_position[4];     /* xyz */
    float body_orientation[4];
    float body_scale[4];        /* xyz */
    float local_position[4];    /* xyz */
    float local_orientation[4];
    float shape_params[4];
    float material[4];
    uint32_t enabled;
    uint32_t reserved[3];
} cn_collider;

/* Host view of the unified particle pool (ParticleSoAHost, include/physics/physics_types.h:858). */
typedef struct cn_particle_soa_view {
    uint32_t count;
    const float *positions_inv_mass;    /* float4[count] */
    const float *velocities;            /* float4[count] */
    const float *radii;                 /* float[count] */
    const uint32_t *kinds;              /* CN_PARTICLE_KIND_* */
    const uint32_t *fluid_material_indices;
    const uint32_t *particle_material_indices;
    const uint32_t *environment_indices;
    const uint32_t *phases;
    const uint32_t *collision_layers;
    const uint32_t *collision_masks;
    /* Constraint adjacency (ParticleSoA adjacencyOffsets / adjacencyCounts / adjacencyIndices,
     * physics_world.cpp:5774-5787): particle i's adjacent particles are
     * adjacency_indices[adjacency_offsets[i] .. adjacency_offsets[i] + adjacency_counts[i]).
     * IsAdjacentParticle (physics_particle_broad_phase_query.hlsli:50-62) keeps them out of the
     * contact pairs of a self-colliding phase group.  NULL offsets: no particle has any. */
    const uint32_t *adjacency_offsets;
    const uint32_t *adjacency_counts;
    uint32_t adjacency_index_count;
    const uint32_t *adjacency_indices;
} cn_particle_soa_view;

/* Scene tables uploaded beside the particles. */
typedef struct cn_scene_tables {
    uint32_t fluid_material_count;
    const cn_fluid_material *fluid_materials;
    uint32_t particle_material_count;
    const float *particle_materials;    /* float4 (friction, restitution, damping, staticFriction) */
    uint32_t collider_count;
    const cn_box_collider *colliders;
    /* static colliders of any shape (spheres, oriented boxes, capsules), after the boxes: collider
     * k >= collider_count is shape_colliders[k - collider_count]; at most 65535 in all.  A particle
     * holds at most 32 candidate colliders at a time (more: cn_validate_meta reports
     * contact_overflows and returns CN_ERR_CAPACITY). */
    uint32_t shape_collider_count;
    const cn_collider *shape_colliders;
} cn_scene_tables;

/* Soft/cloth constraints (physics_particle_types.hlsli:147-205 layouts). */
typedef struct cn_soft_edge {
    uint32_t a, b;
    float rest_length;
    float compliance;
    uint32_t enabled;
    uint32_t reserved[3];
} cn_soft_edge;

typedef struct cn_soft_bend {
    uint32_t p0, p1, p2;
    float rest_angle;
    float compliance;
    uint32_t reserved[3];
} cn_soft_bend;

typedef struct cn_cloth_dihedral {
    uint32_t e0, e1, o0, o1;
    float rest_angle;
    float compliance;
    uint32_t reserved[2];
} cn_cloth_dihedral;

typedef struct cn_soft_tet {
    uint32_t p[4];
    float rest_volume;
    float compliance;
    uint32_t reserved[2];
} cn_soft_tet;

typedef struct cn_constraint_view {
    uint32_t edge_count;
    const cn_soft_edge *edges;
    uint32_t bend_count;
    const cn_soft_bend *bends;
    uint32_t dihedral_count;
    const cn_cloth_dihedral *dihedrals;
    uint32_t tet_count;
    const cn_soft_tet *tets;
} cn_constraint_view;

/* GpuPhysicsSolverConfig (assets/shaders/include/physics/physics_solver_config.hlsli:5-51),
 * 10 float4 records: contact0, contact1, rigid0, rigid1, joints0, joints1, soft0, soft1,
 * fluid0, fluid1.  cn_default_step_params() fills the PhysicsSolverDesc defaults. */
#define CN_SOLVER_CONFIG_FLOATS 40

typedef struct cn_step_params {
    float dt;                        /* frame delta; substep dt = dt / substeps */
    uint32_t substeps;
    uint32_t default_iterations;
    uint32_t fluid_iterations;       /* 0 -> default_iterations */
    uint32_t soft_internal_iterations;
    uint32_t soft_contact_iterations;
    float gravity[4];
    float solver_config[CN_SOLVER_CONFIG_FLOATS];
    float cell_size;                 /* particle grid cell size; 0 -> derived like the reference */
    uint32_t max_fluid_neighbors;    /* neighbor slots per fluid particle (reference: 128) */
    uint32_t key_mode;               /* CN_KEY_MORTON or CN_KEY_HASH */
    uint32_t flags;                  /* CN_STEP_* bits (cn_default_step_params: CN_STEP_CANDIDATE_REUSE) */
} cn_step_params;

/* flags: in fluid-only scenes the fluid iterations after a substep's first take the neighbors from
 * the candidate lists its first iteration builds (neighbors within 1.05 h of the positions at the
 * first iteration) instead of walking the 27 cells again, in every cell block whose particles'
 * displacement since the first iteration plus the largest one in the block's halo stays below
 * 0.05 h; the other blocks walk.  The lists are the same either way (same neighbors, same order),
 * so the results are too.  The candidate lists take 198 bytes per particle, allocated on the first
 * step that asks (left off when they would exceed 4 GiB less 1 MiB).  On in
 * cn_default_step_params. */
#define CN_STEP_CANDIDATE_REUSE 1u

/* Overflow and diagnostic counters of the last cn_step (validateGpuMetaBlocking analog,
 * physics_solver.cpp:1403-1492), plus the fluid-neighbor truncation the reference lacks. */
typedef struct cn_meta {
    uint32_t max_fluid_neighbor_count;    /* largest required (untruncated) in-radius count */
    uint32_t fluid_neighbor_truncations;  /* particles whose list hit max_fluid_neighbors */
    uint32_t max_cell_occupancy;          /* longest run of equal keys over the substeps of the last
                                             cn_step */
    uint32_t contact_overflows;           /* particles with more contact candidates than a list holds
                                             (32), or, in a scene of more than 32 colliders, more
                                             colliders in reach than a slot's collider list holds
                                             (32), summed over the substeps; the surplus ones
                                             are dropped */
    uint64_t fluid_neighbor_total;        /* listed neighbors at iteration 0, summed over the
                                             substeps of the last cn_step (mean K = total /
                                             (particles * substeps)) */
    uint32_t last_particle_count;
    uint32_t lds_tiles;                   /* density-pass cell blocks whose halo was staged in LDS */
    uint32_t run_table_tiles;             /* density-pass cell blocks walked through the run table */
    uint32_t max_displacement2_bits;      /* float bits of the largest (|delta| + slack)^2 */
    uint32_t candidate_reuses;            /* fluid particle iterations that filtered their iteration-0
                                             candidate list instead of walking the 27 cells */
    uint32_t reserved;
} cn_meta;

/* Raw device pointers of the resident scene (PhysicsGpuSceneView analog). */
typedef struct cn_device_views {
    void *positions_inv_mass;   /* float4[count], particle-index order after cn_step */
    void *velocities;           /* float4[count] */
    uint32_t count;
    void *stream;               /* hipStream_t */
} cn_device_views;

typedef struct cn_ctx cn_ctx;

typedef struct cn_device_desc {
    int device_ordinal;
    uint32_t flags;             /* reserved */
} cn_device_desc;

/* -------- context -------- */
uint32_t cn_abi_version(void);
cn_status cn_ctx_create(const cn_device_desc *desc, cn_ctx **out_ctx);
void cn_ctx_destroy(cn_ctx *ctx);
const char *cn_last_error(const cn_ctx *ctx);

/* -------- host authoring helpers (no device work) -------- */
/* PhysicsSolverDesc defaults (include/physics/physics_solver.h:21-143). */
void cn_default_step_params(cn_step_params *params);
/* toFluidSolverMaterial (src/physics/physics_world.cpp:154-263). */
cn_status cn_fluid_material_from_desc(const cn_fluid_material_desc *desc, float particle_radius,
                                      cn_fluid_material *out);
/* recomputeParticleGridCellSize (src/physics/physics_world.cpp:6829-6846). */
float cn_particle_grid_cell_size(uint32_t count, const float *radii,
                                 const uint32_t *fluid_material_indices,
                                 uint32_t fluid_material_count,
                                 const cn_fluid_material *fluid_materials);

/* -------- scene state (PhysicsSceneGpuState analog) -------- */
/* ensureCapacity (physics_scene_gpu_state.cpp:305).  CN_ERR_CAPACITY when the neighbor lists
   (particle_capacity x max_fluid_neighbors 16-bit entries) would exceed 4 GiB less 1 MiB. */
cn_status cn_scene_reserve(cn_ctx *ctx, uint32_t particle_capacity, uint32_t max_fluid_neighbors);
/* uploadWorldState / uploadParticles / uploadSoftTopology (:2348, :3163, :3255). */
cn_status cn_scene_upload(cn_ctx *ctx, const cn_particle_soa_view *particles,
                          const cn_scene_tables *tables, const cn_constraint_view *constraints);
/* PhysicsSolver::step (physics_solver.cpp:336-1401) for the particle path; asynchronous on the
 * context stream. */
cn_status cn_step(cn_ctx *ctx, const cn_step_params *params);
/* -------- scene pool (batched path) --------
 * One independent scene: its particle view, tables and constraints (indices local to the scene). */
typedef struct cn_scene_desc {
    const cn_particle_soa_view *particles;
    const cn_scene_tables *tables;
    const cn_constraint_view *constraints;   /* may be NULL */
} cn_scene_desc;
/* Uploads scene_count independent scenes into one pool (no reference counterpart: the reference
 * steps one world per solver; its batched RL environments loop worlds or co-locate them by
 * environment index, docs/source/guides/batched-environments.md).  Particles of scene k occupy
 * pool indices [first_k, first_k + count_k) in upload order; every scene keeps its own cell size,
 * static colliders and key range, so each scene's results equal those of uploading it alone.
 * cn_step then steps every scene with one launch per pass; cn_readback and cn_device_views_get
 * address the whole pool.  Requires CN_KEY_MORTON when scene_count > 1. */
cn_status cn_batch_upload(cn_ctx *ctx, uint32_t scene_count, const cn_scene_desc *scenes);
/* cn_step for the listed scenes of the pool only: every other scene's state is left bit for bit
 * untouched (serving: one launch per pass for the scenes that have a pending step request).  No
 * reference counterpart. */
cn_status cn_step_scenes(cn_ctx *ctx, const cn_step_params *params, uint32_t count, const uint32_t *scenes);
/* No reference counterpart.  When the scenes listed to cn_step_scenes hold at most compact_below
 * of the pool's particles (default 0.5; 0 never), the step runs on their slots alone: their
 * segments of the pool's slot order are gathered into a compact range, stepped as a pool of those
 * scenes, and put back, so its cost follows the listed particles instead of the pool's.  Results
 * are the same either way.  Such a step is not captured into a graph. */
cn_status cn_tick_policy(cn_ctx *ctx, float compact_below);
/* No reference counterpart.  enable != 0 (the default): in a substep with no particle-particle
 * contact pass between them, the soft-constraint phase (and the particle-rigid contacts solved in
 * it) runs on a second HIP stream of the context beside the fluid passes, joined before the
 * velocity update; they move disjoint particles, so the results are the same either way.
 * enable == 0 runs every pass on the context stream. */
cn_status cn_concurrent_phases(cn_ctx *ctx, int enable);
/* Pool index range of one uploaded scene. */
cn_status cn_batch_scene_range(const cn_ctx *ctx, uint32_t scene, uint32_t *first, uint32_t *count);
/* enable != 0: cn_step captures its kernel sequence into a HIP graph on the first call with given
 * parameters and replays it afterwards (no reference counterpart: the reference records Vulkan
 * command lists per step).  Uploads and reserves drop the cached graphs.  Off by default. */
cn_status cn_graph_enable(cn_ctx *ctx, int enable);

/* Launch shape of the fluid passes: halo particles a workgroup stages in LDS and workgroups of a
 * persistent launch.  Results do not depend on it. */
typedef struct cn_tune_result {
    uint32_t tile_max;            /* chosen halo capacity, particles (256..4096) */
    uint32_t grid_max;            /* chosen workgroups of a persistent launch */
    uint32_t candidates;          /* shapes timed */
    uint32_t reserved;
    float ms_per_step;            /* cn_step time under the chosen shape */
    float default_ms_per_step;    /* under the default shape (4096, 1024) */
} cn_tune_result;
/* Times `steps` cn_step calls of the uploaded scene (or pool) under each candidate shape, starting
 * every time from the current state, keeps the fastest and leaves the state as it found it (no
 * reference counterpart: the reference's dispatch sizes are fixed in its shaders).  A larger
 * tile_max within 1 % of the fastest is preferred: it walks no more dense blocks through the run
 * table when the fluid later compresses.  Allocates a copy of the particle state while it runs. */
cn_status cn_autotune(cn_ctx *ctx, const cn_step_params *params, uint32_t steps, cn_tune_result *out);
/* Sets the launch shape directly (tile_max in [256, 4096], grid_max in [8, 8192]). */
cn_status cn_tune_set(cn_ctx *ctx, uint32_t tile_max, uint32_t grid_max);

/* Blocking readback (physics_scene_gpu_state.cpp:4082-4142). Any pointer may be NULL. */
cn_status cn_readback(cn_ctx *ctx, float *positions_inv_mass, float *previous_positions,
                      float *velocities, uint32_t count);
cn_status cn_device_views_get(cn_ctx *ctx, cn_device_views *out);
/* Overwrites the particle state of scene `scene` (0 for a single uploaded scene): positions with
   inverse masses and velocities, float4 per particle in the scene's particle order (cn_readback's
   order within cn_batch_scene_range); the previous positions become the new positions.  The other
   scenes, the topology, the materials and the colliders stay, so an environment reset in a pool
   needs no re-upload (the reference re-uploads its whole particle set when the world's particle
   revision changes, physics_scene_gpu_state.cpp:2393-2486).  flags: CN_STATE_DEVICE_POINTERS when
   both arrays are device memory, read on the context's stream (e.g. torch tensors ordered before
   it); otherwise host memory, copied before the call returns. */
#define CN_STATE_DEVICE_POINTERS 1u
cn_status cn_scene_set_state(cn_ctx *ctx, uint32_t scene, const float *positions_inv_mass,
                             const float *velocities, uint32_t flags);
/* Dirty-run upload of particle state (uploadContiguousRuns, physics_scene_gpu_state.cpp:255-284):
   overwrites positions with inverse masses and velocities of the uploaded particles
   [first, first + count) in cn_readback's particle order (a pool's particle index space), leaving
   every other particle, the topology, materials and colliders resident; the previous positions of
   the run become its new positions.  A caller that tracks dirty particle indices sends one call per
   contiguous run.  flags as for cn_scene_set_state. */
cn_status cn_particles_update_range(cn_ctx *ctx, uint32_t first, uint32_t count, const float *positions_inv_mass,
                                    const float *velocities, uint32_t flags);
/* Replaces the static box colliders of scene `scene` (0 for a single uploaded scene) with
   `colliders`: the same count as the scene's upload, each of which may be moved, resized,
   re-materialled, enabled or disabled; the visit order is rebuilt, and the next cn_step uses them.
   Particles, topology, materials and the other scenes stay resident (the reference's
   uploadWorldState, physics_scene_gpu_state.cpp:2348-2392, takes the world's collider set
   whenever it uploads).  CN_ERR_INVALID_ARGUMENT if the count differs: adding or removing
   colliders needs a re-upload. */
cn_status cn_scene_set_colliders(cn_ctx *ctx, uint32_t scene, const cn_box_collider *colliders,
                                 uint32_t count);
/* Replaces the material tables of scene `scene` (0 for a single uploaded scene): its fluid
   materials (cn_fluid_material_from_desc records) and particle materials (float4 each), each with
   the count of the scene's upload; a NULL table is left as it is.  The next cn_step uses them
   (the reference's uploadWorldState, physics_scene_gpu_state.cpp:2348-2392, takes the world's
   material tables whenever it uploads).  CN_ERR_UNSUPPORTED if a fluid material's smoothing
   radius changes: that changes the scene's grid cell size, which needs a re-upload. */
cn_status cn_scene_set_materials(cn_ctx *ctx, uint32_t scene, const cn_fluid_material *fluid_materials,
                                 uint32_t fluid_count, const float *particle_materials,
                                 uint32_t particle_count);
/* Replaces the parameters of scene `scene`'s constraints (rest lengths, angles and volumes,
   compliances, edge enable flags): `cv` holds the scene's constraints in the counts and order of
   its upload, each on the same particles (the scene's own indices).  The next cn_step uses them;
   the incidence lists built at upload stay.  CN_ERR_INVALID_ARGUMENT if a count differs,
   CN_ERR_UNSUPPORTED if a constraint names other particles (a topology change needs a re-upload;
   the reference re-uploads its soft topology when its revision changes,
   physics_scene_gpu_state.cpp:3255). */
cn_status cn_scene_set_constraints(cn_ctx *ctx, uint32_t scene, const cn_constraint_view *cv);
/* Render anisotropy (buildFluidRenderAnisotropy, physics_pass_dispatcher.cpp:3098-3138 with
   physics_fluid_build_render_anisotropy.cs.hlsl:34-171, the renderer's fluidAnisotropy1..3
   buffers).  When enabled, every cn_step / cn_step_scenes with fluid ends by writing, per
   particle of the stepped scenes, three float4 (axis, scale): the principal axes of its
   neighbors' weighted covariance on the final positions, by decreasing spread, with scales
   1.1 sqrt(eigenvalue) clamped to [r, 2r]; other particles get the identity frame with scale r.
   Off by default (it changes no particle state).  cn_render_anisotropy_read copies axis k of the
   first `count` particles in particle order to axisk (each may be NULL); CN_ERR_NOT_READY until
   a step with fluid has written them since they were enabled or the scene was uploaded. */
cn_status cn_render_anisotropy_enable(cn_ctx *ctx, int enable);
cn_status cn_render_anisotropy_read(cn_ctx *ctx, float *axis1, float *axis2, float *axis3,
                                    uint32_t count);
/* validateGpuMetaBlocking analog; returns CN_ERR_CAPACITY if a capacity overflowed. */
cn_status cn_validate_meta(cn_ctx *ctx, cn_meta *out);
cn_status cn_synchronize(cn_ctx *ctx);
/* Non-blocking: CN_OK when all work enqueued on the context stream has finished, CN_ERR_NOT_READY
 * while some is still running (no reference counterpart; used by the serving loop). */
cn_status cn_query(cn_ctx *ctx);

/* -------- per-pass entries (host in/out; used by parity tests) -------- */
/* buildParticleBroadPhaseEntries + buildParticleBroadPhaseKeys
 * (physics_particle_build_broad_phase_{entries,keys}.cs.hlsl). */
cn_status cn_pass_broadphase_keys(cn_ctx *ctx, uint32_t count, const float *positions_inv_mass,
                                  float cell_size, uint32_t *keys, int32_t *cells_xyz);
/* sortParticleBroadPhase (physics_pass_dispatcher.cpp:3963-4063): stable ascending sort of
 * (key, index) pairs, in place. */
cn_status cn_pass_sort_pairs(cn_ctx *ctx, uint32_t count, uint32_t *keys, uint32_t *indices);
/* No reference counterpart: the sort of cn_pass_sort_pairs for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * keys and indices are DEVICE arrays of count entries on the context's GPU, sorted in place
 * (the same stable ascending order as cn_pass_sort_pairs).  The call is enqueued on the context
 * stream and returns without waiting (cn_synchronize waits); it calls no hipMalloc/hipFree.  Its
 * working memory is a fixed array in the library's device code, one per GPU
 * (2 * 8 * CN_NOALLOC_MAX_PAIRS bytes plus the sort's digit counts), shared by every call on that
 * GPU: calls that overlap on one GPU (two contexts' streams at once) race on it.  count above
 * CN_NOALLOC_MAX_PAIRS returns CN_ERR_CAPACITY. */
#define CN_NOALLOC_MAX_PAIRS (1u << 20)
cn_status cn_pass_sort_pairs_noalloc(cn_ctx *ctx, uint32_t count, uint32_t *keys, uint32_t *indices);
/* clear/mark/build ParticleCellRanges (physics_particle_*_cell_range*.cs.hlsl); ranges holds
 * capacity uint4 records {key, start, end, 0}, unused {0xffffffff, 0, 0, 0}. */
cn_status cn_pass_cell_ranges(cn_ctx *ctx, uint32_t count, const uint32_t *sorted_keys,
                              uint32_t capacity, uint32_t *ranges);

/* One substep's broadphase and first neighbor build on the uploaded scene, without committing
 * state: predict, keys, sort, cell runs, then the iteration-0 neighbor lists
 * (physics_solver.cpp:580-633 and :876).  sorted_keys/sorted_ids (count entries) are the sorted
 * cell keys and the particle index at each sorted position; counts (count) and lists
 * (count * max_fluid_neighbors, neighbor particle indices) are in particle-index order.
 * Any output may be NULL. */
cn_status cn_pass_substep_broadphase(cn_ctx *ctx, const cn_step_params *params, uint32_t *sorted_keys,
                                     uint32_t *sorted_ids, uint32_t *counts, uint32_t *lists);

/* -------- profiling (HIP events around the dominant kernels of cn_step) -------- */
#define CN_KERNEL_SLOTS 16
typedef struct cn_kernel_stats {
    char name[CN_KERNEL_SLOTS][32];
    uint32_t launches[CN_KERNEL_SLOTS];
    double total_ms[CN_KERNEL_SLOTS];
    uint32_t used;
} cn_kernel_stats;
cn_status cn_profile_enable(cn_ctx *ctx, int enable);
cn_status cn_profile_read(cn_ctx *ctx, cn_kernel_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* CN_HIP_H */
