/*
 * cn_oracle.h -- CPU restatement of CRESSim-Neo's particle substep (TEST INFRASTRUCTURE).
 *
 * This library is the parity oracle and the CPU baseline.  Only tests, the smoke check and the
 * CPU-baseline leg of the benchmark load it; the product library never links or calls it.
 *
 * It follows the reference pass by pass, in the reference's own data order: particle-index
 * order, the 32-bit cell hash, the 32-pass 1-bit split radix sort, the nextPow2(2N) cell-range
 * table searched by binary search, per-iteration neighbor rebuilds with a fixed slot budget per
 * particle.  Every function cites the shader or host file:line it restates.  All arithmetic is
 * IEEE float32 with no contraction (build flags -ffp-contract=off -fno-fast-math).
 */
#ifndef CN_ORACLE_H
#define CN_ORACLE_H

#include <stdint.h>

#include "../include/cn_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cno_scene {
    uint32_t n;
    float *pos;                 /* float4[n] xyz + invMass (in/out) */
    float *prev;                /* float4[n] */
    float *vel;                 /* float4[n] */
    const float *radii;
    const uint32_t *kinds;
    const uint32_t *fluid_mat;
    const uint32_t *part_mat;
    const uint32_t *meta;       /* uint4[n]: environment, phase, layer, mask */
    uint32_t n_fluid_mats;
    const cn_fluid_material *fluid_mats;
    uint32_t n_part_mats;
    const float *part_mats;     /* float4 */
    uint32_t n_colliders;
    const cn_box_collider *colliders;
    /* soft constraints */
    uint32_t n_edges;
    const cn_soft_edge *edges;
    uint32_t n_bends;
    const cn_soft_bend *bends;
    uint32_t n_dihedrals;
    const cn_cloth_dihedral *dihedrals;
    uint32_t n_tets;
    const cn_soft_tet *tets;
    /* per-constraint correction records; they persist across substeps and steps because the tet
     * solve leaves a record untouched when it returns early (caller allocates, zero-initialised) */
    float *edge_corr;           /* 8 floats per edge (two float4) */
    float *bend_corr;           /* 16 floats per bend + dihedral (four float4) */
    float *tet_corr;            /* 16 floats per tet */
    /* optional output: float4[3n], the render anisotropy of the last substep (axis1 of every
     * particle, then axis2, then axis3; w = scale), or NULL */
    float *render_aniso;
    /* constraint adjacency of the particles (ParticleSoA adjacencyOffsets/Counts/Indices,
     * physics_world.cpp:5774-5787), read by IsAdjacentParticle for self-colliding phase groups;
     * NULL offsets: no particle has any */
    const uint32_t *adj_offsets;
    const uint32_t *adj_counts;
    const uint32_t *adj_indices;
    /* static colliders of any shape after the boxes: collider k >= n_colliders is
     * shape_colliders[k - n_colliders] */
    uint32_t n_shape_colliders;
    const cn_collider *shape_colliders;
} cno_scene;

typedef struct cno_stats {
    uint32_t max_required_neighbors;
    uint32_t truncations;
    uint64_t neighbor_total_first_iter;   /* sum of counts of the last substep's iteration 0 */
    uint32_t fluid_count;
    uint32_t contact_overflows;           /* particles with more than CNO_CONTACT_PARTNERS candidates */
    double max_abs_constraint_last;       /* max |C| / scale-free residual of last density pass */
    double mean_constraint_last;
} cno_stats;

/* ---- host authoring ---- */
void cno_default_params(cn_step_params *p);
void cno_fluid_material(const cn_fluid_material_desc *desc, float radius, cn_fluid_material *out);
float cno_grid_cell_size(uint32_t n, const float *radii, const uint32_t *fluid_mat,
                         uint32_t n_mats, const cn_fluid_material *mats);

/* ---- broadphase passes ---- */
void cno_broadphase_entries(uint32_t n, const float *pos, float cell_size, uint32_t *keys,
                            int32_t *cells);
void cno_radix_sort_1bit(uint32_t n, uint32_t *keys, uint32_t *indices);
uint32_t cno_next_pow2(uint32_t v);
void cno_cell_ranges(uint32_t n, const uint32_t *sorted_keys, uint32_t capacity, uint32_t *ranges);
/* returns 1 and fills start/end when found */
int cno_find_cell_range(const uint32_t *ranges, uint32_t capacity, uint32_t key, uint32_t *start,
                        uint32_t *end);

/* ---- static-box boundary ---- */
/* Order in which the static LBVH visits the colliders (reverse Morton order). */
void cno_static_collider_order(uint32_t n_colliders, const cn_box_collider *colliders,
                               uint32_t *order);
/* ComputeFluidBoundaryContactInfo of point p against one resolved static collider: returns 0 when
 * it is disabled; sd = signed distance, nrm = normal, aabb (may be NULL) = broad-phase min, max. */
int cno_collider_contact(const cn_collider *collider, const float *p, float *sd, float *nrm, float *aabb);
/* The particle-rigid contact of a solid particle (position p, radius) against one resolved static
 * collider (physics_particle_rigid_generate_contacts.cs.hlsl:82-176): returns 1 with the world
 * normal and the penetration when the penetration is positive. */
int cno_rigid_contact(const cn_collider *collider, const float *p, float radius, float *nrm, float *pen);
/* counts[n], offsets[n] (exclusive), pairs[capacity] = collider indices; ranges uint2[n] */
uint32_t cno_fluid_boundary_candidates(const cno_scene *s, const uint32_t *order, uint32_t *ranges,
                                       uint32_t *pairs, uint32_t capacity);

/* ---- fluid passes (iteration = iterationIndex) ---- */
void cno_fluid_neighbors(const cno_scene *s, float cell_size, const int32_t *cells,
                         const uint32_t *sorted_keys, const uint32_t *sorted_idx,
                         const uint32_t *ranges_tbl, uint32_t capacity, const float *delta_acc,
                         uint32_t iteration, uint32_t cap, uint32_t *counts, uint32_t *list,
                         uint32_t *required);
void cno_fluid_density(const cno_scene *s, const float *cfg, const float *delta_acc,
                       uint32_t iteration, uint32_t cap, const uint32_t *counts,
                       const uint32_t *list, const uint32_t *bnd_ranges,
                       const uint32_t *bnd_pairs, float *normal_constraint);
void cno_fluid_delta(const cno_scene *s, const float *cfg, float dt, const float *delta_acc,
                     uint32_t iteration, uint32_t cap, const uint32_t *counts,
                     const uint32_t *list, const uint32_t *bnd_ranges, const uint32_t *bnd_pairs,
                     const float *normal_constraint, float *delta_out);
void cno_fluid_apply(const cno_scene *s, const float *cfg, uint32_t iteration,
                     const float *delta_in, float *delta_acc);
void cno_fluid_clamp(const cno_scene *s, const float *delta_acc_in, const uint32_t *bnd_ranges,
                     const uint32_t *bnd_pairs, float *delta_acc_out);
void cno_update_velocities(cno_scene *s, float dt, float *delta_acc);
void cno_project_boundary_velocities(cno_scene *s, const float *cfg, float dt,
                                     const uint32_t *bnd_ranges, const uint32_t *bnd_pairs);
void cno_predict(cno_scene *s, float dt, const float *gravity);
/* render anisotropy over the post-update neighbor lists; out float4[3n] */
void cno_fluid_render_anisotropy(const cno_scene *s, uint32_t cap, const uint32_t *counts,
                                 const uint32_t *list, float *out);
/* vorticity confinement over the post-update neighbor lists; vort float4[n] = (curl, |curl|) */
void cno_fluid_vorticity(const cno_scene *s, uint32_t cap, const uint32_t *counts, const uint32_t *list,
                         float *vort);
void cno_fluid_vorticity_apply(cno_scene *s, float dt, uint32_t cap, const uint32_t *counts,
                               const uint32_t *list, const float *vort);

/* ---- XPBD soft constraints ---- */
/* Cephes single-precision atan2f / acosf (S. L. Moshier), evaluated without contraction; the GPU
 * kernels evaluate the same sequence. */
float cno_atan2f(float y, float x);
float cno_acosf(float x);

/* ---- particle-particle contacts ---- */
#define CNO_CONTACT_PARTNERS 32u
/* candidate partners of every particle in its 27-cell walk order (pcount[n], plist[n * pcap]);
 * returns the number of particles whose candidates exceeded pcap */
uint32_t cno_contact_partners(const cno_scene *s, const int32_t *cells, const uint32_t *sorted_idx,
                              const uint32_t *ranges_tbl, uint32_t capacity, uint32_t pcap,
                              uint32_t *pcount, uint32_t *plist);

/* ---- full step: PhysicsSolver::step for the particle path ---- */
int cno_step(cno_scene *s, const cn_step_params *params, cno_stats *stats);

/* number of OpenMP threads the oracle uses (1 when built without OpenMP) */
int cno_threads(void);
void cno_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
