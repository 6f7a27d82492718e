// cn_internal.h -- declarations shared by the library's translation units (not installed).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cn_device.h"

namespace cn {

constexpr uint32_t kIterStateWords = 64;  // iterations with tiled fluid passes (later ones walk the run table)
constexpr uint32_t kXcds = 8;             // accelerator dies; each has its own L2 and block-queue counter
// iter_state words: block-queue counters of the density and delta passes (kXcds per iteration,
// each on a 128-byte line of its own: the workgroups of every die pop their blocks with atomics
// at the memory side, which serialise per line), the queue counter of the wide-tile delta pass per
// iteration and the wide-block count (all cleared once per substep), then the largest |delta|^2
// bits of each iteration
constexpr uint32_t kQueueStride = 32;  // words between two block-queue counters
constexpr uint32_t kStateDensity = 0;
constexpr uint32_t kStateDelta = kIterStateWords * kXcds * kQueueStride;
constexpr uint32_t kStateWideQueue = 2 * kIterStateWords * kXcds * kQueueStride;
constexpr uint32_t kStateWideCount = kStateWideQueue + kIterStateWords;
constexpr uint32_t kStateMaxDisp = kStateWideCount + 1;
constexpr uint32_t kStatePrevDisp = kStateMaxDisp + kIterStateWords;  // the last iteration's word of the
                                                                        // previous substep
constexpr uint32_t kStateWords = kStatePrevDisp + 1;
// per-block halo descriptor: start slot and exclusive particle offset of each of the 6x6x6 halo
// cells (x fastest), the halo particle total, a purity flag, the block's particle range and the
// cell coordinates of its low corner
constexpr uint32_t kHalo = 216;
constexpr uint32_t kHdStart = 0;
constexpr uint32_t kHdOff = kHalo;              // kHalo + 1 offsets, the last one the total
constexpr uint32_t kHdFlags = 2 * kHalo + 1;
constexpr uint32_t kHdP0 = kHdFlags + 1;
constexpr uint32_t kHdP1 = kHdFlags + 2;
constexpr uint32_t kHdBase = kHdFlags + 3;      // 3 words
constexpr uint32_t kHdScene = kHdFlags + 6;
constexpr uint32_t kHaloStride = 448;
constexpr uint32_t kHdNoFluid = 2u;  // kHdFlags value of a block in a scene without fluid particles
// candidate contact partners a particle keeps per substep; more are counted as overflows
constexpr uint32_t kContactPartners = 32;
constexpr uint32_t kMetaWords = 16;  // device diagnostics words
// candidate list slots of a fluid particle: its neighbors within h + skin at iteration 0
// (skin = kSkinFrac h), reused by a later iteration of the substep while its own displacement since
// iteration 0 plus the largest one in its block's halo stays below the skin
constexpr uint32_t kSuperSlots = 96;
constexpr float kSkinFrac = 0.05f;
constexpr uint16_t kSuperNone = 0xffffu;  // count marker: the candidates did not fit, walk instead
// Per-slot lists of 16-bit entries are stored in groups of kListGroup consecutive slots: entry k
// of slot i sits at (i / kListGroup) * kListGroup * rows + k * kListGroup + i % kListGroup, so one
// wave's row is one 128-byte line and a group's rows are contiguous (few pages per wave).
constexpr uint32_t kListGroup = 64;
// byte offset of entry 0 of slot i
__host__ __device__ inline size_t list_entry_base(uint32_t i, uint32_t rows)
{
    return ((size_t)(i / kListGroup) * kListGroup * rows + (i % kListGroup)) * sizeof(uint16_t);
}
__host__ __device__ inline size_t list_slots(size_t n) { return (n + kListGroup - 1) / kListGroup * kListGroup; }
// A list buffer is addressed with 32-bit byte offsets from its base (one 32-bit add per entry on the
// device), so it stays within 4 GiB less 1 MiB: a list's write position may run up to 512 KiB past
// its particle's last entry without wrapping.
constexpr uint64_t kListBytesMax = 0xffffffffull - 0x100000ull;

// ---- radix sort (cn_sort.hip) ----
size_t sort_scratch_words(uint32_t n);
uint2 *radix_sort_pairs(uint2 *buf_a, uint2 *buf_b, uint32_t n, uint32_t key_bits, uint32_t *scratch,
                        hipStream_t stream);
void launch_pack_pairs(const uint32_t *keys, const uint32_t *vals, uint32_t n, uint2 *out, hipStream_t stream);
void launch_unpack_pairs(const uint2 *in, uint32_t n, uint32_t *keys, uint32_t *vals, hipStream_t stream);

// Device pointers of one resident scene (single scene or one slot pool).
// Device forms of the soft constraints: the caller's records (cn_soft_edge, cn_soft_bend,
// cn_cloth_dihedral, cn_soft_tet) without their padding, with the compliance and the tet rest volume
// clamped at zero on upload (the solves use only max(value, 0)) and a disabled edge marked by a NaN
// compliance, so a solve reads 16-24 bytes per constraint instead of 32.
struct EdgeDev {
    uint32_t a, b;
    float rest_length, compliance;
};
struct BendDev {
    uint32_t p0, p1, p2;
    float rest_angle, compliance;
};
struct DihedralDev {
    uint32_t e0, e1, o0, o1;
    float rest_angle, compliance;
};
struct TetDev {
    uint32_t p[4];
    float rest_volume, compliance;
};

// One constraint's position correction for one of its vertices (12 bytes: the records are written
// by the constraint solves and summed per vertex, so their size is most of the soft phase's traffic).
struct CorrRec {
    float x, y, z;
};

struct SceneBuffers {
    uint32_t n = 0;                 // slots of the per-slot arrays in use
    uint32_t npid = 0;              // particles of the pool (particle-index arrays hold npid entries)
    // particle state in slot order, double-buffered across the per-substep reorder
    float4 *pos[2] = {nullptr, nullptr};
    float4 *vel[2] = {nullptr, nullptr};
    uint32_t *cls[2] = {nullptr, nullptr};
    uint32_t *id[2] = {nullptr, nullptr};
    int cur = 0;
    float4 *prev[2] = {nullptr, nullptr};  // position before the last predict, in slot order
    int4 *cellc = nullptr;          // cell coordinates of each slot (w unused)
    uint32_t *bnd = nullptr;        // candidate colliders: bitmask over the scene's visit order, or a
                                    // list count (kBndList) with the entries in bndl
    uint16_t *bndl = nullptr;       // kBndListCap candidate colliders per slot (scenes of more than
                                    // kMaskColliders colliders; null when the pool has none)
    // broadphase
    uint2 *kv[2] = {nullptr, nullptr};
    uint32_t *sort_scratch = nullptr;
    uint32_t *elem_run = nullptr;
    uint32_t *run_key = nullptr;
    uint32_t *run_start = nullptr;  // n + 1
    int4 *run_info = nullptr;       // (cx, cy, cz, pure)
    uint32_t *bucket = nullptr;     // (1 << bucket_bits) + 1
    uint32_t bucket_bits = 0;
    uint32_t *block_runs = nullptr; // per 1024-element block, then scanned
    uint32_t *run_count = nullptr;  // device scalar
    uint32_t *perm = nullptr;
    // fluid solve
    float4 *dacc[2] = {nullptr, nullptr};
    float4 *curpos = nullptr;       // (x* + delta, inverse mass): neighbor positions of the walk
    float4 *cs = nullptr;           // (x* + delta, density constraint): neighbor records of the delta pass
    float4 *nc = nullptr;           // surface normal * tension (only with surface tension)
    uint32_t *blk_first_run = nullptr;  // first run of each Morton 4x4x4 block, then the run count
    uint32_t *nblk = nullptr;       // [0] Morton blocks, [1] blocks of scenes with fluid (a prefix)
    uint32_t *blk_halo = nullptr;   // kHaloStride words per block, for the first blk_cap blocks
    uint32_t blk_cap = 0;
    uint32_t *wide_blk = nullptr;   // blocks whose halo fits the wide delta tile only (blk_cap)
    uint32_t *iter_state = nullptr; // kStateWords: block-queue counters and max |delta|^2 per iteration
    float *fluid_t2 = nullptr;      // per fluid material: smallest d^2 whose correctly rounded sqrt >= h
    uint16_t *list = nullptr;       // fluid neighbor lists as halo offsets, in slot groups (list_row)
    uint16_t *slist = nullptr;      // candidate lists of iteration 0 as halo offsets, in slot groups
    uint16_t *scnt = nullptr;       // candidate count per slot, or kSuperNone
    float *disp = nullptr;          // per slot: bound on |x* + delta - x*| after the last delta pass
    // constraint adjacency in particle-index order (pool indices): particle i's adjacent
    // particles are adj_idx[adj_off[i] .. adj_off[i] + adj_cnt[i]); null when no scene has any
    uint32_t *adj_off = nullptr;
    uint32_t *adj_cnt = nullptr;
    uint32_t *adj_idx = nullptr;
    uint32_t *lcnt = nullptr;       // fluid neighbor count per slot
    uint32_t *nbr = nullptr;        // neighbor lists of the pass entry, k-major: nbr[k * n + i]
    uint32_t *nbr_cnt = nullptr;
    size_t nbr_alloc = 0;
    uint32_t cap = 0;
    // particle-particle contacts (allocated when the scene has a solid particle)
    uint32_t *ct_list = nullptr;    // candidate partner slots, k-major: ct_list[k * n + i]
    uint32_t *ct_count = nullptr;
    float4 *ct_rec = nullptr;       // (normal, penetration) of each candidate's last contact, k-major
    uint32_t *ct_bits = nullptr;    // bit k: candidate k's last contact is active
    float4 *ct_corr = nullptr;      // summed correction of the last contact pass
    float4 *ct_gen = nullptr;       // position of each solid's last particle-rigid contact generation
    size_t ct_alloc = 0;
    // index-order exports
    float4 *out_pos = nullptr;
    float4 *out_vel = nullptr;
    float4 *out_prev = nullptr;
    float4 *aniso = nullptr;        // render anisotropy: axis1 | axis2 | axis3, n each, particle order
    size_t aniso_alloc = 0;
    // tables
    ParticleClass *classes = nullptr;
    uint32_t nclasses = 0;
    uint32_t fluid_uniform = 0;     // every scene with fluid has one particle class
    cn_fluid_material *fluid_mats = nullptr;
    float4 *part_mats = nullptr;
    Collider *colliders = nullptr;
    uint32_t ncolliders = 0;
    ColGroup *col_groups = nullptr;  // per scene of more than kMaskColliders colliders, from grp0
    SceneInfo *scenes = nullptr;    // one record per scene of the pool
    uint32_t nscenes = 0;
    uint32_t *scene_active = nullptr;  // per scene: 1 = stepped by cn_step, 0 = state left untouched
    uint32_t *pid_scene = nullptr;     // scene of each particle index
    // diagnostics (device): [0] max required neighbors, [1] truncations, [2..3] neighbor total,
    // [4] contact partner overflows, [5] LDS blocks, [6] run-table blocks, [7] max |delta|^2 bits,
    // [8] longest run of equal keys, [9] candidate-list reuses; kMetaWords words
    uint32_t *meta = nullptr;
    // soft constraints
    uint32_t *slot_of = nullptr;
    float4 *posi = nullptr;         // positions of the constrained particles in particle-index order
                                    // during the soft phase of an iteration
    uint32_t *soft_ids = nullptr;   // the constrained particles, ascending
    uint32_t *rigid_pick = nullptr; // per constrained particle: its particle-rigid work of the substep
    uint32_t nsoft = 0;
    EdgeDev *edges = nullptr;
    BendDev *bends = nullptr;
    DihedralDev *dihedrals = nullptr;
    TetDev *tets = nullptr;
    uint32_t ne = 0, nb = 0, nd = 0, nt = 0;
    float *lam = nullptr;           // edges | bends + dihedrals | tets
    CorrRec *corr = nullptr;        // correction records of edges | bends + dihedrals | tets, each group
                                    // in its incidence (CSR) order, then one spare record
    size_t corr_off[3] = {0, 0, 0};
    uint32_t *inc = nullptr;        // three CSR incidence structures back to back, then the record maps
    size_t inc_off[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

// ---- broadphase + fluid launches (cn_kernels.hip) ----
void launch_keys(const SceneBuffers &s, const SubstepConsts &c, hipStream_t st);
void launch_runs(SceneBuffers &s, const SubstepConsts &c, const uint2 *sorted, uint32_t key_bits, hipStream_t st);
void launch_reorder(SceneBuffers &s, const SubstepConsts &c, hipStream_t st);
void launch_blocks(SceneBuffers &s, const SubstepConsts &c, hipStream_t st);
void fluid_tiles(SubstepConsts &c);
void launch_blk_halo(const SceneBuffers &s, const SubstepConsts &c, hipStream_t st);
void launch_fluid_density(const SceneBuffers &s, const SubstepConsts &c, uint32_t it, int din, hipStream_t st);
void launch_fluid_delta(const SceneBuffers &s, const SubstepConsts &c, uint32_t it, int din, hipStream_t st);
void launch_fluid_lists(const SceneBuffers &s, const SubstepConsts &c, hipStream_t st);
// vorticity confinement on the finalized positions and velocities of the substep
void launch_fluid_vorticity(const SceneBuffers &s, const SubstepConsts &c, hipStream_t st);
// render anisotropy of the finalized positions into out (3 n float4, particle-index order)
void launch_fluid_render_aniso(const SceneBuffers &s, const SubstepConsts &c, float4 *out, hipStream_t st);
void launch_finalize(const SceneBuffers &s, const SubstepConsts &c, int din, hipStream_t st);
void launch_export(const SceneBuffers &s, bool state, bool prev, hipStream_t st);
void launch_set_state(SceneBuffers &s, uint32_t first, uint32_t count, const float4 *P, const float4 *V,
                      hipStream_t st);
void launch_fill(uint32_t *p, uint32_t words, uint32_t value, hipStream_t st);

// ---- particle-particle contacts (cn_contact.hip) ----
void launch_contact_partners(const SceneBuffers &s, const SubstepConsts &c, hipStream_t st);
void launch_contact_iteration(const SceneBuffers &s, const SubstepConsts &c, int din, hipStream_t st);
void launch_contact_velocity(const SceneBuffers &s, const SubstepConsts &c, hipStream_t st);
void launch_rigid_velocity(const SceneBuffers &s, const SubstepConsts &c, uint32_t iters, hipStream_t st);
// the particle-rigid contacts of scenes solved in particle-index order, on the soft phase's positions;
// pick: first call of the substep (each particle's slot and candidates are looked up once)
void launch_rigid_pid(const SceneBuffers &s, const SubstepConsts &c, bool pick, hipStream_t st);
void launch_tick_gather(const SceneBuffers &pool, const SceneBuffers &tick, const uint4 *seg, uint32_t nseg,
                        hipStream_t st);
void launch_tick_scatter(const SceneBuffers &pool, const SceneBuffers &tick, const uint4 *seg, uint32_t nseg,
                         hipStream_t st);

// ---- XPBD soft constraints (cn_soft.hip) ----
struct SoftArgs {
    float4 *pos;               // slot order
    float4 *posi;              // particle-index order (constrained particles) while the soft passes run
    const uint32_t *soft_ids;  // constrained particles
    uint32_t nsoft;
    const uint32_t *slot_of;   // particle index -> slot
    const uint32_t *id;        // slot -> particle index
    uint32_t n;
    float dt, relax;
    const EdgeDev *edges;
    uint32_t ne;
    float *lam_e;
    CorrRec *corr_e;           // correction records in incidence order
    const uint32_t *map_e;     // 2 per edge: the position of its vertex records in corr_e
    const BendDev *bends;
    uint32_t nb;
    const DihedralDev *dih;
    uint32_t nd;
    float *lam_b;              // bends then dihedrals
    CorrRec *corr_b;
    const uint32_t *map_b;     // 4 per bend / dihedral (a bend's fourth is the spare record)
    const TetDev *tets;
    uint32_t nt;
    float *lam_t;
    CorrRec *corr_t;
    const uint32_t *map_t;     // 4 per tet
    const uint32_t *inc_e_start;  // incidence ranges: a particle's records are corr_x[start[p], start[p + 1])
    const uint32_t *inc_b_start;
    const uint32_t *inc_t_start;
    const uint32_t *active;    // per scene
    const uint32_t *pid_scene; // particle index -> scene
    uint32_t all_live;         // every scene takes part in the step (no per-particle check)
    uint32_t first;            // first iteration of a substep: every lambda starts at zero (read as
                               // zero, and written by every constraint, also where it does nothing)
};
// gather / scatter: copy the constrained particles' positions from slot order before the solves and
// back after them (needed only around iterations where another pass moves those particles)
// solve = false: only the gather and / or the scatter
void launch_soft_iteration(const SoftArgs &a, bool gather, bool scatter, bool solve, hipStream_t st);

// per-pass entries in the reference's data order (cn_passes.hip)
void launch_ref_keys(const float4 *pos, uint32_t n, float inv_cell, uint32_t *keys, int32_t *cells,
                     hipStream_t st);
void launch_ref_cell_ranges(const uint32_t *sorted_keys, uint32_t n, uint32_t capacity, uint32_t *ranges,
                            hipStream_t st);

}  // namespace cn
