// cn_api.hip -- host side of libcn_hip.so: context, scene state, the substep driver and the C ABI.
//
// The driver follows PhysicsSolver::step (src/physics/physics_solver.cpp:336-1401) for the particle
// path: per substep predict + broadphase, then the position-solve iterations, then the velocity
// update and the fluid boundary velocity projection.  Every pass is a kernel on the context's
// stream; nothing here synchronizes except readback, validation and profiling reads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "cn_internal.h"

using namespace cn;

// Device scratch of cn_pass_sort_pairs_noalloc: two pair buffers of CN_NOALLOC_MAX_PAIRS entries and
// the radix sort's per-tile digit counts for that many pairs (256 digits x 256 tiles + 256 totals).
// Every GPU the library runs on holds its own copy from module load until the process ends.  In
// host code `scratch` is only a stand-in for the symbol: its device address on the current GPU
// comes from hipGetSymbolAddress (noalloc_scratch below), and only that address reaches a kernel.
constexpr size_t kNoAllocPairBytes = (size_t)CN_NOALLOC_MAX_PAIRS * sizeof(uint2);
constexpr size_t kNoAllocCountWords = 256u * (CN_NOALLOC_MAX_PAIRS / 4096u) + 256u;
constexpr size_t kNoAllocScratchBytes = 2 * kNoAllocPairBytes + kNoAllocCountWords * sizeof(uint32_t);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each GPU (null: not resolved yet)

// Resolves the scratch array's address on the current device once per device (called with that
// device current); false when the runtime cannot give it.
bool resolve_noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return false;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return true;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return false;
    g_scratch[device] = p;
    return true;
}

void *noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch[device];
}

}  // namespace

struct cn_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // the soft phase's stream while it runs beside the fluid passes (created on first use, default
    // priority like the context stream), and the events that fork it from and join it to `stream`
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    bool split_phases = true;        // run the soft and fluid phases of a substep concurrently
    bool side_failed = false;        // an event call of the split phases failed in the last step
    std::string err;
    SceneBuffers s;
    uint32_t capacity = 0;       // particle capacity of the allocation
    uint32_t cap_alloc = 0;      // neighbor slots per particle of the allocation
    bool uploaded = false;
    bool stepped = false;        // a cn_step ran since the upload (slot order differs from index order)
    float cell_size = 0.0f;      // derived on upload like the reference world
    uint32_t morton_bits = 10;
    uint32_t bucket_bits_alloc = 4;  // bucket index bits the allocation holds
    float max_h = 0.0f;              // largest fluid smoothing radius of the scene
    bool has_fluid = false;          // any fluid particle (hasFluidWork, physics_solver.cpp:465)
    bool has_tension = false;        // any fluid material with surface tension
    bool has_solid = false;          // any solid particle
    bool any_pairs = false;          // some scene can hold a particle-particle contact pair
    bool any_rigid = false;          // some scene solves particle-rigid contacts in slot order
    bool any_rigid_pid = false;      // some scene solves them in particle-index order (soft phase)
    bool has_vorticity = false;      // a fluid particle's material confines vorticity
    bool aniso_enabled = false;      // cn_step ends with the render anisotropy
    bool aniso_ready = false;        // a step has written it since it was enabled or uploaded
    uint32_t tile_max = 4096;        // halo particles a fluid-pass workgroup may stage in LDS
    uint32_t grid_max = 768;         // workgroups of a persistent fluid launch (three per CU)
    // scene pool
    uint32_t nscenes = 0;
    uint32_t scene_bits = 0;         // key bits above the Morton code holding the scene rank
    uint32_t fluid_scenes = 0;       // scenes with fluid particles (ranks [0, fluid_scenes))
    std::vector<uint32_t> scene_first;  // first particle of each scene, then the total
    std::vector<float> scene_cell;      // derived grid cell size of each scene
    std::vector<uint32_t> scene_col_cap;  // box colliders of each scene: its uploaded box count
    std::vector<std::vector<cn_collider>> scene_shapes;  // each scene's shape colliders, as uploaded
    std::vector<SceneInfo> scene_info;    // host copy of the device scene records
    std::vector<uint32_t> scene_fm;       // per scene: first fluid material, count (pairs)
    std::vector<uint32_t> scene_pm;       // per scene: first particle material, count (pairs)
    std::vector<cn_fluid_material> fm_host;  // host copy of the pool's fluid material table
    std::vector<uint8_t> fm_fluid_used;      // per fluid material: referenced by a fluid particle
    std::vector<uint32_t> scene_con;      // per scene: first edge, bend, dihedral, tet (4 words)
    std::vector<EdgeDev> edges_host;      // host copies of the device constraint records
    std::vector<BendDev> bends_host;
    std::vector<DihedralDev> dih_host;
    std::vector<TetDev> tets_host;
    bool graph_valid = false;         // the cached graphs refer to the current buffers
    bool graphs_enabled = false;
    struct Graph {
        cn_step_params params;
        int cur0, cur1;                  // slot-order parity before and after the step
        bool all_active;                 // recorded for a step of every scene
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    std::vector<uint32_t> active_host;  // last per-scene active flags written to the device
    // compact ticks (cn_step_scenes over a small part of a pool)
    std::vector<uint32_t> slot_base;    // per scene: first slot of its segment in the pool's slot order
    bool slots_by_prefix = false;       // the slot order holds every scene as one segment (a step ran)
    float compact_below = 0.5f;         // compact a tick whose scenes hold at most this share of the pool
    SceneBuffers tick;                  // the compact tick's per-slot arrays (the rest are the pool's)
    uint32_t tick_cap = 0;              // slots they hold
    uint32_t tick_list_cap = 0;         // neighbor slots per particle of tick.list
    uint4 *tick_seg = nullptr;          // per listed scene: pool slot start, compact start, count
    uint32_t tick_seg_cap = 0;
    std::vector<uint4> tick_seg_host;   // its host copy (pageable: staged when the copy is enqueued)
    float4 *stage = nullptr;            // host-provided particle state on its way to the device
    size_t stage_cap = 0;
    // profiling
    bool profiling = false;
    struct Ev {
        int slot;
        hipEvent_t a, b;
    };
    std::vector<Ev> events;           // recorded pairs since the last cn_profile_read
    std::vector<hipEvent_t> ev_pool;  // timing events, created once and reused
    size_t ev_used = 0;
    cn_kernel_stats stats{};
    cn_meta last_meta{};
};

namespace {

cn_status fail(cn_ctx *ctx, cn_status st, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return st;
}

#define CN_HIP(ctx, call)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail((ctx), CN_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
hipError_t dalloc(T **p, size_t count)
{
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * count);
}

void dfree(void *p)
{
    if (p) (void)hipFree(p);
}

uint32_t ceil_log2(uint64_t v)
{
    uint32_t b = 0;
    while ((1ull << b) < v) ++b;
    return b;
}

void free_scene(SceneBuffers &s)
{
    for (int k = 0; k < 2; ++k) {
        dfree(s.pos[k]);
        dfree(s.vel[k]);
        dfree(s.cls[k]);
        dfree(s.id[k]);
        dfree(s.kv[k]);
        dfree(s.dacc[k]);
    }
    dfree(s.prev[0]);
    dfree(s.prev[1]);
    dfree(s.scene_active);
    dfree(s.pid_scene);
    dfree(s.cellc);
    dfree(s.bnd);
    dfree(s.bndl);
    dfree(s.col_groups);
    dfree(s.sort_scratch);
    dfree(s.elem_run);
    dfree(s.run_key);
    dfree(s.run_start);
    dfree(s.run_info);
    dfree(s.bucket);
    dfree(s.block_runs);
    dfree(s.run_count);
    dfree(s.perm);
    dfree(s.nc);
    dfree(s.curpos);
    dfree(s.cs);
    dfree(s.blk_first_run);
    dfree(s.nblk);
    dfree(s.blk_halo);
    dfree(s.wide_blk);
    dfree(s.iter_state);
    dfree(s.fluid_t2);
    dfree(s.list);
    dfree(s.lcnt);
    dfree(s.slist);
    dfree(s.scnt);
    dfree(s.disp);
    dfree(s.adj_off);
    dfree(s.adj_cnt);
    dfree(s.adj_idx);
    dfree(s.nbr);
    dfree(s.nbr_cnt);
    dfree(s.ct_list);
    dfree(s.ct_count);
    dfree(s.ct_rec);
    dfree(s.ct_bits);
    dfree(s.ct_corr);
    dfree(s.ct_gen);
    dfree(s.out_pos);
    dfree(s.out_vel);
    dfree(s.out_prev);
    dfree(s.aniso);
    dfree(s.classes);
    dfree(s.fluid_mats);
    dfree(s.part_mats);
    dfree(s.colliders);
    dfree(s.scenes);
    dfree(s.meta);
    dfree(s.slot_of);
    dfree(s.posi);
    dfree(s.soft_ids);
    dfree(s.rigid_pick);
    dfree(s.edges);
    dfree(s.bends);
    dfree(s.dihedrals);
    dfree(s.tets);
    dfree(s.lam);
    dfree(s.corr);
    dfree(s.inc);
    s = SceneBuffers{};
}

void clear_graphs(cn_ctx *ctx)
{
    for (auto &g : ctx->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    ctx->graphs.clear();
}

void free_constraints(SceneBuffers &s)
{
    dfree(s.edges);
    dfree(s.bends);
    dfree(s.dihedrals);
    dfree(s.tets);
    dfree(s.lam);
    dfree(s.corr);
    dfree(s.inc);
    dfree(s.posi);
    dfree(s.soft_ids);
    dfree(s.rigid_pick);
    s.posi = nullptr;
    s.soft_ids = nullptr;
    s.rigid_pick = nullptr;
    s.nsoft = 0;
    s.edges = nullptr;
    s.bends = nullptr;
    s.dihedrals = nullptr;
    s.tets = nullptr;
    s.lam = nullptr;
    s.corr = nullptr;
    s.inc = nullptr;
    s.ne = s.nb = s.nd = s.nt = 0;
}

// Particle -> incident constraints, ascending constraint order (physics_scene_gpu_state.cpp:3273-3360);
// appends start[n + 1] then ref[] to out and returns the two offsets.
// map (per slots per constraint) receives the CSR position of each constraint vertex; slots without
// one point at the spare record past the last.
template <class Fn>
void build_incidence(uint32_t n, uint32_t groups, const uint32_t *counts, const uint32_t *arity, const uint32_t *offset,
                     Fn parts, std::vector<uint32_t> &out, size_t &start_off, size_t &ref_off,
                     std::vector<uint32_t> &map, uint32_t per, size_t &refs)
{
    std::vector<uint32_t> cnt(n + 1, 0);
    for (uint32_t g = 0; g < groups; ++g)
        for (uint32_t c = 0; c < counts[g]; ++c)
            for (uint32_t k = 0; k < arity[g]; ++k) {
                const uint32_t p = parts(g, c, k);
                if (p < n) ++cnt[p];
            }
    start_off = out.size();
    uint32_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        out.push_back(acc);
        acc += cnt[i];
    }
    out.push_back(acc);
    ref_off = out.size();
    out.resize(out.size() + acc);
    std::vector<uint32_t> cur(n);
    for (uint32_t i = 0; i < n; ++i) cur[i] = out[start_off + i];
    size_t ncons = 0;
    for (uint32_t g = 0; g < groups; ++g) ncons += counts[g];
    map.assign(ncons * per, acc);
    refs = acc;
    for (uint32_t g = 0; g < groups; ++g)
        for (uint32_t c = 0; c < counts[g]; ++c)
            for (uint32_t k = 0; k < arity[g]; ++k) {
                const uint32_t p = parts(g, c, k);
                if (p < n) {
                    map[(size_t)(c + offset[g]) * per + k] = cur[p];
                    out[ref_off + cur[p]++] = ((c + offset[g]) << 2) | k;
                }
            }
}

// device forms of the constraint records, particle indices shifted by base: padding dropped,
// max(compliance, 0) and max(rest volume, 0) taken here, a disabled edge carried as a NaN compliance
float clamp0(float v) { return v > 0.0f ? v : 0.0f; }
EdgeDev edge_dev(const cn_soft_edge &E, uint32_t base)
{
    return {E.a + base, E.b + base, E.rest_length, E.enabled ? clamp0(E.compliance) : std::nanf("")};
}
BendDev bend_dev(const cn_soft_bend &B, uint32_t base)
{
    return {B.p0 + base, B.p1 + base, B.p2 + base, B.rest_angle, clamp0(B.compliance)};
}
DihedralDev dihedral_dev(const cn_cloth_dihedral &D, uint32_t base)
{
    return {D.e0 + base, D.e1 + base, D.o0 + base, D.o1 + base, D.rest_angle, clamp0(D.compliance)};
}
TetDev tet_dev(const cn_soft_tet &T, uint32_t base)
{
    return {{T.p[0] + base, T.p[1] + base, T.p[2] + base, T.p[3] + base}, clamp0(T.rest_volume), clamp0(T.compliance)};
}

// host restatement of the static-collider visit order: enabled colliders, Morton codes of their
// broad-phase AABB centres in the static extent (physics_rigid_morton_codes.cs.hlsl:8-50), stable
// order, walked right child first (physics_particle_emit_fluid_boundary_candidates.cs.hlsl:158-225)
uint32_t expand10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Static colliders resolved to their world pose (physics_math.hlsli with rsqrt taken as 1 / sqrt,
// sums left to right): ComputeFluidBoundaryContactInfo's pose (body pose composed with the
// unscaled local position, physics_fluid_boundary_common.hlsli:21-33) and the broad-phase AABB of
// physics_rigid_update_world_aabbs.cs.hlsl:51-91 (scaled local position, ComputeBodyAabb of
// physics_shape_common.hlsli:49-75, the margin added after).
struct Quat {
    float x, y, z, w;
};
Quat quat_normalize(const float *q)
{
    const float l = ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3];
    if (l <= kEps) return {0.0f, 0.0f, 0.0f, 1.0f};
    const float r = 1.0f / std::sqrt(l);
    return {q[0] * r, q[1] * r, q[2] * r, q[3] * r};
}
Quat quat_mul(const Quat &a, const Quat &b)
{
    return {((a.w * b.x + a.x * b.w) + a.y * b.z) - a.z * b.y, ((a.w * b.y - a.x * b.z) + a.y * b.w) + a.z * b.x,
            ((a.w * b.z + a.x * b.y) - a.y * b.x) + a.z * b.w, ((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z};
}
void quat_rotate_h(const Quat &q, const float *v, float *o)
{
    const float t[3] = {2.0f * (q.y * v[2] - q.z * v[1]), 2.0f * (q.z * v[0] - q.x * v[2]),
                        2.0f * (q.x * v[1] - q.y * v[0])};
    o[0] = (v[0] + q.w * t[0]) + (q.y * t[2] - q.z * t[1]);
    o[1] = (v[1] + q.w * t[1]) + (q.z * t[0] - q.x * t[2]);
    o[2] = (v[2] + q.w * t[2]) + (q.x * t[1] - q.y * t[0]);
}

Collider box_collider(const cn_box_collider &b, uint32_t index)
{
    Collider k{};
    for (int a = 0; a < 3; ++a) {
        k.c[a] = b.center[a];
        k.he[a] = b.half_extents[a];
        k.aabb_min[a] = (b.center[a] - b.half_extents[a]) - kBroadPhaseMargin;
        k.aabb_max[a] = (b.center[a] + b.half_extents[a]) + kBroadPhaseMargin;
    }
    for (int a = 0; a < 4; ++a) k.material[a] = b.material[a];
    k.env = b.environment;
    k.layer = b.layer;
    k.mask = b.mask;
    k.enabled = b.enabled;
    k.index = index;
    k.shape = CN_SHAPE_BOX;
    k.rotated = 0u;
    k.q[3] = 1.0f;
    return k;
}

Collider shape_collider(const cn_collider &g, uint32_t index)
{
    Collider k{};
    const Quat bq = quat_normalize(g.body_orientation), lq = quat_normalize(g.local_orientation);
    const Quat q = [&] {
        const Quat m = quat_mul(bq, lq);
        const float v[4] = {m.x, m.y, m.z, m.w};
        return quat_normalize(v);
    }();
    const float sx = std::fabs(g.body_scale[0]), sy = std::fabs(g.body_scale[1]), sz = std::fabs(g.body_scale[2]);
    const float *prm = g.shape_params;
    float rot[3];
    quat_rotate_h(bq, g.local_position, rot);
    for (int a = 0; a < 3; ++a) k.c[a] = g.body_position[a] + rot[a];
    k.q[0] = q.x;
    k.q[1] = q.y;
    k.q[2] = q.z;
    k.q[3] = q.w;
    k.rotated = (q.x == 0.0f && q.y == 0.0f && q.z == 0.0f && q.w == 1.0f) ? 0u : 1u;
    k.shape = g.shape;
    const float up[3] = {0.0f, 1.0f, 0.0f};
    float axis[3];
    quat_rotate_h(q, up, axis);
    float r = 0.0f, hh = 0.0f;
    if (g.shape == CN_SHAPE_SPHERE) {
        r = prm[0] * std::max(sx, std::max(sy, sz));  // SphereRadius
        k.he[0] = r;
    } else if (g.shape == CN_SHAPE_BOX) {
        k.he[0] = prm[0] * sx;  // BoxHalfExtents
        k.he[1] = prm[1] * sy;
        k.he[2] = prm[2] * sz;
    } else {  // CapsuleSegment
        r = prm[0] * std::max(sx, sz);
        hh = prm[1] * sy;
        k.he[0] = r;
        for (int a = 0; a < 3; ++a) {
            k.sa[a] = k.c[a] - axis[a] * hh;
            k.sb[a] = k.c[a] + axis[a] * hh;
        }
    }
    // the broad-phase AABB from the scaled local position
    const float lps[3] = {g.local_position[0] * g.body_scale[0], g.local_position[1] * g.body_scale[1],
                          g.local_position[2] * g.body_scale[2]};
    quat_rotate_h(bq, lps, rot);
    float ca[3], ext[3];
    for (int a = 0; a < 3; ++a) ca[a] = g.body_position[a] + rot[a];
    if (g.shape == CN_SHAPE_SPHERE) {
        ext[0] = ext[1] = ext[2] = r;
    } else if (g.shape == CN_SHAPE_BOX) {
        const float ex[3] = {1.0f, 0.0f, 0.0f}, ey[3] = {0.0f, 1.0f, 0.0f}, ez[3] = {0.0f, 0.0f, 1.0f};
        float ax[3], ay[3], az[3];
        quat_rotate_h(q, ex, ax);
        quat_rotate_h(q, ey, ay);
        quat_rotate_h(q, ez, az);
        for (int a = 0; a < 3; ++a)
            ext[a] = (std::fabs(ax[a]) * k.he[0] + std::fabs(ay[a]) * k.he[1]) + std::fabs(az[a]) * k.he[2];
    } else {
        float A[3], B[3];
        for (int a = 0; a < 3; ++a) {
            A[a] = ca[a] - axis[a] * hh;
            B[a] = ca[a] + axis[a] * hh;
        }
        const float d[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        const float len = std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
        for (int a = 0; a < 3; ++a) ext[a] = std::fabs(axis[a]) * (0.5f * len) + r;
    }
    for (int a = 0; a < 3; ++a) {
        k.aabb_min[a] = (ca[a] - ext[a]) - kBroadPhaseMargin;
        k.aabb_max[a] = (ca[a] + ext[a]) + kBroadPhaseMargin;
    }
    for (int a = 0; a < 4; ++a) k.material[a] = g.material[a];
    k.env = g.environment;
    k.layer = g.layer;
    k.mask = g.mask;
    k.enabled = g.enabled;
    k.index = index;
    return k;
}

// Host restatement of the static-collider visit order: the enabled colliders (boxes, then shape
// colliders), Morton codes of their broad-phase AABB centres in the static extent
// (physics_rigid_morton_codes.cs.hlsl:8-50), stable order, walked right child first
// (physics_particle_emit_fluid_boundary_candidates.cs.hlsl:158-225).
std::vector<Collider> visit_ordered_colliders(const cn_box_collider *boxes, uint32_t nbox, const cn_collider *shapes,
                                              uint32_t nshape)
{
    std::vector<Collider> all;
    for (uint32_t i = 0; i < nbox; ++i)
        if (boxes[i].enabled) all.push_back(box_collider(boxes[i], i));
    for (uint32_t i = 0; i < nshape; ++i)
        if (shapes[i].enabled) all.push_back(shape_collider(shapes[i], nbox + i));
    float gmn[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
    float gmx[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
    for (const Collider &e : all)
        for (int a = 0; a < 3; ++a) {
            gmn[a] = std::min(gmn[a], e.aabb_min[a]);
            gmx[a] = std::max(gmx[a], e.aabb_max[a]);
        }
    std::vector<std::pair<uint32_t, uint32_t>> code(all.size());  // (Morton code, position)
    for (size_t k = 0; k < all.size(); ++k) {
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) {
            const float size = std::max(gmx[a] - gmn[a], 1.0e-5f);
            const float center = 0.5f * (all[k].aabb_min[a] + all[k].aabb_max[a]);
            float m = (center - gmn[a]) / size;
            m = std::min(std::max(m, 0.0f), 1.0f);
            float sc = m * 1024.0f;
            sc = std::min(std::max(sc, 0.0f), 1023.0f);
            q[a] = static_cast<uint32_t>(sc);
        }
        code[k] = {expand10(q[0]) * 4u + expand10(q[1]) * 2u + expand10(q[2]), (uint32_t)k};
    }
    std::stable_sort(code.begin(), code.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    std::vector<Collider> out;
    for (auto it = code.rbegin(); it != code.rend(); ++it) out.push_back(all[it->second]);
    return out;
}

uint32_t resolve_iters(uint32_t v, uint32_t d) { return v > 0 ? v : d; }

// Smallest float T with sqrt(T) >= h under the correctly rounded IEEE square root (the host's
// std::sqrt and the kernels' __builtin_sqrtf agree), so that for every float d2,
// sqrt(d2) < h  <=>  d2 < T.  The rounded root is monotone, hence the set below h is an interval.
float sqrt_threshold(float h)
{
    if (!(h == h)) return h;
    float x = h * h;
    while (x > 0.0f && std::sqrt(x) >= h) x = std::nextafter(x, 0.0f);
    while (!(std::sqrt(x) >= h)) x = std::nextafter(x, HUGE_VALF);
    return x;
}

struct ProfScope {
    cn_ctx *ctx;
    int slot;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(cn_ctx *c, const char *name) : ctx(c), slot(-1)
    {
        if (!ctx->profiling) return;
        cn_kernel_stats &st = ctx->stats;
        for (uint32_t k = 0; k < st.used; ++k)
            if (std::strncmp(st.name[k], name, 31) == 0) slot = (int)k;
        if (slot < 0 && st.used < CN_KERNEL_SLOTS) {
            slot = (int)st.used++;
            std::snprintf(st.name[slot], 32, "%s", name);
        }
        if (slot < 0) return;
        a = take();
        b = take();
        if (!a || !b) {
            a = nullptr;
            return;
        }
        (void)hipEventRecord(a, ctx->stream);
    }
    // Events without the system-scope fences: they only time work on the context's stream.
    hipEvent_t take()
    {
        if (ctx->ev_used == ctx->ev_pool.size()) {
            hipEvent_t e = nullptr;
            if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
            ctx->ev_pool.push_back(e);
        }
        return ctx->ev_pool[ctx->ev_used++];
    }
    ~ProfScope()
    {
        if (slot < 0 || !a) return;
        (void)hipEventRecord(b, ctx->stream);
        ctx->events.push_back({slot, a, b});
    }
};

uint32_t key_bits_of(const cn_ctx *ctx, const cn_step_params *p)
{
    return p->key_mode == CN_KEY_HASH ? 32u : 3u * ctx->morton_bits + ctx->scene_bits;
}

SubstepConsts make_consts(const cn_ctx *ctx, const SceneBuffers &s, const cn_step_params *p, float dt)
{
    SubstepConsts c{};
    c.dt = dt;
    // 0: each scene's derived cell size (SceneInfo); > 0: the caller's cell size for every scene
    c.inv_cell = p->cell_size > 0.0f ? 1.0f / std::max(p->cell_size, 1.0e-4f) : 0.0f;
    c.gravity[0] = p->gravity[0];
    c.gravity[1] = p->gravity[1];
    c.gravity[2] = p->gravity[2];
    c.n = s.n;
    c.npid = s.npid;
    c.cap = p->max_fluid_neighbors;
    c.key_mode = p->key_mode;
    c.morton_bits = ctx->morton_bits;
    c.bucket_shift = key_bits_of(ctx, p) - s.bucket_bits;
    c.uniform_class = s.nclasses == 1 ? 1u : 0u;
    c.fluid_uniform = s.fluid_uniform;
    c.tension = ctx->has_tension ? 1u : 0u;
    c.tile_max = ctx->tile_max;
    c.blk_cap = s.blk_cap;
    c.grid_max = ctx->grid_max;
    c.superset = (p->flags & CN_STEP_CANDIDATE_REUSE) && s.slist ? 1u : 0u;
    c.fluid_iters = p->fluid_iterations > 0 ? p->fluid_iterations : std::max<uint32_t>(p->default_iterations, 1u);
    c.fluid_scenes = ctx->fluid_scenes;
    fluid_tiles(c);
    c.cfl_guard = 1.001f;
    std::memcpy(c.cfg, p->solver_config, sizeof(c.cfg));
    return c;
}



// predict + keys + sort + runs + reorder for one substep; returns the sorted (key, old slot) pairs
uint2 *broadphase(cn_ctx *ctx, SceneBuffers &s, const SubstepConsts &c, uint32_t key_bits)
{
    {
        ProfScope ps(ctx, "keys");
        launch_keys(s, c, ctx->stream);
    }
    uint2 *sorted;
    {
        ProfScope ps(ctx, "sort");
        sorted = radix_sort_pairs(s.kv[0], s.kv[1], c.n, key_bits, s.sort_scratch, ctx->stream);
    }
    {
        ProfScope ps(ctx, "runs");
        launch_runs(s, c, sorted, key_bits, ctx->stream);
    }
    {
        ProfScope ps(ctx, "reorder");
        launch_reorder(s, c, ctx->stream);
    }
    if (c.key_mode == CN_KEY_MORTON) {
        ProfScope ps(ctx, "blocks");
        launch_blocks(s, c, ctx->stream);
        if (ctx->has_fluid) launch_blk_halo(s, c, ctx->stream);
    }
    return sorted;
}

}  // namespace

extern "C" {

uint32_t cn_abi_version(void) { return CN_ABI_VERSION; }

cn_status cn_ctx_create(const cn_device_desc *desc, cn_ctx **out)
{
    if (!out) return CN_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    cn_ctx *ctx = new cn_ctx();
    ctx->device = desc ? desc->device_ordinal : 0;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= ctx->device) {
        delete ctx;
        return CN_ERR_DEVICE;
    }
    if (hipSetDevice(ctx->device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return CN_ERR_DEVICE;
    }
    if (!resolve_noalloc_scratch(ctx->device)) {  // the no-allocation entry points' scratch on this GPU
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return CN_ERR_DEVICE;
    }
    *out = ctx;
    return CN_OK;
}

void cn_ctx_destroy(cn_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    clear_graphs(ctx);
    free_scene(ctx->s);
    free_scene(ctx->tick);
    dfree(ctx->tick_seg);
    dfree(ctx->stage);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *cn_last_error(const cn_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void cn_default_step_params(cn_step_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    // PhysicsSolverDesc defaults, include/physics/physics_solver.h:21-143, in makeSolverConfig
    // order (src/physics/physics_solver.cpp:100-129)
    p->dt = 1.0f / 60.0f;
    p->substeps = 1;
    p->default_iterations = 20;
    p->gravity[1] = -9.81f;
    const float cfg[CN_SOLVER_CONFIG_FLOATS] = {1.0e-3f, 4.0f,    0.95f,  0.05f,  0.5f,  2.0f,  0.0f,  1.0e-5f,
                                                0.25f,   0.90f,   0.10f,  1.0f,   0.01f, 0.10f, 0.8f,  0.5f,
                                                0.95f,   0.7f,    0.05f,  0.02f,  0.12f, 1.0e-5f, 5.0e-5f, 1.0e-5f,
                                                0.2f,    0.0f,    0.0f,   0.0f,   0.0f,  0.0f,  0.0f,  0.0f,
                                                0.7f,    1.0f,    0.12f,  0.08f,  0.0f,  0.0f,  0.0f,  0.0f};
    std::memcpy(p->solver_config, cfg, sizeof(cfg));
    p->max_fluid_neighbors = 128;  // kDefaultFluidMaxNeighborhood, src/physics/rigid_body_common.h:304
    p->key_mode = CN_KEY_MORTON;
    p->flags = CN_STEP_CANDIDATE_REUSE;
}

// toFluidSolverMaterial and its helpers, src/physics/physics_world.cpp:99-263
cn_status cn_fluid_material_from_desc(const cn_fluid_material_desc *desc, float r, cn_fluid_material *out)
{
    if (!desc || !out) return CN_ERR_INVALID_ARGUMENT;
    const float pi = 3.14159265358979323846f;
    const float rest = std::max(2.0f * r, 1.0e-4f);
    const float h = std::max((2.0f / 0.6f) * r, rest);
    auto w = [&](float d) {
        const float omq = 1.0f - d / h;
        const float k = 15.0f / (pi * h * h * h);
        return k * omq * omq;
    };
    auto dw = [&](float d) {
        const float omq = 1.0f - d / h;
        const float k = 15.0f / (pi * h * h * h);
        return -k * 2.0f * omq / h;
    };
    // tightPack3D: hexagonal-ish pack of spacing `rest` inside radius h, at most 2048 samples
    const int dim = static_cast<int>(std::ceil(h / rest));
    const float row = std::sqrt(0.75f) * rest;
    float rho = 0.0f, dcs = 0.0f, a = 0.0f, b = 0.0f;
    uint32_t count = 0;
    bool full = false;
    for (int z = -dim; z <= dim && !full; ++z)
        for (int y = -dim; y <= dim && !full; ++y)
            for (int x = -dim; x <= dim; ++x) {
                const float sx = static_cast<float>(x) * rest + (((y + z) & 1) != 0 ? rest * 0.5f : 0.0f);
                const float sy = static_cast<float>(y) * row;
                const float sz = static_cast<float>(z) * row;
                const float m2 = sx * sx + sy * sy + sz * sz;
                if (m2 == 0.0f || !(std::sqrt(m2) <= h)) continue;
                if (count == 2048u) {
                    full = true;
                    break;
                }
                ++count;
                const float dist = std::sqrt(sx * sx + sy * sy + sz * sz);
                rho += w(dist);
                const float der = dw(dist);
                dcs += der * der;
                if (sy <= 0.0f && dist > 0.0f) {
                    const float ct = sy / dist;
                    a += der * ct;
                    b -= dist * ct;
                }
            }
    if (rho <= 1.0e-6f) rho = 1.0f;
    if (dcs <= 1.0e-6f) dcs = 1.0f;
    const float sts = (std::fabs(b) > 1.0e-6f && std::fabs(a) > 1.0e-6f) ? std::fabs(a / b) : 1.0f;
    const float inv = 1.0f / rho;
    const float rr = std::min(std::max(rest / h, 1.0e-3f), 0.999f);
    const float rr2 = rr * rr;
    out->rest_density = rho;
    out->inv_rest_density = inv;
    out->smoothing_radius = h;
    out->density_constraint_scale = 1.0f / std::max(dcs, 1.0e-6f);
    out->viscosity = desc->viscosity * inv;
    out->cohesion = desc->cohesion * h;
    out->cohesion1 = -(1.0f + rr) / rr2;
    out->cohesion2 = (rr2 + rr + 1.0f) / rr2;
    out->surface_tension = (desc->surface_tension * inv) / sts;
    out->vorticity_confinement = desc->vorticity_confinement * inv * h;
    out->gravity_scale = desc->gravity_scale;
    out->cfl_radius = desc->cfl_coefficient * h;
    return CN_OK;
}

float cn_particle_grid_cell_size(uint32_t n, const float *radii, const uint32_t *fmi, uint32_t nm,
                                 const cn_fluid_material *fm)
{
    float g = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        g = std::max(g, radii[i] * 2.0f);
        if (fmi && fmi[i] != 0xffffffffu && fmi[i] < nm) g = std::max(g, fm[fmi[i]].smoothing_radius);
    }
    return std::max(g, 1.0e-4f);
}

cn_status cn_scene_reserve(cn_ctx *ctx, uint32_t n, uint32_t cap)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    if (n <= ctx->capacity && cap <= ctx->cap_alloc && ctx->s.pos[0]) return CN_OK;
    if ((uint64_t)list_slots(n ? n : 1) * (cap ? cap : 1) * sizeof(uint16_t) > kListBytesMax)
        return fail(ctx, CN_ERR_CAPACITY, "neighbor lists of n x max_fluid_neighbors entries exceed 4 GiB less 1 MiB");
    CN_HIP(ctx, hipSetDevice(ctx->device));
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_scene(ctx->s);
    free_scene(ctx->tick);
    ctx->tick_cap = 0;
    ctx->uploaded = false;
    ctx->graph_valid = false;
    SceneBuffers &s = ctx->s;
    const size_t N = n ? n : 1;
    for (int k = 0; k < 2; ++k) {
        CN_HIP(ctx, dalloc(&s.pos[k], N));
        CN_HIP(ctx, dalloc(&s.vel[k], N));
        CN_HIP(ctx, dalloc(&s.cls[k], N));
        CN_HIP(ctx, dalloc(&s.id[k], N));
        CN_HIP(ctx, dalloc(&s.kv[k], N));
        CN_HIP(ctx, dalloc(&s.dacc[k], N));
    }
    CN_HIP(ctx, dalloc(&s.prev[0], N));
    CN_HIP(ctx, dalloc(&s.prev[1], N));
    CN_HIP(ctx, dalloc(&s.cellc, N));
    CN_HIP(ctx, dalloc(&s.bnd, N));
    CN_HIP(ctx, dalloc(&s.sort_scratch, sort_scratch_words((uint32_t)N)));
    CN_HIP(ctx, dalloc(&s.elem_run, N));
    CN_HIP(ctx, dalloc(&s.run_key, N));
    CN_HIP(ctx, dalloc(&s.run_start, N + 1));
    CN_HIP(ctx, dalloc(&s.run_info, N));
    const uint32_t bb = std::max<uint32_t>(4u, ceil_log2(N) > 1 ? ceil_log2(N) - 1 : 1);
    s.bucket_bits = std::min<uint32_t>(bb, 30u);
    ctx->bucket_bits_alloc = s.bucket_bits;
    CN_HIP(ctx, dalloc(&s.bucket, (size_t(1) << s.bucket_bits) + 1));
    CN_HIP(ctx, dalloc(&s.block_runs, (N + 1023) / 1024 + 1));
    CN_HIP(ctx, dalloc(&s.run_count, 1));
    CN_HIP(ctx, dalloc(&s.perm, N));
    CN_HIP(ctx, dalloc(&s.nc, N));
    CN_HIP(ctx, dalloc(&s.curpos, N));
    CN_HIP(ctx, dalloc(&s.cs, N));
    CN_HIP(ctx, dalloc(&s.blk_first_run, N + 1));
    CN_HIP(ctx, dalloc(&s.nblk, 2));
    // halo descriptors for up to one cell block per 32 particles; sparser scenes walk the rest of
    // their blocks through the run table
    s.blk_cap = (uint32_t)std::max<size_t>(256, N / 32);
    CN_HIP(ctx, dalloc(&s.blk_halo, (size_t)s.blk_cap * kHaloStride));
    CN_HIP(ctx, dalloc(&s.wide_blk, s.blk_cap));
    CN_HIP(ctx, dalloc(&s.list, list_slots(N) * (cap ? cap : 1)));
    CN_HIP(ctx, dalloc(&s.lcnt, N));
    CN_HIP(ctx, dalloc(&s.iter_state, kStateWords));
    CN_HIP(ctx, hipMemsetAsync(s.iter_state, 0, kStateWords * sizeof(uint32_t), ctx->stream));
    CN_HIP(ctx, dalloc(&s.out_pos, N));
    CN_HIP(ctx, dalloc(&s.out_vel, N));
    CN_HIP(ctx, dalloc(&s.out_prev, N));
    CN_HIP(ctx, dalloc(&s.meta, kMetaWords));
    CN_HIP(ctx, hipMemsetAsync(s.meta, 0, kMetaWords * sizeof(uint32_t), ctx->stream));
    CN_HIP(ctx, hipMemsetAsync(s.dacc[0], 0, N * sizeof(float4), ctx->stream));
    CN_HIP(ctx, hipMemsetAsync(s.dacc[1], 0, N * sizeof(float4), ctx->stream));
    ctx->capacity = n;
    ctx->cap_alloc = cap;
    s.cap = cap;
    return CN_OK;
}

}  // extern "C"

namespace {

// The candidate groups of a scene of more than kMaskColliders colliders: kColGroup consecutive
// colliders of its visit order (the enabled ones, sc[0 .. ncol)) and their AABB union; `slots`
// groups are written (those past the colliders empty), so edits that enable colliders fit.
std::vector<ColGroup> collider_groups(const std::vector<Collider> &sc, uint32_t ncol, uint32_t slots)
{
    std::vector<ColGroup> g(slots);
    for (uint32_t k = 0; k < slots; ++k) {
        ColGroup &G = g[k];
        G.first = k * kColGroup;
        G.count = G.first < ncol ? std::min<uint32_t>(kColGroup, ncol - G.first) : 0u;
        for (int a = 0; a < 3; ++a) {
            G.mn[a] = 3.402823466e+38f;
            G.mx[a] = -3.402823466e+38f;
        }
        for (uint32_t q = G.first; q < G.first + G.count; ++q)
            for (int a = 0; a < 3; ++a) {
                G.mn[a] = std::min(G.mn[a], sc[q].aabb_min[a]);
                G.mx[a] = std::max(G.mx[a], sc[q].aabb_max[a]);
            }
    }
    return g;
}

uint32_t group_slots(uint32_t collider_slots)
{
    return collider_slots > kMaskColliders ? (collider_slots + kColGroup - 1) / kColGroup : 0u;
}

cn_status check_scene(cn_ctx *ctx, const cn_scene_desc &d)
{
    const cn_particle_soa_view *pv = d.particles;
    const cn_scene_tables *tb = d.tables;
    const cn_constraint_view *cv = d.constraints;
    if (!pv || !tb) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null argument");
    const uint32_t n = pv->count;
    if (n == 0) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "empty particle pool");
    if (!pv->positions_inv_mass || !pv->velocities || !pv->radii || !pv->kinds || !pv->particle_material_indices)
        return fail(ctx, CN_ERR_INVALID_ARGUMENT, "missing particle arrays");
    if (cv) {
        if ((cv->edge_count && !cv->edges) || (cv->bend_count && !cv->bends) ||
            (cv->dihedral_count && !cv->dihedrals) || (cv->tet_count && !cv->tets))
            return fail(ctx, CN_ERR_INVALID_ARGUMENT, "constraint count without array");
        for (uint32_t e = 0; e < cv->edge_count; ++e)
            if (cv->edges[e].a >= n || cv->edges[e].b >= n)
                return fail(ctx, CN_ERR_INVALID_ARGUMENT, "edge particle index out of range");
        for (uint32_t b = 0; b < cv->bend_count; ++b)
            if (cv->bends[b].p0 >= n || cv->bends[b].p1 >= n || cv->bends[b].p2 >= n)
                return fail(ctx, CN_ERR_INVALID_ARGUMENT, "bend particle index out of range");
        for (uint32_t q = 0; q < cv->dihedral_count; ++q) {
            const cn_cloth_dihedral &D = cv->dihedrals[q];
            if (D.e0 >= n || D.e1 >= n || D.o0 >= n || D.o1 >= n)
                return fail(ctx, CN_ERR_INVALID_ARGUMENT, "dihedral particle index out of range");
        }
        for (uint32_t t = 0; t < cv->tet_count; ++t)
            for (int k = 0; k < 4; ++k)
                if (cv->tets[t].p[k] >= n) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "tet particle index out of range");
    }
    if ((uint64_t)tb->collider_count + tb->shape_collider_count > kMaxColliders)
        return fail(ctx, CN_ERR_UNSUPPORTED, "more than 65535 static colliders in one scene");
    if (tb->shape_collider_count && !tb->shape_colliders) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null shape colliders");
    for (uint32_t k = 0; k < tb->shape_collider_count; ++k)
        if (tb->shape_colliders[k].shape > CN_SHAPE_CAPSULE)
            return fail(ctx, CN_ERR_INVALID_ARGUMENT, "unknown collider shape");
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t kind = pv->kinds[i];
        const uint32_t fm = pv->fluid_material_indices ? pv->fluid_material_indices[i] : 0xffffffffu;
        if (kind == CN_PARTICLE_KIND_FLUID && fm >= tb->fluid_material_count)
            return fail(ctx, CN_ERR_INVALID_ARGUMENT, "fluid particle with an invalid fluid material index");
        if (pv->particle_material_indices[i] >= tb->particle_material_count)
            return fail(ctx, CN_ERR_INVALID_ARGUMENT, "invalid particle material index");
    }
    if (pv->adjacency_offsets) {  // constraint adjacency: every list inside the index array, in range
        if (!pv->adjacency_counts || (pv->adjacency_index_count && !pv->adjacency_indices))
            return fail(ctx, CN_ERR_INVALID_ARGUMENT, "adjacency offsets without counts or indices");
        for (uint32_t i = 0; i < n; ++i)
            if ((uint64_t)pv->adjacency_offsets[i] + pv->adjacency_counts[i] > pv->adjacency_index_count)
                return fail(ctx, CN_ERR_INVALID_ARGUMENT, "adjacency list past the index array");
        for (uint32_t k = 0; k < pv->adjacency_index_count; ++k)
            if (pv->adjacency_indices[k] >= n)
                return fail(ctx, CN_ERR_INVALID_ARGUMENT, "adjacency particle index out of range");
    }
    return CN_OK;
}

// Uploads a pool of independent scenes: particles, classes, materials and constraints are
// concatenated with index offsets, colliders are visit-ordered per scene, and each scene gets its
// own cell size and key range.  One scene is the plain path.
cn_status upload_pool(cn_ctx *ctx, uint32_t count, const cn_scene_desc *descs)
{
    if (!ctx || !descs || count == 0) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null argument or no scene");
    uint64_t total = 0, refs = 0, ne = 0, nb = 0, nd = 0, nt = 0;
    for (uint32_t k = 0; k < count; ++k) {
        const cn_status st = check_scene(ctx, descs[k]);
        if (st != CN_OK) return st;
        total += descs[k].particles->count;
        if (const cn_constraint_view *cv = descs[k].constraints) {
            ne += cv->edge_count;
            nb += cv->bend_count;
            nd += cv->dihedral_count;
            nt += cv->tet_count;
        }
    }
    refs = 2 * ne + 3 * nb + 4 * nd + 4 * nt;
    if (total >= (1ull << 31)) return fail(ctx, CN_ERR_UNSUPPORTED, "more than 2^31 particles");
    if (ne >= (1u << 30) || nb + nd >= (1u << 30) || nt >= (1u << 30) || refs >= (1ull << 32))
        return fail(ctx, CN_ERR_UNSUPPORTED, "too many constraints");
    const uint32_t n = (uint32_t)total;
    if (n > ctx->capacity || !ctx->s.pos[0]) {
        const cn_status st = cn_scene_reserve(ctx, n, std::max<uint32_t>(ctx->cap_alloc, 128u));
        if (st != CN_OK) return st;
    }
    SceneBuffers &s = ctx->s;
    CN_HIP(ctx, hipSetDevice(ctx->device));
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s.n = n;
    s.npid = n;
    s.cur = 0;
    ctx->slots_by_prefix = false;

    // per-scene tables, concatenated
    std::vector<float> pos(4ull * n), vel(4ull * n);
    std::vector<cn_fluid_material> fmh;
    std::vector<float> pmh;
    std::vector<Collider> col;
    std::vector<ColGroup> groups;  // candidate groups of the scenes of more than kMaskColliders colliders
    std::vector<SceneInfo> sinfo(count);
    std::vector<cn_soft_edge> edges;
    std::vector<cn_soft_bend> bends;
    std::vector<cn_cloth_dihedral> dih;
    std::vector<cn_soft_tet> tets;
    std::map<std::array<uint32_t, 9>, uint32_t> cmap;
    std::vector<ParticleClass> classes;
    std::vector<uint32_t> cls(n);
    // constraint adjacency of the pool (pool particle indices), built only when a scene has one
    bool any_adj = false;
    for (uint32_t k = 0; k < count; ++k) any_adj |= descs[k].particles->adjacency_offsets != nullptr;
    std::vector<uint32_t> adj_off, adj_cnt, adj_idx;
    if (any_adj) {
        adj_off.assign(n, 0u);
        adj_cnt.assign(n, 0u);
    }
    ctx->scene_first.assign(count + 1, 0);
    ctx->scene_cell.assign(count, 0.0f);
    ctx->scene_col_cap.assign(count, 0u);
    ctx->scene_shapes.assign(count, {});
    ctx->scene_fm.assign(2 * (size_t)count, 0u);
    ctx->scene_con.assign(4 * (size_t)count, 0u);
    ctx->scene_pm.assign(2 * (size_t)count, 0u);
    ctx->has_fluid = false;
    ctx->has_solid = false;
    ctx->any_pairs = false;
    ctx->any_rigid = false;
    ctx->any_rigid_pid = false;
    ctx->has_vorticity = false;
    ctx->aniso_ready = false;
    ctx->fm_fluid_used.clear();
    uint32_t base = 0;
    double max_span = 1.0;
    for (uint32_t k = 0; k < count; ++k) {
        const cn_particle_soa_view *pv = descs[k].particles;
        const cn_scene_tables *tb = descs[k].tables;
        const cn_constraint_view *cv = descs[k].constraints;
        const uint32_t m = pv->count;
        const uint32_t fm0 = (uint32_t)fmh.size(), pm0 = (uint32_t)(pmh.size() / 4);
        ctx->scene_fm[2 * k] = fm0;
        ctx->scene_fm[2 * k + 1] = tb->fluid_material_count;
        ctx->scene_pm[2 * k] = pm0;
        ctx->scene_pm[2 * k + 1] = tb->particle_material_count;
        fmh.insert(fmh.end(), tb->fluid_materials, tb->fluid_materials + tb->fluid_material_count);
        ctx->fm_fluid_used.resize(fmh.size(), 0u);
        pmh.insert(pmh.end(), tb->particle_materials, tb->particle_materials + 4 * tb->particle_material_count);
        std::memcpy(pos.data() + 4ull * base, pv->positions_inv_mass, 4ull * m * sizeof(float));
        std::memcpy(vel.data() + 4ull * base, pv->velocities, 4ull * m * sizeof(float));
        std::array<uint32_t, 9> last{};
        uint32_t last_id = 0xffffffffu;
        bool scene_solid = false, scene_fluid = false, two_groups = false, self_group = false;
        uint32_t group0 = 0xffffffffu;
        if (pv->adjacency_offsets)
            for (uint32_t i = 0; i < m; ++i) {
                adj_off[base + i] = (uint32_t)adj_idx.size();
                adj_cnt[base + i] = pv->adjacency_counts[i];
                for (uint32_t q = 0; q < pv->adjacency_counts[i]; ++q)
                    adj_idx.push_back(base + pv->adjacency_indices[pv->adjacency_offsets[i] + q]);
            }
        for (uint32_t i = 0; i < m; ++i) {
            ParticleClass pc{};
            pc.radius = pv->radii[i];
            pc.kind = pv->kinds[i];
            const uint32_t fmi = pv->fluid_material_indices ? pv->fluid_material_indices[i] : 0xffffffffu;
            pc.fluid_mat = fmi < tb->fluid_material_count ? fm0 + fmi : 0xffffffffu;
            pc.part_mat = pm0 + pv->particle_material_indices[i];
            pc.env = pv->environment_indices ? pv->environment_indices[i] : 0u;
            pc.phase = pv->phases ? pv->phases[i] : 0u;
            pc.layer = pv->collision_layers ? pv->collision_layers[i] : 1u;
            pc.mask = pv->collision_masks ? pv->collision_masks[i] : 0xffffffffu;
            pc.scene = k;
            ctx->has_fluid |= pc.kind == CN_PARTICLE_KIND_FLUID;
            if (pc.kind == CN_PARTICLE_KIND_FLUID && fmi < tb->fluid_material_count) {
                ctx->has_vorticity |= std::max(tb->fluid_materials[fmi].vorticity_confinement, 0.0f) > kEps;
                ctx->fm_fluid_used[fm0 + fmi] = 1u;
            }
            ctx->has_solid |= pc.kind == CN_PARTICLE_KIND_SOLID;
            scene_solid |= pc.kind == CN_PARTICLE_KIND_SOLID;
            scene_fluid |= pc.kind == CN_PARTICLE_KIND_FLUID;
            if (pc.kind == CN_PARTICLE_KIND_SOLID) {
                const uint32_t g = pc.phase & 0x7fffffffu;
                if (group0 == 0xffffffffu) group0 = g;
                two_groups |= g != group0;
                self_group |= (pc.phase & 0x80000000u) != 0u;
            }
            std::array<uint32_t, 9> key;
            std::memcpy(key.data(), &pc, sizeof(pc));
            if (last_id != 0xffffffffu && key == last) {
                cls[base + i] = last_id;
                continue;
            }
            auto it = cmap.find(key);
            uint32_t id;
            if (it == cmap.end()) {
                id = (uint32_t)classes.size();
                cmap.emplace(key, id);
                classes.push_back(pc);
            } else {
                id = it->second;
            }
            cls[base + i] = id;
            last = key;
            last_id = id;
        }
        // static colliders in the scene's own LBVH visit order
        const std::vector<Collider> sc =
            visit_ordered_colliders(tb->colliders, tb->collider_count, tb->shape_colliders, tb->shape_collider_count);
        ctx->scene_shapes[k].assign(tb->shape_colliders, tb->shape_colliders + tb->shape_collider_count);
        sinfo[k].col0 = (uint32_t)col.size();
        sinfo[k].ncol = (uint32_t)sc.size();
        sinfo[k].fluid_only = (!scene_solid && (!cv || (cv->edge_count | cv->bend_count | cv->dihedral_count |
                                                         cv->tet_count) == 0u))
                                  ? 1u
                                  : 0u;
        sinfo[k].has_fluid = scene_fluid ? 1u : 0u;
        sinfo[k].pairs = (scene_solid && scene_fluid) || two_groups || self_group ? 1u : 0u;
        ctx->any_pairs |= sinfo[k].pairs != 0u;
        // hasSoftRigidContactWork, physics_solver.cpp:486-488, for a scene whose solid particles
        // can meet a static collider (disabled colliders count: an edit may enable them)
        sinfo[k].rigid = scene_solid && tb->collider_count + tb->shape_collider_count > 0 ? 1u : 0u;
        // a scene without particle-particle pairs whose every particle is constrained solves its
        // particle-rigid contacts in particle-index order, in the soft phase
        if (sinfo[k].rigid && !sinfo[k].pairs && cv) {
            std::vector<uint8_t> held(m, 0u);
            for (uint32_t e = 0; e < cv->edge_count; ++e) held[cv->edges[e].a] = held[cv->edges[e].b] = 1u;
            for (uint32_t b = 0; b < cv->bend_count; ++b)
                held[cv->bends[b].p0] = held[cv->bends[b].p1] = held[cv->bends[b].p2] = 1u;
            for (uint32_t d = 0; d < cv->dihedral_count; ++d) {
                const cn_cloth_dihedral &D = cv->dihedrals[d];
                held[D.e0] = held[D.e1] = held[D.o0] = held[D.o1] = 1u;
            }
            for (uint32_t t = 0; t < cv->tet_count; ++t)
                for (int q = 0; q < 4; ++q) held[cv->tets[t].p[q]] = 1u;
            if (std::all_of(held.begin(), held.end(), [](uint8_t v) { return v != 0u; })) {
                sinfo[k].rigid = 0u;
                sinfo[k].rigid_pid = 1u;
            }
        }
        ctx->any_rigid |= sinfo[k].rigid != 0u;
        ctx->any_rigid_pid |= sinfo[k].rigid_pid != 0u;
        {
            const uint32_t slots = tb->collider_count + tb->shape_collider_count;
            const std::vector<ColGroup> g = collider_groups(sc, (uint32_t)sc.size(), group_slots(slots));
            sinfo[k].grp0 = (uint32_t)groups.size();
            groups.insert(groups.end(), g.begin(), g.end());
        }
        col.insert(col.end(), sc.begin(), sc.end());
        // slots for the disabled colliders too, so cn_scene_set_colliders can enable them in place
        col.resize(col.size() + (tb->collider_count + tb->shape_collider_count - sc.size()));
        ctx->scene_col_cap[k] = tb->collider_count;
        // constraints, shifted to pool indices
        ctx->scene_con[4 * k] = (uint32_t)edges.size();
        ctx->scene_con[4 * k + 1] = (uint32_t)bends.size();
        ctx->scene_con[4 * k + 2] = (uint32_t)dih.size();
        ctx->scene_con[4 * k + 3] = (uint32_t)tets.size();
        if (cv) {
            for (uint32_t e = 0; e < cv->edge_count; ++e) {
                cn_soft_edge E = cv->edges[e];
                E.a += base;
                E.b += base;
                edges.push_back(E);
            }
            for (uint32_t q = 0; q < cv->bend_count; ++q) {
                cn_soft_bend B = cv->bends[q];
                B.p0 += base;
                B.p1 += base;
                B.p2 += base;
                bends.push_back(B);
            }
            for (uint32_t q = 0; q < cv->dihedral_count; ++q) {
                cn_cloth_dihedral D = cv->dihedrals[q];
                D.e0 += base;
                D.e1 += base;
                D.o0 += base;
                D.o1 += base;
                dih.push_back(D);
            }
            for (uint32_t t = 0; t < cv->tet_count; ++t) {
                cn_soft_tet T = cv->tets[t];
                for (int j = 0; j < 4; ++j) T.p[j] += base;
                tets.push_back(T);
            }
        }
        // grid cell size of the scene (recomputeParticleGridCellSize) and its extent in cells
        const float cell = cn_particle_grid_cell_size(m, pv->radii, pv->fluid_material_indices,
                                                      tb->fluid_material_count, tb->fluid_materials);
        ctx->scene_cell[k] = cell;
        sinfo[k].inv_cell = 1.0f / std::max(cell, 1.0e-4f);
        float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        for (uint32_t i = 0; i < m; ++i)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], pv->positions_inv_mass[4 * i + a]);
                hi[a] = std::max(hi[a], pv->positions_inv_mass[4 * i + a]);
            }
        for (const Collider &c : sc)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], c.aabb_min[a]);
                hi[a] = std::max(hi[a], c.aabb_max[a]);
            }
        for (int a = 0; a < 3; ++a) max_span = std::max(max_span, (double)(hi[a] - lo[a]) / (double)cell + 3.0);
        ctx->scene_first[k] = base;
        base += m;
    }
    ctx->scene_first[count] = base;
    ctx->cell_size = ctx->scene_cell[0];
    // key layout: [scene index | 3 x morton_bits]; coordinates beyond 2^morton_bits cells wrap and
    // their runs are filtered per particle
    ctx->scene_bits = count > 1 ? ceil_log2(count) : 0u;
    const uint32_t room = (32u - ctx->scene_bits) / 3u;
    ctx->morton_bits = std::min<uint32_t>(std::min<uint32_t>(10u, room),
                                          std::max<uint32_t>(1u, ceil_log2((uint64_t)std::ceil(max_span))));
    // key prefixes: scenes with fluid first (in upload order), so the fluid passes' blocks are a
    // prefix of the block list
    ctx->fluid_scenes = 0;
    for (uint32_t k = 0; k < count; ++k) ctx->fluid_scenes += sinfo[k].has_fluid;
    {
        uint32_t rf = 0, rs = ctx->fluid_scenes;
        std::vector<uint32_t> by_rank(count);
        for (uint32_t k = 0; k < count; ++k) {
            const uint32_t rank = sinfo[k].has_fluid ? rf++ : rs++;
            sinfo[k].key_prefix = ctx->scene_bits ? rank << (3u * ctx->morton_bits) : 0u;
            by_rank[rank] = k;
        }
        // once sorted, the slots hold the scenes in rank order, each as one segment
        ctx->slot_base.assign(count, 0u);
        uint32_t first = 0;
        for (uint32_t r = 0; r < count; ++r) {
            const uint32_t k = by_rank[r];
            ctx->slot_base[k] = first;
            first += ctx->scene_first[k + 1] - ctx->scene_first[k];
        }
    }
    ctx->nscenes = count;

    dfree(s.classes);
    dfree(s.fluid_mats);
    dfree(s.part_mats);
    dfree(s.colliders);
    dfree(s.scenes);
    dfree(s.fluid_t2);
    dfree(s.scene_active);
    dfree(s.pid_scene);
    s.scene_active = nullptr;
    s.pid_scene = nullptr;
    s.classes = nullptr;
    s.fluid_mats = nullptr;
    s.part_mats = nullptr;
    s.colliders = nullptr;
    s.scenes = nullptr;
    s.fluid_t2 = nullptr;
    dfree(s.adj_off);
    dfree(s.adj_cnt);
    dfree(s.adj_idx);
    s.adj_off = s.adj_cnt = s.adj_idx = nullptr;
    if (any_adj) {
        CN_HIP(ctx, dalloc(&s.adj_off, n));
        CN_HIP(ctx, dalloc(&s.adj_cnt, n));
        CN_HIP(ctx, dalloc(&s.adj_idx, adj_idx.size()));
        CN_HIP(ctx, hipMemcpy(s.adj_off, adj_off.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
        CN_HIP(ctx, hipMemcpy(s.adj_cnt, adj_cnt.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
        if (!adj_idx.empty())
            CN_HIP(ctx, hipMemcpy(s.adj_idx, adj_idx.data(), adj_idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    s.nclasses = (uint32_t)classes.size();
    {
        // the fluid passes take a cell block's particle class from its first particle when no
        // scene with fluid mixes classes (a block holds cells of one scene)
        std::vector<uint32_t> per_scene(count, 0u);
        for (const ParticleClass &pc : classes) ++per_scene[pc.scene];
        s.fluid_uniform = 1u;
        for (uint32_t k = 0; k < count; ++k)
            if (sinfo[k].has_fluid && per_scene[k] != 1u) s.fluid_uniform = 0u;
    }
    CN_HIP(ctx, dalloc(&s.classes, classes.size()));
    CN_HIP(ctx, hipMemcpy(s.classes, classes.data(), classes.size() * sizeof(ParticleClass), hipMemcpyHostToDevice));
    CN_HIP(ctx, dalloc(&s.fluid_mats, fmh.size()));
    if (!fmh.empty())
        CN_HIP(ctx, hipMemcpy(s.fluid_mats, fmh.data(), fmh.size() * sizeof(cn_fluid_material), hipMemcpyHostToDevice));
    CN_HIP(ctx, dalloc(&s.part_mats, pmh.size() / 4));
    if (!pmh.empty()) CN_HIP(ctx, hipMemcpy(s.part_mats, pmh.data(), pmh.size() * sizeof(float), hipMemcpyHostToDevice));
    s.ncolliders = (uint32_t)col.size();
    CN_HIP(ctx, dalloc(&s.colliders, col.size()));
    if (!col.empty())
        CN_HIP(ctx, hipMemcpy(s.colliders, col.data(), col.size() * sizeof(Collider), hipMemcpyHostToDevice));
    // scenes of more than kMaskColliders colliders: their groups, and a candidate list per slot
    dfree(s.col_groups);
    dfree(s.bndl);
    s.col_groups = nullptr;
    s.bndl = nullptr;
    if (!groups.empty()) {
        CN_HIP(ctx, dalloc(&s.col_groups, groups.size()));
        CN_HIP(ctx, hipMemcpy(s.col_groups, groups.data(), groups.size() * sizeof(ColGroup), hipMemcpyHostToDevice));
        CN_HIP(ctx, dalloc(&s.bndl, (size_t)std::max<uint32_t>(ctx->capacity, n) * kBndListCap));
    }
    s.nscenes = count;
    CN_HIP(ctx, dalloc(&s.scenes, sinfo.size()));
    CN_HIP(ctx, hipMemcpy(s.scenes, sinfo.data(), sinfo.size() * sizeof(SceneInfo), hipMemcpyHostToDevice));
    ctx->scene_info = sinfo;
    ctx->fm_host = fmh;
    {
        const std::vector<uint32_t> ones(count, 1u);
        CN_HIP(ctx, dalloc(&s.scene_active, count));
        CN_HIP(ctx, hipMemcpy(s.scene_active, ones.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice));
        ctx->active_host.assign(count, 1u);
        std::vector<uint32_t> ps(n);
        for (uint32_t k = 0; k < count; ++k)
            for (uint32_t i = ctx->scene_first[k]; i < ctx->scene_first[k + 1]; ++i) ps[i] = k;
        CN_HIP(ctx, dalloc(&s.pid_scene, n));
        CN_HIP(ctx, hipMemcpy(s.pid_scene, ps.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    ctx->max_h = 1.0e-4f;
    ctx->has_tension = false;
    std::vector<float> t2(fmh.size());
    for (size_t k = 0; k < fmh.size(); ++k) {
        const cn_fluid_material &m = fmh[k];
        ctx->max_h = std::max(ctx->max_h, m.smoothing_radius);
        ctx->has_tension |= (m.surface_tension > 0.0f ? m.surface_tension : 0.0f) > kEps;
        t2[k] = sqrt_threshold(m.smoothing_radius > 1.0e-4f ? m.smoothing_radius : 1.0e-4f);
    }
    CN_HIP(ctx, dalloc(&s.fluid_t2, t2.size()));
    if (!t2.empty()) CN_HIP(ctx, hipMemcpy(s.fluid_t2, t2.data(), t2.size() * sizeof(float), hipMemcpyHostToDevice));


    std::vector<uint32_t> ids(n);
    for (uint32_t i = 0; i < n; ++i) ids[i] = i;
    CN_HIP(ctx, hipMemcpy(s.pos[0], pos.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    CN_HIP(ctx, hipMemcpy(s.vel[0], vel.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    CN_HIP(ctx, hipMemcpy(s.cls[0], cls.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    CN_HIP(ctx, hipMemcpy(s.id[0], ids.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    CN_HIP(ctx, hipMemcpy(s.out_pos, pos.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    CN_HIP(ctx, hipMemcpy(s.out_vel, vel.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    CN_HIP(ctx, hipMemcpy(s.out_prev, pos.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    CN_HIP(ctx, hipMemcpy(s.prev[0], pos.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    // a fresh scene starts with fresh iteration state (no displacement carried into the first
    // substep's candidate lists)
    CN_HIP(ctx, hipMemset(s.iter_state, 0, kStateWords * sizeof(uint32_t)));

    // contact state: particle-particle candidate lists and the last contacts, kContactPartners
    // per particle, and the particle-rigid generation positions (only a scene with a solid
    // particle has contacts)
    if (ctx->has_solid && s.ct_alloc < n) {
        dfree(s.ct_list);
        dfree(s.ct_count);
        dfree(s.ct_rec);
        dfree(s.ct_bits);
        dfree(s.ct_corr);
        dfree(s.ct_gen);
        s.ct_list = s.ct_count = s.ct_bits = nullptr;
        s.ct_rec = s.ct_corr = s.ct_gen = nullptr;
        s.ct_alloc = 0;
        const size_t m = std::max<size_t>(n, ctx->capacity);
        CN_HIP(ctx, dalloc(&s.ct_list, m * kContactPartners));
        CN_HIP(ctx, dalloc(&s.ct_rec, m * kContactPartners));
        CN_HIP(ctx, dalloc(&s.ct_count, m));
        CN_HIP(ctx, dalloc(&s.ct_bits, m));
        CN_HIP(ctx, dalloc(&s.ct_corr, m));
        CN_HIP(ctx, dalloc(&s.ct_gen, m));
        s.ct_alloc = m;
    }

    // soft constraints, their incidence lists and the particle -> slot map
    free_constraints(s);
    if (!edges.empty() || !bends.empty() || !dih.empty() || !tets.empty()) {
        s.ne = (uint32_t)edges.size();
        s.nb = (uint32_t)bends.size();
        s.nd = (uint32_t)dih.size();
        s.nt = (uint32_t)tets.size();
        CN_HIP(ctx, dalloc(&s.edges, s.ne));
        CN_HIP(ctx, dalloc(&s.bends, s.nb));
        CN_HIP(ctx, dalloc(&s.dihedrals, s.nd));
        CN_HIP(ctx, dalloc(&s.tets, s.nt));
        {
            // device forms: padding dropped, max(compliance, 0) and max(rest volume, 0) taken here,
            // a disabled edge carried as a NaN compliance
            std::vector<EdgeDev> de(s.ne);
            for (uint32_t e = 0; e < s.ne; ++e) de[e] = edge_dev(edges[e], 0);
            std::vector<BendDev> db(s.nb);
            for (uint32_t q = 0; q < s.nb; ++q) db[q] = bend_dev(bends[q], 0);
            std::vector<DihedralDev> dd(s.nd);
            for (uint32_t q = 0; q < s.nd; ++q) dd[q] = dihedral_dev(dih[q], 0);
            std::vector<TetDev> dt(s.nt);
            for (uint32_t t = 0; t < s.nt; ++t) dt[t] = tet_dev(tets[t], 0);
            if (s.ne) CN_HIP(ctx, hipMemcpy(s.edges, de.data(), s.ne * sizeof(EdgeDev), hipMemcpyHostToDevice));
            if (s.nb) CN_HIP(ctx, hipMemcpy(s.bends, db.data(), s.nb * sizeof(BendDev), hipMemcpyHostToDevice));
            if (s.nd) CN_HIP(ctx, hipMemcpy(s.dihedrals, dd.data(), s.nd * sizeof(DihedralDev), hipMemcpyHostToDevice));
            if (s.nt) CN_HIP(ctx, hipMemcpy(s.tets, dt.data(), s.nt * sizeof(TetDev), hipMemcpyHostToDevice));
            ctx->edges_host = std::move(de);
            ctx->bends_host = std::move(db);
            ctx->dih_host = std::move(dd);
            ctx->tets_host = std::move(dt);
        }
        const size_t nlam = (size_t)s.ne + s.nb + s.nd + s.nt;
        CN_HIP(ctx, dalloc(&s.lam, nlam));
        std::vector<uint32_t> inc, map_e, map_b, map_t;
        size_t refs[3] = {0, 0, 0};
        {
            const uint32_t cnt[1] = {s.ne}, ar[1] = {2}, off[1] = {0};
            build_incidence(n, 1, cnt, ar, off, [&](uint32_t, uint32_t c, uint32_t k) { return k == 0 ? edges[c].a : edges[c].b; },
                            inc, s.inc_off[0], s.inc_off[1], map_e, 2u, refs[0]);
        }
        {
            const uint32_t cnt[2] = {s.nb, s.nd}, ar[2] = {3, 4}, off[2] = {0, s.nb};
            build_incidence(n, 2, cnt, ar, off,
                            [&](uint32_t g, uint32_t c, uint32_t k) {
                                if (g == 0) {
                                    const cn_soft_bend &B = bends[c];
                                    return k == 0 ? B.p0 : (k == 1 ? B.p1 : B.p2);
                                }
                                const cn_cloth_dihedral &D = dih[c];
                                return k == 0 ? D.e0 : (k == 1 ? D.e1 : (k == 2 ? D.o0 : D.o1));
                            },
                            inc, s.inc_off[2], s.inc_off[3], map_b, 4u, refs[1]);
        }
        {
            const uint32_t cnt[1] = {s.nt}, ar[1] = {4}, off[1] = {0};
            build_incidence(n, 1, cnt, ar, off, [&](uint32_t, uint32_t c, uint32_t k) { return tets[c].p[k]; }, inc,
                            s.inc_off[4], s.inc_off[5], map_t, 4u, refs[2]);
        }
        // record maps after the incidence structures; records of each group in CSR order + a spare
        s.inc_off[6] = inc.size();
        inc.insert(inc.end(), map_e.begin(), map_e.end());
        s.inc_off[7] = inc.size();
        inc.insert(inc.end(), map_b.begin(), map_b.end());
        s.inc_off[8] = inc.size();
        inc.insert(inc.end(), map_t.begin(), map_t.end());
        s.corr_off[0] = 0;
        s.corr_off[1] = refs[0] + 1;
        s.corr_off[2] = s.corr_off[1] + refs[1] + 1;
        const size_t ncorr = s.corr_off[2] + refs[2] + 1;
        CN_HIP(ctx, dalloc(&s.corr, ncorr));
        CN_HIP(ctx, hipMemset(s.corr, 0, ncorr * sizeof(CorrRec)));
        // constrained particles: any incidence in the three structures
        {
            std::vector<uint32_t> soft;
            for (uint32_t i = 0; i < n; ++i) {
                bool any = false;
                for (int g = 0; g < 3; ++g) any |= inc[s.inc_off[2 * g] + i + 1] != inc[s.inc_off[2 * g] + i];
                if (any) soft.push_back(i);
            }
            s.nsoft = (uint32_t)soft.size();
            CN_HIP(ctx, dalloc(&s.soft_ids, soft.size()));
            CN_HIP(ctx, dalloc(&s.rigid_pick, soft.size()));
            if (!soft.empty())
                CN_HIP(ctx, hipMemcpy(s.soft_ids, soft.data(), soft.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
            CN_HIP(ctx, dalloc(&s.posi, n));
        }
        CN_HIP(ctx, dalloc(&s.inc, inc.size()));
        CN_HIP(ctx, hipMemcpy(s.inc, inc.data(), inc.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        dfree(s.slot_of);
        s.slot_of = nullptr;
        CN_HIP(ctx, dalloc(&s.slot_of, std::max<uint32_t>(ctx->capacity, n)));
        CN_HIP(ctx, hipMemcpy(s.slot_of, ids.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    }

    const uint32_t kb = 3u * ctx->morton_bits + ctx->scene_bits;
    const uint32_t bb = std::max<uint32_t>(4u, ceil_log2(n) > 1 ? ceil_log2(n) - 1 : 1);
    // bucket bits cannot exceed the key bits of either key mode; the allocation fits the largest
    s.bucket_bits = std::min<uint32_t>(std::min<uint32_t>(bb, kb), ctx->bucket_bits_alloc);
    ctx->uploaded = true;
    ctx->stepped = false;
    ctx->graph_valid = false;
    return CN_OK;
}

}  // namespace

extern "C" {

cn_status cn_scene_upload(cn_ctx *ctx, const cn_particle_soa_view *pv, const cn_scene_tables *tb,
                          const cn_constraint_view *cv)
{
    const cn_scene_desc d{pv, tb, cv};
    return upload_pool(ctx, 1, &d);
}

cn_status cn_batch_upload(cn_ctx *ctx, uint32_t scene_count, const cn_scene_desc *scenes)
{
    return upload_pool(ctx, scene_count, scenes);
}

cn_status cn_batch_scene_range(const cn_ctx *ctx, uint32_t scene, uint32_t *first, uint32_t *count)
{
    if (!ctx || !ctx->uploaded || scene >= ctx->nscenes) return CN_ERR_INVALID_ARGUMENT;
    if (first) *first = ctx->scene_first[scene];
    if (count) *count = ctx->scene_first[scene + 1] - ctx->scene_first[scene];
    return CN_OK;
}

}  // extern "C"

namespace {

// Enqueues one cn_step (every substep) on the context stream over the buffer set s (the pool's,
// or a compact tick's view of the listed scenes); advances s's slot-order parity.
// every scene of the context takes part in the step (the active flags last written are all set)
bool all_active(const cn_ctx *ctx)
{
    for (uint32_t f : ctx->active_host)
        if (!f) return false;
    return true;
}

void record_step(cn_ctx *ctx, SceneBuffers &s, const cn_step_params *p)
{
    const uint32_t substeps = std::max<uint32_t>(p->substeps, 1u);
    const uint32_t def = std::max<uint32_t>(p->default_iterations, 1u);
    const uint32_t fluid_iters = resolve_iters(p->fluid_iterations, def);
    const float dt = p->dt / static_cast<float>(substeps);
    const SubstepConsts c = make_consts(ctx, s, p, dt);
    const uint32_t key_bits = key_bits_of(ctx, p);
    launch_fill(s.meta, kMetaWords, 0u, ctx->stream);
    const uint32_t soft_iters = resolve_iters(p->soft_internal_iterations, def);
    const uint32_t contact_iters = resolve_iters(p->soft_contact_iterations, def);
    const uint32_t max_iters = std::max(std::max(fluid_iters, soft_iters), std::max(contact_iters, def));
    const bool has_soft = s.ne + s.nb + s.nd + s.nt > 0;
    // contacts need a solid particle (fluid pairs are never candidates, and static bodies take
    // solid particles only); a pool where no scene can hold a contact skips them, and so does each
    // such scene of a pool: their zero corrections change nothing but the sign of zero coordinates
    const bool has_contacts =
        contact_iters > 0 && s.ct_alloc >= s.n && ((s.n > 1 && ctx->any_pairs) || ctx->any_rigid);
    // particle-rigid contacts of the scenes that solve them in particle-index order, inside the soft
    // phase; the constrained positions then stay in particle-index order from the substep's first
    // soft pass to its last one, unless a slot-order contact pass moves particles in between
    const bool rigid_pid = contact_iters > 0 && has_soft && ctx->any_rigid_pid && s.ct_alloc >= s.n;
    const uint32_t pid_iters = std::max(soft_iters, rigid_pid ? contact_iters : 0u);
    // The soft phase (with the particle-rigid contacts solved inside it) moves constrained particles
    // only and the fluid passes fluid particles only; without a slot-order contact pass between
    // them they are two independent chains within a substep, so the soft chain runs on a second
    // stream beside the fluid passes (a bandwidth-bound chain next to a VALU-bound one) and joins
    // the context stream before the velocity update.
    const bool split = has_soft && ctx->has_fluid && !has_contacts && ctx->side && !ctx->profiling;
    hipStream_t soft_st = split ? ctx->side : ctx->stream;
    SoftArgs sa{};
    if (has_soft) {
        sa.slot_of = s.slot_of;
        sa.n = s.npid;
        sa.dt = dt;
        sa.relax = p->solver_config[kCfgSoftRelax];
        sa.edges = s.edges;
        sa.ne = s.ne;
        sa.bends = s.bends;
        sa.nb = s.nb;
        sa.dih = s.dihedrals;
        sa.nd = s.nd;
        sa.tets = s.tets;
        sa.nt = s.nt;
        sa.lam_e = s.lam;
        sa.lam_b = s.lam + s.ne;
        sa.lam_t = s.lam + s.ne + s.nb + s.nd;
        sa.posi = s.posi;
        sa.soft_ids = s.soft_ids;
        sa.nsoft = s.nsoft;
        sa.corr_e = s.corr + s.corr_off[0];
        sa.corr_b = s.corr + s.corr_off[1];
        sa.corr_t = s.corr + s.corr_off[2];
        sa.map_e = s.inc + s.inc_off[6];
        sa.map_b = s.inc + s.inc_off[7];
        sa.map_t = s.inc + s.inc_off[8];
        sa.inc_e_start = s.inc + s.inc_off[0];
        sa.inc_b_start = s.inc + s.inc_off[2];
        sa.inc_t_start = s.inc + s.inc_off[4];
        sa.active = s.scene_active;
        sa.pid_scene = s.pid_scene;
        sa.all_live = all_active(ctx) ? 1u : 0u;
    }
    for (uint32_t sub = 0; sub < substeps; ++sub) {
        broadphase(ctx, s, c, key_bits);
        if (split && (hipEventRecord(ctx->fork_ev, ctx->stream) != hipSuccess ||
                      hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0) != hipSuccess))
            ctx->side_failed = true;
        if (has_soft) {
            sa.pos = s.pos[s.cur];
            sa.id = s.id[s.cur];
        }
        if (has_contacts) {
            ProfScope ps(ctx, "contact_partners");
            launch_contact_partners(s, c, ctx->stream);
        }
        int din = 0;
        for (uint32_t it = 0; it < max_iters; ++it) {
            if (ctx->has_fluid && it < fluid_iters) {
                {
                    ProfScope ps(ctx, "fluid_density");
                    launch_fluid_density(s, c, it, din, ctx->stream);
                }
                {
                    ProfScope ps(ctx, "fluid_delta");
                    launch_fluid_delta(s, c, it, din, ctx->stream);
                }
                din = 1 - din;
            }
            if (has_soft && it < pid_iters) {
                // Without slot-order contact passes nothing but the soft phase moves a constrained
                // particle within the substep (the fluid passes list fluid particles only), so the
                // positions stay in particle-index order from the first soft iteration to the last.
                ProfScope ps(ctx, "soft");
                const bool gather = it == 0 || has_contacts, scatter = it + 1 == pid_iters || has_contacts;
                sa.first = it == 0 ? 1u : 0u;
                if (rigid_pid && it < contact_iters) {
                    launch_soft_iteration(sa, gather, false, it < soft_iters, soft_st);
                    launch_rigid_pid(s, c, it == 0, soft_st);
                    if (scatter) launch_soft_iteration(sa, false, true, false, soft_st);
                } else {
                    launch_soft_iteration(sa, gather, scatter, it < soft_iters, soft_st);
                }
            }
            if (has_contacts && it < contact_iters) {
                ProfScope ps(ctx, "contacts");
                launch_contact_iteration(s, c, din, ctx->stream);
            }
        }
        if (split && (hipEventRecord(ctx->join_ev, ctx->side) != hipSuccess ||
                      hipStreamWaitEvent(ctx->stream, ctx->join_ev, 0) != hipSuccess))
            ctx->side_failed = true;
        {
            ProfScope ps(ctx, "finalize");
            launch_finalize(s, c, din, ctx->stream);
        }
        // physics_solver.cpp:1261-1283; with dt <= epsilon every velocity is zero after the update,
        // so the curl and the confinement vanish
        if (ctx->has_fluid && ctx->has_vorticity && c.dt > kEps) {
            ProfScope ps(ctx, "fluid_vorticity");
            launch_fluid_vorticity(s, c, ctx->stream);
        }
        // physics_solver.cpp:1285-1290; each substep overwrites the previous one's output
        if (ctx->has_fluid && ctx->aniso_enabled && sub + 1 == substeps) {
            ProfScope ps(ctx, "fluid_render_anisotropy");
            launch_fluid_render_aniso(s, c, s.aniso, ctx->stream);
        }
        if (has_contacts && ctx->any_pairs) {
            ProfScope ps(ctx, "contact_velocity");
            for (uint32_t v = 0; v < contact_iters; ++v) launch_contact_velocity(s, c, ctx->stream);
        }
        // physics_solver.cpp:1301-1309, after the particle-particle velocity iterations
        if ((has_contacts && ctx->any_rigid) || rigid_pid) {
            ProfScope ps(ctx, "rigid_velocity");
            launch_rigid_velocity(s, c, contact_iters, ctx->stream);
        }
    }
    {
        ProfScope ps(ctx, "export");
        launch_export(s, true, false, ctx->stream);
    }
}

// The soft phase's stream and the fork / join events, on first use.
cn_status ensure_side(cn_ctx *ctx)
{
    if (!ctx->split_phases || ctx->side) return CN_OK;
    CN_HIP(ctx, hipStreamCreate(&ctx->side));
    CN_HIP(ctx, hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
    CN_HIP(ctx, hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming));
    ctx->graph_valid = false;
    return CN_OK;
}

// The per-slot arrays of a compact tick for m slots (grown to at least twice the last size, and
// to a sixteenth of the pool, so a serving loop settles on one allocation): the arrays the step
// reads and writes in slot order.  The tables, constraints and their state, the particle-index
// arrays, the diagnostics and the iteration state are the pool's.
cn_status ensure_tick(cn_ctx *ctx, uint32_t m)
{
    const SceneBuffers &s = ctx->s;
    SceneBuffers &t = ctx->tick;
    const bool want_ct = s.ct_alloc > 0, want_sl = s.slist != nullptr, want_bl = s.bndl != nullptr;
    if (t.pos[0] && m <= ctx->tick_cap && ctx->tick_list_cap == ctx->cap_alloc && (t.ct_alloc > 0) == want_ct &&
        (t.slist != nullptr) == want_sl && (t.bndl != nullptr) == want_bl)
        return CN_OK;
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_scene(t);
    ctx->tick_cap = 0;
    const size_t N = std::min<size_t>(std::max<size_t>({(size_t)m, 2 * (size_t)ctx->tick_cap, s.n / 16, 1}), s.n);
    for (int k = 0; k < 2; ++k) {
        CN_HIP(ctx, dalloc(&t.pos[k], N));
        CN_HIP(ctx, dalloc(&t.vel[k], N));
        CN_HIP(ctx, dalloc(&t.cls[k], N));
        CN_HIP(ctx, dalloc(&t.id[k], N));
        CN_HIP(ctx, dalloc(&t.kv[k], N));
        CN_HIP(ctx, dalloc(&t.dacc[k], N));
        CN_HIP(ctx, dalloc(&t.prev[k], N));
        CN_HIP(ctx, hipMemsetAsync(t.dacc[k], 0, N * sizeof(float4), ctx->stream));
    }
    CN_HIP(ctx, dalloc(&t.cellc, N));
    CN_HIP(ctx, dalloc(&t.bnd, N));
    if (want_bl) CN_HIP(ctx, dalloc(&t.bndl, N * kBndListCap));
    CN_HIP(ctx, dalloc(&t.sort_scratch, sort_scratch_words((uint32_t)N)));
    CN_HIP(ctx, dalloc(&t.elem_run, N));
    CN_HIP(ctx, dalloc(&t.run_key, N));
    CN_HIP(ctx, dalloc(&t.run_start, N + 1));
    CN_HIP(ctx, dalloc(&t.run_info, N));
    t.bucket_bits = std::min<uint32_t>(std::max<uint32_t>(4u, ceil_log2(N) > 1 ? ceil_log2(N) - 1 : 1), 30u);
    CN_HIP(ctx, dalloc(&t.bucket, (size_t(1) << t.bucket_bits) + 1));
    CN_HIP(ctx, dalloc(&t.block_runs, (N + 1023) / 1024 + 1));
    CN_HIP(ctx, dalloc(&t.run_count, 1));
    CN_HIP(ctx, dalloc(&t.perm, N));
    CN_HIP(ctx, dalloc(&t.nc, N));
    CN_HIP(ctx, dalloc(&t.curpos, N));
    CN_HIP(ctx, dalloc(&t.cs, N));
    CN_HIP(ctx, dalloc(&t.blk_first_run, N + 1));
    CN_HIP(ctx, dalloc(&t.nblk, 2));
    t.blk_cap = (uint32_t)std::max<size_t>(256, N / 32);
    CN_HIP(ctx, dalloc(&t.blk_halo, (size_t)t.blk_cap * kHaloStride));
    CN_HIP(ctx, dalloc(&t.wide_blk, t.blk_cap));
    CN_HIP(ctx, dalloc(&t.list, list_slots(N) * std::max<uint32_t>(ctx->cap_alloc, 1u)));
    CN_HIP(ctx, dalloc(&t.lcnt, N));
    if (want_sl) {
        CN_HIP(ctx, dalloc(&t.slist, list_slots(N) * kSuperSlots));
        CN_HIP(ctx, dalloc(&t.scnt, N));
        CN_HIP(ctx, dalloc(&t.disp, N));
    }
    if (want_ct) {
        CN_HIP(ctx, dalloc(&t.ct_list, N * kContactPartners));
        CN_HIP(ctx, dalloc(&t.ct_rec, N * kContactPartners));
        CN_HIP(ctx, dalloc(&t.ct_count, N));
        CN_HIP(ctx, dalloc(&t.ct_bits, N));
        CN_HIP(ctx, dalloc(&t.ct_corr, N));
        CN_HIP(ctx, dalloc(&t.ct_gen, N));
        t.ct_alloc = N;
    }
    ctx->tick_cap = (uint32_t)N;
    ctx->tick_list_cap = ctx->cap_alloc;
    return CN_OK;
}

// The buffer set of a compact tick of m slots: the pool's, with every per-slot array replaced by
// the tick's.
SceneBuffers tick_view(const cn_ctx *ctx, uint32_t m, const cn_step_params *p)
{
    SceneBuffers v = ctx->s;
    const SceneBuffers &t = ctx->tick;
    v.n = m;
    for (int k = 0; k < 2; ++k) {
        v.pos[k] = t.pos[k];
        v.vel[k] = t.vel[k];
        v.cls[k] = t.cls[k];
        v.id[k] = t.id[k];
        v.kv[k] = t.kv[k];
        v.dacc[k] = t.dacc[k];
        v.prev[k] = t.prev[k];
    }
    v.cur = 0;
    v.cellc = t.cellc;
    v.bnd = t.bnd;
    v.bndl = t.bndl;
    v.sort_scratch = t.sort_scratch;
    v.elem_run = t.elem_run;
    v.run_key = t.run_key;
    v.run_start = t.run_start;
    v.run_info = t.run_info;
    v.bucket = t.bucket;
    const uint32_t bb = std::max<uint32_t>(4u, ceil_log2(m) > 1 ? ceil_log2(m) - 1 : 1);
    v.bucket_bits = std::min<uint32_t>(std::min<uint32_t>(bb, key_bits_of(ctx, p)), t.bucket_bits);
    v.block_runs = t.block_runs;
    v.run_count = t.run_count;
    v.perm = t.perm;
    v.nc = t.nc;
    v.curpos = t.curpos;
    v.cs = t.cs;
    v.blk_first_run = t.blk_first_run;
    v.nblk = t.nblk;
    v.blk_halo = t.blk_halo;
    v.blk_cap = t.blk_cap;
    v.wide_blk = t.wide_blk;
    v.list = t.list;
    v.lcnt = t.lcnt;
    v.slist = t.slist;
    v.scnt = t.scnt;
    v.disp = t.disp;
    v.nbr = nullptr;
    v.nbr_cnt = nullptr;
    v.nbr_alloc = 0;
    v.ct_list = t.ct_list;
    v.ct_count = t.ct_count;
    v.ct_rec = t.ct_rec;
    v.ct_bits = t.ct_bits;
    v.ct_corr = t.ct_corr;
    v.ct_gen = t.ct_gen;
    v.ct_alloc = t.ct_alloc;
    return v;
}

// One step of the listed scenes (flags) on a compact slot range: their segments of the pool's slot
// order gathered, stepped as a pool of those scenes alone, and put back; m = their particles.  The
// scenes' results are those of the full-pool step (a pool steps each scene as it steps alone).
cn_status compact_step(cn_ctx *ctx, const cn_step_params *p, const std::vector<uint32_t> &flags, uint32_t m)
{
    std::vector<uint32_t> act;
    for (uint32_t k = 0; k < ctx->nscenes; ++k)
        if (flags[k]) act.push_back(k);
    std::sort(act.begin(), act.end(), [&](uint32_t a, uint32_t b) { return ctx->slot_base[a] < ctx->slot_base[b]; });
    std::vector<uint4> &seg = ctx->tick_seg_host;
    seg.resize(act.size());
    uint32_t at = 0;
    for (size_t q = 0; q < act.size(); ++q) {
        const uint32_t k = act[q], cnt = ctx->scene_first[k + 1] - ctx->scene_first[k];
        if ((uint64_t)ctx->slot_base[k] + cnt > ctx->s.n) return fail(ctx, CN_ERR_DEVICE, "scene segment out of range");
        seg[q] = make_uint4(ctx->slot_base[k], at, cnt, 0u);
        at += cnt;
    }
    if (at != m || m == 0) return fail(ctx, CN_ERR_DEVICE, "compact tick: segment sizes disagree");
    const cn_status st = ensure_tick(ctx, m);
    if (st != CN_OK) return st;
    if (ctx->tick_seg_cap < seg.size()) {
        CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
        dfree(ctx->tick_seg);
        ctx->tick_seg = nullptr;
        ctx->tick_seg_cap = 0;
        CN_HIP(ctx, dalloc(&ctx->tick_seg, std::max<size_t>(seg.size(), ctx->nscenes)));
        ctx->tick_seg_cap = std::max<uint32_t>((uint32_t)seg.size(), ctx->nscenes);
    }
    CN_HIP(ctx, hipMemcpyAsync(ctx->tick_seg, seg.data(), seg.size() * sizeof(uint4), hipMemcpyHostToDevice,
                               ctx->stream));
    SceneBuffers v = tick_view(ctx, m, p);
    launch_tick_gather(ctx->s, v, ctx->tick_seg, (uint32_t)seg.size(), ctx->stream);
    record_step(ctx, v, p);
    launch_tick_scatter(ctx->s, v, ctx->tick_seg, (uint32_t)seg.size(), ctx->stream);
    ctx->stepped = true;
    if (ctx->side_failed) {
        ctx->side_failed = false;
        return fail(ctx, CN_ERR_DEVICE, "event record / wait between the context's streams failed");
    }
    CN_HIP(ctx, hipGetLastError());
    return CN_OK;
}

}  // namespace

extern "C" {

static cn_status step_impl(cn_ctx *ctx, const cn_step_params *p);

// Writes the per-scene active flags to the device when they differ from the last ones written.
static cn_status set_active(cn_ctx *ctx, const std::vector<uint32_t> &flags)
{
    if (flags == ctx->active_host) return CN_OK;
    ctx->active_host = flags;
    CN_HIP(ctx, hipMemcpyAsync(ctx->s.scene_active, ctx->active_host.data(), flags.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice, ctx->stream));
    return CN_OK;
}

cn_status cn_step(cn_ctx *ctx, const cn_step_params *p)
{
    if (!ctx || !p) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "cn_step before cn_scene_upload");
    CN_HIP(ctx, hipSetDevice(ctx->device));
    const cn_status st = set_active(ctx, std::vector<uint32_t>(ctx->nscenes, 1u));
    return st != CN_OK ? st : step_impl(ctx, p);
}

cn_status cn_step_scenes(cn_ctx *ctx, const cn_step_params *p, uint32_t count, const uint32_t *scenes)
{
    if (!ctx || !p || (count && !scenes)) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "cn_step_scenes before an upload");
    std::vector<uint32_t> flags(ctx->nscenes, 0u);
    for (uint32_t k = 0; k < count; ++k) {
        if (scenes[k] >= ctx->nscenes) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "scene index out of range");
        flags[scenes[k]] = 1u;
    }
    CN_HIP(ctx, hipSetDevice(ctx->device));
    const cn_status st = set_active(ctx, flags);
    return st != CN_OK ? st : step_impl(ctx, p);
}

static cn_status step_impl(cn_ctx *ctx, const cn_step_params *p)
{
    if (!ctx || !p) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "cn_step before cn_scene_upload");
    if (p->key_mode != CN_KEY_MORTON && p->key_mode != CN_KEY_HASH)
        return fail(ctx, CN_ERR_INVALID_ARGUMENT, "unknown key_mode");
    if (p->key_mode == CN_KEY_HASH && ctx->nscenes > 1)
        return fail(ctx, CN_ERR_UNSUPPORTED, "hash keys hold one scene; a scene pool needs CN_KEY_MORTON");
    if (p->max_fluid_neighbors > ctx->cap_alloc)
        return fail(ctx, CN_ERR_CAPACITY, "max_fluid_neighbors exceeds the neighbor slots of cn_scene_reserve");
    CN_HIP(ctx, hipSetDevice(ctx->device));
    // candidate lists on first use, for scenes with fluid, when they fit the 32-bit list addressing
    if ((p->flags & CN_STEP_CANDIDATE_REUSE) && !ctx->s.slist && ctx->has_fluid &&
        (uint64_t)list_slots(ctx->capacity) * kSuperSlots * sizeof(uint16_t) <= kListBytesMax) {
        CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
        CN_HIP(ctx, dalloc(&ctx->s.slist, list_slots(ctx->capacity) * kSuperSlots));
        CN_HIP(ctx, dalloc(&ctx->s.scnt, std::max<uint32_t>(ctx->capacity, 1u)));
        CN_HIP(ctx, dalloc(&ctx->s.disp, std::max<uint32_t>(ctx->capacity, 1u)));
        ctx->graph_valid = false;
    }
    // the vorticity and anisotropy passes' neighbor lists on first use (the pass entry's list
    // buffer, sized for every slot and the reserved neighbor count)
    if (ctx->has_fluid && (ctx->has_vorticity || ctx->aniso_enabled) &&
        ctx->s.nbr_alloc < (size_t)ctx->capacity * std::max<uint32_t>(ctx->cap_alloc, 1u)) {
        CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
        SceneBuffers &s = ctx->s;
        dfree(s.nbr);
        dfree(s.nbr_cnt);
        s.nbr = nullptr;
        s.nbr_cnt = nullptr;
        s.nbr_alloc = 0;
        const size_t slots = (size_t)ctx->capacity * std::max<uint32_t>(ctx->cap_alloc, 1u);
        CN_HIP(ctx, dalloc(&s.nbr, slots));
        CN_HIP(ctx, dalloc(&s.nbr_cnt, std::max<uint32_t>(ctx->capacity, 1u)));
        s.nbr_alloc = slots;
        ctx->graph_valid = false;
    }
    if (ctx->aniso_enabled && ctx->s.aniso_alloc < ctx->s.n) {  // output on first use
        CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
        dfree(ctx->s.aniso);
        ctx->s.aniso = nullptr;
        ctx->s.aniso_alloc = 0;
        CN_HIP(ctx, dalloc(&ctx->s.aniso, 3ull * ctx->s.n));
        ctx->s.aniso_alloc = ctx->s.n;
        ctx->graph_valid = false;
    }
    if (ctx->aniso_enabled && ctx->has_fluid) ctx->aniso_ready = true;
    {
        const cn_status st = ensure_side(ctx);
        if (st != CN_OK) return st;
    }
    // a step of a small part of a pool runs on the listed scenes' slots only
    if (ctx->compact_below > 0.0f && ctx->nscenes > 1 && ctx->slots_by_prefix && p->key_mode == CN_KEY_MORTON) {
        uint64_t m = 0;
        for (uint32_t k = 0; k < ctx->nscenes; ++k)
            if (ctx->active_host[k]) m += ctx->scene_first[k + 1] - ctx->scene_first[k];
        if (m > 0 && (double)m <= (double)ctx->compact_below * ctx->s.n)
            return compact_step(ctx, p, ctx->active_host, (uint32_t)m);
    }
    if (!ctx->graphs_enabled || ctx->profiling) {
        record_step(ctx, ctx->s, p);
        ctx->stepped = true;
        ctx->slots_by_prefix = true;
        if (ctx->side_failed) {
            ctx->side_failed = false;
            return fail(ctx, CN_ERR_DEVICE, "event record / wait between the context's streams failed");
        }
        CN_HIP(ctx, hipGetLastError());
        return CN_OK;
    }
    // graph path: one captured graph per (parameters, starting slot-order parity), replayed
    if (!ctx->graph_valid) {
        clear_graphs(ctx);
        ctx->graph_valid = true;
    }
    const int cur0 = ctx->s.cur;
    const bool every = all_active(ctx);
    for (auto &g : ctx->graphs)
        if (g.cur0 == cur0 && g.all_active == every && std::memcmp(&g.params, p, sizeof(*p)) == 0) {
            CN_HIP(ctx, hipGraphLaunch(g.exec, ctx->stream));
            ctx->s.cur = g.cur1;
            ctx->stepped = true;
            ctx->slots_by_prefix = true;
            return CN_OK;
        }
    cn_ctx::Graph g{};
    g.params = *p;
    g.cur0 = cur0;
    g.all_active = every;
    CN_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    record_step(ctx, ctx->s, p);
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &g.graph);
    if (ce != hipSuccess || ctx->side_failed) {
        ctx->side_failed = false;
        if (ce == hipSuccess) (void)hipGraphDestroy(g.graph);
        ctx->s.cur = cur0;
        return fail(ctx, CN_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
    }
    g.cur1 = ctx->s.cur;
    const hipError_t ie = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g.graph);
        ctx->s.cur = cur0;
        return fail(ctx, CN_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ie));
    }
    if (ctx->graphs.size() >= 8) clear_graphs(ctx);
    ctx->graphs.push_back(g);
    CN_HIP(ctx, hipGraphLaunch(g.exec, ctx->stream));
    ctx->stepped = true;
    ctx->slots_by_prefix = true;
    return CN_OK;
}

cn_status cn_tick_policy(cn_ctx *ctx, float compact_below)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    if (!(compact_below >= 0.0f && compact_below <= 1.0f))
        return fail(ctx, CN_ERR_INVALID_ARGUMENT, "compact_below outside [0, 1]");
    ctx->compact_below = compact_below;
    return CN_OK;
}

cn_status cn_concurrent_phases(cn_ctx *ctx, int enable)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    if ((enable != 0) != ctx->split_phases) ctx->graph_valid = false;
    ctx->split_phases = enable != 0;
    if (!ctx->split_phases && ctx->side) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->side);
        (void)hipEventDestroy(ctx->fork_ev);
        (void)hipEventDestroy(ctx->join_ev);
        (void)hipStreamDestroy(ctx->side);
        ctx->side = nullptr;
        ctx->fork_ev = ctx->join_ev = nullptr;
    }
    return CN_OK;
}

cn_status cn_graph_enable(cn_ctx *ctx, int enable)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    ctx->graphs_enabled = enable != 0;
    return CN_OK;
}

cn_status cn_tune_set(cn_ctx *ctx, uint32_t tile_max, uint32_t grid_max)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    if (tile_max < 256 || tile_max > 4096 || grid_max < kXcds || grid_max > 8192)
        return fail(ctx, CN_ERR_INVALID_ARGUMENT, "tile_max outside [256, 4096] or grid_max outside [8, 8192]");
    if (tile_max != ctx->tile_max || grid_max != ctx->grid_max) ctx->graph_valid = false;
    ctx->tile_max = tile_max;
    ctx->grid_max = grid_max;
    return CN_OK;
}

// Times cn_step under each candidate launch shape from the same saved state, keeps the fastest
// and puts the state back.  Results do not depend on the shape: a block whose halo exceeds the
// tile is walked through the run table with the same arithmetic in the same order.
cn_status cn_autotune(cn_ctx *ctx, const cn_step_params *p, uint32_t steps, cn_tune_result *out)
{
    if (!ctx || !p) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "cn_autotune before cn_scene_upload");
    if (p->max_fluid_neighbors > ctx->cap_alloc)
        return fail(ctx, CN_ERR_CAPACITY, "max_fluid_neighbors exceeds the neighbor slots of cn_scene_reserve");
    if (p->key_mode != CN_KEY_MORTON && p->key_mode != CN_KEY_HASH)
        return fail(ctx, CN_ERR_INVALID_ARGUMENT, "unknown key_mode");
    if (p->key_mode == CN_KEY_HASH && ctx->nscenes > 1)
        return fail(ctx, CN_ERR_UNSUPPORTED, "hash keys hold one scene; a scene pool needs CN_KEY_MORTON");
    steps = std::max<uint32_t>(steps, 1u);
    CN_HIP(ctx, hipSetDevice(ctx->device));
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const cn_status st0 = set_active(ctx, std::vector<uint32_t>(ctx->nscenes, 1u));
    if (st0 != CN_OK) return st0;
    const cn_status st1 = ensure_side(ctx);
    if (st1 != CN_OK) return st1;
    SceneBuffers &s = ctx->s;
    const size_t n = s.n;
    // saved state: slot-order particles, previous positions, exports and diagnostics
    float4 *sp = nullptr, *sv = nullptr, *spr = nullptr, *so = nullptr;
    uint32_t *sc = nullptr, *si = nullptr, *sm = nullptr;
    auto release = [&]() {
        dfree(sp);
        dfree(sv);
        dfree(spr);
        dfree(so);
        dfree(sc);
        dfree(si);
        dfree(sm);
    };
    const int cur0 = s.cur;
    const bool stepped0 = ctx->stepped, prefix0 = ctx->slots_by_prefix;
    const uint32_t tile0 = ctx->tile_max, grid0 = ctx->grid_max;
    if (dalloc(&sp, n) != hipSuccess || dalloc(&sv, n) != hipSuccess || dalloc(&spr, n) != hipSuccess ||
        dalloc(&so, 3 * n) != hipSuccess || dalloc(&sc, n) != hipSuccess || dalloc(&si, n) != hipSuccess ||
        dalloc(&sm, kMetaWords) != hipSuccess) {
        release();
        return fail(ctx, CN_ERR_OUT_OF_MEMORY, "cn_autotune: no room for the saved state");
    }
    auto copy = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    };
    hipError_t e = hipSuccess;
    auto save = [&]() {
        const size_t f4 = n * sizeof(float4), u4 = n * sizeof(uint32_t);
        if (e == hipSuccess) e = copy(sp, s.pos[cur0], f4);
        if (e == hipSuccess) e = copy(sv, s.vel[cur0], f4);
        if (e == hipSuccess) e = copy(spr, s.prev[cur0], f4);
        if (e == hipSuccess) e = copy(so, s.out_pos, f4);
        if (e == hipSuccess) e = copy(so + n, s.out_vel, f4);
        if (e == hipSuccess) e = copy(so + 2 * n, s.out_prev, f4);
        if (e == hipSuccess) e = copy(sc, s.cls[cur0], u4);
        if (e == hipSuccess) e = copy(si, s.id[cur0], u4);
        if (e == hipSuccess) e = copy(sm, s.meta, kMetaWords * sizeof(uint32_t));
    };
    auto restore = [&]() {
        const size_t f4 = n * sizeof(float4), u4 = n * sizeof(uint32_t);
        s.cur = cur0;
        ctx->stepped = stepped0;
        ctx->slots_by_prefix = prefix0;
        if (e == hipSuccess) e = copy(s.pos[cur0], sp, f4);
        if (e == hipSuccess) e = copy(s.vel[cur0], sv, f4);
        if (e == hipSuccess) e = copy(s.prev[cur0], spr, f4);
        if (e == hipSuccess) e = copy(s.out_pos, so, f4);
        if (e == hipSuccess) e = copy(s.out_vel, so + n, f4);
        if (e == hipSuccess) e = copy(s.out_prev, so + 2 * n, f4);
        if (e == hipSuccess) e = copy(s.cls[cur0], sc, u4);
        if (e == hipSuccess) e = copy(s.id[cur0], si, u4);
        if (e == hipSuccess) e = copy(s.meta, sm, kMetaWords * sizeof(uint32_t));
    };
    save();
    hipEvent_t ea = nullptr, eb = nullptr;
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ea, hipEventDisableSystemFence);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&eb, hipEventDisableSystemFence);
    const uint32_t tiles[] = {1024, 2048, 3072, 4096};
    const uint32_t grids[] = {256, 512, 768, 1024};
    float best = 3.4e38f, dflt = 0.0f;
    uint32_t best_tile = tile0, best_grid = grid0, tried = 0;
    for (uint32_t tile : tiles)
        for (uint32_t grid : grids) {
            if (e != hipSuccess) break;
            ctx->tile_max = tile;
            ctx->grid_max = grid;
            record_step(ctx, ctx->s, p);  // warm-up from the saved state
            restore();
            if (e == hipSuccess) e = hipEventRecord(ea, ctx->stream);
            for (uint32_t k = 0; k < steps; ++k) record_step(ctx, ctx->s, p);
            if (e == hipSuccess) e = hipEventRecord(eb, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(eb);
            if (e == hipSuccess) e = hipGetLastError();
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, ea, eb);
            restore();
            ms /= (float)steps;
            ++tried;
            if (tile == 4096 && grid == 1024) dflt = ms;
            // tiles are tried in ascending order; a larger one within 1 % of the best is kept: it
            // sends no more dense blocks to the run-table walk, and a fluid that compresses after
            // the timed steps grows such blocks
            if (tile > best_tile && best < 3.4e38f ? ms < best * 1.01f : ms < best) {
                best = ms;
                best_tile = tile;
                best_grid = grid;
            }
        }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (ea) (void)hipEventDestroy(ea);
    if (eb) (void)hipEventDestroy(eb);
    release();
    ctx->graph_valid = false;
    if (e != hipSuccess) {
        ctx->tile_max = tile0;
        ctx->grid_max = grid0;
        return fail(ctx, CN_ERR_DEVICE, std::string("cn_autotune: ") + hipGetErrorString(e));
    }
    ctx->tile_max = best_tile;
    ctx->grid_max = best_grid;
    if (out) {
        out->tile_max = best_tile;
        out->grid_max = best_grid;
        out->candidates = tried;
        out->reserved = 0;
        out->ms_per_step = best;
        out->default_ms_per_step = dflt;
    }
    return CN_OK;
}

cn_status cn_readback(cn_ctx *ctx, float *pos, float *prev, float *vel, uint32_t count)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "no scene");
    if (count > ctx->s.n) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "count exceeds particle count");
    CN_HIP(ctx, hipSetDevice(ctx->device));
    if (prev && ctx->stepped) launch_export(ctx->s, false, true, ctx->stream);
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (pos) CN_HIP(ctx, hipMemcpy(pos, ctx->s.out_pos, count * sizeof(float4), hipMemcpyDeviceToHost));
    if (prev) CN_HIP(ctx, hipMemcpy(prev, ctx->s.out_prev, count * sizeof(float4), hipMemcpyDeviceToHost));
    if (vel) CN_HIP(ctx, hipMemcpy(vel, ctx->s.out_vel, count * sizeof(float4), hipMemcpyDeviceToHost));
    return CN_OK;
}

cn_status cn_render_anisotropy_enable(cn_ctx *ctx, int enable)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    if (ctx->aniso_enabled != (enable != 0)) ctx->graph_valid = false;
    ctx->aniso_enabled = enable != 0;
    ctx->aniso_ready = false;
    return CN_OK;
}

cn_status cn_render_anisotropy_read(cn_ctx *ctx, float *axis1, float *axis2, float *axis3, uint32_t count)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "no scene");
    if (count > ctx->s.n) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "count exceeds particle count");
    if (!ctx->aniso_ready)
        return fail(ctx, CN_ERR_NOT_READY, "no fluid step has written the render anisotropy since it was enabled");
    CN_HIP(ctx, hipSetDevice(ctx->device));
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float *dst[3] = {axis1, axis2, axis3};
    for (int k = 0; k < 3; ++k)
        if (dst[k] && count)
            CN_HIP(ctx, hipMemcpy(dst[k], ctx->s.aniso + (size_t)k * ctx->s.n, count * sizeof(float4),
                                  hipMemcpyDeviceToHost));
    return CN_OK;
}

cn_status cn_scene_set_state(cn_ctx *ctx, uint32_t scene, const float *positions_inv_mass, const float *velocities,
                             uint32_t flags)
{
    if (!ctx || !positions_inv_mass || !velocities) return CN_ERR_INVALID_ARGUMENT;
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "no scene");
    if (scene >= ctx->nscenes) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "scene index out of range");
    const uint32_t first = ctx->scene_first[scene], count = ctx->scene_first[scene + 1] - first;
    return cn_particles_update_range(ctx, first, count, positions_inv_mass, velocities, flags);
}

cn_status cn_particles_update_range(cn_ctx *ctx, uint32_t first, uint32_t count, const float *positions_inv_mass,
                                    const float *velocities, uint32_t flags)
{
    if (!ctx || (count && (!positions_inv_mass || !velocities))) return CN_ERR_INVALID_ARGUMENT;
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "no scene");
    if ((uint64_t)first + count > ctx->s.n) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "particle run out of range");
    if (count == 0) return CN_OK;
    CN_HIP(ctx, hipSetDevice(ctx->device));
    const float4 *P = reinterpret_cast<const float4 *>(positions_inv_mass);
    const float4 *V = reinterpret_cast<const float4 *>(velocities);
    const bool host = !(flags & CN_STATE_DEVICE_POINTERS);
    if (host) {
        if (ctx->stage_cap < 2 * (size_t)count) {
            CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
            dfree(ctx->stage);
            ctx->stage = nullptr;
            ctx->stage_cap = 0;
            CN_HIP(ctx, dalloc(&ctx->stage, 2 * (size_t)count));
            ctx->stage_cap = 2 * (size_t)count;
        }
        CN_HIP(ctx, hipMemcpyAsync(ctx->stage, P, count * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
        CN_HIP(ctx, hipMemcpyAsync(ctx->stage + count, V, count * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
        P = ctx->stage;
        V = ctx->stage + count;
    }
    launch_set_state(ctx->s, first, count, P, V, ctx->stream);
    CN_HIP(ctx, hipGetLastError());
    // host arrays (and the stage) may be reused as soon as the call returns
    if (host) CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CN_OK;
}

cn_status cn_scene_set_colliders(cn_ctx *ctx, uint32_t scene, const cn_box_collider *colliders, uint32_t count)
{
    if (!ctx || (count && !colliders)) return CN_ERR_INVALID_ARGUMENT;
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "no scene");
    if (scene >= ctx->nscenes) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "scene index out of range");
    const uint32_t cap = ctx->scene_col_cap[scene];
    if (count != cap)
        return fail(ctx, CN_ERR_INVALID_ARGUMENT, "collider count differs from the scene's upload; re-upload to add or remove colliders");
    if (count == 0) return CN_OK;
    CN_HIP(ctx, hipSetDevice(ctx->device));
    const std::vector<cn_collider> &shapes = ctx->scene_shapes[scene];
    std::vector<Collider> sc = visit_ordered_colliders(colliders, count, shapes.data(), (uint32_t)shapes.size());
    const uint32_t ncol = (uint32_t)sc.size();
    const size_t slots = cap + shapes.size();  // the scene's collider slots: boxes and shape colliders
    sc.resize(slots);
    SceneInfo &si = ctx->scene_info[scene];
    si.ncol = ncol;
    CN_HIP(ctx, hipMemcpyAsync(ctx->s.colliders + si.col0, sc.data(), slots * sizeof(Collider), hipMemcpyHostToDevice,
                               ctx->stream));
    const std::vector<ColGroup> g = collider_groups(sc, ncol, group_slots((uint32_t)slots));
    if (!g.empty())
        CN_HIP(ctx, hipMemcpyAsync(ctx->s.col_groups + si.grp0, g.data(), g.size() * sizeof(ColGroup),
                                   hipMemcpyHostToDevice, ctx->stream));
    CN_HIP(ctx, hipMemcpyAsync(ctx->s.scenes + scene, &si, sizeof(SceneInfo), hipMemcpyHostToDevice, ctx->stream));
    // the host sources may go as soon as the call returns
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CN_OK;
}

cn_status cn_scene_set_materials(cn_ctx *ctx, uint32_t scene, const cn_fluid_material *fluid_materials,
                                 uint32_t fluid_count, const float *particle_materials, uint32_t particle_count)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "no scene");
    if (scene >= ctx->nscenes) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "scene index out of range");
    const uint32_t f0 = ctx->scene_fm[2 * scene], fn = ctx->scene_fm[2 * scene + 1];
    const uint32_t p0 = ctx->scene_pm[2 * scene], pn = ctx->scene_pm[2 * scene + 1];
    if (fluid_materials && fluid_count != fn)
        return fail(ctx, CN_ERR_INVALID_ARGUMENT, "fluid material count differs from the scene's upload");
    if (particle_materials && particle_count != pn)
        return fail(ctx, CN_ERR_INVALID_ARGUMENT, "particle material count differs from the scene's upload");
    if (fluid_materials)
        for (uint32_t k = 0; k < fn; ++k)
            if (std::memcmp(&fluid_materials[k].smoothing_radius, &ctx->fm_host[f0 + k].smoothing_radius,
                            sizeof(float)) != 0)
                return fail(ctx, CN_ERR_UNSUPPORTED, "a new smoothing radius changes the scene's grid; re-upload");
    CN_HIP(ctx, hipSetDevice(ctx->device));
    if (fluid_materials && fn) {
        std::copy(fluid_materials, fluid_materials + fn, ctx->fm_host.begin() + f0);
        CN_HIP(ctx, hipMemcpyAsync(ctx->s.fluid_mats + f0, fluid_materials, fn * sizeof(cn_fluid_material),
                                   hipMemcpyHostToDevice, ctx->stream));
    }
    if (particle_materials && pn)
        CN_HIP(ctx, hipMemcpyAsync(ctx->s.part_mats + p0, particle_materials, pn * sizeof(float4),
                                   hipMemcpyHostToDevice, ctx->stream));
    // the flags that pick kernels are recomputed as the upload computes them (tension over the
    // whole table, vorticity over the materials a fluid particle references; the passes still
    // test each particle's own material), and a captured graph holds the old choice
    bool tension = false, vort = false;
    for (size_t k = 0; k < ctx->fm_host.size(); ++k) {
        const cn_fluid_material &m = ctx->fm_host[k];
        tension |= (m.surface_tension > 0.0f ? m.surface_tension : 0.0f) > kEps;
        if (ctx->fm_fluid_used[k]) vort |= std::max(m.vorticity_confinement, 0.0f) > kEps;
    }
    if (tension != ctx->has_tension || vort != ctx->has_vorticity) ctx->graph_valid = false;
    ctx->has_tension = tension;
    ctx->has_vorticity = vort;
    // the host tables may go as soon as the call returns
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CN_OK;
}

extern "C++" {
// Converts and writes records [first, first + count) of one constraint type (host copy and device).
template <class Dev, class Src, class Conv>
static cn_status set_records(cn_ctx *ctx, std::vector<Dev> &host, Dev *dev, uint32_t first, uint32_t count,
                             const Src *src, uint32_t base, Conv conv)
{
    if (!count) return CN_OK;
    for (uint32_t q = 0; q < count; ++q) host[first + q] = conv(src[q], base);
    CN_HIP(ctx, hipMemcpyAsync(dev + first, host.data() + first, count * sizeof(Dev), hipMemcpyHostToDevice,
                               ctx->stream));
    return CN_OK;
}
}  // extern "C++"

cn_status cn_scene_set_constraints(cn_ctx *ctx, uint32_t scene, const cn_constraint_view *cv)
{
    if (!ctx || !cv) return CN_ERR_INVALID_ARGUMENT;
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "no scene");
    if (scene >= ctx->nscenes) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "scene index out of range");
    const uint32_t *c0 = &ctx->scene_con[4 * scene];
    const bool last = scene + 1 == ctx->nscenes;
    const uint32_t *c1 = last ? nullptr : &ctx->scene_con[4 * (scene + 1)];
    const SceneBuffers &s = ctx->s;
    const uint32_t ne = (last ? s.ne : c1[0]) - c0[0], nb = (last ? s.nb : c1[1]) - c0[1];
    const uint32_t nd = (last ? s.nd : c1[2]) - c0[2], nt = (last ? s.nt : c1[3]) - c0[3];
    if (cv->edge_count != ne || cv->bend_count != nb || cv->dihedral_count != nd || cv->tet_count != nt)
        return fail(ctx, CN_ERR_INVALID_ARGUMENT, "constraint counts differ from the scene's upload");
    if ((ne && !cv->edges) || (nb && !cv->bends) || (nd && !cv->dihedrals) || (nt && !cv->tets))
        return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null constraint array");
    // every record keeps its particles (the incidence lists and record maps built at upload stay
    // valid); all are checked before any is written
    const uint32_t base = ctx->scene_first[scene];
    for (uint32_t q = 0; q < ne; ++q) {
        const EdgeDev a = edge_dev(cv->edges[q], base), &b = ctx->edges_host[c0[0] + q];
        if (a.a != b.a || a.b != b.b) return fail(ctx, CN_ERR_UNSUPPORTED, "an edge changed its particles; re-upload");
    }
    for (uint32_t q = 0; q < nb; ++q) {
        const BendDev a = bend_dev(cv->bends[q], base), &b = ctx->bends_host[c0[1] + q];
        if (a.p0 != b.p0 || a.p1 != b.p1 || a.p2 != b.p2)
            return fail(ctx, CN_ERR_UNSUPPORTED, "a bend changed its particles; re-upload");
    }
    for (uint32_t q = 0; q < nd; ++q) {
        const DihedralDev a = dihedral_dev(cv->dihedrals[q], base), &b = ctx->dih_host[c0[2] + q];
        if (a.e0 != b.e0 || a.e1 != b.e1 || a.o0 != b.o0 || a.o1 != b.o1)
            return fail(ctx, CN_ERR_UNSUPPORTED, "a dihedral changed its particles; re-upload");
    }
    for (uint32_t q = 0; q < nt; ++q) {
        const TetDev a = tet_dev(cv->tets[q], base), &b = ctx->tets_host[c0[3] + q];
        if (std::memcmp(a.p, b.p, sizeof(a.p)) != 0)
            return fail(ctx, CN_ERR_UNSUPPORTED, "a tet changed its particles; re-upload");
    }
    CN_HIP(ctx, hipSetDevice(ctx->device));
    cn_status st = set_records(ctx, ctx->edges_host, s.edges, c0[0], ne, cv->edges, base, edge_dev);
    if (st == CN_OK) st = set_records(ctx, ctx->bends_host, s.bends, c0[1], nb, cv->bends, base, bend_dev);
    if (st == CN_OK) st = set_records(ctx, ctx->dih_host, s.dihedrals, c0[2], nd, cv->dihedrals, base, dihedral_dev);
    if (st == CN_OK) st = set_records(ctx, ctx->tets_host, s.tets, c0[3], nt, cv->tets, base, tet_dev);
    // the host sources may go as soon as the call returns
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return st;
}

cn_status cn_device_views_get(cn_ctx *ctx, cn_device_views *out)
{
    if (!ctx || !out) return CN_ERR_INVALID_ARGUMENT;
    out->positions_inv_mass = ctx->s.out_pos;
    out->velocities = ctx->s.out_vel;
    out->count = ctx->s.n;
    out->stream = ctx->stream;
    return CN_OK;
}

cn_status cn_query(cn_ctx *ctx)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    const hipError_t e = hipStreamQuery(ctx->stream);
    if (e == hipSuccess) return CN_OK;
    if (e == hipErrorNotReady) return CN_ERR_NOT_READY;
    return fail(ctx, CN_ERR_DEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

cn_status cn_synchronize(cn_ctx *ctx)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    CN_HIP(ctx, hipSetDevice(ctx->device));
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CN_OK;
}

cn_status cn_validate_meta(cn_ctx *ctx, cn_meta *out)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "no scene");
    CN_HIP(ctx, hipSetDevice(ctx->device));
    uint32_t m[kMetaWords];
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CN_HIP(ctx, hipMemcpy(m, ctx->s.meta, sizeof(m), hipMemcpyDeviceToHost));
    cn_meta r{};
    r.max_fluid_neighbor_count = m[0];
    r.fluid_neighbor_truncations = m[1];
    uint64_t tot;
    std::memcpy(&tot, &m[2], sizeof(tot));
    r.fluid_neighbor_total = tot;
    r.last_particle_count = ctx->s.n;
    r.contact_overflows = m[4];
    r.max_cell_occupancy = m[8];
    r.lds_tiles = m[5];
    r.run_table_tiles = m[6];
    r.max_displacement2_bits = m[7];
    r.candidate_reuses = m[9];
    if (out) *out = r;
    if (r.fluid_neighbor_truncations)
        return fail(ctx, CN_ERR_CAPACITY, "fluid neighbor list truncated at max_fluid_neighbors");
    if (r.contact_overflows)
        return fail(ctx, CN_ERR_CAPACITY, "particle with more contact candidates than its list holds");
    return CN_OK;
}

cn_status cn_pass_broadphase_keys(cn_ctx *ctx, uint32_t n, const float *pos, float cell, uint32_t *keys,
                                  int32_t *cells)
{
    if (!ctx || !pos || !keys || !cells) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) return CN_OK;
    CN_HIP(ctx, hipSetDevice(ctx->device));
    float4 *dp = nullptr;
    uint32_t *dk = nullptr;
    int32_t *dc = nullptr;
    CN_HIP(ctx, dalloc(&dp, n));
    CN_HIP(ctx, dalloc(&dk, n));
    CN_HIP(ctx, dalloc(&dc, 3ull * n));
    CN_HIP(ctx, hipMemcpy(dp, pos, n * sizeof(float4), hipMemcpyHostToDevice));
    launch_ref_keys(dp, n, 1.0f / std::max(cell, 1.0e-4f), dk, dc, ctx->stream);
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CN_HIP(ctx, hipMemcpy(keys, dk, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    CN_HIP(ctx, hipMemcpy(cells, dc, 3ull * n * sizeof(int32_t), hipMemcpyDeviceToHost));
    dfree(dp);
    dfree(dk);
    dfree(dc);
    return CN_OK;
}

cn_status cn_pass_sort_pairs(cn_ctx *ctx, uint32_t n, uint32_t *keys, uint32_t *indices)
{
    if (!ctx || !keys || !indices) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null argument");
    if (n <= 1) return CN_OK;
    CN_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<uint2> h(n);
    for (uint32_t i = 0; i < n; ++i) h[i] = make_uint2(keys[i], indices[i]);
    uint2 *a = nullptr, *b = nullptr;
    uint32_t *scr = nullptr;
    CN_HIP(ctx, dalloc(&a, n));
    CN_HIP(ctx, dalloc(&b, n));
    CN_HIP(ctx, dalloc(&scr, sort_scratch_words(n)));
    CN_HIP(ctx, hipMemcpy(a, h.data(), n * sizeof(uint2), hipMemcpyHostToDevice));
    uint2 *res = radix_sort_pairs(a, b, n, 32u, scr, ctx->stream);
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CN_HIP(ctx, hipMemcpy(h.data(), res, n * sizeof(uint2), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) {
        keys[i] = h[i].x;
        indices[i] = h[i].y;
    }
    dfree(a);
    dfree(b);
    dfree(scr);
    return CN_OK;
}

cn_status cn_pass_sort_pairs_noalloc(cn_ctx *ctx, uint32_t n, uint32_t *keys, uint32_t *indices)
{
    if (!ctx || (n && (!keys || !indices))) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null argument");
    if (n > CN_NOALLOC_MAX_PAIRS)
        return fail(ctx, CN_ERR_CAPACITY,
                    "cn_pass_sort_pairs_noalloc: " + std::to_string(n) + " pairs exceed its scratch of " +
                        std::to_string(CN_NOALLOC_MAX_PAIRS) + " (use cn_pass_sort_pairs)");
    if (n <= 1) return CN_OK;
    CN_HIP(ctx, hipSetDevice(ctx->device));
    char *base = static_cast<char *>(noalloc_scratch(ctx->device));
    if (!base) return fail(ctx, CN_ERR_DEVICE, "cn_pass_sort_pairs_noalloc: scratch address not resolved");
    uint2 *a = reinterpret_cast<uint2 *>(base);
    uint2 *b = reinterpret_cast<uint2 *>(base + kNoAllocPairBytes);
    uint32_t *counts = reinterpret_cast<uint32_t *>(base + 2 * kNoAllocPairBytes);
    // the sort writes every count and total it reads, so nothing left by an earlier call is read
    launch_pack_pairs(keys, indices, n, a, ctx->stream);
    const uint2 *res = radix_sort_pairs(a, b, n, 32u, counts, ctx->stream);
    launch_unpack_pairs(res, n, keys, indices, ctx->stream);
    CN_HIP(ctx, hipGetLastError());
    return CN_OK;
}

cn_status cn_pass_cell_ranges(cn_ctx *ctx, uint32_t n, const uint32_t *sorted_keys, uint32_t capacity,
                              uint32_t *ranges)
{
    if (!ctx || !ranges || (n && !sorted_keys)) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null argument");
    if (capacity == 0) return CN_OK;
    CN_HIP(ctx, hipSetDevice(ctx->device));
    uint32_t *dk = nullptr, *dr = nullptr;
    CN_HIP(ctx, dalloc(&dk, n));
    CN_HIP(ctx, dalloc(&dr, 4ull * capacity + n));
    if (n) CN_HIP(ctx, hipMemcpy(dk, sorted_keys, n * sizeof(uint32_t), hipMemcpyHostToDevice));
    launch_ref_cell_ranges(dk, n, capacity, dr, ctx->stream);
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CN_HIP(ctx, hipMemcpy(ranges, dr, 4ull * capacity * sizeof(uint32_t), hipMemcpyDeviceToHost));
    dfree(dk);
    dfree(dr);
    return CN_OK;
}

cn_status cn_pass_substep_broadphase(cn_ctx *ctx, const cn_step_params *p, uint32_t *sorted_keys,
                                     uint32_t *sorted_ids, uint32_t *counts, uint32_t *lists)
{
    if (!ctx || !p) return fail(ctx, CN_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->uploaded) return fail(ctx, CN_ERR_NOT_READY, "no scene");
    if (p->max_fluid_neighbors > ctx->cap_alloc)
        return fail(ctx, CN_ERR_CAPACITY, "max_fluid_neighbors exceeds the neighbor slots of cn_scene_reserve");
    CN_HIP(ctx, hipSetDevice(ctx->device));
    SceneBuffers &s = ctx->s;
    const uint32_t n = s.n;
    const uint32_t substeps = std::max<uint32_t>(p->substeps, 1u);
    const SubstepConsts c = make_consts(ctx, s, p, p->dt / static_cast<float>(substeps));
    const size_t slots = (size_t)n * std::max<uint32_t>(c.cap, 1u);
    if (s.nbr_alloc < slots) {
        dfree(s.nbr);
        dfree(s.nbr_cnt);
        s.nbr = nullptr;
        s.nbr_cnt = nullptr;
        s.nbr_alloc = 0;
        CN_HIP(ctx, dalloc(&s.nbr, slots));
        CN_HIP(ctx, dalloc(&s.nbr_cnt, std::max<size_t>(n, ctx->capacity)));
        s.nbr_alloc = slots;
    }
    const int old_cur = s.cur;
    uint2 *sorted = broadphase(ctx, s, c, key_bits_of(ctx, p));
    ctx->slots_by_prefix = true;
    launch_fluid_lists(s, c, ctx->stream);
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<uint2> kv(n);
    std::vector<uint32_t> id(n), cnt(n);
    CN_HIP(ctx, hipMemcpy(kv.data(), sorted, n * sizeof(uint2), hipMemcpyDeviceToHost));
    CN_HIP(ctx, hipMemcpy(id.data(), s.id[s.cur], n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    CN_HIP(ctx, hipMemcpy(cnt.data(), s.nbr_cnt, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    const uint32_t cap = p->max_fluid_neighbors;
    if (sorted_keys)
        for (uint32_t t = 0; t < n; ++t) sorted_keys[t] = kv[t].x;
    if (sorted_ids) std::memcpy(sorted_ids, id.data(), n * sizeof(uint32_t));
    if (counts || lists) {
        std::vector<uint32_t> nb((size_t)n * (cap ? cap : 1));
        if (cap) CN_HIP(ctx, hipMemcpy(nb.data(), s.nbr, (size_t)n * cap * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (uint32_t t = 0; t < n; ++t) {
            const uint32_t d = id[t];
            const uint32_t ct = cnt[t];
            if (counts) counts[d] = ct;
            if (lists)
                for (uint32_t k = 0; k < ct && k < cap; ++k) lists[(size_t)d * cap + k] = id[nb[(size_t)k * n + t]];
        }
    }
    s.cur = old_cur;
    return CN_OK;
}

cn_status cn_profile_enable(cn_ctx *ctx, int enable)
{
    if (!ctx) return CN_ERR_INVALID_ARGUMENT;
    ctx->profiling = enable != 0;
    return CN_OK;
}

cn_status cn_profile_read(cn_ctx *ctx, cn_kernel_stats *out)
{
    if (!ctx || !out) return CN_ERR_INVALID_ARGUMENT;
    CN_HIP(ctx, hipSetDevice(ctx->device));
    CN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &e : ctx->events) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            ctx->stats.total_ms[e.slot] += ms;
            ctx->stats.launches[e.slot] += 1;
        }
    }
    ctx->events.clear();
    ctx->ev_used = 0;
    *out = ctx->stats;
    return CN_OK;
}

}  // extern "C"
