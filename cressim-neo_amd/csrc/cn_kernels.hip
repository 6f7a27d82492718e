// cn_kernels.hip -- broadphase and PBF fluid kernels of the particle substep (gfx950).
//
// Data layout: particle state lives in "slot" order, the cell-sorted order of the previous
// substep (slot t holds particle id[t]).  Every substep re-sorts the slots by cell key and by
// particle index inside a cell, then gathers the state into the new order, so a cell's particles
// are contiguous and the 27-cell neighbor walk reads short contiguous runs from L2 instead of
// scattered 16-byte records.  The neighbor list of a particle is, exactly as in the reference,
// the 27 cells in dz,dy,dx order with each cell's particles in ascending particle index
// (physics_fluid_build_neighbors.cs.hlsl:67-141), so results do not depend on the slot order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cn_internal.h"

namespace cn {

static inline uint32_t grid1(uint32_t n) { return (n + kBlock - 1) / kBlock; }

// physics_soft_predict.cs.hlsl:12-47
__device__ __forceinline__ void predict(float4 &p, float4 &v, float4 &prev, float gs, const SubstepConsts &c)
{
    prev = make_float4(p.x, p.y, p.z, 0.0f);
    if (p.w <= kEps) {
        v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float dts = c.dt * gs;
    v.x = v.x + c.gravity[0] * dts;
    v.y = v.y + c.gravity[1] * dts;
    v.z = v.z + c.gravity[2] * dts;
    v.w = 0.0f;
    p.x = p.x + v.x * c.dt;
    p.y = p.y + v.y * c.dt;
    p.z = p.z + v.z * c.dt;
}

__device__ __forceinline__ float gravity_scale(const ParticleClass &pc, const cn_fluid_material *fm)
{
    return pc.kind == CN_PARTICLE_KIND_FLUID ? fm[pc.fluid_mat].gravity_scale : 1.0f;
}

// physics_particle_build_broad_phase_entries.cs.hlsl:19-24
__device__ __forceinline__ int3 cell_of(float x, float y, float z, float inv_cell)
{
    return make_int3((int)floorf(x * inv_cell), (int)floorf(y * inv_cell), (int)floorf(z * inv_cell));
}

// ---------------------------------------------------------------------------------------------
// keys: predict (without storing) and emit (cell key, slot) pairs
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float scene_inv_cell(const SceneInfo &sc, const SubstepConsts &c)
{
    return c.inv_cell > 0.0f ? c.inv_cell : sc.inv_cell;
}

__global__ __launch_bounds__(kBlock) void k_keys(const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                                 const uint32_t *__restrict__ cls,
                                                 const ParticleClass *__restrict__ classes,
                                                 const cn_fluid_material *__restrict__ fm,
                                                 const SceneInfo *__restrict__ scenes,
                                                 const uint32_t *__restrict__ active, SubstepConsts c,
                                                 uint2 *__restrict__ kv)
{
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= c.n) return;
    float4 p = pos[s], v = vel[s], pv;
    const ParticleClass pc = classes[c.uniform_class ? 0u : cls[s]];
    if (active[pc.scene]) predict(p, v, pv, gravity_scale(pc, fm), c);
    const SceneInfo sc = scenes[pc.scene];
    const int3 g = cell_of(p.x, p.y, p.z, scene_inv_cell(sc, c));
    kv[s] = make_uint2(cell_key(g.x, g.y, g.z, sc.key_prefix, c.key_mode, c.morton_bits), s);
}

void launch_keys(const SceneBuffers &s, const SubstepConsts &c, hipStream_t st)
{
    hipLaunchKernelGGL(k_keys, dim3(grid1(c.n)), dim3(kBlock), 0, st, s.pos[s.cur], s.vel[s.cur], s.cls[s.cur],
                       s.classes, s.fluid_mats, s.scenes, s.scene_active, c, s.kv[0]);
}

// ---------------------------------------------------------------------------------------------
// runs of equal keys: flags -> block counts -> scan -> run table, bucket index, rank fix-up
// ---------------------------------------------------------------------------------------------
constexpr int kRunItems = 4;
constexpr int kRunTile = kBlock * kRunItems;

__device__ __forceinline__ bool run_flag(const uint2 *kv, uint32_t i, uint32_t n)
{
    return i < n && (i == 0 || kv[i].x != kv[i - 1].x);
}

__global__ __launch_bounds__(kBlock) void k_run_count(const uint2 *__restrict__ kv, uint32_t n,
                                                      uint32_t *__restrict__ block_runs)
{
    __shared__ uint32_t sh[kBlock];
    const uint32_t base = blockIdx.x * kRunTile + threadIdx.x * kRunItems;
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < kRunItems; ++k) cnt += run_flag(kv, base + k, n) ? 1u : 0u;
    uint32_t total;
    block_exclusive_scan(cnt, sh, total);
    if (threadIdx.x == 0) block_runs[blockIdx.x] = total;
}

// single workgroup: exclusive scan of the per-block run counts; total -> *run_count
__global__ __launch_bounds__(1024) void k_scan_block_runs(uint32_t *__restrict__ v, uint32_t nb,
                                                          uint32_t *__restrict__ run_count)
{
    __shared__ uint32_t sh[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nb + 1023u) / 1024u;
    const uint32_t b0 = t * per;
    uint32_t local = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (b0 + k < nb) local += v[b0 + k];
    sh[t] = local;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t add = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    uint32_t run = sh[t] - local;
    for (uint32_t k = 0; k < per; ++k)
        if (b0 + k < nb) {
            const uint32_t x = v[b0 + k];
            v[b0 + k] = run;
            run += x;
        }
    if (t == 1023) *run_count = sh[1023];
}

__global__ __launch_bounds__(kBlock) void k_run_emit(const uint2 *__restrict__ kv, uint32_t n,
                                                     const uint32_t *__restrict__ block_off,
                                                     uint32_t *__restrict__ elem_run, uint32_t *__restrict__ run_key,
                                                     uint32_t *__restrict__ run_start)
{
    __shared__ uint32_t sh[kBlock];
    const uint32_t base = blockIdx.x * kRunTile + threadIdx.x * kRunItems;
    bool f[kRunItems];
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < kRunItems; ++k) {
        f[k] = run_flag(kv, base + k, n);
        cnt += f[k] ? 1u : 0u;
    }
    uint32_t total;
    uint32_t incl = block_off[blockIdx.x] + block_exclusive_scan(cnt, sh, total);
#pragma unroll
    for (int k = 0; k < kRunItems; ++k) {
        const uint32_t i = base + k;
        if (i >= n) break;
        if (f[k]) {
            run_key[incl] = kv[i].x;
            run_start[incl] = i;
            ++incl;
        }
        elem_run[i] = incl - 1u;
        if (i == n - 1) run_start[incl] = n;
    }
}

// bucket[b] = first run whose key >> shift is >= b (b in [0, nbuckets]); lookups scan
// runs [bucket[b], bucket[b + 1]).
__global__ __launch_bounds__(kBlock) void k_bucket_fill(const uint32_t *__restrict__ run_key,
                                                        const uint32_t *__restrict__ run_count, uint32_t shift,
                                                        uint32_t nbuckets, uint32_t *__restrict__ bucket)
{
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    if (b > nbuckets) return;
    const uint32_t R = *run_count;
    uint32_t lo = 0, hi = R;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if ((run_key[mid] >> shift) < b) lo = mid + 1u;
        else hi = mid;
    }
    bucket[b] = lo;
}

// Stable order inside a run is by particle index: rank = #run members with a smaller id.
__global__ __launch_bounds__(kBlock) void k_rank_fix(const uint2 *__restrict__ kv, uint32_t n,
                                                     const uint32_t *__restrict__ elem_run,
                                                     const uint32_t *__restrict__ run_start,
                                                     const uint32_t *__restrict__ id_old, uint32_t *__restrict__ perm)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = elem_run[i];
    const uint32_t s = run_start[r], e = run_start[r + 1];
    const uint32_t o = min(kv[i].y, n - 1u);
    if (e - s == 1u) {
        perm[i] = o;
        return;
    }
    const uint32_t me = id_old[o];
    uint32_t rank = 0;
    for (uint32_t q = s; q < e; ++q) rank += id_old[min(kv[q].y, n - 1u)] < me ? 1u : 0u;
    perm[s + rank] = o;
}

void launch_runs(SceneBuffers &s, const SubstepConsts &c, const uint2 *sorted, uint32_t key_bits, hipStream_t st)
{
    const uint32_t n = c.n;
    const uint32_t nb = (n + kRunTile - 1) / kRunTile;
    hipLaunchKernelGGL(k_run_count, dim3(nb), dim3(kBlock), 0, st, sorted, n, s.block_runs);
    hipLaunchKernelGGL(k_scan_block_runs, dim3(1), dim3(1024), 0, st, s.block_runs, nb, s.run_count);
    hipLaunchKernelGGL(k_run_emit, dim3(nb), dim3(kBlock), 0, st, sorted, n, s.block_runs, s.elem_run, s.run_key,
                       s.run_start);
    const uint32_t nbuckets = 1u << s.bucket_bits;
    const uint32_t shift = key_bits - s.bucket_bits;
    hipLaunchKernelGGL(k_bucket_fill, dim3((nbuckets + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s.run_key,
                       s.run_count, shift, nbuckets, s.bucket);
    hipLaunchKernelGGL(k_rank_fix, dim3(grid1(n)), dim3(kBlock), 0, st, sorted, n, s.elem_run, s.run_start,
                       s.id[s.cur], s.perm);
}

// ---------------------------------------------------------------------------------------------
// reorder: gather state into the new slot order, predict, cell coordinates, boundary candidates
// ---------------------------------------------------------------------------------------------

// Candidate colliders of a particle as a bitmask over the scene's colliders in LBVH visit order:
// a fluid particle's fluid-boundary candidates (physics_particle_count_fluid_boundary_candidates
// .cs.hlsl:20-28 query radius, :43-85 filters), a solid particle's particle-rigid candidates
// (physics_particle_count_particle_rigid_candidates.cs.hlsl:112-151: the box of its radius; a
// static body takes solid particles only; every collider is a body of its own).
// The half width of a particle's candidate query box: 0 when it takes no candidates.
__device__ __forceinline__ float candidate_query(const ParticleClass &pc, const cn_fluid_material *fm)
{
    if (pc.kind == CN_PARTICLE_KIND_FLUID) {
        const float r = fmaxd(pc.radius, 0.0f);
        const cn_fluid_material m = fm[pc.fluid_mat];
        const float support = fmaxd(m.smoothing_radius, r);
        const float reuse = fmaxd(fmaxd(m.cfl_radius, r), support);
        return support + reuse;
    }
    return pc.kind == CN_PARTICLE_KIND_SOLID ? pc.radius : 0.0f;
}

__device__ __forceinline__ bool candidate_ok(const Collider &cc, const ParticleClass &pc, float x0, float y0, float z0,
                                             float x1, float y1, float z1)
{
    const bool ov = x0 <= cc.aabb_max[0] && x1 >= cc.aabb_min[0] && y0 <= cc.aabb_max[1] && y1 >= cc.aabb_min[1] &&
                    z0 <= cc.aabb_max[2] && z1 >= cc.aabb_min[2];
    return ov && cc.enabled && cc.env == pc.env && (pc.mask & cc.layer) != 0u && (cc.mask & pc.layer) != 0u;
}

// A particle's candidate colliders in the scene's LBVH visit order: a bitmask over the scene's
// colliders for a scene of at most kMaskColliders, else the count (with kBndList) of the entries
// written to the slot's list, the colliders found group by group; more than kBndListCap are
// counted in meta[4] and the first kBndListCap kept.
__device__ __forceinline__ uint32_t boundary_candidates(const ParticleClass &pc, const cn_fluid_material *fm,
                                                        const Collider *__restrict__ col, uint32_t ncol,
                                                        const ColGroup *__restrict__ grp, uint16_t *__restrict__ list,
                                                        uint32_t *__restrict__ meta, float x, float y, float z)
{
    if (ncol == 0 || (pc.kind != CN_PARTICLE_KIND_FLUID && pc.kind != CN_PARTICLE_KIND_SOLID)) return 0u;
    const float qr = candidate_query(pc, fm);
    const float x0 = x - qr, y0 = y - qr, z0 = z - qr;
    const float x1 = x + qr, y1 = y + qr, z1 = z + qr;
    if (ncol <= kMaskColliders) {
        uint32_t mask = 0;
        for (uint32_t k = 0; k < ncol; ++k)
            if (candidate_ok(col[k], pc, x0, y0, z0, x1, y1, z1)) mask |= 1u << k;
        return mask;
    }
    uint32_t n = 0;
    const uint32_t ngrp = (ncol + kColGroup - 1) / kColGroup;
    for (uint32_t gi = 0; gi < ngrp; ++gi) {
        const ColGroup g = grp[gi];
        if (!(x0 <= g.mx[0] && x1 >= g.mn[0] && y0 <= g.mx[1] && y1 >= g.mn[1] && z0 <= g.mx[2] && z1 >= g.mn[2]))
            continue;
        for (uint32_t k = g.first; k < g.first + g.count; ++k) {
            if (!candidate_ok(col[k], pc, x0, y0, z0, x1, y1, z1)) continue;
            if (n < kBndListCap) list[n] = (uint16_t)k;
            ++n;
        }
    }
    if (n > kBndListCap) atomicAdd(meta + 4, 1u);
    return n ? kBndList | min(n, kBndListCap) : 0u;
}

__global__ __launch_bounds__(kBlock) void k_reorder(
    const uint32_t *__restrict__ perm, const float4 *__restrict__ pos_o, const float4 *__restrict__ vel_o,
    const uint32_t *__restrict__ cls_o, const uint32_t *__restrict__ id_o, float4 *__restrict__ pos_n,
    float4 *__restrict__ vel_n, uint32_t *__restrict__ cls_n, uint32_t *__restrict__ id_n,
    const float4 *__restrict__ prev_o, float4 *__restrict__ prev, float4 *__restrict__ cur,
    int4 *__restrict__ cellc, uint32_t *__restrict__ bnd, uint16_t *__restrict__ bndl,
    uint32_t *__restrict__ slot_of, const ParticleClass *__restrict__ classes,
    const cn_fluid_material *__restrict__ fm, const Collider *__restrict__ col, const ColGroup *__restrict__ grp,
    const SceneInfo *__restrict__ scenes, const uint32_t *__restrict__ active, uint32_t *__restrict__ meta,
    SubstepConsts c)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= c.n) return;
    const uint32_t o = min(perm[t], c.n - 1u);  // perm is a permutation; the clamp only bounds reads
    float4 p = pos_o[o], v = vel_o[o], pv;
    const uint32_t k = cls_o[o];
    const ParticleClass pc = classes[c.uniform_class ? 0u : k];
    // a scene this step leaves untouched keeps its state; its particles only change slot
    if (active[pc.scene]) predict(p, v, pv, gravity_scale(pc, fm), c);
    else pv = prev_o[o];
    pos_n[t] = p;
    cur[t] = p;  // solver position x* + delta with delta = 0 at the first iteration
    vel_n[t] = v;
    prev[t] = pv;
    cls_n[t] = k;
    const uint32_t pid = id_o[o];
    id_n[t] = pid;
    if (slot_of) slot_of[min(pid, c.npid - 1u)] = t;
    const SceneInfo sc = scenes[pc.scene];
    const int3 g = cell_of(p.x, p.y, p.z, scene_inv_cell(sc, c));
    // w: the scene's key prefix, from which neighbor-cell keys are formed
    cellc[t] = make_int4(g.x, g.y, g.z, (int)sc.key_prefix);
    // a scene with more than kMaskColliders colliders has groups, and the pool a list per slot
    bnd[t] = boundary_candidates(pc, fm, col + sc.col0, bndl ? sc.ncol : min(sc.ncol, kMaskColliders),
                                 grp + sc.grp0, bndl + (size_t)t * kBndListCap, meta, p.x, p.y, p.z);
}

// (cx, cy, cz, pure): pure runs hold a single cell; mixed runs (hash collisions or wrapped
// Morton coordinates) are filtered per particle like the reference (:94-98).
__global__ __launch_bounds__(kBlock) void k_run_info(const uint32_t *__restrict__ run_start,
                                                     const uint32_t *__restrict__ run_count,
                                                     const int4 *__restrict__ cellc, int4 *__restrict__ run_info,
                                                     uint32_t *__restrict__ meta)
{
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    uint32_t occ = 0;
    if (r < *run_count) {
        const uint32_t s = run_start[r], e = run_start[r + 1];
        occ = e - s;
        const int4 c0 = cellc[s];
        int pure = 1;
        for (uint32_t q = s + 1; q < e; ++q) {
            const int4 cq = cellc[q];
            if (cq.x != c0.x || cq.y != c0.y || cq.z != c0.z) {
                pure = 0;
                break;
            }
        }
        run_info[r] = make_int4(c0.x, c0.y, c0.z, pure);
    }
    // longest run: wave maximum, and an atomic only when it beats the value already stored
    for (int o = 32; o > 0; o >>= 1) occ = max(occ, (uint32_t)__shfl_xor((int)occ, o));
    if ((threadIdx.x & 63u) == 0 && occ > __builtin_nontemporal_load(meta + 8)) atomicMax(meta + 8, occ);
}

void launch_reorder(SceneBuffers &s, const SubstepConsts &c, hipStream_t st)
{
    const int o = s.cur, nw = 1 - s.cur;
    hipLaunchKernelGGL(k_reorder, dim3(grid1(c.n)), dim3(kBlock), 0, st, s.perm, s.pos[o], s.vel[o], s.cls[o],
                       s.id[o], s.pos[nw], s.vel[nw], s.cls[nw], s.id[nw], s.prev[o], s.prev[nw], s.curpos, s.cellc,
                       s.bnd, s.bndl,
                       s.ne + s.nb + s.nd + s.nt ? s.slot_of : nullptr, s.classes, s.fluid_mats, s.colliders,
                       s.col_groups, s.scenes, s.scene_active, s.meta, c);
    hipLaunchKernelGGL(k_run_info, dim3(grid1(c.n)), dim3(kBlock), 0, st, s.run_start, s.run_count, s.cellc,
                       s.run_info, s.meta);
    s.cur = nw;
}

// ---------------------------------------------------------------------------------------------
// Morton blocks: runs grouped by key >> 6 (aligned 4x4x4 cells), the tiles of the LDS walk
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ bool block_flag(const uint32_t *run_key, uint32_t r, uint32_t R)
{
    return r < R && (r == 0 || (run_key[r] >> 6) != (run_key[r - 1] >> 6));
}

__global__ __launch_bounds__(kBlock) void k_blk_count(const uint32_t *__restrict__ run_key,
                                                      const uint32_t *__restrict__ run_count,
                                                      uint32_t *__restrict__ block_cnt, uint32_t *__restrict__ nfluid)
{
    __shared__ uint32_t sh[kBlock];
    if (blockIdx.x == 0 && threadIdx.x == 0) *nfluid = 0u;
    const uint32_t R = *run_count;
    const uint32_t base = blockIdx.x * kRunTile + threadIdx.x * kRunItems;
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < kRunItems; ++k) cnt += block_flag(run_key, base + k, R) ? 1u : 0u;
    uint32_t total;
    block_exclusive_scan(cnt, sh, total);
    if (threadIdx.x == 0) block_cnt[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_blk_emit(const uint32_t *__restrict__ run_key,
                                                     const uint32_t *__restrict__ run_count,
                                                     const uint32_t *__restrict__ block_off,
                                                     uint32_t *__restrict__ blk_first_run,
                                                     uint32_t *__restrict__ nblk, uint32_t *__restrict__ work,
                                                     uint32_t nwork, uint32_t last_word, uint32_t prev_word,
                                                     uint32_t prefix_shift, uint32_t fluid_scenes)
{
    __shared__ uint32_t sh[kBlock];
    const uint32_t R = *run_count;
    const uint32_t base = blockIdx.x * kRunTile + threadIdx.x * kRunItems;
    bool f[kRunItems];
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < kRunItems; ++k) {
        f[k] = block_flag(run_key, base + k, R);
        cnt += f[k] ? 1u : 0u;
    }
    uint32_t total;
    uint32_t idx = block_off[blockIdx.x] + block_exclusive_scan(cnt, sh, total);
#pragma unroll
    for (int k = 0; k < kRunItems; ++k) {
        const uint32_t r = base + k;
        if (r >= R) break;
        if (f[k]) {
            // scenes with fluid hold the lowest key prefixes, so their blocks come first
            if ((uint32_t)((uint64_t)run_key[r] >> prefix_shift) < fluid_scenes) atomicMax(nblk + 1, idx + 1);
            blk_first_run[idx++] = r;
        }
        if (r == R - 1) {
            blk_first_run[idx] = R;
            *nblk = idx;
        }
    }
    // the substep's queue counters, cleared by every workgroup's share (the max-displacement words
    // past nwork are kept)
    for (uint32_t w = blockIdx.x * kBlock + threadIdx.x; w < nwork; w += gridDim.x * kBlock) work[w] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) work[prev_word] = work[last_word];
}

void launch_blocks(SceneBuffers &s, const SubstepConsts &c, hipStream_t st)
{
    const uint32_t nb = (c.n + kRunTile - 1) / kRunTile;
    hipLaunchKernelGGL(k_blk_count, dim3(nb), dim3(kBlock), 0, st, s.run_key, s.run_count, s.block_runs, s.nblk + 1);
    hipLaunchKernelGGL(k_scan_block_runs, dim3(1), dim3(1024), 0, st, s.block_runs, nb, s.nblk);
    const uint32_t last = kStateMaxDisp + std::min<uint32_t>(std::max<uint32_t>(c.fluid_iters, 1u), kIterStateWords) - 1;
    hipLaunchKernelGGL(k_blk_emit, dim3(nb), dim3(kBlock), 0, st, s.run_key, s.run_count, s.block_runs,
                       s.blk_first_run, s.nblk, s.iter_state, kStateMaxDisp, last, kStatePrevDisp, 3u * c.morton_bits,
                       c.fluid_scenes);
}

// ---------------------------------------------------------------------------------------------
// finalize: physics_particle_update_velocities.cs.hlsl:13-44 then
// physics_fluid_project_boundary_velocities.cs.hlsl:26-118 on the same particle
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_finalize(float4 *__restrict__ pos, float4 *__restrict__ vel,
                                                     const float4 *__restrict__ prev, const uint32_t *__restrict__ cls,
                                                     const float4 *__restrict__ dacc, const uint32_t *__restrict__ bnd,
                                                     const uint16_t *__restrict__ bndl,
                                                     const ParticleClass *__restrict__ classes,
                                                     const float4 *__restrict__ part_mats,
                                                     const Collider *__restrict__ col,
                                                     const SceneInfo *__restrict__ scenes,
                                                     const uint32_t *__restrict__ active, SubstepConsts c)
{
    const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= c.n) return;
    const ParticleClass pc = classes[c.uniform_class ? 0u : cls[i]];
    if (!active[pc.scene]) return;
    float4 p = pos[i];
    float vx, vy, vz;
    const float dt = c.dt;
    const bool fluid = pc.kind == CN_PARTICLE_KIND_FLUID;
    const float4 pm = part_mats[pc.part_mat];
    if (p.w <= kEps || dt <= kEps) {
        vx = vy = vz = 0.0f;
    } else {
        const float4 pv = prev[i];
        if (fluid) {
            const float4 d = dacc[i];
            p.x = p.x + d.x;
            p.y = p.y + d.y;
            p.z = p.z + d.z;
            pos[i] = p;
        }
        const float damping = fmaxd(pm.z, 0.0f);
        const float scale = fmaxd(1.0f - damping * dt, 0.0f);
        vx = fdiv(p.x - pv.x, dt) * scale;
        vy = fdiv(p.y - pv.y, dt) * scale;
        vz = fdiv(p.z - pv.z, dt) * scale;
    }
    const SceneInfo sc = scenes[pc.scene];
    if (fluid && sc.ncol > 0) {
        const float r = fmaxd(pc.radius, 0.0f);
        const float slop = c.cfg[kCfgContactSlop];
        const float damping = fmaxd(pm.z, 0.0f);
        float best = 3.402823466e+38f, bnx = 0.0f, bny = 1.0f, bnz = 0.0f, bmf = 0.0f, bms = 0.0f;
        bool found = false;
        const Collider *scol = col + sc.col0;
        each_candidate(bnd[i], bndl, i, [&](uint32_t k) {
            float sd, nx, ny, nz;
            if (!collider_contact(scol[k], p.x, p.y, p.z, sd, nx, ny, nz)) return;
            const float cdist = r + 2.0f * slop;
            if (sd > cdist) return;
            if (sd < best) {
                best = sd;
                bnx = nx;
                bny = ny;
                bnz = nz;
                bmf = fsqrt(fmaxd(0.0f, pm.x) * fmaxd(0.0f, scol[k].material[0]));
                bms = fsqrt(fmaxd(0.0f, pm.w) * fmaxd(0.0f, scol[k].material[3]));
                found = true;
            }
        });
        if (found) {
            const float pen = fmaxd(r - best, 0.0f);
            const float cdist = fmaxd(r + 2.0f * slop, 1.0e-5f);
            const float cw = fsat(fdiv(cdist - best, cdist));
            float rx = vx, ry = vy, rz = vz;
            const float nv = dot3(rx, ry, rz, bnx, bny, bnz);
            if (nv < 0.0f) {
                rx = rx - bnx * nv;
                ry = ry - bny * nv;
                rz = rz - bnz * nv;
            }
            const float rn = dot3(rx, ry, rz, bnx, bny, bnz);
            float tx = rx - bnx * rn, ty = ry - bny * rn, tz = rz - bnz * rn;
            if (dt > kEps) {
                // ComputePositionFrictionDelta, physics_rigid_solver_shared.hlsli:50-67
                const float ex = tx * dt, ey = ty * dt, ez = tz * dt;
                const float penf = fmaxd(pen, slop);
                const float td = len3(ex, ey, ez);
                float fx = 0.0f, fy = 0.0f, fz = 0.0f;
                if (!(td <= c.cfg[kCfgPosFrictionMin] || penf <= kEps)) {
                    const float sl = fsat(bms) * penf;
                    if (td <= sl) {
                        fx = ex;
                        fy = ey;
                        fz = ez;
                    } else {
                        const float ks = fmind(fdiv(fsat(bmf) * penf, td), 1.0f);
                        fx = ex * ks;
                        fy = ey * ks;
                        fz = ez * ks;
                    }
                }
                tx = tx - fdiv(fx, dt);
                ty = ty - fdiv(fy, dt);
                tz = tz - fdiv(fz, dt);
            }
            const float tds = fmaxd(1.0f - damping * cw * dt, 0.0f);
            tx = tx * tds;
            ty = ty * tds;
            tz = tz * tds;
            const float rn2 = fmaxd(dot3(rx, ry, rz, bnx, bny, bnz), 0.0f);
            vx = bnx * rn2 + tx;
            vy = bny * rn2 + ty;
            vz = bnz * rn2 + tz;
        }
    }
    vel[i] = make_float4(vx, vy, vz, 0.0f);
}

void launch_finalize(const SceneBuffers &s, const SubstepConsts &c, int din, hipStream_t st)
{
    hipLaunchKernelGGL(k_finalize, dim3(grid1(c.n)), dim3(kBlock), 0, st, s.pos[s.cur], s.vel[s.cur], s.prev[s.cur],
                       s.cls[s.cur], s.dacc[din], s.bnd, s.bndl, s.classes, s.part_mats, s.colliders, s.scenes,
                       s.scene_active, c);
}

// slot order -> particle-index order
__global__ __launch_bounds__(kBlock) void k_export(const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                                   const float4 *__restrict__ prev, const uint32_t *__restrict__ id,
                                                   uint32_t n, uint32_t npid, float4 *__restrict__ opos,
                                                   float4 *__restrict__ ovel, float4 *__restrict__ oprev)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t d = min(id[t], npid - 1u);
    if (opos) opos[d] = pos[t];
    if (ovel) ovel[d] = vel[t];
    if (oprev) oprev[d] = prev[t];
}

void launch_export(const SceneBuffers &s, bool state, bool prev, hipStream_t st)
{
    hipLaunchKernelGGL(k_export, dim3(grid1(s.n)), dim3(kBlock), 0, st, s.pos[s.cur], s.vel[s.cur], s.prev[s.cur],
                       s.id[s.cur], s.n, s.npid, state ? s.out_pos : nullptr, state ? s.out_vel : nullptr,
                       prev ? s.out_prev : nullptr);
}

// overwrite the state of the particles [first, first + count) (particle-index order) from P, V
__global__ __launch_bounds__(kBlock) void k_set_state(const uint32_t *__restrict__ id, uint32_t n, uint32_t first,
                                                      uint32_t count, const float4 *__restrict__ P,
                                                      const float4 *__restrict__ V, float4 *__restrict__ pos,
                                                      float4 *__restrict__ vel, float4 *__restrict__ prev,
                                                      float4 *__restrict__ opos, float4 *__restrict__ ovel,
                                                      float4 *__restrict__ oprev)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t d = id[t];
    if (d < first || d - first >= count) return;
    const float4 p = P[d - first], v = V[d - first];
    pos[t] = p;
    vel[t] = v;
    prev[t] = p;
    opos[d] = p;
    ovel[d] = v;
    oprev[d] = p;
}

void launch_set_state(SceneBuffers &s, uint32_t first, uint32_t count, const float4 *P, const float4 *V,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_set_state, dim3(grid1(s.n)), dim3(kBlock), 0, st, s.id[s.cur], s.n, first, count, P, V,
                       s.pos[s.cur], s.vel[s.cur], s.prev[s.cur], s.out_pos, s.out_vel, s.out_prev);
}

// ---------------------------------------------------------------------------------------------
// compact ticks: the slots of the listed scenes of a pool, gathered into a compact slot range and
// put back after the step.  A pool's slot order holds each scene as one segment (the scene's key
// prefix orders the sort), so a listed scene is the segment seg.x .. seg.x + seg.z of the pool and
// seg.y .. seg.y + seg.z of the compact range; segments are in slot order.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t tick_pool_slot(const uint4 *__restrict__ seg, uint32_t nseg, uint32_t j)
{
    uint32_t lo = 0, hi = nseg;  // the last segment whose compact start is <= j
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (seg[mid].y <= j) lo = mid;
        else hi = mid;
    }
    return seg[lo].x + (j - seg[lo].y);
}

__global__ __launch_bounds__(kBlock) void k_tick_gather(const uint4 *__restrict__ seg, uint32_t nseg, uint32_t m,
                                                        const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                                        const float4 *__restrict__ prev, const uint32_t *__restrict__ cls,
                                                        const uint32_t *__restrict__ id, float4 *__restrict__ tpos,
                                                        float4 *__restrict__ tvel, float4 *__restrict__ tprev,
                                                        uint32_t *__restrict__ tcls, uint32_t *__restrict__ tid)
{
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= m) return;
    const uint32_t t = tick_pool_slot(seg, nseg, j);
    tpos[j] = pos[t];
    tvel[j] = vel[t];
    tprev[j] = prev[t];
    tcls[j] = cls[t];
    tid[j] = id[t];
}

__global__ __launch_bounds__(kBlock) void k_tick_scatter(const uint4 *__restrict__ seg, uint32_t nseg, uint32_t m,
                                                         uint32_t npid, const float4 *__restrict__ tpos,
                                                         const float4 *__restrict__ tvel,
                                                         const float4 *__restrict__ tprev,
                                                         const uint32_t *__restrict__ tcls,
                                                         const uint32_t *__restrict__ tid, float4 *__restrict__ pos,
                                                         float4 *__restrict__ vel, float4 *__restrict__ prev,
                                                         uint32_t *__restrict__ cls, uint32_t *__restrict__ id,
                                                         uint32_t *__restrict__ slot_of)
{
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= m) return;
    const uint32_t t = tick_pool_slot(seg, nseg, j);
    const uint32_t pid = tid[j];
    pos[t] = tpos[j];
    vel[t] = tvel[j];
    prev[t] = tprev[j];
    cls[t] = tcls[j];
    id[t] = pid;
    if (slot_of) slot_of[min(pid, npid - 1u)] = t;
}

void launch_tick_gather(const SceneBuffers &pool, const SceneBuffers &tick, const uint4 *seg, uint32_t nseg,
                        hipStream_t st)
{
    const int o = pool.cur, w = tick.cur;
    hipLaunchKernelGGL(k_tick_gather, dim3(grid1(tick.n)), dim3(kBlock), 0, st, seg, nseg, tick.n, pool.pos[o],
                       pool.vel[o], pool.prev[o], pool.cls[o], pool.id[o], tick.pos[w], tick.vel[w], tick.prev[w],
                       tick.cls[w], tick.id[w]);
}

void launch_tick_scatter(const SceneBuffers &pool, const SceneBuffers &tick, const uint4 *seg, uint32_t nseg,
                         hipStream_t st)
{
    const int o = tick.cur, w = pool.cur;
    hipLaunchKernelGGL(k_tick_scatter, dim3(grid1(tick.n)), dim3(kBlock), 0, st, seg, nseg, tick.n, pool.npid,
                       tick.pos[o], tick.vel[o], tick.prev[o], tick.cls[o], tick.id[o], pool.pos[w], pool.vel[w],
                       pool.prev[w], pool.cls[w], pool.id[w],
                       pool.ne + pool.nb + pool.nd + pool.nt ? pool.slot_of : nullptr);
}

// word fill as a kernel (stays a kernel node when the step is captured into a graph)
__global__ __launch_bounds__(kBlock) void k_fill(uint32_t *__restrict__ p, uint32_t words, uint32_t value)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < words) p[i] = value;
}

void launch_fill(uint32_t *p, uint32_t words, uint32_t value, hipStream_t st)
{
    if (words) hipLaunchKernelGGL(k_fill, dim3(grid1(words)), dim3(kBlock), 0, st, p, words, value);
}

}  // namespace cn
