// cn_fluid.hip -- PBF fluid iterations on gfx950: neighbor walk + density constraint, neighbor
// filter + density constraint, and delta + apply + clamp.
//
// Reference passes (assets/shaders/physics/fluid/solver/): physics_fluid_build_neighbors.cs.hlsl
// :21-147, physics_fluid_compute_density_lambda.cs.hlsl:26-115,
// physics_fluid_compute_delta_positions.cs.hlsl:27-157, physics_fluid_apply_delta_positions.cs.hlsl
// :11-31, physics_fluid_clamp_boundary.cs.hlsl:23-80.  The reference rebuilds each fluid particle's
// neighbor list every iteration: the fluid particles within h of its current position x* + delta,
// found by walking the 27 cells around its broadphase cell in dz,dy,dx order, each cell's particles
// in ascending particle index, keeping the first max_fluid_neighbors.
//
// Here the first iteration of a substep walks the cells once and stores a superset list (radius
// h + skin, same walk order) and every later iteration filters that list by the current distance
// < h, which yields the same neighbors in the same order as long as no particle has moved more than
// skin / 2 since the walk.  The delta pass tracks the largest displacement; when the bound could be
// crossed the next iteration walks the cells again (exact lists), so results never depend on the
// skin.  The walk itself is tiled: a workgroup takes one aligned 4x4x4 block of cells (one Morton
// block), stages the current positions of its 6x6x6-cell halo in LDS and walks from there.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cn_internal.h"

namespace cn {

constexpr uint32_t kExact = 0x80000000u;  // nbr_cnt flag: the stored list is exact, not a superset
constexpr int kHalo = 216;                // 6 x 6 x 6 cells around a 4 x 4 x 4 block
constexpr int kTileMax = 2560;            // halo particles a workgroup stages in LDS

struct FluidArgs {
    const float4 *pos;
    const uint32_t *cls;
    const int4 *cellc;
    const uint32_t *bnd;
    const float4 *din;
    float4 *dout;
    float4 *cur;
    float4 *cs;
    float4 *nc;
    uint32_t *nbr;
    uint32_t *nbr_cnt;
    const uint32_t *run_key;
    const uint32_t *run_start;
    const int4 *run_info;
    const uint32_t *bucket;
    uint32_t bucket_shift;
    const uint32_t *blk_first_run;
    const uint32_t *nblk;
    uint32_t *iter_state;
    const ParticleClass *classes;
    const cn_fluid_material *fm;
    const Collider *col;
    uint32_t *meta;
    uint32_t it;
};

__device__ __forceinline__ uint32_t *work_counter(const FluidArgs &a) { return a.iter_state + min(a.it, kIterStateWords - 1); }
__device__ __forceinline__ uint32_t *max_disp(const FluidArgs &a, uint32_t it) { return a.iter_state + kIterStateWords + it; }
__device__ __forceinline__ uint32_t *walk_flag(const FluidArgs &a, uint32_t it) { return a.iter_state + 2 * kIterStateWords + it; }

// True when iteration a.it must build its lists by walking the cells: the first iteration, after a
// fallback earlier in the substep, or when the displacement bound of the superset lists is spent.
__device__ __forceinline__ bool need_walk(const FluidArgs &a, const SubstepConsts &c)
{
    if (a.it == 0 || a.it >= kIterStateWords) return true;
    if (*walk_flag(a, a.it - 1) != 0u) return true;
    return !(__uint_as_float(*max_disp(a, a.it - 1)) < c.fallback_d2);
}

__device__ __forceinline__ int find_run(uint32_t key, const FluidArgs &a)
{
    const uint32_t b = key >> a.bucket_shift;
    uint32_t r = a.bucket[b];
    const uint32_t e = a.bucket[b + 1];
    while (r < e && a.run_key[r] < key) ++r;
    return (r < e && a.run_key[r] == key) ? (int)r : -1;
}

// ShouldProcessFluidNeighbor, physics_fluid_neighbor_build.hlsli:48-70 (the index test is the caller's)
__device__ __forceinline__ bool neighbor_ok(const ParticleClass &self, const ParticleClass &o)
{
    return o.kind == CN_PARTICLE_KIND_FLUID && o.env == self.env && (self.mask & o.layer) != 0u &&
           (o.mask & self.layer) != 0u;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// One particle's neighbor walk / filter state.
struct Walk {
    uint32_t i;
    float sx, sy, sz;   // x* + delta of the particle
    float h, rl;        // smoothing radius, list radius (h or h + skin)
    uint32_t lcap;      // list slots (cap for exact lists, the allocation for supersets)
    bool tension;
    bool acc;           // accumulate density and counts (false on the list-only second pass)
    ParticleClass self;
    float density, snx, sny, snz;
    uint32_t lcnt, inr, req;
};

// One candidate of the walk, in walk order.  The density sum runs over the first `cap` in-radius
// neighbors in that order, exactly the reference's list.
template <bool UNIFORM>
__device__ __forceinline__ void visit(Walk &w, uint32_t q, const float4 pj, uint32_t qcls, bool store,
                                      const FluidArgs &a, const SubstepConsts &c)
{
    if (q == w.i) return;
    if (!UNIFORM && !neighbor_ok(w.self, a.classes[qcls])) return;
    if (pj.w <= kEps) return;
    const float dx = w.sx - pj.x, dy = w.sy - pj.y, dz = w.sz - pj.z;
    const float dist = len3(dx, dy, dz);
    if (store && dist < w.rl) {
        if (w.lcnt < w.lcap) a.nbr[(size_t)w.lcnt * c.n + w.i] = q;
        ++w.lcnt;
    }
    if (!w.acc || dist >= w.h) return;
    ++w.req;
    if (w.inr >= c.cap) return;
    ++w.inr;
    w.density = w.density + spiky_w(dist, w.h);
    if (w.tension) {
        // FluidSpikyKernelGradient, physics_fluid_common.hlsli:44-53
        float gx = 0.0f, gy = 0.0f, gz = 0.0f;
        if (!(dist <= kEps || dist >= w.h)) {
            const float dw = spiky_dw(dist, w.h);
            gx = dw * fdiv(dx, dist);
            gy = dw * fdiv(dy, dist);
            gz = dw * fdiv(dz, dist);
        }
        w.snx = w.snx + gx;
        w.sny = w.sny + gy;
        w.snz = w.snz + gz;
    }
}

// 27-cell walk through the run table (hash keys, fallback tiles, single exact re-walks).
template <bool UNIFORM>
__device__ void walk_global(Walk &w, int4 ci, bool store, const FluidArgs &a, const SubstepConsts &c)
{
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int tx = ci.x + dx, ty = ci.y + dy, tz = ci.z + dz;
                // a cell whose hash equals the empty-table marker is never found
                // (physics_fluid_build_neighbors.cs.hlsl:81-85 with FindParticleCellRange)
                if (cell_hash(tx, ty, tz) == kInvalid) continue;
                const int r = find_run(cell_key(tx, ty, tz, c.key_mode, c.morton_bits), a);
                if (r < 0) continue;
                const int4 info = a.run_info[r];
                const bool pure = info.w != 0;
                if (pure && (info.x != tx || info.y != ty || info.z != tz)) continue;
                const uint32_t qe = a.run_start[r + 1];
                for (uint32_t q = a.run_start[r]; q < qe; ++q) {
                    if (!pure) {
                        const int4 cq = a.cellc[q];
                        if (cq.x != tx || cq.y != ty || cq.z != tz) continue;
                    }
                    visit<UNIFORM>(w, q, a.cur[q], UNIFORM ? 0u : a.cls[q], store, a, c);
                }
            }
}

struct MetaAcc {
    uint32_t req = 0, trunc = 0, listed = 0;
};

__device__ __forceinline__ void begin(Walk &w, uint32_t i, const ParticleClass &pc, const cn_fluid_material &m,
                                      const float4 s, bool exact, const SubstepConsts &c)
{
    w.i = i;
    w.sx = s.x;
    w.sy = s.y;
    w.sz = s.z;
    w.h = fmaxd(m.smoothing_radius, 1.0e-4f);
    w.rl = exact ? w.h : w.h + c.skin;
    w.lcap = exact ? c.cap : c.list_cap;
    w.tension = fmaxd(m.surface_tension, 0.0f) > kEps;
    w.acc = true;
    w.self = pc;
    w.density = w.snx = w.sny = w.snz = 0.0f;
    w.lcnt = w.inr = w.req = 0;
}

// Boundary ghost density and the density constraint
// (physics_fluid_compute_density_lambda.cs.hlsl:97-114, physics_fluid_boundary_common.hlsli:112-138).
__device__ __forceinline__ void finish(const Walk &w, const cn_fluid_material &m, uint32_t list_word, bool write_list,
                                       MetaAcc &ma, const FluidArgs &a, const SubstepConsts &c)
{
    float density = w.density;
    const uint32_t bm = a.bnd[w.i];
    const float bscale = c.cfg[kCfgFluidBndDensity];
    for (uint32_t k = 0; k < c.ncolliders; ++k) {
        if (!((bm >> k) & 1u)) continue;
        float gx, gy, gz;
        if (!ghost_delta(a.col[k], w.sx, w.sy, w.sz, w.h, gx, gy, gz)) continue;
        const float gd = len3(gx, gy, gz);
        if (gd >= w.h) continue;
        density = density + spiky_w(gd, w.h) * bscale;
    }
    const float rho0 = fmaxd(m.rest_density, 1.0f);
    const float C = fmaxd(density - rho0, -rho0 * fmaxd(c.cfg[kCfgFluidMaxUnder], 0.0f)) * m.density_constraint_scale;
    a.cs[w.i] = make_float4(w.sx, w.sy, w.sz, C);
    if (w.tension) {
        const float st = fmaxd(m.surface_tension, 0.0f);
        a.nc[w.i] = make_float4(w.snx * st, w.sny * st, w.snz * st, C);
    }
    if (write_list) a.nbr_cnt[w.i] = list_word;
    ma.req = max(ma.req, w.req);
    ma.trunc += w.req > c.cap ? 1u : 0u;
    ma.listed += w.inr;
}

__device__ __forceinline__ void inactive(uint32_t i, const FluidArgs &a)
{
    const float4 s = a.cur[i];
    a.cs[i] = make_float4(s.x, s.y, s.z, 0.0f);
    a.nbr_cnt[i] = kExact;
}

// Walk one particle through the run table and store its list (superset on the first iteration).
template <bool UNIFORM>
__device__ void walk_particle_global(uint32_t i, bool exact, MetaAcc &ma, const FluidArgs &a, const SubstepConsts &c)
{
    const ParticleClass pc = a.classes[UNIFORM ? 0u : a.cls[i]];
    const float4 s = a.cur[i];
    if (pc.kind != CN_PARTICLE_KIND_FLUID || s.w <= kEps) {
        inactive(i, a);
        return;
    }
    const cn_fluid_material m = a.fm[pc.fluid_mat];
    Walk w;
    begin(w, i, pc, m, s, exact, c);
    const bool self_ok = !UNIFORM || neighbor_ok(pc, pc);
    const int4 ci = a.cellc[i];
    if (c.cap > 0 && self_ok) {
        walk_global<UNIFORM>(w, ci, true, a, c);
        if (w.lcnt > w.lcap) {  // superset overflow: a second, list-only pass stores the exact list
            w.rl = w.h;
            w.lcap = c.cap;
            w.lcnt = 0;
            w.acc = false;
            walk_global<UNIFORM>(w, ci, true, a, c);
            exact = true;
        }
    }
    const uint32_t word = min(w.lcnt, w.lcap) | (exact ? kExact : 0u);
    finish(w, m, word, true, ma, a, c);
}

__device__ __forceinline__ void flush_meta(const MetaAcc &ma, const FluidArgs &a)
{
    const uint32_t wmax = wave_max(ma.req);
    const uint32_t wtr = wave_sum(ma.trunc);
    const uint32_t wls = wave_sum(ma.listed);
    if ((threadIdx.x & 63u) == 0) {
        if (wmax) atomicMax(&a.meta[0], wmax);
        if (wtr) atomicAdd(&a.meta[1], wtr);
        if (a.it == 0 && wls) atomicAdd((unsigned long long *)&a.meta[2], (unsigned long long)wls);
    }
}

__device__ __forceinline__ void walk_bookkeeping(bool walk, const FluidArgs &a)
{
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.it < kIterStateWords) {
        *max_disp(a, a.it) = 0u;
        *walk_flag(a, a.it) = (a.it > 0 && walk) ? 1u : 0u;
        if (a.it > 0 && walk) atomicAdd(&a.meta[4], 1u);
    }
}

// Thread-per-particle walk (hash keys).
template <bool UNIFORM>
__global__ __launch_bounds__(kBlock) void k_fluid_walk_global(FluidArgs a, SubstepConsts c)
{
    const bool walk = need_walk(a, c);
    walk_bookkeeping(walk, a);
    if (!walk) return;
    MetaAcc ma;
    const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i < c.n) walk_particle_global<UNIFORM>(i, a.it > 0, ma, a, c);
    flush_meta(ma, a);
}

// Tiled walk (Morton keys): persistent workgroups take Morton blocks from a work counter.
template <bool UNIFORM>
__global__ __launch_bounds__(kBlock) void k_fluid_walk_tiled(FluidArgs a, SubstepConsts c)
{
    __shared__ float4 lcur[kTileMax];
    __shared__ uint32_t lcls[UNIFORM ? 1 : kTileMax];
    __shared__ uint32_t hstart[kHalo];
    __shared__ uint32_t hoff[kHalo + 1];
    __shared__ uint32_t scan[kBlock];
    __shared__ uint32_t s_blk;
    __shared__ int s_base[3];

    const bool walk = need_walk(a, c);
    walk_bookkeeping(walk, a);
    if (!walk) return;
    const bool exact = a.it > 0;
    const uint32_t tid = threadIdx.x;
    const uint32_t nblk = *a.nblk;
    MetaAcc ma;
    for (;;) {
        if (tid == 0) s_blk = atomicAdd(work_counter(a), 1u);
        __syncthreads();
        const uint32_t b = s_blk;
        if (b >= nblk) break;
        const uint32_t r0 = a.blk_first_run[b], r1 = a.blk_first_run[b + 1];
        const uint32_t p0 = a.run_start[r0], p1 = a.run_start[r1];
        if (tid == 0) {
            const int4 c0 = a.cellc[p0];
            s_base[0] = c0.x & ~3;
            s_base[1] = c0.y & ~3;
            s_base[2] = c0.z & ~3;
        }
        __syncthreads();
        const int bx = s_base[0], by = s_base[1], bz = s_base[2];
        bool bad = false;
        for (uint32_t i = p0 + tid; i < p1; i += kBlock) {
            const int4 ci = a.cellc[i];
            bad |= (uint32_t)(ci.x - bx) > 3u || (uint32_t)(ci.y - by) > 3u || (uint32_t)(ci.z - bz) > 3u;
        }
        uint32_t cnt = 0, st = 0;
        if (tid < (uint32_t)kHalo) {
            const int tx = bx - 1 + (int)(tid % 6u), ty = by - 1 + (int)((tid / 6u) % 6u), tz = bz - 1 + (int)(tid / 36u);
            if (cell_hash(tx, ty, tz) != kInvalid) {
                const int r = find_run(cell_key(tx, ty, tz, c.key_mode, c.morton_bits), a);
                if (r >= 0) {
                    const int4 info = a.run_info[r];
                    if (info.w == 0) bad = true;
                    else if (info.x == tx && info.y == ty && info.z == tz) {
                        st = a.run_start[r];
                        cnt = a.run_start[r + 1] - st;
                    }
                }
            }
            hstart[tid] = st;
        }
        const int any_bad = __syncthreads_or(bad ? 1 : 0);
        uint32_t total;
        const uint32_t off = block_exclusive_scan(cnt, scan, total);
        if (tid < (uint32_t)kHalo) hoff[tid] = off;
        if (tid == 0) hoff[kHalo] = total;
        __syncthreads();
        const bool lds = !any_bad && total <= (uint32_t)kTileMax;
        if (tid == 0) atomicAdd(&a.meta[lds ? 5 : 6], 1u);
        if (lds) {
            for (uint32_t f = tid; f < total; f += kBlock) {
                uint32_t lo = 0, hi = kHalo;  // last h with hoff[h] <= f
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (hoff[mid] <= f) lo = mid;
                    else hi = mid;
                }
                const uint32_t q = hstart[lo] + (f - hoff[lo]);
                lcur[f] = a.cur[q];
                if (!UNIFORM) lcls[f] = a.cls[q];
            }
        }
        __syncthreads();
        for (uint32_t i = p0 + tid; i < p1; i += kBlock) {
            if (!lds) {
                walk_particle_global<UNIFORM>(i, exact, ma, a, c);
                continue;
            }
            const ParticleClass pc = a.classes[UNIFORM ? 0u : a.cls[i]];
            const float4 s = a.cur[i];
            if (pc.kind != CN_PARTICLE_KIND_FLUID || s.w <= kEps) {
                inactive(i, a);
                continue;
            }
            const cn_fluid_material m = a.fm[pc.fluid_mat];
            Walk w;
            begin(w, i, pc, m, s, exact, c);
            const bool self_ok = !UNIFORM || neighbor_ok(pc, pc);
            const int4 ci = a.cellc[i];
            bool ex = exact;
            if (c.cap > 0 && self_ok) {
                const int rx = ci.x - bx + 1, ry = ci.y - by + 1, rz = ci.z - bz + 1;
                for (int pass = 0; pass < 2; ++pass) {
                    for (int dz = -1; dz <= 1; ++dz)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dx = -1; dx <= 1; ++dx) {
                                const int hh = (rx + dx) + 6 * (ry + dy) + 36 * (rz + dz);
                                const uint32_t f0 = hoff[hh], f1 = hoff[hh + 1];
                                const uint32_t qb = hstart[hh] - f0;
                                for (uint32_t f = f0; f < f1; ++f)
                                    visit<UNIFORM>(w, qb + f, lcur[f], UNIFORM ? 0u : lcls[f], true, a, c);
                            }
                    if (w.lcnt <= w.lcap) break;
                    // superset overflow: a second, list-only pass stores the exact list
                    w.rl = w.h;
                    w.lcap = c.cap;
                    w.lcnt = 0;
                    w.acc = false;
                    ex = true;
                }
            }
            const uint32_t word = min(w.lcnt, w.lcap) | (ex ? kExact : 0u);
            finish(w, m, word, true, ma, a, c);
        }
        __syncthreads();
    }
    flush_meta(ma, a);
}

// Later iterations: filter the stored superset list by the current distance.
template <bool UNIFORM>
__global__ __launch_bounds__(kBlock) void k_fluid_filter(FluidArgs a, SubstepConsts c)
{
    if (need_walk(a, c)) return;
    MetaAcc ma;
    const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i < c.n) {
        const ParticleClass pc = a.classes[UNIFORM ? 0u : a.cls[i]];
        const float4 s = a.cur[i];
        if (pc.kind != CN_PARTICLE_KIND_FLUID || s.w <= kEps) {
            inactive(i, a);
        } else {
            const uint32_t word = a.nbr_cnt[i];
            if (word & kExact) {
                walk_particle_global<UNIFORM>(i, true, ma, a, c);
            } else {
                const cn_fluid_material m = a.fm[pc.fluid_mat];
                Walk w;
                begin(w, i, pc, m, s, true, c);
                for (uint32_t k = 0; k < word; ++k) {
                    const uint32_t q = a.nbr[(size_t)k * c.n + i];
                    visit<true>(w, q, a.cur[q], 0u, false, a, c);
                }
                finish(w, m, word, false, ma, a, c);
            }
        }
    }
    flush_meta(ma, a);
}

// physics_fluid_compute_delta_positions.cs.hlsl:27-157 over the stored list (in-radius entries in
// list order up to the slot budget), then physics_fluid_apply_delta_positions.cs.hlsl:11-31 and
// physics_fluid_clamp_boundary.cs.hlsl:23-80 on the same particle.
template <bool UNIFORM>
__global__ __launch_bounds__(kBlock) void k_fluid_delta(FluidArgs a, SubstepConsts c)
{
    const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    float d2 = 0.0f;
    if (i < c.n) {
        const ParticleClass pc = a.classes[UNIFORM ? 0u : a.cls[i]];
        if (pc.kind != CN_PARTICLE_KIND_FLUID) {
            a.dout[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        } else {
            const float4 p = a.pos[i];
            const bool use_d = a.it > 0;
            const float4 di = use_d ? a.din[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float ox = 0.0f, oy = 0.0f, oz = 0.0f, ow = 0.0f;
            if (p.w > kEps) {
                const float4 csi = a.cs[i];
                const float sx = csi.x, sy = csi.y, sz = csi.z, Ci = csi.w;
                const cn_fluid_material m = a.fm[pc.fluid_mat];
                const float h = fmaxd(m.smoothing_radius, 1.0e-4f);
                const float cohesion = c.dt * fmaxd(m.cohesion, 0.0f);
                const float viscosity = c.dt * fmaxd(m.viscosity, 0.0f);
                const float st = fmaxd(m.surface_tension, 0.0f);
                const float cfl = fmaxd(m.cfl_radius, 0.0f);
                const bool tension = st > kEps;
                // relDelta = delta_i - delta_j is zero on the first iteration, and the CFL term can
                // only act when |relDelta| >= cfl, i.e. never while 2 max|delta| < cfl
                bool cfl_on = false;
                if (use_d && cfl > kEps) {
                    const float md2 = a.it - 1 < kIterStateWords ? __uint_as_float(*max_disp(a, a.it - 1)) : 3.4e38f;
                    cfl_on = !(4.0f * md2 * c.cfl_guard < cfl * cfl);
                }
                const bool visc_on = use_d && viscosity > kEps;
                const bool need_rel = cfl_on || visc_on;
                float4 nci = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (tension) nci = a.nc[i];
                float dpx = 0.0f, dpy = 0.0f, dpz = 0.0f, wsum = 0.0f;
                const uint32_t cnt = a.nbr_cnt[i] & ~kExact;
                uint32_t inr = 0;
                for (uint32_t k = 0; k < cnt; ++k) {
                    const uint32_t q = a.nbr[(size_t)k * c.n + i];
                    const float4 cq = a.cs[q];
                    const float ddx = sx - cq.x, ddy = sy - cq.y, ddz = sz - cq.z;
                    const float dist = len3(ddx, ddy, ddz);
                    if (dist >= h) continue;
                    if (inr >= c.cap) break;
                    ++inr;
                    if (dist <= kEps) continue;
                    const float nx = fdiv(ddx, dist), ny = fdiv(ddy, dist), nz = fdiv(ddz, dist);
                    const float der = spiky_dw(dist, h);
                    const float sc = -0.5f * (Ci + cq.w) * der;
                    float pdx = sc * nx, pdy = sc * ny, pdz = sc * nz;
                    const bool same = UNIFORM || pc.fluid_mat == a.classes[a.cls[q]].fluid_mat;
                    if (same) {
                        float rx = 0.0f, ry = 0.0f, rz = 0.0f;
                        if (need_rel) {
                            const float4 dq = a.din[q];
                            rx = di.x - dq.x;
                            ry = di.y - dq.y;
                            rz = di.z - dq.z;
                        }
                        if (cohesion > kEps) {
                            const float ck = cohesion * cohesion_kernel(dist, h, m.cohesion1, m.cohesion2);
                            pdx = pdx - nx * ck;
                            pdy = pdy - ny * ck;
                            pdz = pdz - nz * ck;
                        }
                        if (visc_on) {
                            const float vs = 1.0f - fdiv(1.0f, 1.0f + viscosity * spiky_w(dist, h));
                            pdx = pdx - vs * rx;
                            pdy = pdy - vs * ry;
                            pdz = pdz - vs * rz;
                        }
                        if (cfl_on) {
                            const float proj = dot3(rx, ry, rz, nx, ny, nz);
                            if (proj < -cfl) {
                                const float pcf = proj + cfl;
                                pdx = pdx - nx * pcf * 0.5f;
                                pdy = pdy - ny * pcf * 0.5f;
                                pdz = pdz - nz * pcf * 0.5f;
                            }
                        }
                        if (tension) {
                            const float4 ncq = a.nc[q];
                            pdx = pdx - c.dt * (nci.x - ncq.x);
                            pdy = pdy - c.dt * (nci.y - ncq.y);
                            pdz = pdz - c.dt * (nci.z - ncq.z);
                        }
                    }
                    dpx = dpx + pdx;
                    dpy = dpy + pdy;
                    dpz = dpz + pdz;
                    wsum = wsum + 1.0f;
                }
                // ComputeFluidBoundaryDeltaFromCandidates, physics_fluid_boundary_common.hlsli:140-167
                float bx = 0.0f, by = 0.0f, bz = 0.0f;
                const float scn = fmaxd(Ci, 0.0f);
                const uint32_t bm = a.bnd[i];
                for (uint32_t k = 0; k < c.ncolliders; ++k) {
                    if (!((bm >> k) & 1u)) continue;
                    float gx, gy, gz;
                    if (!ghost_delta(a.col[k], sx, sy, sz, h, gx, gy, gz)) continue;
                    const float gd = len3(gx, gy, gz);
                    if (gd <= kEps || gd >= h) continue;
                    const float f = -scn * spiky_dw(gd, h);
                    bx = bx + f * fdiv(gx, gd);
                    by = by + f * fdiv(gy, gd);
                    bz = bz + f * fdiv(gz, gd);
                }
                const float solve = c.cfg[kCfgFluidSolve], bdel = c.cfg[kCfgFluidBndDelta];
                ox = dpx * solve + bx * bdel;
                oy = dpy * solve + by * bdel;
                oz = dpz * solve + bz * bdel;
                ow = wsum;
            }
            // apply
            const float wgt = fmaxd(ow * c.cfg[kCfgFluidRelax], 1.0f);
            const float ax = di.x + fdiv(ox, wgt), ay = di.y + fdiv(oy, wgt), az = di.z + fdiv(oz, wgt);
            // clamp
            const float r = fmaxd(pc.radius, 0.0f);
            float cx = p.x + ax, cy = p.y + ay, cz = p.z + az;
            const uint32_t bm = a.bnd[i];
            for (int sweep = 0; sweep < 2; ++sweep) {
                bool changed = false;
                for (uint32_t k = 0; k < c.ncolliders; ++k) {
                    if (!((bm >> k) & 1u)) continue;
                    float sd, nx, ny, nz;
                    if (!box_contact(a.col[k], cx, cy, cz, sd, nx, ny, nz)) continue;
                    if (sd >= r) continue;
                    const float push = r - sd;
                    cx = cx + nx * push;
                    cy = cy + ny * push;
                    cz = cz + nz * push;
                    changed = true;
                }
                if (!changed) break;
            }
            const float ex = cx - p.x, ey = cy - p.y, ez = cz - p.z;
            a.dout[i] = make_float4(ex, ey, ez, 0.0f);
            const float nx = p.x + ex, ny = p.y + ey, nz = p.z + ez;
            a.cur[i] = make_float4(nx, ny, nz, p.w);
            // displacement bound of the superset lists: |delta| plus a few ulps of the position
            const float slack = fmaxd(fmaxd(fabsf(nx), fabsf(ny)), fabsf(nz)) * 4.76837158e-7f + 1.0e-30f;
            const float dl = len3(ex, ey, ez) + slack;
            d2 = dl * dl;
        }
    }
    if (a.it < kIterStateWords) {
        const uint32_t wm = wave_max(__float_as_uint(d2));
        if ((threadIdx.x & 63u) == 0 && wm) {
            atomicMax(max_disp(a, a.it), wm);
            atomicMax(&a.meta[7], wm);
        }
    }
}

static FluidArgs fluid_args(const SceneBuffers &s, const SubstepConsts &c, uint32_t it, int din)
{
    FluidArgs a;
    const int cur = s.cur;
    a.pos = s.pos[cur];
    a.cls = s.cls[cur];
    a.cellc = s.cellc;
    a.bnd = s.bnd;
    a.din = s.dacc[din];
    a.dout = s.dacc[1 - din];
    a.cur = s.curpos;
    a.cs = s.cs;
    a.nc = s.nc;
    a.nbr = s.nbr;
    a.nbr_cnt = s.nbr_cnt;
    a.run_key = s.run_key;
    a.run_start = s.run_start;
    a.run_info = s.run_info;
    a.bucket = s.bucket;
    a.bucket_shift = c.bucket_shift;
    a.blk_first_run = s.blk_first_run;
    a.nblk = s.nblk;
    a.iter_state = s.iter_state;
    a.classes = s.classes;
    a.fm = s.fluid_mats;
    a.col = s.colliders;
    a.meta = s.meta;
    a.it = it;
    return a;
}

static uint32_t grid_of(uint32_t n) { return (n + kBlock - 1) / kBlock; }

void launch_fluid_density(const SceneBuffers &s, const SubstepConsts &c, uint32_t it, int din, hipStream_t st)
{
    launch_fluid_walk(s, c, it, din, st);
    launch_fluid_filter(s, c, it, din, st);
}

void launch_fluid_filter(const SceneBuffers &s, const SubstepConsts &c, uint32_t it, int din, hipStream_t st)
{
    const FluidArgs a = fluid_args(s, c, it, din);
    if (it > 0) {
        if (c.uniform_class)
            hipLaunchKernelGGL(k_fluid_filter<true>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c);
        else
            hipLaunchKernelGGL(k_fluid_filter<false>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c);
    }
}

void launch_fluid_walk(const SceneBuffers &s, const SubstepConsts &c, uint32_t it, int din, hipStream_t st)
{
    const FluidArgs a = fluid_args(s, c, it, din);
    const bool tiled = c.key_mode == CN_KEY_MORTON && it + 1 < kIterStateWords;
    if (tiled) {
        // persistent grid: enough workgroups to fill the chip, each pulling Morton blocks
        const uint32_t g = std::min<uint32_t>(grid_of(c.n), 256u * 6u);
        if (c.uniform_class)
            hipLaunchKernelGGL(k_fluid_walk_tiled<true>, dim3(g), dim3(kBlock), 0, st, a, c);
        else
            hipLaunchKernelGGL(k_fluid_walk_tiled<false>, dim3(g), dim3(kBlock), 0, st, a, c);
    } else {
        if (c.uniform_class)
            hipLaunchKernelGGL(k_fluid_walk_global<true>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c);
        else
            hipLaunchKernelGGL(k_fluid_walk_global<false>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c);
    }
}

void launch_fluid_delta(const SceneBuffers &s, const SubstepConsts &c, uint32_t it, int din, hipStream_t st)
{
    const FluidArgs a = fluid_args(s, c, it, din);
    if (c.uniform_class)
        hipLaunchKernelGGL(k_fluid_delta<true>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c);
    else
        hipLaunchKernelGGL(k_fluid_delta<false>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c);
}

}  // namespace cn
