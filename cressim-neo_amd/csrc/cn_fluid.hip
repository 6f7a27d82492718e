// cn_fluid.hip -- PBF fluid iterations on gfx950: the density-constraint pass and the
// delta + apply + clamp pass, both tiled over aligned 4x4x4-cell Morton blocks whose 6x6x6-cell
// halo is staged in LDS, and the exact neighbor-list pass of the per-pass entry.
//
// Reference passes (assets/shaders/physics/fluid/solver/): physics_fluid_build_neighbors.cs.hlsl
// :21-147, physics_fluid_compute_density_lambda.cs.hlsl:26-115,
// physics_fluid_compute_delta_positions.cs.hlsl:27-157, physics_fluid_apply_delta_positions.cs.hlsl
// :11-31, physics_fluid_clamp_boundary.cs.hlsl:23-80.  Every iteration the reference builds each
// fluid particle's neighbor list -- the fluid particles within h of its current position x* + delta,
// found by walking the 27 cells around its broadphase cell in dz, dy, dx order, each cell's
// particles in ascending slot, keeping the first max_fluid_neighbors -- and its density and delta
// passes iterate that list.
//
// Here a workgroup takes one aligned 4x4x4-cell block and stages its 6x6x6-cell halo in LDS, laid
// out cell by cell with x fastest, so the three dx cells of one (dz, dy) row are one contiguous LDS
// range and the 27-cell walk is nine range loops.  The density pass walks them with a loop body
// that only tests and compacts: the list it writes holds 16-bit halo offsets, not slots.  The sums
// then run over the compacted list, where every lane of a wave has work, and so does the delta
// pass, which stages the density records of the same halo and reads the list in the same offsets.
// The in-radius test sqrt(d2) < h is evaluated as d2 < T(h), T being the smallest float whose
// correctly rounded square root is >= h (the rounded square root is monotone), so the root is taken
// for listed neighbors only.  Blocks that cannot be staged (impure cells, a halo over the LDS
// capacity) and hash-key scenes walk the run table and sum in the walk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "cn_internal.h"
#include "cn_walk.h"

namespace cn {

// Record position of particles that no neighbor walk may list (non-fluid or zero inverse mass):
// every squared distance to it is +inf, which fails d2 < T.
constexpr float kFarAway = __builtin_huge_valf();

struct FluidArgs {
    const float4 *pos;        // x* (slot order)
    const uint32_t *cls;
    const int4 *cellc;
    const uint32_t *bnd;
    const uint16_t *bndl;
    const float4 *din;        // accumulated delta of the previous iteration
    float4 *dout;
    float4 *cur;              // (x* + delta, inverse mass) at the start of the iteration
    float4 *cs;               // density-pass records: (x* + delta, C)
    float4 *nc;               // surface normal * tension, C
    uint32_t *nbr;
    uint32_t *nbr_cnt;
    uint16_t *list;           // listed neighbors as halo offsets in slot groups (list_entry_base)
    uint32_t *lcnt;           // listed neighbor count
    uint16_t *slist;          // iteration-0 candidates as halo offsets in slot groups, LDS blocks
    uint16_t *scnt;           // candidate count, or kSuperNone
    float *disp;              // per slot: bound on the displacement since iteration 0 (null: not kept)
    const uint32_t *run_key;
    const uint32_t *run_start;
    const int4 *run_info;
    const uint32_t *bucket;
    uint32_t bucket_shift;
    const uint32_t *blk_first_run;
    const uint32_t *nblk;
    uint32_t *blk_halo;
    uint32_t *wide_blk;
    uint32_t *iter_state;
    const ParticleClass *classes;
    const cn_fluid_material *fm;
    const float *t2;          // per fluid material: d2 < t2 <=> sqrt(d2) < h
    const Collider *col;
    const SceneInfo *scenes;
    const uint32_t *active;   // per scene: stepped this cn_step
    uint32_t *meta;
    uint32_t it;
    uint32_t tile;            // halo particles this launch can stage in LDS
};

__device__ __forceinline__ RunTable run_table(const FluidArgs &a)
{
    return RunTable{a.run_key, a.run_start, a.run_info, a.bucket, a.bucket_shift, a.cellc};
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

// ---------------------------------------------------------------------------------------------
// candidate walks, in the reference's order: f(k) per candidate until f returns false
// ---------------------------------------------------------------------------------------------

// The same walk over a staged halo (hc = the particle's halo cell): nine LDS ranges.
template <class F>
__device__ __forceinline__ void walk_rows(const uint32_t *hoff, int hc, F &&f)
{
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
            const int h0 = hc - 1 + 6 * dy + 36 * dz;
            const uint32_t f1 = hoff[h0 + 3];
            for (uint32_t k = hoff[h0]; k < f1; ++k)
                if (!f(k)) return;
        }
}

// Halo cell holding staged particle f: the last cell whose offset is <= f.
__device__ __forceinline__ uint32_t halo_cell_of(const uint32_t *hoff, uint32_t f)
{
    uint32_t lo = 0, hi = kHalo;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (hoff[mid] <= f) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Block queue: the blocks are split into kXcds contiguous chunks, one per die, so neighboring
// blocks (which share halo cells) run under the same L2; a workgroup drains its die's chunk, then
// helps the others.  Called by one thread; returns kInvalid when every chunk is drained.
struct BlockQueue {
    uint32_t die, tries;
    __device__ explicit BlockQueue(uint32_t wg) : die(wg % kXcds), tries(0) {}
    __device__ uint32_t next(uint32_t *q, uint32_t nblk)
    {
        while (tries < kXcds) {
            const uint32_t b0 = (uint32_t)(((uint64_t)nblk * die) / kXcds);
            const uint32_t b1 = (uint32_t)(((uint64_t)nblk * (die + 1)) / kXcds);
            const uint32_t t = atomicAdd(q + die * kQueueStride, 1u);
            if (t < b1 - b0) return b0 + t;
            die = (die + 1) % kXcds;
            ++tries;
        }
        return kInvalid;
    }
};

struct MetaAcc {
    uint32_t req = 0, trunc = 0, listed = 0, lds = 0, runs = 0, reuse = 0;
};

__device__ __forceinline__ void flush_meta(const MetaAcc &ma, const FluidArgs &a)
{
    const uint32_t wmax = wave_max(ma.req);
    const uint32_t wtr = wave_sum(ma.trunc);
    const uint32_t wls = wave_sum(ma.listed);
    const uint32_t wru = wave_sum(ma.reuse);
    if ((threadIdx.x & 63u) == 0) {
        if (wmax) atomicMax(&a.meta[0], wmax);
        if (wtr) atomicAdd(&a.meta[1], wtr);
        if (wru) atomicAdd(&a.meta[9], wru);
        if (a.it == 0 && wls) atomicAdd((unsigned long long *)&a.meta[2], (unsigned long long)wls);
    }
    if (threadIdx.x == 0) {
        if (ma.lds) atomicAdd(&a.meta[5], ma.lds);
        if (ma.runs) atomicAdd(&a.meta[6], ma.runs);
    }
}

// The persistent tiled passes gather one block's counters at a time into workgroup totals in LDS
// (so no per-lane accumulator lives across blocks) and flush those once at the end.
constexpr uint32_t kWgMetaWords = 6;  // largest request, truncations, listed, reuses, LDS blocks, run blocks

__device__ __forceinline__ void meta_to_wg(const MetaAcc &ma, uint32_t *s_meta)
{
    const uint32_t wmax = wave_max(ma.req);
    const uint32_t wtr = wave_sum(ma.trunc);
    const uint32_t wls = wave_sum(ma.listed);
    const uint32_t wru = wave_sum(ma.reuse);
    if ((threadIdx.x & 63u) == 0) {
        if (wmax) atomicMax(&s_meta[0], wmax);
        if (wtr) atomicAdd(&s_meta[1], wtr);
        if (wls) atomicAdd(&s_meta[2], wls);
        if (wru) atomicAdd(&s_meta[3], wru);
    }
    if (threadIdx.x == 0) {
        s_meta[4] += ma.lds;
        s_meta[5] += ma.runs;
    }
}

__device__ __forceinline__ void flush_wg_meta(const uint32_t *s_meta, const FluidArgs &a)
{
    if (threadIdx.x != 0) return;
    if (s_meta[0]) atomicMax(&a.meta[0], s_meta[0]);
    if (s_meta[1]) atomicAdd(&a.meta[1], s_meta[1]);
    if (s_meta[3]) atomicAdd(&a.meta[9], s_meta[3]);
    if (a.it == 0 && s_meta[2]) atomicAdd((unsigned long long *)&a.meta[2], (unsigned long long)s_meta[2]);
    if (s_meta[4]) atomicAdd(&a.meta[5], s_meta[4]);
    if (s_meta[5]) atomicAdd(&a.meta[6], s_meta[5]);
}

// ---------------------------------------------------------------------------------------------
// density pass: physics_fluid_build_neighbors.cs.hlsl:88-131 (the list filter) with
// physics_fluid_compute_density_lambda.cs.hlsl:26-115 (density, surface normal, constraint)
// ---------------------------------------------------------------------------------------------
struct DensitySum {
    float density = 0.0f, snx = 0.0f, sny = 0.0f, snz = 0.0f;
    uint32_t inr = 0, req = 0;
};

// Quotients of one neighbor pair, held negated where that saves the final negation of
// fdiv_rcp: nn = -(dx, dy, dz) / dist, nq = -(dist / h), omq = 1 - dist / h (as 1 + nq) and the spiky
// derivative der = (-2 k (1 - q)) / h (FluidSpikyKernelDerivative, physics_fluid_common.hlsli:31-42;
// 0 unless eps < dist < h), computed as -((2 k omq) / h).  The consumers take -(a b) = (-a) b and
// x - (-y) = x + y, which are exact, so every result has the bits of the plain forms.  FAST divides
// through reciprocals (fdiv_rcp: 1 / dist once per pair, 1 / h once per particle); quot_fast_ok
// says when that is exact, and the other particles take the IEEE divisions.
struct PairQuot {
    float nnx, nny, nnz, nq, omq, der;  // nq = -(dist / h)
};

// fdiv_rcp is exact for every quotient of pair_quot when h <= 1024 and each coordinate of the
// particle has magnitude >= 2^-62: a nonzero coordinate difference is then >= 2^-86 (both values
// are multiples of 2^-86, or the difference exceeds half the larger one), the live distances lie
// in (eps, h), and with h in [1e-4, 1024] the spiky numerator 2 k (1 - q) lies in
// [2^-51, 2^43] (k = 15 / (pi h^3), 1 - q >= 2^-24 when nonzero), so every dividend is 0 or
// >= 2^-96 and every quotient lies in [2^-100, 2^100).
__device__ __forceinline__ bool quot_fast_ok(float h, float sx, float sy, float sz)
{
    return h <= 1024.0f && fminf(fminf(fabsf(sx), fabsf(sy)), fabsf(sz)) >= 0x1p-62f;
}

// LIVE: the caller adds the pair's terms only for eps < dist < h (a listed neighbor that is not
// coincident), so the zero the derivative takes outside that range is never read.
template <bool FAST, bool LIVE = false>
__device__ __forceinline__ void pair_quot(PairQuot &p, float dx, float dy, float dz, float dist, float h, float ih,
                                          float k)
{
    float &nq = p.nq;
    if (FAST) {
        const float y = frcp_exact(dist);
        p.nnx = fdiv_rcp_neg(dx, dist, y);
        p.nny = fdiv_rcp_neg(dy, dist, y);
        p.nnz = fdiv_rcp_neg(dz, dist, y);
        nq = fdiv_rcp_neg(dist, h, ih);
    } else {
        p.nnx = fdiv(-dx, dist);
        p.nny = fdiv(-dy, dist);
        p.nnz = fdiv(-dz, dist);
        nq = -fdiv(dist, h);
    }
    p.omq = __fadd_rn(1.0f, nq);
    const float num = __fmul_rn(__fmul_rn(2.0f, k), p.omq);
    const float der = FAST ? fdiv_rcp_neg(num, h, ih) : -fdiv(num, h);
    p.der = LIVE ? der : ((dist <= kEps || dist >= h) ? 0.0f : der);
}

struct DensitySelf {
    ParticleClass pc;
    float4 s;
    float h, t2, kw, ih;  // ih = 1 / h
    bool tension;
    bool fd;          // the neighbor quotients may use fdiv_rcp (pair_quot)
    bool sup;         // candidate lists apply (fluid-only scene, tiled pass)
    float ts2, allow; // candidate threshold; bound on D_i + D_j under which they stay complete
};

// Returns false (and writes the unreachable record) for particles that take no part in the solve.
// sup_on: candidate lists apply to fluid-only scenes (only the fluid passes move their particles
// within a substep, and they record every displacement).
__device__ __forceinline__ bool density_begin(DensitySelf &x, uint32_t i, const float4 s, uint32_t ic,
                                              const FluidArgs &a, bool sup_on = false)
{
    x.pc = a.classes[ic];
    x.s = s;
    x.sup = false;
    if (x.pc.kind != CN_PARTICLE_KIND_FLUID || s.w <= kEps) {
        a.cs[i] = make_float4(kFarAway, kFarAway, kFarAway, 0.0f);
        return false;
    }
    const cn_fluid_material &m = a.fm[x.pc.fluid_mat];
    x.h = fmaxd(m.smoothing_radius, 1.0e-4f);
    x.t2 = a.t2[x.pc.fluid_mat];
    x.kw = spiky_coef(x.h);
    x.ih = fdiv(1.0f, x.h);
    x.fd = quot_fast_ok(x.h, s.x, s.y, s.z);
    x.tension = fmaxd(m.surface_tension, 0.0f) > kEps;
    if (sup_on && a.scenes[x.pc.scene].fluid_only) {
        // a pair j within h of particle i at iteration k was within h + D_i + D_j at iteration 0
        // (D: displacement since iteration 0, x* to x* + delta), so the candidates within
        // h + skin hold every later list while D_i + D_j < skin; D_j is bounded by the largest D in
        // the block's halo, which holds all of i's 27 cells.  The margins cover the rounding of the
        // squared distances (a few ulps) and of the recorded displacements.
        const float skin = kSkinFrac * x.h;
        const float hs = x.h + skin;
        x.sup = true;
        x.ts2 = hs * hs * 1.00001f;
        x.allow = 0.98f * skin;
    }
    return true;
}

__device__ __forceinline__ bool density_walks(const DensitySelf &x, bool uniform, const SubstepConsts &c)
{
    return c.cap > 0 && (!uniform || neighbor_ok(x.pc, x.pc));
}

// The list filter: candidate pj (class qcls) is a neighbor.
template <bool UNIFORM>
__device__ __forceinline__ bool density_test(const DensitySelf &x, const float4 pj, uint32_t qcls, const FluidArgs &a)
{
    if (!UNIFORM && !neighbor_ok(x.pc, a.classes[qcls])) return false;
    if (pj.w <= kEps) return false;
    const float dx = x.s.x - pj.x, dy = x.s.y - pj.y, dz = x.s.z - pj.z;
    return dot3(dx, dy, dz, dx, dy, dz) < x.t2;
}

// Density and (negated) surface-normal terms of one listed neighbor.
struct DensityTerm {
    float w, ngx, ngy, ngz;
};

// TENSION = false: the particle's material has no surface tension (x.tension false), so the normal
// terms are left out at compile time.
template <bool TENSION = true>
__device__ __forceinline__ DensityTerm density_term(const DensitySelf &x, const float4 pj)
{
    DensityTerm t;
    const float dx = x.s.x - pj.x, dy = x.s.y - pj.y, dz = x.s.z - pj.z;
    const float dist = fsqrt_nearzero_free(dot3(dx, dy, dz, dx, dy, dz));
    float omq = __fadd_rn(1.0f, fdiv_rcp_neg(dist, x.h, x.ih));  // 1 - dist / h
    if (!x.fd) omq = __fsub_rn(1.0f, fdiv(dist, x.h));
    // every term the callers add is a neighbor's (dist < h): the kernel's range test is left out
    t.w = spiky_w_om_in(omq, x.kw);
    t.ngx = t.ngy = t.ngz = 0.0f;
    if (TENSION && x.tension && !(dist <= kEps || dist >= x.h)) {
        // FluidSpikyKernelGradient, physics_fluid_common.hlsli:44-53, negated: der * -(d / dist)
        PairQuot p;
        pair_quot<true>(p, dx, dy, dz, dist, x.h, x.ih, x.kw);
        if (!x.fd) pair_quot<false>(p, dx, dy, dz, dist, x.h, x.ih, x.kw);
        t.ngx = p.der * p.nnx;
        t.ngy = p.der * p.nny;
        t.ngz = p.der * p.nnz;
    }
    return t;
}

template <bool TENSION = true>
__device__ __forceinline__ void density_add(DensitySum &d, const DensitySelf &x, const DensityTerm &t)
{
    d.density = d.density + t.w;
    if (TENSION && x.tension) {
        d.snx = d.snx - t.ngx;
        d.sny = d.sny - t.ngy;
        d.snz = d.snz - t.ngz;
    }
}

// Boundary ghost density (physics_fluid_boundary_common.hlsli:112-138), the constraint and the record.
__device__ __forceinline__ void density_end(uint32_t i, uint32_t bw, const DensitySum &d, const DensitySelf &x,
                                            MetaAcc &ma, const FluidArgs &a, const SubstepConsts &c)
{
    float density = d.density;
    const float bscale = c.cfg[kCfgFluidBndDensity];
    const uint32_t col0 = a.scenes[x.pc.scene].col0;
    each_candidate(bw, a.bndl, i, [&](uint32_t k) {
        float gx, gy, gz;
        if (!ghost_delta(a.col[col0 + k], x.s.x, x.s.y, x.s.z, x.h, gx, gy, gz)) return;
        const float gd = len3(gx, gy, gz);
        if (gd >= x.h) return;
        density = density + spiky_w_k(gd, x.h, x.kw) * bscale;
    });
    // the material is read again here rather than held in registers across the walk
    const cn_fluid_material &m = a.fm[x.pc.fluid_mat];
    const float rho0 = fmaxd(m.rest_density, 1.0f);
    const float C =
        fmaxd(density - rho0, -rho0 * fmaxd(c.cfg[kCfgFluidMaxUnder], 0.0f)) * m.density_constraint_scale;
    a.cs[i] = make_float4(x.s.x, x.s.y, x.s.z, C);
    if (x.tension) {
        const float st = fmaxd(m.surface_tension, 0.0f);
        a.nc[i] = make_float4(d.snx * st, d.sny * st, d.snz * st, C);
    }
    ma.req = max(ma.req, d.req);
    ma.trunc += d.req > c.cap ? 1u : 0u;
    ma.listed += d.inr;
}

// Run-table walk with the sums taken in the walk.
// UNIFORM: every particle the walk meets has class ucls.
template <bool UNIFORM>
__device__ __forceinline__ void density_global(uint32_t i, MetaAcc &ma, const FluidArgs &a, const SubstepConsts &c,
                                               uint32_t ucls = 0u)
{
    if (!a.active[a.classes[UNIFORM ? ucls : a.cls[i]].scene]) return;
    DensitySelf x;
    if (!density_begin(x, i, a.cur[i], UNIFORM ? ucls : a.cls[i], a)) return;
    DensitySum d;
    if (density_walks(x, UNIFORM, c))
        walk_runs(a.cellc[i], run_table(a), c, [&](uint32_t q) {
            if (q == i) return true;
            const float4 pj = a.cur[q];
            if (!density_test<UNIFORM>(x, pj, UNIFORM ? 0u : a.cls[q], a)) return true;
            ++d.req;
            if (d.inr < c.cap) {
                ++d.inr;
                density_add(d, x, density_term(x, pj));
            }
            return true;
        });
    density_end(i, a.bnd[i], d, x, ma, a, c);
}

// ---------------------------------------------------------------------------------------------
// delta pass: physics_fluid_compute_delta_positions.cs.hlsl:27-157, then
// physics_fluid_apply_delta_positions.cs.hlsl:11-31 and physics_fluid_clamp_boundary.cs.hlsl:23-80
// on the same particle.
// ---------------------------------------------------------------------------------------------
// relDelta = delta_i - delta_j is zero on the first iteration, and the CFL term can only act when
// |relDelta| >= cfl, i.e. never while 2 max|delta| < cfl (max|delta| of the previous iteration).
__device__ __forceinline__ bool cfl_term_on(const cn_fluid_material &m, const FluidArgs &a, const SubstepConsts &c)
{
    const float cfl = fmaxd(m.cfl_radius, 0.0f);
    if (!(a.it > 0 && cfl > kEps)) return false;
    const float md2 = a.it - 1 < kIterStateWords ? __uint_as_float(a.iter_state[kStateMaxDisp + a.it - 1]) : 3.4e38f;
    return !(4.0f * md2 * c.cfl_guard < cfl * cfl);
}
__device__ __forceinline__ bool visc_term_on(const cn_fluid_material &m, const FluidArgs &a, const SubstepConsts &c)
{
    return a.it > 0 && c.dt * fmaxd(m.viscosity, 0.0f) > kEps;
}

struct DeltaSelf {
    ParticleClass pc;
    float4 di, nci;
    float sx, sy, sz, Ci, h, t2, kw, ih, cohesion, viscosity, cfl, dt;
    bool tension, visc_on, cfl_on, need_rel, fd;
};

struct DeltaTerm {
    float x, y, z;
    bool live;  // the neighbor is not coincident (dist > eps): it adds a term and a weight
};

// One listed neighbor's term (physics_fluid_compute_delta_positions.cs.hlsl:60-140).
template <bool UNIFORM>
__device__ __forceinline__ DeltaTerm delta_term(const DeltaSelf &x, const float4 cq, const float4 dq, const float4 ncq,
                                                uint32_t qcls, const FluidArgs &a)
{
    DeltaTerm t;
    const float ddx = x.sx - cq.x, ddy = x.sy - cq.y, ddz = x.sz - cq.z;
    const float dist = fsqrt_nearzero_free(dot3(ddx, ddy, ddz, ddx, ddy, ddz));
    t.live = !(dist <= kEps);
    PairQuot p;
    // a listed neighbor (dist < h) whose term counts only when live (dist > eps)
    pair_quot<true, true>(p, ddx, ddy, ddz, dist, x.h, x.ih, x.kw);
    if (!x.fd) pair_quot<false, true>(p, ddx, ddy, ddz, dist, x.h, x.ih, x.kw);
    // sc * n with sc = -0.5 (Ci + Cj) der, as (0.5 (Ci + Cj) der) * -n
    const float msc = 0.5f * (x.Ci + cq.w) * p.der;
    float pdx = msc * p.nnx, pdy = msc * p.nny, pdz = msc * p.nnz;
    const bool same = UNIFORM || x.pc.fluid_mat == a.classes[qcls].fluid_mat;
    if (same) {
        float rx = 0.0f, ry = 0.0f, rz = 0.0f;
        if (x.need_rel) {
            rx = x.di.x - dq.x;
            ry = x.di.y - dq.y;
            rz = x.di.z - dq.z;
        }
        if (x.cohesion > kEps) {
            // the cohesion coefficients are read here (a rare path) rather than held across the loop
            const cn_fluid_material &m = a.fm[x.pc.fluid_mat];
            const float ck = x.cohesion * cohesion_q(dist, x.h, -p.nq, m.cohesion1, m.cohesion2);
            pdx = pdx + p.nnx * ck;
            pdy = pdy + p.nny * ck;
            pdz = pdz + p.nnz * ck;
        }
        if (x.visc_on) {
            const float vs = 1.0f - frcp_exact_ge1(1.0f + x.viscosity * spiky_w_om_in(p.omq, x.kw));
            pdx = pdx - vs * rx;
            pdy = pdy - vs * ry;
            pdz = pdz - vs * rz;
        }
        if (x.cfl_on) {
            // proj = r . n = -(r . nn); proj < -cfl, pcf = proj + cfl = cfl - (r . nn)
            const float nproj = dot3(rx, ry, rz, p.nnx, p.nny, p.nnz);
            if (nproj > x.cfl) {
                const float pcf = x.cfl - nproj;
                pdx = pdx + p.nnx * pcf * 0.5f;
                pdy = pdy + p.nny * pcf * 0.5f;
                pdz = pdz + p.nnz * pcf * 0.5f;
            }
        }
        if (x.tension) {
            pdx = pdx - x.dt * (x.nci.x - ncq.x);
            pdy = pdy - x.dt * (x.nci.y - ncq.y);
            pdz = pdz - x.dt * (x.nci.z - ncq.z);
        }
    }
    t.x = pdx;
    t.y = pdy;
    t.z = pdz;
    return t;
}

struct DeltaSum {
    float x = 0.0f, y = 0.0f, z = 0.0f, w = 0.0f;
};

__device__ __forceinline__ void delta_add(DeltaSum &d, const DeltaTerm &t)
{
    if (!t.live) return;
    d.x = d.x + t.x;
    d.y = d.y + t.y;
    d.z = d.z + t.z;
    d.w = d.w + 1.0f;
}

// A particle's own inputs of the delta pass, loaded together at the top of its loop iteration so
// their round trips to memory overlap (rather than one after another as the code reaches them).
struct DeltaPre {
    float4 p, di;
    uint32_t cls, bnd;
};

template <bool UNIFORM>
__device__ __forceinline__ DeltaPre delta_pre(uint32_t i, const FluidArgs &a, uint32_t ucls = 0u)
{
    DeltaPre r;
    r.p = a.pos[i];
    r.di = a.it > 0 ? a.din[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    r.cls = UNIFORM ? ucls : a.cls[i];
    r.bnd = a.bnd[i];
    return r;
}

// sum(x, d) adds the neighbor terms in list order.  own addresses the particle's record in
// rcs / rnc (LDS offset or slot).  Returns (|delta| + slack)^2 of the particle.
template <bool UNIFORM, class Sum>
__device__ __forceinline__ float delta_particle(uint32_t i, uint32_t own, const DeltaPre &pre, Sum &&sum,
                                                const float4 *rcs, const float4 *rnc, const FluidArgs &a,
                                                const SubstepConsts &c)
{
    const ParticleClass pc = a.classes[pre.cls];
    if (pc.kind != CN_PARTICLE_KIND_FLUID) {
        a.dout[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (a.disp) a.disp[i] = 0.0f;
        return 0.0f;
    }
    const float4 p = pre.p;
    const float4 di = pre.di;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, ow = 0.0f;
    if (p.w > kEps) {
        DeltaSelf x;
        x.pc = pc;
        x.di = di;
        const float4 csi = rcs[own];
        x.sx = csi.x;
        x.sy = csi.y;
        x.sz = csi.z;
        x.Ci = csi.w;
        const cn_fluid_material &m = a.fm[pc.fluid_mat];
        x.h = fmaxd(m.smoothing_radius, 1.0e-4f);
        x.t2 = a.t2[pc.fluid_mat];
        x.kw = spiky_coef(x.h);
        x.ih = fdiv(1.0f, x.h);
        x.fd = quot_fast_ok(x.h, x.sx, x.sy, x.sz);
        x.dt = c.dt;
        x.cohesion = c.dt * fmaxd(m.cohesion, 0.0f);
        x.viscosity = c.dt * fmaxd(m.viscosity, 0.0f);
        x.cfl = fmaxd(m.cfl_radius, 0.0f);
        x.tension = fmaxd(m.surface_tension, 0.0f) > kEps;
        x.cfl_on = cfl_term_on(m, a, c);
        x.visc_on = visc_term_on(m, a, c);
        x.need_rel = x.cfl_on || x.visc_on;
        x.nci = x.tension ? rnc[own] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        DeltaSum d;
        if (c.cap > 0 && (!UNIFORM || neighbor_ok(pc, pc))) sum(x, d);
        // ComputeFluidBoundaryDeltaFromCandidates, physics_fluid_boundary_common.hlsli:140-167
        float bx = 0.0f, by = 0.0f, bz = 0.0f;
        const float scn = fmaxd(x.Ci, 0.0f);
        const uint32_t col0 = a.scenes[pc.scene].col0;
        each_candidate(pre.bnd, a.bndl, i, [&](uint32_t k) {
            float gx, gy, gz;
            if (!ghost_delta(a.col[col0 + k], x.sx, x.sy, x.sz, x.h, gx, gy, gz)) return;
            const float gd = len3(gx, gy, gz);
            if (gd <= kEps || gd >= x.h) return;
            const float f = -scn * spiky_dw_k(gd, x.h, x.kw);
            bx = bx + f * fdiv(gx, gd);
            by = by + f * fdiv(gy, gd);
            bz = bz + f * fdiv(gz, gd);
        });
        const float solve = c.cfg[kCfgFluidSolve], bdel = c.cfg[kCfgFluidBndDelta];
        ox = d.x * solve + bx * bdel;
        oy = d.y * solve + by * bdel;
        oz = d.z * solve + bz * bdel;
        ow = d.w;
    }
    // apply
    const float wgt = fmaxd(ow * c.cfg[kCfgFluidRelax], 1.0f);
    const float ax = di.x + fdiv(ox, wgt), ay = di.y + fdiv(oy, wgt), az = di.z + fdiv(oz, wgt);
    // clamp
    const float r = fmaxd(pc.radius, 0.0f);
    float cx = p.x + ax, cy = p.y + ay, cz = p.z + az;
    const uint32_t bw = pre.bnd;
    const uint32_t bcol0 = a.scenes[pc.scene].col0;
    for (int sweep = 0; sweep < 2; ++sweep) {
        bool changed = false;
        each_candidate(bw, a.bndl, i, [&](uint32_t k) {
            float sd, nx, ny, nz;
            if (!collider_contact(a.col[bcol0 + k], cx, cy, cz, sd, nx, ny, nz)) return;
            if (sd >= r) return;
            const float push = r - sd;
            cx = cx + nx * push;
            cy = cy + ny * push;
            cz = cz + nz * push;
            changed = true;
        });
        if (!changed) break;
    }
    const float ex = cx - p.x, ey = cy - p.y, ez = cz - p.z;
    a.dout[i] = make_float4(ex, ey, ez, 0.0f);
    const float nx = p.x + ex, ny = p.y + ey, nz = p.z + ez;
    a.cur[i] = make_float4(nx, ny, nz, p.w);
    // |delta| plus a few ulps of the position: bounds the next iteration's relDelta for the CFL skip
    const float slack = fmaxd(fmaxd(fabsf(nx), fabsf(ny)), fabsf(nz)) * 4.76837158e-7f + 1.0e-30f;
    const float dl = len3(ex, ey, ez) + slack;
    if (a.disp) a.disp[i] = dl;
    return dl * dl;
}

// Run-table walk: the list filter and the cap applied in the walk.
template <bool UNIFORM>
__device__ __forceinline__ float delta_global(uint32_t i, const FluidArgs &a, const SubstepConsts &c,
                                              uint32_t ucls = 0u)
{
    if (!a.active[a.classes[UNIFORM ? ucls : a.cls[i]].scene]) return 0.0f;
    const int4 ci = a.cellc[i];
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return delta_particle<UNIFORM>(
        i, i, delta_pre<UNIFORM>(i, a, ucls),
        [&](const DeltaSelf &x, DeltaSum &d) {
            uint32_t inr = 0;
            walk_runs(ci, run_table(a), c, [&](uint32_t q) {
                if (q == i) return true;
                const uint32_t qc = UNIFORM ? 0u : a.cls[q];
                if (!UNIFORM && !neighbor_ok(x.pc, a.classes[qc])) return true;
                const float4 cq = a.cs[q];
                const float ddx = x.sx - cq.x, ddy = x.sy - cq.y, ddz = x.sz - cq.z;
                if (!(dot3(ddx, ddy, ddz, ddx, ddy, ddz) < x.t2)) return true;
                if (inr >= c.cap) return false;
                ++inr;
                delta_add(d, delta_term<UNIFORM>(x, cq, x.need_rel ? a.din[q] : zero, x.tension ? a.nc[q] : zero, qc,
                                                 a));
                return true;
            });
        },
        a.cs, a.nc, a, c);
}

__device__ __forceinline__ void flush_disp(float d2, const FluidArgs &a)
{
    if (a.it >= kIterStateWords) return;
    const uint32_t wm = wave_max(__float_as_uint(d2));
    if ((threadIdx.x & 63u) == 0 && wm) {
        atomicMax(a.iter_state + kStateMaxDisp + a.it, wm);
        atomicMax(&a.meta[7], wm);
    }
}

// ---------------------------------------------------------------------------------------------
// per-block halo descriptors, once per substep
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_blk_halo(FluidArgs a, SubstepConsts c)
{
    __shared__ uint32_t scan[kBlock];
    const uint32_t nb = min(a.nblk[1], c.blk_cap);
    const uint32_t tid = threadIdx.x;
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const uint32_t p0 = a.run_start[a.blk_first_run[b]], p1 = a.run_start[a.blk_first_run[b + 1]];
        const uint32_t scene = a.classes[a.cls[p0]].scene;
        uint32_t *hd = a.blk_halo + (size_t)b * kHaloStride;
        if (!a.scenes[scene].has_fluid) {  // a scene without fluid: no fluid pass visits the block
            if (tid == 0) {
                hd[kHdFlags] = kHdNoFluid;
                hd[kHdP0] = p0;
                hd[kHdP1] = p1;
                hd[kHdScene] = scene;
            }
            continue;
        }
        const int4 c0 = a.cellc[p0];
        const int bx = c0.x & ~3, by = c0.y & ~3, bz = c0.z & ~3;
        // particles of a wrapped Morton block that lie outside the aligned block, particles whose own
        // cell hashes to the empty-table marker (their cell has no halo range), or halo cells shared
        // with other cells leave the block to the run-table walk
        bool bad = false;
        for (uint32_t i = p0 + tid; i < p1; i += kBlock) {
            const int4 ci = a.cellc[i];
            bad |= (uint32_t)(ci.x - bx) > 3u || (uint32_t)(ci.y - by) > 3u || (uint32_t)(ci.z - bz) > 3u;
            bad |= cell_hash(ci.x, ci.y, ci.z) == kInvalid;
        }
        uint32_t cnt = 0, st = 0;
        if (tid < kHalo) {
            const int tx = bx - 1 + (int)(tid % 6u), ty = by - 1 + (int)((tid / 6u) % 6u), tz = bz - 1 + (int)(tid / 36u);
            if (cell_hash(tx, ty, tz) != kInvalid) {
                const int r = find_run(cell_key(tx, ty, tz, (uint32_t)c0.w, c.key_mode, c.morton_bits), run_table(a));
                if (r >= 0) {
                    const int4 info = a.run_info[r];
                    if (info.w == 0) bad = true;
                    else if (info.x == tx && info.y == ty && info.z == tz) {
                        st = a.run_start[r];
                        cnt = a.run_start[r + 1] - st;
                    }
                }
            }
        }
        const int any_bad = __syncthreads_or(bad ? 1 : 0);
        uint32_t total;
        const uint32_t off = block_exclusive_scan(cnt, scan, total);
        if (tid < kHalo) {
            hd[kHdStart + tid] = st;
            hd[kHdOff + tid] = off;
        }
        if (tid == 0) {
            // blocks the delta pass stages only in the wide tile get their own launch
            if (!any_bad && total > c.tile_delta && total <= c.tile_wide)
                a.wide_blk[atomicAdd(a.iter_state + kStateWideCount, 1u)] = b;
            hd[kHdOff + kHalo] = total;
            hd[kHdFlags] = any_bad ? 0u : 1u;
            hd[kHdP0] = p0;
            hd[kHdP1] = p1;
            hd[kHdBase] = (uint32_t)bx;
            hd[kHdBase + 1] = (uint32_t)by;
            hd[kHdBase + 2] = (uint32_t)bz;
            hd[kHdScene] = scene;
        }
    }
}

constexpr int kTileBlock = 512;  // threads of a tiled fluid workgroup: one pass over a block's particles
// waves per SIMD the tiled passes are compiled for (at most 80 VGPRs): three 8-wave workgroups per
// CU, which their LDS budget (kDynLdsBytes) also allows
constexpr int kTileWaves = 6;
constexpr uint32_t kTermStep = 4;  // listed neighbors whose loads are issued together in the term loops
constexpr uint32_t kDeltaStep = 4;  // the same in the delta pass
constexpr uint32_t kTilePad = 4;   // spare entries after each staged density array (walk loads past a range)

// Common head of the tiled passes: claim a block (next() runs on thread 0), decide LDS or run
// table, load the halo table.  skip_to is a halo size above the LDS capacity up to which a block is
// left to the wide launch (0: none).
struct TileState {
    uint32_t blk;
    uint32_t p0, p1;
    bool lds, skip;
    int bx, by, bz;
};

template <class Next>
__device__ __forceinline__ bool tile_begin(Next &&next, uint32_t *s_blk, uint32_t *hoff, uint32_t *hstart,
                                           TileState &ts, uint32_t skip_to, const FluidArgs &a,
                                           const SubstepConsts &c)
{
    const uint32_t tid = threadIdx.x;
    if (tid == 0) *s_blk = next();
    __syncthreads();
    const uint32_t b = *s_blk;
    if (b == kInvalid) return false;
    ts.blk = b;
    ts.lds = false;
    ts.skip = false;
    if (b < c.blk_cap) {
        const uint32_t *hd = a.blk_halo + (size_t)b * kHaloStride;
        ts.p0 = hd[kHdP0];
        ts.p1 = hd[kHdP1];
        if (!a.active[hd[kHdScene]] || hd[kHdFlags] == kHdNoFluid) {  // untouched this step, or no fluid
            ts.skip = true;
            __syncthreads();
            return true;
        }
        const uint32_t total = hd[kHdOff + kHalo];
        const bool pure = hd[kHdFlags] != 0u;
        ts.lds = pure && total <= a.tile;
        ts.skip = pure && !ts.lds && total <= skip_to;
        if (ts.lds) {
            for (uint32_t t = tid; t < kHalo; t += blockDim.x) {
                hstart[t] = hd[kHdStart + t];
                hoff[t] = hd[kHdOff + t];
            }
            if (tid == 0) hoff[kHalo] = hd[kHdOff + kHalo];
            ts.bx = (int)hd[kHdBase];
            ts.by = (int)hd[kHdBase + 1];
            ts.bz = (int)hd[kHdBase + 2];
        }
    } else {
        ts.p0 = a.run_start[a.blk_first_run[b]];
        ts.p1 = a.run_start[a.blk_first_run[b + 1]];
        const uint32_t scene = a.classes[a.cls[ts.p0]].scene;
        ts.skip = !a.active[scene] || !a.scenes[scene].has_fluid;
    }
    __syncthreads();
    return true;
}

__device__ __forceinline__ int halo_cell(const TileState &ts, const int4 ci)
{
    return (ci.x - ts.bx + 1) + 6 * (ci.y - ts.by + 1) + 36 * (ci.z - ts.bz + 1);
}

// ---------------------------------------------------------------------------------------------
// LDS-path loops
// ---------------------------------------------------------------------------------------------

// The density walk over the staged halo: tests every candidate in walk order and compacts those
// within the radius (squared distance below thr2: the neighbor threshold t2, or ts2 for the
// candidate list of a substep) as halo offsets into a list, four candidates per step with their
// LDS loads issued first.  Returns the untruncated count.  Staged particles with zero inverse mass
// carry an unreachable position, so the distance test also rejects them.
//
// Cells that no candidate inside can reach are left out of the walk (the order of the others is
// unchanged): a candidate's position at the substep start lies in its cell and has moved by at most
// the largest displacement since, so when the particle's distance to the cell's box exceeds the
// search radius plus that displacement (in cell units, with slack for the rounding of the cell
// assignment), none of the cell's particles can be within the radius.
struct RowCull {
    float ux, uy, uz;  // the particle's current position in cell units
    int cx, cy, cz;    // its cell
    float lim2;        // squared limit in cell units; +inf walks every cell
};

__device__ __forceinline__ RowCull row_cull(const float4 s, const int4 ci, float inv, float radius, float dmax)
{
    RowCull rc;
    rc.ux = s.x * inv;
    rc.uy = s.y * inv;
    rc.uz = s.z * inv;
    rc.cx = ci.x;
    rc.cy = ci.y;
    rc.cz = ci.z;
    const float slack = 1.0e-4f + 1.0e-6f * fmaxd(fmaxd(fabsf(rc.ux), fabsf(rc.uy)), fabsf(rc.uz));
    const float lim = (radius + dmax) * inv * 1.00001f + slack;
    rc.lim2 = dmax < 3.0e38f && lim < 1.0e18f ? lim * lim : __builtin_huge_valf();
    return rc;
}

__device__ __forceinline__ float gap_below(float u, int c) { return fmaxd(u - (float)c, 0.0f); }
__device__ __forceinline__ float gap_above(float u, int c) { return fmaxd((float)(c + 1) - u, 0.0f); }

// A particle's list in a grouped list buffer (list_entry_base): the buffer (a kernel argument, so a
// scalar base) and the byte offset of entry 0; entry k sits 2 * kListGroup * k bytes further.  The
// offsets are 32-bit (list buffers stay below 4 GiB), so an address is one 32-bit add from the base.
constexpr uint32_t kEntryStride = 2u * kListGroup;
struct ListRef {
    char *base;
    uint32_t off;
    __device__ __forceinline__ uint16_t &entry(uint32_t k) const
    {
        return *reinterpret_cast<uint16_t *>(base + (off + k * kEntryStride));
    }
};

__device__ __forceinline__ ListRef list_ref(uint16_t *buf, uint32_t i, uint32_t rows)
{
    return ListRef{reinterpret_cast<char *>(buf), (uint32_t)list_entry_base(i, rows)};
}

// List entries are byte offsets into a staged array of 16-byte records (halo index x 16; the tile
// holds at most 4,096 halo particles, so they fit 16 bits), so the LDS address of a listed
// neighbor's record needs no scaling; staged arrays of 4-byte items take the entry / 4.
constexpr uint32_t kEntryScale = 16;
static_assert(4096u * kEntryScale <= 65536u, "a byte-offset entry of the largest tile fits 16 bits");
template <class T>
__device__ __forceinline__ const T &at16(const T *base, uint32_t e)
{
    static_assert(sizeof(T) == 16, "16-byte records");
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + e);
}
__device__ __forceinline__ uint32_t at4(const uint32_t *base, uint32_t e)
{
    return *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(base) + (e >> 2));
}

// Appends listed halo offsets to a particle's list, storing the first cap.  The write position is
// the byte offset of the next entry and advances one entry stride per push, so a push is a compare,
// a store and an add; n() is the untruncated count.  The position may run past entry cap by one
// stride per walked candidate (at most the tile, 2^12 entries: 512 KiB), which the list buffer's
// limit leaves room for below 4 GiB.
struct ListOut {
    char *base;
    uint32_t a0, a, lim;  // byte offsets of entry 0, the next entry and entry cap
    __device__ __forceinline__ ListOut(const ListRef &r, uint32_t cap)
        : base(r.base), a0(r.off), a(r.off), lim(r.off + cap * kEntryStride)
    {
    }
    __device__ __forceinline__ void push(uint32_t f)
    {
        if (a < lim) *reinterpret_cast<uint16_t *>(base + a) = (uint16_t)f;
        a += kEntryStride;
    }
    __device__ __forceinline__ uint32_t n() const { return (a - a0) / kEntryStride; }
};

// One contiguous range [k, f1) of the walk, four candidates per step: the loads read through a
// pointer with immediate offsets and may run up to three entries past f1 (the staged arrays carry
// kTilePad spare entries), which the u < f1 - k test then drops.
template <bool UNIFORM>
__device__ __forceinline__ void scan_range(uint32_t k, uint32_t f1, const float4 *lpos, const uint32_t *lcls,
                                           const DensitySelf &x, float thr2, const FluidArgs &a, ListOut &lo)
{
    const float4 *pp = lpos + k;
    const uint32_t *pc = lcls + k;
    uint32_t kb = k * kEntryScale;  // the entry of candidate k
    for (; k < f1; k += 4, kb += 4 * kEntryScale, pp += 4, pc += 4) {
        float4 pj[4];
        uint32_t qc[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            pj[u] = pp[u];
            qc[u] = UNIFORM ? 0u : pc[u];
        }
        const uint32_t rem = f1 - k;  // >= 1
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const float dx = x.s.x - pj[u].x, dy = x.s.y - pj[u].y, dz = x.s.z - pj[u].z;
            const float d2 = dot3(dx, dy, dz, dx, dy, dz);
            bool cand = (d2 < thr2) & (u == 0 || u < rem);
            if (!UNIFORM) cand = cand && neighbor_ok(x.pc, a.classes[qc[u]]);
            if (cand) lo.push(kb + u * kEntryScale);
        }
    }
}

template <bool UNIFORM>
__device__ __forceinline__ uint32_t scan_rows(const uint32_t *hoff, int hc, uint32_t fs, const float4 *lpos,
                                              const uint32_t *lcls, const DensitySelf &x, float thr2,
                                              const ListRef &li, uint32_t cap, const FluidArgs &a,
                                              const RowCull &rc)
{
    ListOut lo(li, cap);
    const float gxl = gap_below(rc.ux, rc.cx), gxr = gap_above(rc.ux, rc.cx);
    // rows of three cells along x, z outer, y inner; the centre row (r = 4) is walked as the
    // ranges before and after the particle itself
#pragma unroll 1
    for (int r = 0; r < 9; ++r) {
        const int ry = r % 3 - 1, rz = r / 3 - 1;
        const float gy = ry < 0 ? gap_below(rc.uy, rc.cy) : (ry > 0 ? gap_above(rc.uy, rc.cy) : 0.0f);
        const float gz = rz < 0 ? gap_below(rc.uz, rc.cz) : (rz > 0 ? gap_above(rc.uz, rc.cz) : 0.0f);
        const float perp = gy * gy + gz * gz;
        if (!(perp < rc.lim2)) continue;
        const int h0 = hc - 1 + 6 * ry + 36 * rz;
        const uint32_t f1 = hoff[h0 + (gxr * gxr + perp < rc.lim2 ? 3 : 2)];
        const uint32_t k0 = hoff[h0 + (gxl * gxl + perp < rc.lim2 ? 0 : 1)];
        if (r == 4) {
            scan_range<UNIFORM>(k0, fs, lpos, lcls, x, thr2, a, lo);
            scan_range<UNIFORM>(fs + 1, f1, lpos, lcls, x, thr2, a, lo);
        } else {
            scan_range<UNIFORM>(k0, f1, lpos, lcls, x, thr2, a, lo);
        }
    }
    return lo.n();
}

// List entries k .. k + kTermStep - 1 of a list of cnt >= 1 entries (the last one repeated past
// the end, so the LDS reads they drive stay in range).
template <uint32_t N>
__device__ __forceinline__ void list_group(const ListRef &li, uint32_t k, uint32_t cnt, uint32_t (&e)[N])
{
#pragma unroll
    for (uint32_t u = 0; u < N; ++u) e[u] = li.entry(min(k + u, cnt - 1));
}

// The neighbors from the substep's candidate list (complete at iteration 0, and later while the
// displacements stay under the skin, checked by the caller): the in-radius test at the current
// positions, the listed neighbors written to the list in candidate order (= walk order), and the
// density terms of the first cap summed in the same pass.  A step's four candidates are loaded first; each term is then
// evaluated before the test decides whether it is added, so every lane runs the same instructions.
// (Filtering first and summing over the written list afterwards measured slower.)  Returns the
// untruncated neighbor count.
template <bool TENSION>
__device__ __forceinline__ uint32_t density_super(const ListRef &ls, uint32_t scount, const float4 *lpos,
                                                  const DensitySelf &x, const ListRef &li, uint32_t cap,
                                                  DensitySum &d)
{
    ListOut lo(li, cap);
    uint32_t en[kTermStep];
    list_group(ls, 0, scount, en);
    for (uint32_t k = 0; k < scount; k += kTermStep) {
        float4 pj[kTermStep];
        uint32_t e[kTermStep];
#pragma unroll
        for (uint32_t u = 0; u < kTermStep; ++u) {
            e[u] = en[u];
            pj[u] = at16(lpos, e[u]);
        }
        // the next step's entries while this step's terms are evaluated
        list_group(ls, k + kTermStep, scount, en);
#pragma unroll
        for (uint32_t u = 0; u < kTermStep; ++u) {
            const DensityTerm t = density_term<TENSION>(x, pj[u]);
            const float dx = x.s.x - pj[u].x, dy = x.s.y - pj[u].y, dz = x.s.z - pj[u].z;
            if ((k + u < scount) & (dot3(dx, dy, dz, dx, dy, dz) < x.t2)) {
                if (lo.a < lo.lim) density_add<TENSION>(d, x, t);
                lo.push(e[u]);
            }
        }
    }
    return lo.n();
}

// Density sums over a compacted list, kTermStep entries per step (loads first, sums in list order).
template <bool TENSION>
__device__ __forceinline__ void density_list(const ListRef &li, uint32_t cnt, const float4 *lpos,
                                             const DensitySelf &x, DensitySum &d)
{
    // the next group's entries are loaded while the current group's terms are evaluated
    uint32_t en[kTermStep];
    if (cnt > 0) list_group(li, 0, cnt, en);
    for (uint32_t k = 0; k < cnt; k += kTermStep) {
        uint32_t e[kTermStep];
#pragma unroll
        for (uint32_t u = 0; u < kTermStep; ++u) e[u] = en[u];
        list_group(li, k + kTermStep, cnt, en);
        DensityTerm t[kTermStep];
#pragma unroll
        for (uint32_t u = 0; u < kTermStep; ++u) t[u] = density_term<TENSION>(x, at16(lpos, e[u]));
#pragma unroll
        for (uint32_t u = 0; u < kTermStep; ++u)
            if (k + u < cnt) density_add<TENSION>(d, x, t[u]);
    }
}

// ---------------------------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------------------------
// MODE: kDensPlain walks every iteration; with candidate lists, kDensBuild (a substep's first
// iteration) walks the wider radius into the candidate lists and takes the neighbors from them,
// and kDensReuse (a later iteration) takes them from the candidate lists in every block whose
// particles' displacements allow it and walks the other blocks.
constexpr int kDensPlain = 0, kDensBuild = 1, kDensReuse = 2;
// UCLS: the particle classes of a tiled fluid launch -- kClsMixed per particle, kClsOne every
// particle class 0, kClsBlock one class per scene with fluid, read from the block's first particle.
constexpr int kClsMixed = 0, kClsOne = 1, kClsBlock = 2;
template <int UCLS, int MODE>
__global__ __launch_bounds__(kTileBlock) __attribute__((amdgpu_waves_per_eu(kTileWaves))) void k_fluid_density_tiled(
    FluidArgs a, SubstepConsts c)
{
    constexpr bool UNIFORM = UCLS != kClsMixed;
    extern __shared__ float4 dyn[];
    __shared__ uint32_t hoff[kHalo + 1];
    __shared__ uint32_t hstart[kHalo];
    __shared__ uint32_t s_blk;
    __shared__ uint32_t s_meta[kWgMetaWords];
    __shared__ uint32_t s_hmax;  // largest displacement bound in the staged halo (float bits)
    __shared__ uint32_t s_omax;  // the same over the block's own particles with mass (one class)
    __shared__ uint32_t s_obad;  // one of those has no complete candidate list (one class)
    constexpr bool SUP = MODE == kDensBuild || MODE == kDensReuse;  // candidate lists apply
    constexpr bool REUSE = MODE == kDensReuse;
    float4 *lpos = dyn;
    uint32_t *lcls = reinterpret_cast<uint32_t *>(dyn + a.tile + kTilePad);
    const uint32_t tid = threadIdx.x;
    // largest |delta|^2 of the previous iteration: bounds how far any fluid particle has moved
    // from its iteration-0 position
    const float md2 = a.it == 0 ? 0.0f : __uint_as_float(a.iter_state[kStateMaxDisp + a.it - 1]);
    // the density launch resets the displacement word its delta pass raises
    if (blockIdx.x == 0 && tid == 0) a.iter_state[kStateMaxDisp + a.it] = 0u;
    const uint32_t nblk = a.nblk[1];  // the blocks of scenes with fluid
    uint32_t *queue = a.iter_state + kStateDensity + a.it * kXcds * kQueueStride;
    if (tid < kWgMetaWords) s_meta[tid] = 0u;  // ordered before any use by tile_begin's barrier
    if (tid == 0) s_hmax = s_omax = s_obad = 0u;
    BlockQueue bq(blockIdx.x);
    TileState ts;
    auto next = [&]() -> uint32_t { return bq.next(queue, nblk); };
    while (tile_begin(next, &s_blk, hoff, hstart, ts, 0u, a, c)) {
        if (ts.skip) {
            __syncthreads();
            continue;
        }
        MetaAcc ma;
        // the block's class (wave-uniform: the class and what derives from it sit in scalar
        // registers)
        const uint32_t bcls = UCLS == kClsBlock ? __builtin_amdgcn_readfirstlane(a.cls[ts.p0]) : 0u;
        if (ts.lds) {
            const uint32_t total = hoff[kHalo];
            // a halo of one particle class (the common case in a pool of single-material scenes)
            // takes the class filter out of the walk, as a pool of one class does
            const uint32_t ref = UNIFORM ? bcls : a.cls[ts.p0];
            bool mixed = false, obad = false;
            uint32_t dmb = 0, dmo = 0;  // displacement bounds are non-negative floats: their bits order them
            for (uint32_t f = tid; f < total; f += kTileBlock) {
                const uint32_t hc = halo_cell_of(hoff, f);
                const uint32_t q = hstart[hc] + (f - hoff[hc]);
                float4 v = a.cur[q];
                const bool mass = v.w > kEps;
                if (!mass) v.x = v.y = v.z = kFarAway;
                lpos[f] = v;
                if (REUSE) {
                    const uint32_t dq = __float_as_uint(a.disp[q]);
                    dmb = max(dmb, dq);
                    // one class: the block's own particles' part of the reuse check rides along
                    if (UNIFORM && mass && q >= ts.p0 && q < ts.p1) {
                        dmo = max(dmo, dq);
                        obad |= a.scnt[q] == kSuperNone;
                    }
                }
                if (!UNIFORM) {
                    const uint32_t qc = a.cls[q];
                    lcls[f] = qc;
                    mixed |= qc != ref;
                }
            }
            if (REUSE) {
                dmb = wave_max(dmb);
                if ((tid & 63u) == 0u) atomicMax(&s_hmax, dmb);
                if (UNIFORM) {
                    dmo = wave_max(dmo);
                    if ((tid & 63u) == 0u) atomicMax(&s_omax, dmo);
                    if (obad) s_obad = 1u;
                }
            }
            const int any_mixed = __syncthreads_or(mixed ? 1 : 0);  // also the barrier after staging
            const bool bu = UNIFORM || !any_mixed;
            bool reuse = false;
            if (REUSE && UNIFORM) {
                // One class: the check below for every particle at once, from the staging loop.
                // D_i + hmax < allow holds for every particle with mass exactly when it holds for
                // their largest D_i (rounded addition is monotone).
                const ParticleClass pc = a.classes[bcls];
                const float h = fmaxd(a.fm[pc.fluid_mat].smoothing_radius, 1.0e-4f);
                reuse = pc.kind != CN_PARTICLE_KIND_FLUID || c.cap == 0u ||
                        (a.scenes[pc.scene].fluid_only && s_obad == 0u &&
                         __uint_as_float(s_omax) + __uint_as_float(s_hmax) < 0.98f * (kSkinFrac * h));
            } else if (REUSE) {
                // The candidates of iteration 0 (within h + skin of x*) hold particle i's list while
                // D_i + D_j < skin for every j in its 27 cells (D: displacement since iteration 0);
                // D_j is bounded by the halo's largest.  A block where some particle fails, or has
                // no complete candidate list, is walked whole (one decision per block keeps the
                // walk out of the waves of the others).
                const float hmax = __uint_as_float(s_hmax);
                bool bad = false;
                for (uint32_t i = ts.p0 + tid; i < ts.p1; i += kTileBlock) {
                    const int hc = halo_cell(ts, a.cellc[i]);
                    const uint32_t fs = hoff[hc] + (i - hstart[hc]);
                    const ParticleClass pc = a.classes[UNIFORM ? bcls : lcls[fs]];
                    if (pc.kind != CN_PARTICLE_KIND_FLUID || !(lpos[fs].w > kEps) || c.cap == 0u) continue;
                    const float h = fmaxd(a.fm[pc.fluid_mat].smoothing_radius, 1.0e-4f);
                    bad |= !a.scenes[pc.scene].fluid_only || a.scnt[i] == kSuperNone ||
                           !(a.disp[i] + hmax < 0.98f * (kSkinFrac * h));
                }
                reuse = !__syncthreads_or(bad ? 1 : 0);
            }
            // the walk at the neighbor radius (thr2 = t2, radius h) or the candidate radius (ts2) into l
            auto scan = [&](auto bu_tag, const DensitySelf &x, int4 ci, int hc, uint32_t fs, float thr2,
                            float radius, const ListRef &l, uint32_t cap) {
                constexpr bool BU = decltype(bu_tag)::value;
                const float inv = c.inv_cell > 0.0f ? c.inv_cell : a.scenes[x.pc.scene].inv_cell;
                // displacement since the cells were assigned: none at iteration 0; later the largest
                // |delta| of the previous iteration, where only the fluid passes move fluid particles
                const float dmax = a.it == 0 ? 0.0f : (a.scenes[x.pc.scene].fluid_only ? fsqrt(md2) : 3.4e38f);
                const RowCull rc = row_cull(x.s, ci, inv, radius, dmax);
                return scan_rows<BU>(hoff, hc, fs, lpos, lcls, x, thr2, l, cap, a, rc);
            };
            auto super = [&](const DensitySelf &x, const ListRef &ls, uint32_t scount, const ListRef &li,
                             DensitySum &d) {
                return x.tension ? density_super<true>(ls, scount, lpos, x, li, c.cap, d)
                                 : density_super<false>(ls, scount, lpos, x, li, c.cap, d);
            };
            // BLK: the block's halo holds one particle class (bref, wave-uniform), so the class and
            // everything the particle derives from it (material, radius, thresholds) load into
            // scalar registers; a pool of single-material scenes takes this path in its general
            // instantiation
            const uint32_t bref = __builtin_amdgcn_readfirstlane(ref);
            auto particles = [&](auto blk_tag) {
                constexpr bool BLK = decltype(blk_tag)::value;
                for (uint32_t i = ts.p0 + tid; i < ts.p1; i += kTileBlock) {
                    const int4 ci = a.cellc[i];
                    const uint32_t bw = a.bnd[i];  // loaded beside the cell, long before its use
                    const int hc = halo_cell(ts, ci);
                    const uint32_t fs = hoff[hc] + (i - hstart[hc]);
                    DensitySelf x;
                    uint32_t cnt = 0;
                    if (density_begin(x, i, lpos[fs], UNIFORM ? bcls : (BLK ? bref : lcls[fs]), a, SUP)) {
                        DensitySum d;
                        const ListRef li = list_ref(a.list, i, c.cap);
                        if (density_walks(x, BLK || bu, c)) {
                            bool done = false;
                            if (REUSE && reuse) {
                                ++ma.reuse;
                                d.req = super(x, list_ref(a.slist, i, kSuperSlots), a.scnt[i], li, d);
                                done = true;
                            } else if (MODE == kDensBuild && x.sup) {
                                // the candidates within h + skin, then the neighbors among them; a
                                // particle with more candidates than the list holds walks for its
                                // neighbors (and never reuses: its count is marked)
                                const ListRef ls = list_ref(a.slist, i, kSuperSlots);
                                const float rs = fsqrt(x.ts2);
                                const uint32_t sreq = BLK || bu ? scan(std::true_type{}, x, ci, hc, fs, x.ts2, rs, ls, kSuperSlots)
                                                                : scan(std::false_type{}, x, ci, hc, fs, x.ts2, rs, ls, kSuperSlots);
                                a.scnt[i] = sreq <= kSuperSlots ? (uint16_t)sreq : kSuperNone;
                                if (sreq <= kSuperSlots) {
                                    d.req = super(x, ls, sreq, li, d);
                                    done = true;
                                }
                            }
                            if (!done) {
                                d.req = BLK || bu ? scan(std::true_type{}, x, ci, hc, fs, x.t2, x.h, li, c.cap)
                                                  : scan(std::false_type{}, x, ci, hc, fs, x.t2, x.h, li, c.cap);
                                d.inr = min(d.req, c.cap);
                                if (x.tension)
                                    density_list<true>(li, d.inr, lpos, x, d);
                                else
                                    density_list<false>(li, d.inr, lpos, x, d);
                            }
                            d.inr = min(d.req, c.cap);
                        }
                        cnt = d.inr;
                        density_end(i, bw, d, x, ma, a, c);
                    }
                    a.lcnt[i] = cnt;
                }
            };
            if (!UNIFORM && bu)
                particles(std::true_type{});
            else
                particles(std::false_type{});
            if (tid == 0) ++ma.lds;
        } else {
            for (uint32_t i = ts.p0 + tid; i < ts.p1; i += kTileBlock) density_global<UNIFORM>(i, ma, a, c, bcls);
            if (tid == 0) ++ma.runs;
        }
        meta_to_wg(ma, s_meta);
        __syncthreads();
        if (tid == 0) s_hmax = s_omax = s_obad = 0u;  // read again only after the next block's staging barrier
    }
    flush_wg_meta(s_meta, a);
}

// WIDE = false: every block, the previous delta staged in LDS beside the records; blocks whose halo
// only fits without it are skipped and left to the WIDE = true launch, which takes the listed wide
// blocks and reads the previous delta from HBM through a staged slot map.
template <int UCLS, bool WIDE>
__global__ __launch_bounds__(kTileBlock) __attribute__((amdgpu_waves_per_eu(kTileWaves))) void k_fluid_delta_tiled(
    FluidArgs a, SubstepConsts c)
{
    constexpr bool UNIFORM = UCLS != kClsMixed;
    extern __shared__ float4 dyn[];
    __shared__ uint32_t hoff[kHalo + 1];
    __shared__ uint32_t hstart[kHalo];
    __shared__ uint32_t s_blk;
    const bool use_d = a.it > 0;
    const bool tension = c.tension != 0u;
    float4 *lcs = dyn;
    float4 *ldin = lcs + a.tile;
    float4 *lnc = ldin + (use_d && !WIDE ? a.tile : 0u);
    uint32_t *lcls = reinterpret_cast<uint32_t *>(lnc + (tension ? a.tile : 0u));
    uint32_t *lslot = lcls + (UNIFORM ? 0u : a.tile);
    const uint32_t tid = threadIdx.x;
    const uint32_t nblk = a.nblk[1];
    uint32_t *queue = a.iter_state + kStateDelta + a.it * kXcds * kQueueStride;
    uint32_t *wqueue = a.iter_state + kStateWideQueue + a.it;
    const uint32_t nwide = WIDE ? a.iter_state[kStateWideCount] : 0u;
    BlockQueue bq(blockIdx.x);
    TileState ts;
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float d2max = 0.0f;
    auto next = [&]() -> uint32_t {
        if (!WIDE) return bq.next(queue, nblk);
        const uint32_t t = atomicAdd(wqueue, 1u);
        return t < nwide ? a.wide_blk[t] : kInvalid;
    };
    while (tile_begin(next, &s_blk, hoff, hstart, ts, WIDE || !use_d ? 0u : c.tile_wide, a, c)) {
        if (ts.skip) {
            __syncthreads();
            continue;
        }
        // the block's class
        const uint32_t bcls = UCLS == kClsBlock ? __builtin_amdgcn_readfirstlane(a.cls[ts.p0]) : 0u;
        if (ts.lds) {
            const uint32_t total = hoff[kHalo];
            const uint32_t ref = UNIFORM ? bcls : a.cls[ts.p0];
            // the neighbors' previous deltas are read only by the viscosity and CFL terms: staged
            // only when one of the block's fluid particles has one of them on
            bool rel = false;
            if (!WIDE && use_d) {
                if (UNIFORM) {
                    const ParticleClass pc0 = a.classes[bcls];
                    rel = pc0.kind == CN_PARTICLE_KIND_FLUID &&
                          (cfl_term_on(a.fm[pc0.fluid_mat], a, c) || visc_term_on(a.fm[pc0.fluid_mat], a, c));
                } else {
                    bool mine = false;
                    for (uint32_t i = ts.p0 + tid; i < ts.p1; i += kTileBlock) {
                        const ParticleClass pc = a.classes[a.cls[i]];
                        mine |= pc.kind == CN_PARTICLE_KIND_FLUID &&
                                (cfl_term_on(a.fm[pc.fluid_mat], a, c) || visc_term_on(a.fm[pc.fluid_mat], a, c));
                    }
                    rel = __syncthreads_or(mine ? 1 : 0) != 0;
                }
            }
            bool mixed = false;
            for (uint32_t f = tid; f < total; f += kTileBlock) {
                const uint32_t hc = halo_cell_of(hoff, f);
                const uint32_t q = hstart[hc] + (f - hoff[hc]);
                lcs[f] = a.cs[q];
                if (rel) ldin[f] = a.din[q];
                if (WIDE) lslot[f] = q;
                if (tension) lnc[f] = a.nc[q];
                if (!UNIFORM) {
                    const uint32_t qc = a.cls[q];
                    lcls[f] = qc;
                    mixed |= qc != ref;
                }
            }
            // one particle class in the halo: every listed neighbor has the particle's material
            const int any_mixed = __syncthreads_or(mixed ? 1 : 0);  // also the barrier after staging
            const bool bu = UNIFORM || !any_mixed;
            // BLK: one particle class in the halo (bref, wave-uniform): the class-derived values of
            // each particle load into scalar registers (the general instantiation in a pool of
            // single-material scenes)
            const uint32_t bref = __builtin_amdgcn_readfirstlane(ref);
            auto particles = [&](auto blk_tag) {
                constexpr bool BLK = decltype(blk_tag)::value;
                for (uint32_t i = ts.p0 + tid; i < ts.p1; i += kTileBlock) {
                    const int4 ci = a.cellc[i];
                    const DeltaPre pre = delta_pre<UNIFORM || BLK>(i, a, UNIFORM ? bcls : bref);
                    const uint32_t cnt = a.lcnt[i];
                    const int hc = halo_cell(ts, ci);
                    const uint32_t fs = hoff[hc] + (i - hstart[hc]);
                    const ListRef li = list_ref(a.list, i, c.cap);
                    const float d2 = delta_particle<UNIFORM>(
                        i, fs, pre,
                        [&](const DeltaSelf &x, DeltaSum &d) {
                            // the listed neighbors, kDeltaStep per step: loads first (the next group's
                            // entries while this group's terms are evaluated), terms independent, sums
                            // in list order
                            uint32_t en[kDeltaStep];
                            if (cnt > 0) list_group(li, 0, cnt, en);
                            for (uint32_t k = 0; k < cnt; k += kDeltaStep) {
                                uint32_t e[kDeltaStep];
    #pragma unroll
                                for (uint32_t u = 0; u < kDeltaStep; ++u) e[u] = en[u];
                                list_group(li, k + kDeltaStep, cnt, en);
                                float4 dq[kDeltaStep];
    #pragma unroll
                                for (uint32_t u = 0; u < kDeltaStep; ++u)
                                    dq[u] = !x.need_rel ? zero : (WIDE ? a.din[at4(lslot, e[u])] : at16(ldin, e[u]));
                                DeltaTerm t[kDeltaStep];
    #pragma unroll
                                for (uint32_t u = 0; u < kDeltaStep; ++u)
                                    t[u] = BLK || bu
                                               ? delta_term<true>(x, at16(lcs, e[u]), dq[u],
                                                                  x.tension ? at16(lnc, e[u]) : zero, 0u, a)
                                               : delta_term<false>(x, at16(lcs, e[u]), dq[u],
                                                                   x.tension ? at16(lnc, e[u]) : zero, at4(lcls, e[u]), a);
    #pragma unroll
                                for (uint32_t u = 0; u < kDeltaStep; ++u)
                                    if (k + u < cnt) delta_add(d, t[u]);
                            }
                        },
                        lcs, lnc, a, c);
                    d2max = fmaxd(d2max, d2);
                }
            };
            if (!UNIFORM && bu)
                particles(std::true_type{});
            else
                particles(std::false_type{});
        } else {
            for (uint32_t i = ts.p0 + tid; i < ts.p1; i += kTileBlock)
                d2max = fmaxd(d2max, delta_global<UNIFORM>(i, a, c, bcls));
        }
        __syncthreads();
    }
    flush_disp(d2max, a);
}

// Thread per particle through the run table: hash keys, and iterations past the tracked ones.
template <bool UNIFORM>
__global__ __launch_bounds__(kBlock) void k_fluid_density_global(FluidArgs a, SubstepConsts c)
{
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.it < kIterStateWords) a.iter_state[kStateMaxDisp + a.it] = 0u;
    MetaAcc ma;
    const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i < c.n) density_global<UNIFORM>(i, ma, a, c);
    flush_meta(ma, a);
}

template <bool UNIFORM>
__global__ __launch_bounds__(kBlock) void k_fluid_delta_global(FluidArgs a, SubstepConsts c)
{
    const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    float d2 = 0.0f;
    if (i < c.n) d2 = delta_global<UNIFORM>(i, a, c);
    flush_disp(d2, a);
}

// Exact neighbor lists of iteration 0 (the per-pass entry): the first cap in-radius neighbors in
// walk order, k-major, and their count.
template <bool UNIFORM>
__global__ __launch_bounds__(kBlock) void k_fluid_lists(FluidArgs a, SubstepConsts c)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= c.n) return;
    const ParticleClass pc = a.classes[UNIFORM ? 0u : a.cls[i]];
    const float4 s = a.cur[i];
    uint32_t cnt = 0;
    if (pc.kind == CN_PARTICLE_KIND_FLUID && s.w > kEps && c.cap > 0 && (!UNIFORM || neighbor_ok(pc, pc))) {
        const float t2 = a.t2[pc.fluid_mat];
        walk_runs(a.cellc[i], run_table(a), c, [&](uint32_t q) {
            if (q == i) return true;
            if (!UNIFORM && !neighbor_ok(pc, a.classes[a.cls[q]])) return true;
            const float4 pj = a.cur[q];
            if (pj.w <= kEps) return true;
            const float dx = s.x - pj.x, dy = s.y - pj.y, dz = s.z - pj.z;
            if (!(dot3(dx, dy, dz, dx, dy, dz) < t2)) return true;
            a.nbr[(size_t)cnt * c.n + i] = q;
            return ++cnt < c.cap;
        });
    }
    a.nbr_cnt[i] = cnt;
}

// Vorticity confinement after the velocity update: physics_fluid_compute_vorticity.cs.hlsl:16-81
// (APPLY = false: vort[i] = (curl, |curl|), zero for particles that do not confine) and
// physics_fluid_apply_vorticity_confinement.cs.hlsl:16-94 (APPLY = true).  Both walk the
// post-update neighbor list (physics_solver.cpp:1262): the first cap in-radius neighbors in walk
// order on the final positions, of which the same-material ones contribute.  The curl pass walks
// the run table and, where a list buffer is given (nbr, neighbor-major), stores the list; the
// confinement pass then reads it instead of walking again.  Inactive scenes are skipped whole
// (their neighbors are their own).
template <bool UNIFORM, bool APPLY>
__global__ __launch_bounds__(kBlock) void k_fluid_vorticity(FluidArgs a, SubstepConsts c, float4 *__restrict__ vel,
                                                            float4 *__restrict__ vort)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= c.n) return;
    const ParticleClass pc = a.classes[UNIFORM ? 0u : a.cls[i]];
    if (!a.active[pc.scene]) return;
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (pc.kind != CN_PARTICLE_KIND_FLUID) {
        if (!APPLY) vort[i] = zero;
        return;
    }
    const cn_fluid_material &m = a.fm[pc.fluid_mat];
    const float conf = fmaxd(m.vorticity_confinement, 0.0f);
    const float4 s = a.pos[i];
    if (conf <= kEps || s.w <= kEps) {
        if (!APPLY) vort[i] = zero;
        return;
    }
    float4 self = APPLY ? vort[i] : vel[i];
    if (APPLY && self.w <= kEps) return;
    const float h = fmaxd(m.smoothing_radius, 1.0e-4f);
    const float t2 = a.t2[pc.fluid_mat];
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    // one listed neighbor's term (same material, inside the kernel's support)
    auto term = [&](uint32_t q, uint32_t qmat, float dx, float dy, float dz) {
        if (qmat != pc.fluid_mat) return;
        const float dist = len3(dx, dy, dz);
        if (dist <= kEps || dist >= h) return;
        // FluidSpikyKernelGradient, physics_fluid_common.hlsli:45-54
        const float dw = spiky_dw(dist, h);
        const float gx = dw * fdiv(dx, dist), gy = dw * fdiv(dy, dist), gz = dw * fdiv(dz, dist);
        if (APPLY) {
            const float mj = vort[q].w;
            ax = ax + mj * gx;
            ay = ay + mj * gy;
            az = az + mj * gz;
        } else {
            const float4 vj = vel[q];
            const float wx = vj.x - self.x, wy = vj.y - self.y, wz = vj.z - self.z;
            ax = ax + (wy * gz - wz * gy);
            ay = ay + (wz * gx - wx * gz);
            az = az + (wx * gy - wy * gx);
        }
    };
    if (c.cap > 0 && (!UNIFORM || neighbor_ok(pc, pc))) {
        if (APPLY && a.nbr) {
            // the list the curl pass wrote: the same neighbors in the same order (positions unchanged)
            const uint32_t cnt = a.nbr_cnt[i];
            for (uint32_t k = 0; k < cnt; ++k) {
                const uint32_t q = a.nbr[(size_t)k * c.n + i];
                const float4 pj = a.pos[q];
                term(q, UNIFORM ? pc.fluid_mat : a.classes[a.cls[q]].fluid_mat, s.x - pj.x, s.y - pj.y, s.z - pj.z);
            }
        } else {
            uint32_t cnt = 0;
            walk_runs(a.cellc[i], run_table(a), c, [&](uint32_t q) {
                if (q == i) return true;
                const uint32_t qmat = UNIFORM ? pc.fluid_mat : a.classes[a.cls[q]].fluid_mat;
                if (!UNIFORM && !neighbor_ok(pc, a.classes[a.cls[q]])) return true;
                const float4 pj = a.pos[q];
                if (pj.w <= kEps) return true;
                const float dx = s.x - pj.x, dy = s.y - pj.y, dz = s.z - pj.z;
                if (!(dot3(dx, dy, dz, dx, dy, dz) < t2)) return true;
                if (!APPLY && a.nbr) a.nbr[(size_t)cnt * c.n + i] = q;  // for the confinement pass
                term(q, qmat, dx, dy, dz);
                return ++cnt < c.cap;
            });
            if (!APPLY && a.nbr) a.nbr_cnt[i] = cnt;
        }
    }
    if (!APPLY) {
        vort[i] = make_float4(ax, ay, az, len3(ax, ay, az));
        return;
    }
    const float gl = len3(ax, ay, az);
    if (gl <= kEps) return;
    const float nx = fdiv(ax, gl), ny = fdiv(ay, gl), nz = fdiv(az, gl);
    const float fx = conf * (ny * self.z - nz * self.y);
    const float fy = conf * (nz * self.x - nx * self.z);
    const float fz = conf * (nx * self.y - ny * self.x);
    const float4 v = vel[i];
    vel[i] = make_float4(v.x + c.dt * fx, v.y + c.dt * fy, v.z + c.dt * fz, 0.0f);
}

// FluidAnisotropyJacobiRotate, physics_fluid_anisotropy_common.hlsli:28-74 (rsqrt read as
// 1/sqrt): the rotation in the (P, Q) plane that zeroes A[P][Q], accumulated into R.
template <int P, int Q>
__device__ __forceinline__ void jacobi_rotate(float (&A)[3][3], float (&R)[3][3])
{
    constexpr int K = 3 - P - Q;
    const float apq = A[P][Q];
    if (fabsf(apq) <= 1.0e-20f) return;
    const float d = fdiv(A[P][P] - A[Q][Q], 2.0f * apq);
    float t = fdiv(1.0f, fabsf(d) + fsqrt(d * d + 1.0f));
    if (d < 0.0f) t = -t;
    const float cs = fdiv(1.0f, fsqrt(t * t + 1.0f));
    const float sn = t * cs;
    A[P][P] = A[P][P] + t * apq;
    A[Q][Q] = A[Q][Q] - t * apq;
    A[P][Q] = 0.0f;
    A[Q][P] = 0.0f;
    const float akp = cs * A[K][P] + sn * A[K][Q];
    const float akq = -sn * A[K][P] + cs * A[K][Q];
    A[K][P] = akp;
    A[P][K] = akp;
    A[K][Q] = akq;
    A[Q][K] = akq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float rkp = cs * R[k][P] + sn * R[k][Q];
        const float rkq = -sn * R[k][P] + cs * R[k][Q];
        R[k][P] = rkp;
        R[k][Q] = rkq;
    }
}

template <int A_, int B_>
__device__ __forceinline__ void sort_pair(float (&ev)[3], float (&ax)[3][3])
{
    if (ev[A_] < ev[B_]) {
        const float t = ev[A_];
        ev[A_] = ev[B_];
        ev[B_] = t;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float u = ax[A_][c];
            ax[A_][c] = ax[B_][c];
            ax[B_][c] = u;
        }
    }
}

// Render anisotropy: physics_fluid_build_render_anisotropy.cs.hlsl:34-171 with
// physics_fluid_anisotropy_common.hlsli:9-156 (rcp and rsqrt read as correctly rounded 1/x and
// 1/sqrt), over the post-update list walked again, twice: the weighted mean of the listed
// neighbors, then their weighted covariance; four largest-pivot Jacobi rotations, axes sorted by
// decreasing eigenvalue, scales 1.1 sqrt(eigenvalue) clamped to [r, 2r].  A pinned fluid particle
// has no list (the reference reads a count its build never wrote there) and gets the isotropic
// frame.  Written in particle-index order: out[k * n + particle] for axis k.
template <bool UNIFORM>
__global__ __launch_bounds__(kBlock) void k_fluid_render_aniso(FluidArgs a, SubstepConsts c,
                                                               const uint32_t *__restrict__ id,
                                                               float4 *__restrict__ out)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= c.n) return;
    const ParticleClass pc = a.classes[UNIFORM ? 0u : a.cls[i]];
    if (!a.active[pc.scene]) return;
    const size_t pid = id[i];
    const size_t n = c.npid;
    const float radius = fmaxd(pc.radius, 1.0e-4f);
    float ax[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
    float sc[3] = {radius, radius, radius};
    if (pc.kind == CN_PARTICLE_KIND_FLUID) {
        const cn_fluid_material &m = a.fm[pc.fluid_mat];
        const float support = fmaxd(m.smoothing_radius, radius);
        const float inv = fdiv(1.0f, support);
        const float iso = fmaxd(1.0f * radius, radius);
        const float4 s = a.pos[i];
        const float t2 = a.t2[pc.fluid_mat];
        const bool walk = s.w > kEps && c.cap > 0 && (!UNIFORM || neighbor_ok(pc, pc));
        // f(neighbor position, weight) per contributing entry of the list
        auto weigh = [&](const float4 &pj, float d2, auto &&f) {
            if (d2 > 0.0f) {
                const float dist = fsqrt(d2);
                if (dist < support) {
                    const float x = dist * inv;
                    f(pj, 1.0f - x * x * x);
                }
            }
        };
        // first = true: walk the run table (storing the list where a list buffer is given);
        // otherwise read that list back: the same neighbors in the same order
        auto each = [&](bool first, auto &&f) {
            if (!first && a.nbr) {
                const uint32_t cnt = a.nbr_cnt[i];
                for (uint32_t k = 0; k < cnt; ++k) {
                    const float4 pj = a.pos[a.nbr[(size_t)k * c.n + i]];
                    const float dx = s.x - pj.x, dy = s.y - pj.y, dz = s.z - pj.z;
                    weigh(pj, dot3(dx, dy, dz, dx, dy, dz), f);
                }
                return;
            }
            uint32_t cnt = 0;
            walk_runs(a.cellc[i], run_table(a), c, [&](uint32_t q) {
                if (q == i) return true;
                if (!UNIFORM && !neighbor_ok(pc, a.classes[a.cls[q]])) return true;
                const float4 pj = a.pos[q];
                if (pj.w <= kEps) return true;
                const float dx = s.x - pj.x, dy = s.y - pj.y, dz = s.z - pj.z;
                const float d2 = dot3(dx, dy, dz, dx, dy, dz);
                if (!(d2 < t2)) return true;
                if (first && a.nbr) a.nbr[(size_t)cnt * c.n + i] = q;
                weigh(pj, d2, f);
                return ++cnt < c.cap;
            });
            if (first && a.nbr) a.nbr_cnt[i] = cnt;
        };
        float mx = 0.0f, my = 0.0f, mz = 0.0f, wsum = 0.0f;
        if (walk)
            each(true, [&](const float4 &pj, float w) {
                wsum = wsum + w;
                mx = mx + pj.x * w;
                my = my + pj.y * w;
                mz = mz + pj.z * w;
            });
        if (!(wsum <= 0.0f)) {
            mx = fdiv(mx, wsum);
            my = fdiv(my, wsum);
            mz = fdiv(mz, wsum);
            float C[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
            each(false, [&](const float4 &pj, float w) {
                const float cc[3] = {pj.x - mx, pj.y - my, pj.z - mz};
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int q = 0; q < 3; ++q) C[r][q] = C[r][q] + w * cc[r] * cc[q];
            });
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int q = 0; q < 3; ++q) C[r][q] = fdiv(C[r][q], wsum);
            C[0][1] = C[1][0] = 0.5f * (C[0][1] + C[1][0]);
            C[0][2] = C[2][0] = 0.5f * (C[0][2] + C[2][0]);
            C[1][2] = C[2][1] = 0.5f * (C[1][2] + C[2][1]);
            // FluidAnisotropyEigenDecomposition, physics_fluid_anisotropy_common.hlsli:76-122
            float R[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
            for (int it = 0; it < 4; ++it) {
                int pair = 0;
                float mxv = fabsf(C[0][1]);
                const float a02 = fabsf(C[0][2]);
                if (a02 > mxv) {
                    pair = 1;
                    mxv = a02;
                }
                const float a12 = fabsf(C[1][2]);
                if (a12 > mxv) {
                    pair = 2;
                    mxv = a12;
                }
                if (mxv < 1.0e-15f) break;
                if (pair == 0)
                    jacobi_rotate<0, 1>(C, R);
                else if (pair == 1)
                    jacobi_rotate<0, 2>(C, R);
                else
                    jacobi_rotate<1, 2>(C, R);
            }
            float ev[3] = {fmaxd(C[0][0], 0.0f), fmaxd(C[1][1], 0.0f), fmaxd(C[2][2], 0.0f)};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float vx = R[0][k], vy = R[1][k], vz = R[2][k];
                const float l2 = dot3(vx, vy, vz, vx, vy, vz);
                if (l2 > 1.0e-8f) {
                    const float rl = fdiv(1.0f, fsqrt(l2));
                    ax[k][0] = vx * rl;
                    ax[k][1] = vy * rl;
                    ax[k][2] = vz * rl;
                }
            }
            sort_pair<0, 1>(ev, ax);
            sort_pair<0, 2>(ev, ax);
            sort_pair<1, 2>(ev, ax);
#pragma unroll
            for (int k = 0; k < 3; ++k) sc[k] = fmind(fmaxd(fsqrt(ev[k]) * 1.1f, 1.0f * radius), 2.0f * radius);
        } else {
            sc[0] = sc[1] = sc[2] = iso;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k * n + pid] = make_float4(ax[k][0], ax[k][1], ax[k][2], sc[k]);
}

// ---------------------------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------------------------
static FluidArgs fluid_args(const SceneBuffers &s, const SubstepConsts &c, uint32_t it, int din)
{
    FluidArgs a;
    const int cur = s.cur;
    a.pos = s.pos[cur];
    a.cls = s.cls[cur];
    a.cellc = s.cellc;
    a.bnd = s.bnd;
    a.bndl = s.bndl;
    a.din = s.dacc[din];
    a.dout = s.dacc[1 - din];
    a.cur = s.curpos;
    a.cs = s.cs;
    a.nc = s.nc;
    a.nbr = s.nbr;
    a.nbr_cnt = s.nbr_cnt;
    a.list = s.list;
    a.lcnt = s.lcnt;
    a.slist = s.slist;
    a.scnt = s.scnt;
    a.disp = c.superset ? s.disp : nullptr;
    a.run_key = s.run_key;
    a.run_start = s.run_start;
    a.run_info = s.run_info;
    a.bucket = s.bucket;
    a.bucket_shift = c.bucket_shift;
    a.blk_first_run = s.blk_first_run;
    a.nblk = s.nblk;
    a.blk_halo = s.blk_halo;
    a.wide_blk = s.wide_blk;
    a.iter_state = s.iter_state;
    a.classes = s.classes;
    a.fm = s.fluid_mats;
    a.t2 = s.fluid_t2;
    a.col = s.colliders;
    a.scenes = s.scenes;
    a.active = s.scene_active;
    a.meta = s.meta;
    a.it = it;
    a.tile = 0;
    return a;
}

static uint32_t grid_of(uint32_t n) { return (n + kBlock - 1) / kBlock; }

// LDS budget of one tiled workgroup: the dynamic tile plus the static halo table stays under a third
// of the CU's 160 KiB, so three workgroups share a CU
constexpr uint32_t kDynLdsBytes = 50u * 1024u;

static bool tiled(const SubstepConsts &c, uint32_t it) { return c.key_mode == CN_KEY_MORTON && it < kIterStateWords; }

// persistent grid: workgroups in flight on all dies, each pulling cell blocks from the queues
static uint32_t persistent_grid(const SubstepConsts &c)
{
    return std::min<uint32_t>(std::max<uint32_t>(c.grid_max, kXcds), std::max<uint32_t>(kXcds, (c.n + 127u) / 128u));
}

// Halo capacities: the density pass stages positions (and classes, kTilePad spare entries each),
// the delta pass records, previous deltas, normals and classes, the wide delta pass the same
// without previous deltas but with a slot map.  density >= wide >= delta, so every block the delta
// passes stage was staged (and listed in halo offsets) by the density pass.
void fluid_tiles(SubstepConsts &c)
{
    const uint32_t cls = c.fluid_uniform ? 0u : 4u, nc = c.tension ? 16u : 0u;
    c.tile_dens = std::min<uint32_t>(c.tile_max, kDynLdsBytes / (16u + cls) - kTilePad);
    c.tile_delta = std::min<uint32_t>(c.tile_max, kDynLdsBytes / (32u + nc + cls));
    c.tile_wide = std::min<uint32_t>(c.tile_max, kDynLdsBytes / (20u + nc + cls));
}

void launch_blk_halo(const SceneBuffers &s, const SubstepConsts &c, hipStream_t st)
{
    const FluidArgs a = fluid_args(s, c, 0, 0);
    const uint32_t g = std::min<uint32_t>(std::max<uint32_t>(c.blk_cap, 1u), 4096u);
    hipLaunchKernelGGL(k_blk_halo, dim3(g), dim3(kBlock), 0, st, a, c);
}

// One of a tiled pass's class instantiations: one class in the launch, one per scene with fluid,
// or classes per particle.
template <void (*ONE)(FluidArgs, SubstepConsts), void (*BLOCK)(FluidArgs, SubstepConsts),
          void (*MIXED)(FluidArgs, SubstepConsts)>
static void launch_tiled(const SubstepConsts &c, dim3 grid, dim3 block, size_t lds, hipStream_t st, const FluidArgs &a)
{
    if (c.uniform_class)
        hipLaunchKernelGGL(ONE, grid, block, lds, st, a, c);
    else if (c.fluid_uniform)
        hipLaunchKernelGGL(BLOCK, grid, block, lds, st, a, c);
    else
        hipLaunchKernelGGL(MIXED, grid, block, lds, st, a, c);
}

void launch_fluid_density(const SceneBuffers &s, const SubstepConsts &c, uint32_t it, int din, hipStream_t st)
{
    FluidArgs a = fluid_args(s, c, it, din);
    if (tiled(c, it)) {
        const uint32_t per = 16u + (c.fluid_uniform ? 0u : 4u);
        a.tile = c.tile_dens;
        const size_t lds = (size_t)(a.tile + kTilePad) * per;
        const dim3 grid(persistent_grid(c)), block(kTileBlock);
        if (c.superset && it == 0) {
            launch_tiled<k_fluid_density_tiled<kClsOne, kDensBuild>, k_fluid_density_tiled<kClsBlock, kDensBuild>,
                         k_fluid_density_tiled<kClsMixed, kDensBuild>>(c, grid, block, lds, st, a);
        } else if (c.superset) {
            launch_tiled<k_fluid_density_tiled<kClsOne, kDensReuse>, k_fluid_density_tiled<kClsBlock, kDensReuse>,
                         k_fluid_density_tiled<kClsMixed, kDensReuse>>(c, grid, block, lds, st, a);
        } else {
            launch_tiled<k_fluid_density_tiled<kClsOne, kDensPlain>, k_fluid_density_tiled<kClsBlock, kDensPlain>,
                         k_fluid_density_tiled<kClsMixed, kDensPlain>>(c, grid, block, lds, st, a);
        }
    } else {
        if (c.uniform_class)
            hipLaunchKernelGGL(k_fluid_density_global<true>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c);
        else
            hipLaunchKernelGGL(k_fluid_density_global<false>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c);
    }
}

void launch_fluid_delta(const SceneBuffers &s, const SubstepConsts &c, uint32_t it, int din, hipStream_t st)
{
    FluidArgs a = fluid_args(s, c, it, din);
    if (tiled(c, it)) {
        const uint32_t cls = c.fluid_uniform ? 0u : 4u, nc = c.tension ? 16u : 0u;
        // first iteration: no previous delta, so the delta pass stages as much as the wide one
        a.tile = it > 0 ? c.tile_delta : c.tile_wide;
        size_t lds = (size_t)a.tile * (16u + (it > 0 ? 16u : 0u) + nc + cls);
        const dim3 grid(persistent_grid(c)), block(kTileBlock);
        launch_tiled<k_fluid_delta_tiled<kClsOne, false>, k_fluid_delta_tiled<kClsBlock, false>,
                     k_fluid_delta_tiled<kClsMixed, false>>(c, grid, block, lds, st, a);
        if (it > 0 && c.tile_wide > c.tile_delta) {
            a.tile = c.tile_wide;
            lds = (size_t)a.tile * (20u + nc + cls);
            const dim3 wgrid(persistent_grid(c));
            launch_tiled<k_fluid_delta_tiled<kClsOne, true>, k_fluid_delta_tiled<kClsBlock, true>,
                         k_fluid_delta_tiled<kClsMixed, true>>(c, wgrid, block, lds, st, a);
        }
    } else {
        if (c.uniform_class)
            hipLaunchKernelGGL(k_fluid_delta_global<true>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c);
        else
            hipLaunchKernelGGL(k_fluid_delta_global<false>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c);
    }
}

void launch_fluid_vorticity(const SceneBuffers &s, const SubstepConsts &c, hipStream_t st)
{
    const FluidArgs a = fluid_args(s, c, 0, 0);
    float4 *vel = s.vel[s.cur];
    float4 *vort = s.cs;  // free between the last delta pass of a substep and the next density pass
    if (c.uniform_class) {
        hipLaunchKernelGGL((k_fluid_vorticity<true, false>), dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c, vel, vort);
        hipLaunchKernelGGL((k_fluid_vorticity<true, true>), dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c, vel, vort);
    } else {
        hipLaunchKernelGGL((k_fluid_vorticity<false, false>), dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c, vel, vort);
        hipLaunchKernelGGL((k_fluid_vorticity<false, true>), dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c, vel, vort);
    }
}

void launch_fluid_render_aniso(const SceneBuffers &s, const SubstepConsts &c, float4 *out, hipStream_t st)
{
    const FluidArgs a = fluid_args(s, c, 0, 0);
    if (c.uniform_class)
        hipLaunchKernelGGL(k_fluid_render_aniso<true>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c, s.id[s.cur], out);
    else
        hipLaunchKernelGGL(k_fluid_render_aniso<false>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c, s.id[s.cur], out);
}

void launch_fluid_lists(const SceneBuffers &s, const SubstepConsts &c, hipStream_t st)
{
    const FluidArgs a = fluid_args(s, c, 0, 0);
    if (c.uniform_class)
        hipLaunchKernelGGL(k_fluid_lists<true>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c);
    else
        hipLaunchKernelGGL(k_fluid_lists<false>, dim3(grid_of(c.n)), dim3(kBlock), 0, st, a, c);
}

}  // namespace cn
