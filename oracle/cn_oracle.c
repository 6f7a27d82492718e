/*
 * cn_oracle.c -- CPU restatement of CRESSim-Neo's particle substep.
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle and the CPU baseline.  The product library never
 * links, loads or calls this file.  See cn_oracle.h for the conventions.  Each function names the
 * reference file:line it restates (paths relative to the reference repository root).
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -fno-fast-math (oracle/Makefile).
 */
#include "cn_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define EPS 1.0e-6f             /* kEpsilon, assets/shaders/include/physics/core/physics_base.hlsli:7 */
#define INVALID 0xffffffffu     /* kInvalidIndex, physics_base.hlsli:6 */
#define KERNEL_PI 3.14159265359f /* literal of physics_fluid_common.hlsli:26 */
#define BROADPHASE_MARGIN 0.05f  /* kBroadPhaseMargin, physics_shape_common.hlsli:10 */

/* solver config indices (physics_solver_config.hlsli:19-51) */
#define CFG_CONTACT_SLOP 0
#define CFG_POS_FRICTION_MIN 7
#define CFG_SOFT_RELAX 24
#define CFG_FLUID_SOLVE 32
#define CFG_FLUID_RELAX 33
#define CFG_FLUID_BND_DENSITY 34
#define CFG_FLUID_BND_DELTA 35
#define CFG_FLUID_MAX_UNDER 36

static float maxf(float a, float b) { return a > b ? a : b; }
static float minf(float a, float b) { return a < b ? a : b; }
static float saturatef(float a) { return minf(maxf(a, 0.0f), 1.0f); }
static float dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static float len3(const float *a) { return sqrtf(dot3(a, a)); }

int cno_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void cno_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------------------------
 * Host authoring
 * ------------------------------------------------------------------------------------------ */

/* PhysicsSolverDesc defaults, include/physics/physics_solver.h:21-143, packed by makeSolverConfig
 * (src/physics/physics_solver.cpp:100-129). */
void cno_default_params(cn_step_params *p)
{
    memset(p, 0, sizeof(*p));
    p->dt = 1.0f / 60.0f;
    p->substeps = 1;
    p->default_iterations = 20;
    p->gravity[0] = 0.0f;
    p->gravity[1] = -9.81f;
    p->gravity[2] = 0.0f;
    const float cfg[CN_SOLVER_CONFIG_FLOATS] = {
        1.0e-3f, 4.0f, 0.95f, 0.05f,        /* contact0 */
        0.5f, 2.0f, 0.0f, 1.0e-5f,          /* contact1 */
        0.25f, 0.90f, 0.10f, 1.0f,          /* rigid0 */
        0.01f, 0.10f, 0.8f, 0.5f,           /* rigid1 */
        0.95f, 0.7f, 0.05f, 0.02f,          /* joints0 */
        0.12f, 1.0e-5f, 5.0e-5f, 1.0e-5f,   /* joints1 */
        0.2f, 0.0f, 0.0f, 0.0f,             /* soft0 */
        0.0f, 0.0f, 0.0f, 0.0f,             /* soft1 */
        0.7f, 1.0f, 0.12f, 0.08f,           /* fluid0 */
        0.0f, 0.0f, 0.0f, 0.0f};            /* fluid1 */
    memcpy(p->solver_config, cfg, sizeof(cfg));
    p->max_fluid_neighbors = 128;  /* kDefaultFluidMaxNeighborhood, rigid_body_common.h:304 */
    p->key_mode = CN_KEY_HASH;
}

/* fluidSpikyWeight / fluidSpikyWeightDerivative, src/physics/physics_world.cpp:99-112 */
static const float kHostPi = 3.14159265358979323846f;
static float host_spiky_w(float d, float h)
{
    const float omq = 1.0f - d / h;
    const float k = 15.0f / (kHostPi * h * h * h);
    return k * omq * omq;
}
static float host_spiky_dw(float d, float h)
{
    const float omq = 1.0f - d / h;
    const float k = 15.0f / (kHostPi * h * h * h);
    return -k * 2.0f * omq / h;
}

/* tightPack3D + calculateFluidSolveScales, src/physics/physics_world.cpp:140-231 */
static void fluid_solve_scales(float rest_distance, float support, float *rest_density,
                               float *density_scale, float *surface_scale)
{
    const int dim = (int)ceilf(support / rest_distance);
    const float row_scale = sqrtf(0.75f) * rest_distance;
    float rho = 0.0f, dcs = 0.0f, a = 0.0f, b = 0.0f;
    uint32_t count = 0;
    for (int z = -dim; z <= dim; ++z)
        for (int y = -dim; y <= dim; ++y)
            for (int x = -dim; x <= dim; ++x) {
                const float sx =
                    (float)x * rest_distance + ((((y + z) & 1) != 0) ? rest_distance * 0.5f : 0.0f);
                const float sy = (float)y * row_scale;
                const float sz = (float)z * row_scale;
                const float m2 = sx * sx + sy * sy + sz * sz;
                if (m2 == 0.0f)
                    continue;
                if (!(sqrtf(m2) <= support))
                    continue;
                if (count == 2048u)
                    goto done;
                ++count;
                const float dist = sqrtf(sx * sx + sy * sy + sz * sz);
                rho += host_spiky_w(dist, support);
                const float der = host_spiky_dw(dist, support);
                dcs += der * der;
                if (sy <= 0.0f && dist > 0.0f) {
                    const float ct = sy / dist;
                    a += der * ct;
                    b -= dist * ct;
                }
            }
done:
    if (rho <= 1.0e-6f)
        rho = 1.0f;
    if (dcs <= 1.0e-6f)
        dcs = 1.0f;
    *rest_density = rho;
    *density_scale = dcs;
    *surface_scale = (fabsf(b) > 1.0e-6f && fabsf(a) > 1.0e-6f) ? fabsf(a / b) : 1.0f;
}

/* toFluidSolverMaterial, src/physics/physics_world.cpp:233-263 */
void cno_fluid_material(const cn_fluid_material_desc *desc, float radius, cn_fluid_material *out)
{
    const float rest_distance = maxf(2.0f * radius, 1.0e-4f);
    const float h = maxf((2.0f / 0.6f) * radius, rest_distance);
    float rho, dcs, sts;
    fluid_solve_scales(rest_distance, h, &rho, &dcs, &sts);
    const float inv_rho = 1.0f / rho;
    float rr = rest_distance / h;
    rr = rr < 1.0e-3f ? 1.0e-3f : (rr > 0.999f ? 0.999f : rr);
    const float rr2 = rr * rr;
    out->rest_density = rho;
    out->inv_rest_density = inv_rho;
    out->smoothing_radius = h;
    out->density_constraint_scale = 1.0f / maxf(dcs, 1.0e-6f);
    out->viscosity = desc->viscosity * inv_rho;
    out->cohesion = desc->cohesion * h;
    out->cohesion1 = -(1.0f + rr) / rr2;
    out->cohesion2 = (rr2 + rr + 1.0f) / rr2;
    out->surface_tension = (desc->surface_tension * inv_rho) / sts;
    out->vorticity_confinement = desc->vorticity_confinement * inv_rho * h;
    out->gravity_scale = desc->gravity_scale;
    out->cfl_radius = desc->cfl_coefficient * h;
}

/* recomputeParticleGridCellSize, src/physics/physics_world.cpp:6829-6846 */
float cno_grid_cell_size(uint32_t n, const float *radii, const uint32_t *fluid_mat, uint32_t n_mats,
                         const cn_fluid_material *mats)
{
    float g = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        g = maxf(g, radii[i] * 2.0f);
        if (fluid_mat && fluid_mat[i] != INVALID && fluid_mat[i] < n_mats)
            g = maxf(g, mats[fluid_mat[i]].smoothing_radius);
    }
    return maxf(g, 1.0e-4f);
}

/* ------------------------------------------------------------------------------------------
 * Predict: assets/shaders/physics/soft/solver/physics_soft_predict.cs.hlsl:12-47
 * ------------------------------------------------------------------------------------------ */
void cno_predict(cno_scene *s, float dt, const float *g)
{
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        float *p = s->pos + 4 * i;
        float *pv = s->prev + 4 * i;
        float *v = s->vel + 4 * i;
        pv[0] = p[0];
        pv[1] = p[1];
        pv[2] = p[2];
        pv[3] = 0.0f;
        if (p[3] <= EPS) {
            v[0] = v[1] = v[2] = v[3] = 0.0f;
            continue;
        }
        float gs = 1.0f;
        if (s->kinds[i] == CN_PARTICLE_KIND_FLUID)
            gs = s->fluid_mats[s->fluid_mat[i]].gravity_scale;
        const float dts = dt * gs;
        const float vx = v[0] + g[0] * dts;
        const float vy = v[1] + g[1] * dts;
        const float vz = v[2] + g[2] * dts;
        v[0] = vx;
        v[1] = vy;
        v[2] = vz;
        v[3] = 0.0f;
        p[0] = p[0] + vx * dt;
        p[1] = p[1] + vy * dt;
        p[2] = p[2] + vz * dt;
    }
}

/* ------------------------------------------------------------------------------------------
 * Broadphase
 * ------------------------------------------------------------------------------------------ */

/* ComputeParticleGridCellKey, assets/shaders/include/physics/particle/physics_particle_grid.hlsli:4-10 */
static uint32_t cell_hash(int32_t gx, int32_t gy, int32_t gz)
{
    uint32_t seed = (uint32_t)gx * 73856093u;
    seed ^= (uint32_t)gy * 19349663u;
    seed ^= (uint32_t)gz * 83492791u;
    return seed;
}

/* physics_particle_build_broad_phase_entries.cs.hlsl:9-35 + ..._keys.cs.hlsl:8-23 */
void cno_broadphase_entries(uint32_t n, const float *pos, float cell_size, uint32_t *keys,
                            int32_t *cells)
{
    const float safe = maxf(cell_size, 1.0e-4f);
    const float inv = 1.0f / safe;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const int32_t gx = (int32_t)floorf(pos[4 * i + 0] * inv);
        const int32_t gy = (int32_t)floorf(pos[4 * i + 1] * inv);
        const int32_t gz = (int32_t)floorf(pos[4 * i + 2] * inv);
        cells[3 * i + 0] = gx;
        cells[3 * i + 1] = gy;
        cells[3 * i + 2] = gz;
        keys[i] = cell_hash(gx, gy, gz);
    }
}

/* Exclusive scan, the result of dispatchExclusiveScanWithCpuCount
 * (src/physics/physics_pass_dispatcher.cpp:445-602, physics_scan_block.cs.hlsl:10-52,
 * physics_scan_add_offsets.cs.hlsl:7-22): out[i] = sum(in[0..i-1]) modulo 2^32. */
static void exclusive_scan(uint32_t n, const uint32_t *in, uint32_t *out)
{
    uint32_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = acc;
        acc += in[i];
    }
}

/* dispatchSoftRadixSortPass, src/physics/physics_pass_dispatcher.cpp:3963-4063, with
 * physics_radix_classify.cs.hlsl:7-16, physics_radix_finalize.cs.hlsl:8-24 and
 * physics_radix_scatter.cs.hlsl:10-23: 32 one-bit split passes over (key, index) pairs. */
void cno_radix_sort_1bit(uint32_t n, uint32_t *keys, uint32_t *indices)
{
    if (n == 0)
        return;
    uint32_t *k2 = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *i2 = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *flags = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *offs = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *ki = keys, *ii = indices, *ko = k2, *io = i2;
    for (uint32_t bit = 0; bit < 32u; ++bit) {
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < (int64_t)n; ++t)
            flags[t] = (ki[t] >> bit) & 1u;
        exclusive_scan(n, flags, offs);
        const uint32_t total_ones = offs[n - 1] + flags[n - 1];
        const uint32_t total_zeros = n - total_ones;
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < (int64_t)n; ++t) {
            const uint32_t dst =
                flags[t] == 0u ? (uint32_t)t - offs[t] : total_zeros + offs[t];
            ko[dst] = ki[t];
            io[dst] = ii[t];
        }
        uint32_t *tk = ki; ki = ko; ko = tk;
        uint32_t *ti = ii; ii = io; io = ti;
    }
    /* 32 passes: the result ends in the original buffers */
    free(k2);
    free(i2);
    free(flags);
    free(offs);
}

/* nextPowerOfTwo, src/physics/physics_solver.cpp:22-36 */
uint32_t cno_next_pow2(uint32_t v)
{
    if (v <= 1u)
        return 1u;
    --v;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    return v + 1u;
}

/* clearParticleCellRanges / markParticleCellRangeStarts / scan / buildParticleCellRanges:
 * physics_particle_clear_cell_ranges.cs.hlsl:6-21, ..._mark_cell_range_starts.cs.hlsl:7-28,
 * ..._build_cell_ranges.cs.hlsl:10-58; driver physics_pass_dispatcher.cpp:989-1085. */
void cno_cell_ranges(uint32_t n, const uint32_t *sk, uint32_t capacity, uint32_t *ranges)
{
    for (uint32_t r = 0; r < capacity; ++r) {
        ranges[4 * r + 0] = INVALID;
        ranges[4 * r + 1] = 0;
        ranges[4 * r + 2] = 0;
        ranges[4 * r + 3] = 0;
    }
    uint32_t run = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (i != 0 && sk[i - 1] == sk[i])
            continue;
        uint32_t end = i + 1;
        while (end < n && sk[end] == sk[i])
            ++end;
        if (run < capacity) {
            ranges[4 * run + 0] = sk[i];
            ranges[4 * run + 1] = i;
            ranges[4 * run + 2] = end;
            ranges[4 * run + 3] = 0;
        }
        ++run;
    }
}

/* FindParticleCellRange, assets/shaders/include/physics/fluid/physics_fluid_neighbor_build.hlsli:6-46.
 * Returns the record at the lower bound; the caller treats key == 0xffffffff as missing. */
int cno_find_cell_range(const uint32_t *ranges, uint32_t capacity, uint32_t key, uint32_t *start,
                        uint32_t *end)
{
    if (capacity == 0)
        return 0;
    uint32_t lo = 0, hi = capacity;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (ranges[4 * mid] < key)
            lo = mid + 1u;
        else
            hi = mid;
    }
    if (lo < capacity && ranges[4 * lo] == key) {
        if (ranges[4 * lo] == INVALID)
            return 0;
        *start = ranges[4 * lo + 1];
        *end = ranges[4 * lo + 2];
        return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * Static box boundary
 * ------------------------------------------------------------------------------------------ */

/* ExpandBits / Morton3D, assets/shaders/physics/rigid/broad_phase/physics_rigid_morton_codes.cs.hlsl:8-26 */
static uint32_t expand_bits(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
static float clampf(float x, float lo, float hi) { return minf(maxf(x, lo), hi); }
static uint32_t morton3d(float x, float y, float z)
{
    x = clampf(x * 1024.0f, 0.0f, 1023.0f);
    y = clampf(y * 1024.0f, 0.0f, 1023.0f);
    z = clampf(z * 1024.0f, 0.0f, 1023.0f);
    return expand_bits((uint32_t)x) * 4u + expand_bits((uint32_t)y) * 2u + expand_bits((uint32_t)z);
}

/* A static collider resolved to its world pose: ComputeFluidBoundaryContactInfo's pose (the body
 * pose composed with the unscaled local position, physics_fluid_boundary_common.hlsli:21-33) and
 * the broad-phase AABB of physics_rigid_update_world_aabbs.cs.hlsl:51-91 (scaled local position,
 * ComputeBodyAabb of physics_shape_common.hlsli:49-75, the margin added after).  Quaternion helpers
 * follow physics_math.hlsli with rsqrt taken as 1 / sqrt, sums left to right. */
typedef struct ocol {
    uint32_t shape, rotated, enabled, env, layer, mask;
    float c[3], he[3], q[4], sa[3], sb[3], mn[3], mx[3], material[4];
} ocol;

static void quat_normalize(const float *q, float *o)
{
    const float l = ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3];
    if (l <= EPS) {
        o[0] = o[1] = o[2] = 0.0f;
        o[3] = 1.0f;
        return;
    }
    const float r = 1.0f / sqrtf(l);
    for (int a = 0; a < 4; ++a)
        o[a] = q[a] * r;
}
static void quat_mul(const float *a, const float *b, float *o)
{
    o[0] = ((a[3] * b[0] + a[0] * b[3]) + a[1] * b[2]) - a[2] * b[1];
    o[1] = ((a[3] * b[1] - a[0] * b[2]) + a[1] * b[3]) + a[2] * b[0];
    o[2] = ((a[3] * b[2] + a[0] * b[1]) - a[1] * b[0]) + a[2] * b[3];
    o[3] = ((a[3] * b[3] - a[0] * b[0]) - a[1] * b[1]) - a[2] * b[2];
}
/* QuaternionRotate (physics_math.hlsli:33-37) */
static void quat_rotate(const float *q, const float *v, float *o)
{
    const float t[3] = {2.0f * (q[1] * v[2] - q[2] * v[1]), 2.0f * (q[2] * v[0] - q[0] * v[2]),
                        2.0f * (q[0] * v[1] - q[1] * v[0])};
    o[0] = (v[0] + q[3] * t[0]) + (q[1] * t[2] - q[2] * t[1]);
    o[1] = (v[1] + q[3] * t[1]) + (q[2] * t[0] - q[0] * t[2]);
    o[2] = (v[2] + q[3] * t[2]) + (q[0] * t[1] - q[1] * t[0]);
}

static void ocol_box(const cn_box_collider *b, ocol *o)
{
    memset(o, 0, sizeof(*o));
    o->shape = CN_SHAPE_BOX;
    o->q[3] = 1.0f;
    for (int a = 0; a < 3; ++a) {
        o->c[a] = b->center[a];
        o->he[a] = b->half_extents[a];
        o->mn[a] = (b->center[a] - b->half_extents[a]) - BROADPHASE_MARGIN;
        o->mx[a] = (b->center[a] + b->half_extents[a]) + BROADPHASE_MARGIN;
    }
    memcpy(o->material, b->material, sizeof(o->material));
    o->enabled = b->enabled;
    o->env = b->environment;
    o->layer = b->layer;
    o->mask = b->mask;
}

static void ocol_shape(const cn_collider *g, ocol *o)
{
    memset(o, 0, sizeof(*o));
    float bq[4], lq[4], m[4], rot[3];
    quat_normalize(g->body_orientation, bq);
    quat_normalize(g->local_orientation, lq);
    quat_mul(bq, lq, m);
    quat_normalize(m, o->q);
    const float *q = o->q;
    o->rotated = !(q[0] == 0.0f && q[1] == 0.0f && q[2] == 0.0f && q[3] == 1.0f);
    o->shape = g->shape;
    const float sx = fabsf(g->body_scale[0]), sy = fabsf(g->body_scale[1]), sz = fabsf(g->body_scale[2]);
    const float *prm = g->shape_params;
    quat_rotate(bq, g->local_position, rot);
    for (int a = 0; a < 3; ++a)
        o->c[a] = g->body_position[a] + rot[a];
    const float up[3] = {0.0f, 1.0f, 0.0f};
    float axis[3];
    quat_rotate(q, up, axis);
    float r = 0.0f, hh = 0.0f;
    if (g->shape == CN_SHAPE_SPHERE) { /* SphereRadius */
        r = prm[0] * maxf(sx, maxf(sy, sz));
        o->he[0] = r;
    } else if (g->shape == CN_SHAPE_BOX) { /* BoxHalfExtents */
        o->he[0] = prm[0] * sx;
        o->he[1] = prm[1] * sy;
        o->he[2] = prm[2] * sz;
    } else { /* CapsuleSegment */
        r = prm[0] * maxf(sx, sz);
        hh = prm[1] * sy;
        o->he[0] = r;
        for (int a = 0; a < 3; ++a) {
            o->sa[a] = o->c[a] - axis[a] * hh;
            o->sb[a] = o->c[a] + axis[a] * hh;
        }
    }
    const float lps[3] = {g->local_position[0] * g->body_scale[0], g->local_position[1] * g->body_scale[1],
                          g->local_position[2] * g->body_scale[2]};
    quat_rotate(bq, lps, rot);
    float ca[3], ext[3];
    for (int a = 0; a < 3; ++a)
        ca[a] = g->body_position[a] + rot[a];
    if (g->shape == CN_SHAPE_SPHERE) {
        ext[0] = ext[1] = ext[2] = r;
    } else if (g->shape == CN_SHAPE_BOX) {
        const float ex[3] = {1.0f, 0.0f, 0.0f}, ey[3] = {0.0f, 1.0f, 0.0f}, ez[3] = {0.0f, 0.0f, 1.0f};
        float ax[3], ay[3], az[3];
        quat_rotate(q, ex, ax);
        quat_rotate(q, ey, ay);
        quat_rotate(q, ez, az);
        for (int a = 0; a < 3; ++a)
            ext[a] = (fabsf(ax[a]) * o->he[0] + fabsf(ay[a]) * o->he[1]) + fabsf(az[a]) * o->he[2];
    } else {
        float A[3], B[3];
        for (int a = 0; a < 3; ++a) {
            A[a] = ca[a] - axis[a] * hh;
            B[a] = ca[a] + axis[a] * hh;
        }
        const float d[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        const float len = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
        for (int a = 0; a < 3; ++a)
            ext[a] = fabsf(axis[a]) * (0.5f * len) + r;
    }
    for (int a = 0; a < 3; ++a) {
        o->mn[a] = (ca[a] - ext[a]) - BROADPHASE_MARGIN;
        o->mx[a] = (ca[a] + ext[a]) + BROADPHASE_MARGIN;
    }
    memcpy(o->material, g->material, sizeof(o->material));
    o->enabled = g->enabled;
    o->env = g->environment;
    o->layer = g->layer;
    o->mask = g->mask;
}

static uint32_t n_cols(const cno_scene *s) { return s->n_colliders + s->n_shape_colliders; }

/* collider k of the scene: the boxes, then the shape colliders */
static void ocol_get(const cno_scene *s, uint32_t k, ocol *o)
{
    if (k < s->n_colliders)
        ocol_box(&s->colliders[k], o);
    else
        ocol_shape(&s->shape_colliders[k - s->n_colliders], o);
}

/* Visit order of the static LBVH traversal used by the fluid-boundary candidate passes
 * (physics_particle_count_fluid_boundary_candidates.cs.hlsl:30-86: stack, right child popped
 * first).  The static tree is a Karras LBVH over the Morton codes of the enabled static
 * colliders (physics_pass_dispatcher.cpp:4163-4256, physics_rigid_morton_codes.cs.hlsl,
 * physics_rigid_bvh_hierarchy.cs.hlsl), whose in-order leaf sequence is the stable Morton order;
 * a right-first depth-first walk therefore meets the leaves in reverse stable Morton order. */
static void order_ocols(uint32_t nc, const ocol *c, uint32_t *order)
{
    uint32_t ne = 0;
    uint32_t *elem = (uint32_t *)malloc(sizeof(uint32_t) * (nc ? nc : 1));
    for (uint32_t i = 0; i < nc; ++i)
        if (c[i].enabled)
            elem[ne++] = i;
    float gmin[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
    float gmax[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
    for (uint32_t e = 0; e < ne; ++e)
        for (int a = 0; a < 3; ++a) {
            gmin[a] = minf(gmin[a], c[elem[e]].mn[a]);
            gmax[a] = maxf(gmax[a], c[elem[e]].mx[a]);
        }
    uint32_t *codes = (uint32_t *)malloc(sizeof(uint32_t) * (ne ? ne : 1));
    uint32_t *eidx = (uint32_t *)malloc(sizeof(uint32_t) * (ne ? ne : 1));
    for (uint32_t e = 0; e < ne; ++e) {
        float m[3];
        for (int a = 0; a < 3; ++a) {
            const float size = maxf(gmax[a] - gmin[a], 1.0e-5f);
            const float center = 0.5f * (c[elem[e]].mn[a] + c[elem[e]].mx[a]);
            m[a] = saturatef((center - gmin[a]) / size);
        }
        codes[e] = morton3d(m[0], m[1], m[2]);
        eidx[e] = e;
    }
    cno_radix_sort_1bit(ne, codes, eidx);
    for (uint32_t k = 0; k < ne; ++k)
        order[k] = elem[eidx[ne - 1 - k]];
    for (uint32_t k = ne; k < nc; ++k)
        order[k] = INVALID;
    free(codes);
    free(eidx);
    free(elem);
}

void cno_static_collider_order(uint32_t nc, const cn_box_collider *c, uint32_t *order)
{
    ocol *oc = (ocol *)malloc(sizeof(ocol) * (nc ? nc : 1));
    for (uint32_t i = 0; i < nc; ++i)
        ocol_box(&c[i], &oc[i]);
    order_ocols(nc, oc, order);
    free(oc);
}

/* the visit order of every collider of the scene (boxes and shape colliders) */
static void scene_collider_order(const cno_scene *s, uint32_t *order)
{
    const uint32_t nc = n_cols(s);
    ocol *oc = (ocol *)malloc(sizeof(ocol) * (nc ? nc : 1));
    for (uint32_t i = 0; i < nc; ++i)
        ocol_get(s, i, &oc[i]);
    order_ocols(nc, oc, order);
    free(oc);
}

/* The box branch of ComputeFluidBoundaryContactInfo in the box frame (local position l) with
 * ClosestFaceNormalLocal (physics_shape_common.hlsli:100-108). */
static void box_local_contact(const float *he, const float *local, float *sd, float *nrm)
{
    float d[3];
    for (int a = 0; a < 3; ++a) {
        const float cl = clampf(local[a], -he[a], he[a]);
        d[a] = local[a] - cl;
    }
    const float dl = len3(d);
    if (dl > EPS) {
        nrm[0] = d[0] / dl;
        nrm[1] = d[1] / dl;
        nrm[2] = d[2] / dl;
        *sd = dl;
    } else {
        float dist[3];
        for (int a = 0; a < 3; ++a)
            dist[a] = he[a] - fabsf(local[a]);
        nrm[0] = nrm[1] = nrm[2] = 0.0f;
        if (dist[0] <= dist[1] && dist[0] <= dist[2])
            nrm[0] = local[0] >= 0.0f ? 1.0f : -1.0f;
        else if (dist[1] <= dist[2])
            nrm[1] = local[1] >= 0.0f ? 1.0f : -1.0f;
        else
            nrm[2] = local[2] >= 0.0f ? 1.0f : -1.0f;
        *sd = -minf(dist[0], minf(dist[1], dist[2]));
    }
}

/* ComputeFluidBoundaryContactInfo, physics_fluid_boundary_common.hlsli:10-84, for a sphere, a box
 * or a capsule; an identity-oriented box takes its frame without a rotation (the rotation by the
 * identity quaternion changes at most the sign of a zero). */
static int collider_contact(const ocol *c, const float *p, float *sd, float *nrm)
{
    nrm[0] = 0.0f;
    nrm[1] = 1.0f;
    nrm[2] = 0.0f;
    *sd = 0.0f;
    if (!c->enabled)
        return 0;
    if (c->shape == CN_SHAPE_BOX) {
        float local[3];
        for (int a = 0; a < 3; ++a)
            local[a] = p[a] - c->c[a];
        if (!c->rotated) {
            box_local_contact(c->he, local, sd, nrm);
            return 1;
        }
        const float qc[4] = {-c->q[0], -c->q[1], -c->q[2], c->q[3]};
        float u[3], m[3];
        quat_rotate(qc, local, u); /* QuaternionInverseRotate */
        box_local_contact(c->he, u, sd, m);
        quat_rotate(c->q, m, nrm);
        return 1;
    }
    float d[3];
    if (c->shape == CN_SHAPE_SPHERE) {
        for (int a = 0; a < 3; ++a)
            d[a] = p[a] - c->c[a];
    } else {
        const float g[3] = {c->sb[0] - c->sa[0], c->sb[1] - c->sa[1], c->sb[2] - c->sa[2]};
        const float lsq = dot3(g, g);
        float t = 0.0f;
        if (lsq > EPS) {
            const float w[3] = {p[0] - c->sa[0], p[1] - c->sa[1], p[2] - c->sa[2]};
            t = saturatef(dot3(w, g) / lsq);
        }
        for (int a = 0; a < 3; ++a)
            d[a] = p[a] - (c->sa[a] + t * g[a]); /* lerp(a, b, t) = a + t (b - a) */
    }
    const float dist = len3(d);
    if (dist > EPS) {
        nrm[0] = d[0] / dist;
        nrm[1] = d[1] / dist;
        nrm[2] = d[2] / dist;
    }
    *sd = dist - c->he[0];
    return 1;
}

int cno_collider_contact(const cn_collider *g, const float *p, float *sd, float *nrm, float *aabb)
{
    ocol o;
    ocol_shape(g, &o);
    if (aabb)
        for (int a = 0; a < 3; ++a) {
            aabb[a] = o.mn[a];
            aabb[3 + a] = o.mx[a];
        }
    return collider_contact(&o, p, sd, nrm);
}

static int collider_contact_k(const cno_scene *s, uint32_t k, const float *p, float *sd, float *nrm)
{
    ocol o;
    ocol_get(s, k, &o);
    return collider_contact(&o, p, sd, nrm);
}

/* Count + scan + emit of buildFluidBoundaryCandidatePairs (physics_pass_dispatcher.cpp:1305-1407,
 * physics_particle_{count,emit,finalize}_fluid_boundary_candidates.cs.hlsl).  ranges[2i] =
 * offset, ranges[2i+1] = count, written for every particle (the reference writes a range only
 * when count > 0; a stale range only names colliders that lie beyond the 2h query box and so
 * contributes nothing).  Returns the required pair count. */
uint32_t cno_fluid_boundary_candidates(const cno_scene *s, const uint32_t *order, uint32_t *ranges,
                                       uint32_t *pairs, uint32_t capacity)
{
    const uint32_t n = s->n;
    uint32_t *counts = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    uint32_t *offs = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    uint32_t nvis = 0;
    const uint32_t ncol = n_cols(s);
    ocol *oc = (ocol *)malloc(sizeof(ocol) * (ncol ? ncol : 1));
    for (uint32_t k = 0; k < ncol; ++k)
        ocol_get(s, k, &oc[k]);
    while (nvis < ncol && order[nvis] != INVALID)
        ++nvis;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1)
            exclusive_scan(n, counts, offs);
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
            const uint32_t i = (uint32_t)ii;
            if (s->kinds[i] != CN_PARTICLE_KIND_FLUID) {
                if (pass == 0)
                    counts[i] = 0;
                else {
                    ranges[2 * i] = offs[i];
                    ranges[2 * i + 1] = 0;
                }
                continue;
            }
            const float *p = s->pos + 4 * i;
            const uint32_t *md = s->meta + 4 * i;
            const float r = maxf(s->radii[i], 0.0f);
            const cn_fluid_material *m = &s->fluid_mats[s->fluid_mat[i]];
            const float support = maxf(m->smoothing_radius, r);
            const float reuse = maxf(maxf(m->cfl_radius, r), support);
            const float qr = support + reuse;
            float qmin[3], qmax[3];
            for (int a = 0; a < 3; ++a) {
                qmin[a] = p[a] - qr;
                qmax[a] = p[a] + qr;
            }
            uint32_t cnt = 0, w = pass ? offs[i] : 0;
            for (uint32_t k = 0; k < nvis; ++k) {
                const ocol *c = &oc[order[k]];
                const int overlap = qmin[0] <= c->mx[0] && qmax[0] >= c->mn[0] && qmin[1] <= c->mx[1] &&
                                    qmax[1] >= c->mn[1] && qmin[2] <= c->mx[2] && qmax[2] >= c->mn[2];
                if (!overlap)
                    continue;
                if (!(c->enabled && c->env == md[0] && (md[3] & c->layer) != 0u && (c->mask & md[2]) != 0u))
                    continue;
                if (pass == 1) {
                    if (w >= capacity)
                        break;
                    pairs[w++] = order[k];
                }
                ++cnt;
            }
            if (pass == 0)
                counts[i] = cnt;
            else {
                ranges[2 * i] = offs[i];
                ranges[2 * i + 1] = counts[i];
            }
        }
    }
    const uint32_t required = n ? offs[n - 1] + counts[n - 1] : 0;
    free(counts);
    free(offs);
    free(oc);
    return required;
}

/* ------------------------------------------------------------------------------------------
 * Fluid passes
 * ------------------------------------------------------------------------------------------ */

/* FluidSpikyKernel / FluidSpikyKernelDerivative / FluidCohesionKernel,
 * assets/shaders/include/physics/fluid/physics_fluid_common.hlsli:17-65 */
static float spiky_w(float d, float h)
{
    if (d >= h)
        return 0.0f;
    const float omq = 1.0f - d / h;
    const float h3 = h * h * h;
    const float k = 15.0f / (KERNEL_PI * h3);
    return k * omq * omq;
}
static float spiky_dw(float d, float h)
{
    if (d <= EPS || d >= h)
        return 0.0f;
    const float omq = 1.0f - d / h;
    const float h3 = h * h * h;
    const float k = 15.0f / (KERNEL_PI * h3);
    return -2.0f * k * omq / h;
}
static float cohesion_kernel(float d, float h, float c1, float c2)
{
    if (d <= EPS || d >= h)
        return 0.0f;
    const float q = d / h;
    return c1 * q * q * q + c2 * q * q - 1.0f;
}

/* ShouldProcessFluidNeighbor, physics_fluid_neighbor_build.hlsli:48-70 */
static int should_process(const cno_scene *s, uint32_t i, uint32_t env, uint32_t layer,
                          uint32_t mask, uint32_t j)
{
    if (j == i)
        return 0;
    if (s->kinds[j] != CN_PARTICLE_KIND_FLUID)
        return 0;
    const uint32_t *mj = s->meta + 4 * j;
    if (mj[0] != env)
        return 0;
    return (mask & mj[2]) != 0u && (mj[3] & layer) != 0u;
}

/* physics_fluid_build_neighbors.cs.hlsl:21-147.  required[i] counts in-radius neighbors before
 * the slot budget (the reference truncates silently, :143-146). */
void cno_fluid_neighbors(const cno_scene *s, float cell_size, const int32_t *cells,
                         const uint32_t *sk, const uint32_t *si, const uint32_t *tbl,
                         uint32_t capacity, const float *dacc, uint32_t it, uint32_t cap,
                         uint32_t *counts, uint32_t *list, uint32_t *required)
{
    (void)cell_size;
    (void)sk;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)s->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        if (s->kinds[i] != CN_PARTICLE_KIND_FLUID)
            continue;
        const float *pi = s->pos + 4 * i;
        if (pi[3] <= EPS)
            continue;
        if (cap == 0) {
            counts[i] = 0;
            if (required)
                required[i] = 0;
            continue;
        }
        float sp[3];
        for (int a = 0; a < 3; ++a)
            sp[a] = pi[a] + (it > 0 ? dacc[4 * i + a] : 0.0f);
        const float h = maxf(s->fluid_mats[s->fluid_mat[i]].smoothing_radius, 1.0e-4f);
        const uint32_t env = s->meta[4 * i], layer = s->meta[4 * i + 2], mask = s->meta[4 * i + 3];
        const int32_t *ci = cells + 3 * i;
        uint32_t cnt = 0, req = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int32_t tx = ci[0] + dx, ty = ci[1] + dy, tz = ci[2] + dz;
                    uint32_t start, end;
                    if (!cno_find_cell_range(tbl, capacity, cell_hash(tx, ty, tz), &start, &end))
                        continue;
                    for (uint32_t q = start; q < end; ++q) {
                        const uint32_t j = si[q];
                        const int32_t *cj = cells + 3 * j;
                        if (cj[0] != tx || cj[1] != ty || cj[2] != tz)
                            continue;
                        if (!should_process(s, i, env, layer, mask, j))
                            continue;
                        const float *pj = s->pos + 4 * j;
                        if (pj[3] <= EPS)
                            continue;
                        float d[3];
                        for (int a = 0; a < 3; ++a)
                            d[a] = sp[a] - (pj[a] + (it > 0 ? dacc[4 * j + a] : 0.0f));
                        if (len3(d) >= h)
                            continue;
                        ++req;
                        if (cnt >= cap)
                            continue;
                        list[(uint64_t)i * cap + cnt] = j;
                        ++cnt;
                    }
                }
        counts[i] = cnt;
        if (required)
            required[i] = req;
    }
}

/* ComputeFluidBoundaryGhostDelta, physics_fluid_boundary_common.hlsli:90-110 */
static int ghost_delta(const cno_scene *s, uint32_t collider, const float *p, float h, float *g)
{
    float sd, n[3];
    g[0] = g[1] = g[2] = 0.0f;
    if (!collider_contact_k(s, collider, p, &sd, n))
        return 0;
    if (sd < 0.0f || sd >= h)
        return 0;
    const float t = 2.0f * sd;
    g[0] = n[0] * t;
    g[1] = n[1] * t;
    g[2] = n[2] * t;
    return 1;
}

/* physics_fluid_compute_density_lambda.cs.hlsl:26-115 with
 * AccumulateFluidBoundaryDensityFromCandidates (physics_fluid_boundary_common.hlsli:112-138) */
void cno_fluid_density(const cno_scene *s, const float *cfg, const float *dacc, uint32_t it,
                       uint32_t cap, const uint32_t *counts, const uint32_t *list,
                       const uint32_t *br, const uint32_t *bp, float *nc)
{
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        float *o = nc + 4 * i;
        const float *pi = s->pos + 4 * i;
        if (s->kinds[i] != CN_PARTICLE_KIND_FLUID || pi[3] <= EPS) {
            o[0] = o[1] = o[2] = o[3] = 0.0f;
            continue;
        }
        float sp[3];
        for (int a = 0; a < 3; ++a)
            sp[a] = pi[a] + (it > 0 ? dacc[4 * i + a] : 0.0f);
        const cn_fluid_material *m = &s->fluid_mats[s->fluid_mat[i]];
        const float rho0 = maxf(m->rest_density, 1.0f);
        const float dcs = m->density_constraint_scale;
        const float h = maxf(m->smoothing_radius, 1.0e-4f);
        const float st = maxf(m->surface_tension, 0.0f);
        float density = 0.0f, sn[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t k = 0; k < counts[i]; ++k) {
            const uint32_t j = list[(uint64_t)i * cap + k];
            const float *pj = s->pos + 4 * j;
            float d[3];
            for (int a = 0; a < 3; ++a)
                d[a] = sp[a] - (pj[a] + (it > 0 ? dacc[4 * j + a] : 0.0f));
            const float dist = len3(d);
            if (dist >= h)
                continue;
            density += spiky_w(dist, h);
            if (st > EPS) {
                /* FluidSpikyKernelGradient, physics_fluid_common.hlsli:44-53 */
                float gr[3] = {0.0f, 0.0f, 0.0f};
                if (!(dist <= EPS || dist >= h)) {
                    const float dw = spiky_dw(dist, h);
                    for (int a = 0; a < 3; ++a)
                        gr[a] = dw * (d[a] / dist);
                }
                for (int a = 0; a < 3; ++a)
                    sn[a] += gr[a];
            }
        }
        for (uint32_t k = 0; k < br[2 * i + 1]; ++k) {
            float g[3];
            if (!ghost_delta(s, bp[br[2 * i] + k], sp, h, g))
                continue;
            const float dist = len3(g);
            if (dist >= h)
                continue;
            density += spiky_w(dist, h) * cfg[CFG_FLUID_BND_DENSITY];
        }
        const float C =
            maxf(density - rho0, -rho0 * maxf(cfg[CFG_FLUID_MAX_UNDER], 0.0f)) * dcs;
        if (st > EPS) {
            o[0] = sn[0] * st;
            o[1] = sn[1] * st;
            o[2] = sn[2] * st;
        } else {
            o[0] = o[1] = o[2] = 0.0f;
        }
        o[3] = C;
    }
}

/* physics_fluid_compute_delta_positions.cs.hlsl:27-157 with
 * ComputeFluidBoundaryDeltaFromCandidates (physics_fluid_boundary_common.hlsli:140-167) */
void cno_fluid_delta(const cno_scene *s, const float *cfg, float dt, const float *dacc,
                     uint32_t it, uint32_t cap, const uint32_t *counts, const uint32_t *list,
                     const uint32_t *br, const uint32_t *bp, const float *nc, float *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        float *o = out + 4 * i;
        const float *pi = s->pos + 4 * i;
        if (s->kinds[i] != CN_PARTICLE_KIND_FLUID || pi[3] <= EPS) {
            o[0] = o[1] = o[2] = o[3] = 0.0f;
            continue;
        }
        float sd_[3], sp[3];
        for (int a = 0; a < 3; ++a) {
            sd_[a] = it > 0 ? dacc[4 * i + a] : 0.0f;
            sp[a] = pi[a] + sd_[a];
        }
        const float Ci = nc[4 * i + 3];
        const uint32_t mi = s->fluid_mat[i];
        const cn_fluid_material *m = &s->fluid_mats[mi];
        const float h = maxf(m->smoothing_radius, 1.0e-4f);
        const float cohesion = dt * maxf(m->cohesion, 0.0f);
        const float viscosity = dt * maxf(m->viscosity, 0.0f);
        const float st = maxf(m->surface_tension, 0.0f);
        const float cfl = maxf(m->cfl_radius, 0.0f);
        float dp[3] = {0.0f, 0.0f, 0.0f}, wsum = 0.0f;
        float sn[3] = {0.0f, 0.0f, 0.0f};
        if (st > EPS)
            for (int a = 0; a < 3; ++a)
                sn[a] = nc[4 * i + a];
        for (uint32_t k = 0; k < counts[i]; ++k) {
            const uint32_t j = list[(uint64_t)i * cap + k];
            const float *pj = s->pos + 4 * j;
            if (pj[3] <= EPS)
                continue;
            float nd[3], d[3];
            for (int a = 0; a < 3; ++a) {
                nd[a] = it > 0 ? dacc[4 * j + a] : 0.0f;
                d[a] = sp[a] - (pj[a] + nd[a]);
            }
            const float dist = len3(d);
            if (dist >= h || dist <= EPS)
                continue;
            const float Cj = nc[4 * j + 3];
            float nij[3], pd[3];
            for (int a = 0; a < 3; ++a)
                nij[a] = d[a] / dist;
            const float der = spiky_dw(dist, h);
            const float sc = -0.5f * (Ci + Cj) * der;
            for (int a = 0; a < 3; ++a)
                pd[a] = sc * nij[a];
            if (mi == s->fluid_mat[j]) {
                float rel[3];
                for (int a = 0; a < 3; ++a)
                    rel[a] = sd_[a] - nd[a];
                if (cohesion > EPS) {
                    const float ck = cohesion * cohesion_kernel(dist, h, m->cohesion1, m->cohesion2);
                    for (int a = 0; a < 3; ++a)
                        pd[a] -= nij[a] * ck;
                }
                if (viscosity > EPS) {
                    const float vs = 1.0f - (1.0f / (1.0f + viscosity * spiky_w(dist, h)));
                    for (int a = 0; a < 3; ++a)
                        pd[a] -= vs * rel[a];
                }
                if (cfl > EPS) {
                    const float proj = dot3(rel, nij);
                    if (proj < -cfl) {
                        const float pc = proj + cfl;
                        for (int a = 0; a < 3; ++a)
                            pd[a] -= nij[a] * pc * 0.5f;
                    }
                }
                if (st > EPS) {
                    for (int a = 0; a < 3; ++a)
                        pd[a] -= dt * (sn[a] - nc[4 * j + a]);
                }
            }
            for (int a = 0; a < 3; ++a)
                dp[a] += pd[a];
            wsum += 1.0f;
        }
        float bd[3] = {0.0f, 0.0f, 0.0f};
        const float sc = maxf(Ci, 0.0f);
        for (uint32_t k = 0; k < br[2 * i + 1]; ++k) {
            float g[3];
            if (!ghost_delta(s, bp[br[2 * i] + k], sp, h, g))
                continue;
            const float dist = len3(g);
            if (dist <= EPS || dist >= h)
                continue;
            const float der = spiky_dw(dist, h);
            const float f = -sc * der;
            for (int a = 0; a < 3; ++a)
                bd[a] += f * (g[a] / dist);
        }
        for (int a = 0; a < 3; ++a)
            o[a] = dp[a] * cfg[CFG_FLUID_SOLVE] + bd[a] * cfg[CFG_FLUID_BND_DELTA];
        o[3] = wsum;
    }
}

/* physics_fluid_apply_delta_positions.cs.hlsl:11-31 */
void cno_fluid_apply(const cno_scene *s, const float *cfg, uint32_t it, const float *din,
                     float *dacc)
{
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        if (s->kinds[i] != CN_PARTICLE_KIND_FLUID)
            continue;
        const float *d = din + 4 * i;
        const float w = maxf(d[3] * cfg[CFG_FLUID_RELAX], 1.0f);
        float *acc = dacc + 4 * i;
        for (int a = 0; a < 3; ++a)
            acc[a] = (it > 0 ? acc[a] : 0.0f) + d[a] / w;
        acc[3] = 0.0f;
    }
}

/* physics_fluid_clamp_boundary.cs.hlsl:23-80 */
void cno_fluid_clamp(const cno_scene *s, const float *din, const uint32_t *br, const uint32_t *bp,
                     float *dout)
{
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        if (s->kinds[i] != CN_PARTICLE_KIND_FLUID)
            continue;
        const float *b = s->pos + 4 * i;
        const float r = maxf(s->radii[i], 0.0f);
        float cur[3];
        for (int a = 0; a < 3; ++a)
            cur[a] = b[a] + din[4 * i + a];
        for (int sweep = 0; sweep < 2; ++sweep) {
            int changed = 0;
            for (uint32_t k = 0; k < br[2 * i + 1]; ++k) {
                float sd, n[3];
                if (!collider_contact_k(s, bp[br[2 * i] + k], cur, &sd, n))
                    continue;
                if (sd >= r)
                    continue;
                const float push = r - sd;
                for (int a = 0; a < 3; ++a)
                    cur[a] += n[a] * push;
                changed = 1;
            }
            if (!changed)
                break;
        }
        for (int a = 0; a < 3; ++a)
            dout[4 * i + a] = cur[a] - b[a];
        dout[4 * i + 3] = 0.0f;
    }
}

/* physics_particle_update_velocities.cs.hlsl:13-44 */
void cno_update_velocities(cno_scene *s, float dt, float *dacc)
{
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        float *p = s->pos + 4 * i;
        float *v = s->vel + 4 * i;
        if (p[3] <= EPS || dt <= EPS) {
            v[0] = v[1] = v[2] = v[3] = 0.0f;
            continue;
        }
        const float *pv = s->prev + 4 * i;
        float x[3] = {p[0], p[1], p[2]};
        if (s->kinds[i] == CN_PARTICLE_KIND_FLUID) {
            for (int a = 0; a < 3; ++a)
                x[a] = x[a] + dacc[4 * i + a];
            p[0] = x[0];
            p[1] = x[1];
            p[2] = x[2];
            dacc[4 * i] = dacc[4 * i + 1] = dacc[4 * i + 2] = dacc[4 * i + 3] = 0.0f;
        }
        const float damping = maxf(s->part_mats[4 * s->part_mat[i] + 2], 0.0f);
        const float scale = maxf(1.0f - damping * dt, 0.0f);
        for (int a = 0; a < 3; ++a)
            v[a] = ((x[a] - pv[a]) / dt) * scale;
        v[3] = 0.0f;
    }
}

/* ComputePositionFrictionDelta, physics_rigid_solver_shared.hlsli:50-67 */
static void friction_delta(const float *cfg, const float *disp, float pen, float kin, float stat,
                           float *out)
{
    const float td = len3(disp);
    out[0] = out[1] = out[2] = 0.0f;
    if (td <= cfg[CFG_POS_FRICTION_MIN] || pen <= EPS)
        return;
    const float sl = saturatef(stat) * pen;
    if (td <= sl) {
        out[0] = disp[0];
        out[1] = disp[1];
        out[2] = disp[2];
        return;
    }
    const float ks = minf(saturatef(kin) * pen / td, 1.0f);
    for (int a = 0; a < 3; ++a)
        out[a] = disp[a] * ks;
}

/* physics_fluid_project_boundary_velocities.cs.hlsl:26-118 with CombineContactMaterial and
 * ProjectOntoContactTangent (physics_rigid_solver_shared.hlsli:37-48) */
void cno_project_boundary_velocities(cno_scene *s, const float *cfg, float dt, const uint32_t *br,
                                     const uint32_t *bp)
{
    const float slop = cfg[CFG_CONTACT_SLOP];
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        if (s->kinds[i] != CN_PARTICLE_KIND_FLUID)
            continue;
        const float *p = s->pos + 4 * i;
        const float r = maxf(s->radii[i], 0.0f);
        float *vv = s->vel + 4 * i;
        float v[3] = {vv[0], vv[1], vv[2]};
        const float *pm = s->part_mats + 4 * s->part_mat[i];
        const float damping = maxf(pm[2], 0.0f);
        float best = 3.402823466e+38f, bn[3] = {0.0f, 1.0f, 0.0f}, bm[3] = {0.0f, 0.0f, 0.0f};
        int found = 0;
        for (uint32_t k = 0; k < br[2 * i + 1]; ++k) {
            const uint32_t ci = bp[br[2 * i] + k];
            float sd, n[3];
            ocol co;
            ocol_get(s, ci, &co);
            if (!collider_contact(&co, p, &sd, n))
                continue;
            const float cd = r + 2.0f * slop;
            if (sd > cd)
                continue;
            if (sd < best) {
                best = sd;
                bn[0] = n[0];
                bn[1] = n[1];
                bn[2] = n[2];
                const float *cm = co.material;
                bm[0] = sqrtf(maxf(0.0f, pm[0]) * maxf(0.0f, cm[0]));
                bm[1] = maxf(maxf(0.0f, pm[1]), maxf(0.0f, cm[1]));
                bm[2] = sqrtf(maxf(0.0f, pm[3]) * maxf(0.0f, cm[3]));
                found = 1;
            }
        }
        if (!found)
            continue;
        const float pen = maxf(r - best, 0.0f);
        const float cd = maxf(r + 2.0f * slop, 1.0e-5f);
        const float cw = saturatef((cd - best) / cd);
        float rv[3] = {v[0], v[1], v[2]};
        const float nv = dot3(rv, bn);
        if (nv < 0.0f)
            for (int a = 0; a < 3; ++a)
                rv[a] -= bn[a] * nv;
        float tv[3];
        const float rn = dot3(rv, bn);
        for (int a = 0; a < 3; ++a)
            tv[a] = rv[a] - bn[a] * rn;
        if (dt > EPS) {
            float disp[3], fd[3];
            for (int a = 0; a < 3; ++a)
                disp[a] = tv[a] * dt;
            friction_delta(cfg, disp, maxf(pen, slop), bm[0], bm[2], fd);
            for (int a = 0; a < 3; ++a)
                tv[a] -= fd[a] / dt;
        }
        const float tds = maxf(1.0f - damping * cw * dt, 0.0f);
        for (int a = 0; a < 3; ++a)
            tv[a] *= tds;
        const float rn2 = maxf(dot3(rv, bn), 0.0f);
        for (int a = 0; a < 3; ++a)
            vv[a] = bn[a] * rn2 + tv[a];
        vv[3] = 0.0f;
    }
}

/* ------------------------------------------------------------------------------------------
 * XPBD soft constraints
 * ------------------------------------------------------------------------------------------ */

/* Cephes single-precision atanf / atan2f / asinf / acosf (S. L. Moshier, Cephes Math Library),
 * written out with one rounding per operation. */
#define CEPHES_PIF 3.141592653589793238f
#define CEPHES_PIO2F 1.5707963267948966192f
#define CEPHES_PIO4F 0.7853981633974483096f
static float cephes_atanf(float xx)
{
    float x = xx, y, z;
    int neg = 0;
    if (x < 0.0f) {
        neg = 1;
        x = -x;
    }
    if (x > 2.414213562373095f) {
        y = CEPHES_PIO2F;
        x = -(1.0f / x);
    } else if (x > 0.4142135623730950f) {
        y = CEPHES_PIO4F;
        x = (x - 1.0f) / (x + 1.0f);
    } else {
        y = 0.0f;
    }
    z = x * x;
    const float p = (((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z -
                     3.33329491539e-1f) * z * x + x;
    y = y + p;
    return neg ? -y : y;
}

float cno_atan2f(float y, float x)
{
    int code = 0;
    if (x < 0.0f)
        code = 2;
    if (y < 0.0f)
        code |= 1;
    if (x == 0.0f) {
        if (code & 1)
            return -CEPHES_PIO2F;
        if (y == 0.0f)
            return 0.0f;
        return CEPHES_PIO2F;
    }
    if (y == 0.0f)
        return (code & 2) ? CEPHES_PIF : 0.0f;
    float w = 0.0f;
    if (code == 2)
        w = CEPHES_PIF;
    else if (code == 3)
        w = -CEPHES_PIF;
    return w + cephes_atanf(y / x);
}

static float cephes_asinf(float xx)
{
    float a = fabsf(xx), x, z;
    if (a > 1.0f)
        return NAN;
    if (a < 1.0e-4f)
        return xx;
    int flag = 0;
    if (a > 0.5f) {
        z = 0.5f * (1.0f - a);
        x = sqrtf(z);
        flag = 1;
    } else {
        x = a;
        z = x * x;
    }
    z = ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z +
         1.6666752422e-1f) * z * x + x;
    if (flag) {
        z = z + z;
        z = CEPHES_PIO2F - z;
    }
    return xx < 0.0f ? -z : z;
}

float cno_acosf(float x)
{
    if (x < -1.0f || x > 1.0f)
        return NAN;
    if (x < -0.5f)
        return CEPHES_PIF - 2.0f * cephes_asinf(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f)
        return 2.0f * cephes_asinf(sqrtf(0.5f * (1.0f - x)));
    return CEPHES_PIO2F - cephes_asinf(x);
}

static void cross3(const float *a, const float *b, float *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static void sub3(const float *a, const float *b, float *o)
{
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}

/* FluidSpikyKernelGradient, physics_fluid_common.hlsli:45-54 */
static void spiky_grad(const float *d, float dist, float h, float *g)
{
    g[0] = g[1] = g[2] = 0.0f;
    if (dist <= EPS || dist >= h)
        return;
    const float dw = spiky_dw(dist, h);
    for (int a = 0; a < 3; ++a)
        g[a] = dw * (d[a] / dist);
}

/* physics_fluid_compute_vorticity.cs.hlsl:16-81 over the post-update neighbor lists: per fluid
 * particle of a confining material, curl = sum over same-material neighbors of
 * (v_j - v_i) x gradW(x_i - x_j), stored as (curl, |curl|); zero for every other particle. */
void cno_fluid_vorticity(const cno_scene *s, uint32_t cap, const uint32_t *counts, const uint32_t *list,
                         float *vort)
{
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        float *o = vort + 4 * (size_t)i;
        o[0] = o[1] = o[2] = o[3] = 0.0f;
        if (s->kinds[i] != CN_PARTICLE_KIND_FLUID)
            continue;
        const cn_fluid_material *m = &s->fluid_mats[s->fluid_mat[i]];
        if (maxf(m->vorticity_confinement, 0.0f) <= EPS)
            continue;
        const float *pi = s->pos + 4 * (size_t)i;
        if (pi[3] <= EPS)
            continue;
        const float *vi = s->vel + 4 * (size_t)i;
        const float h = maxf(m->smoothing_radius, 1.0e-4f);
        float curl[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t k = 0; k < counts[i]; ++k) {
            const uint32_t j = list[(uint64_t)i * cap + k];
            if (s->fluid_mat[j] != s->fluid_mat[i])
                continue;
            const float *pj = s->pos + 4 * (size_t)j;
            float d[3], g[3], w[3], c[3];
            sub3(pi, pj, d);
            const float dist = len3(d);
            if (dist <= EPS || dist >= h)
                continue;
            sub3(s->vel + 4 * (size_t)j, vi, w);
            spiky_grad(d, dist, h, g);
            cross3(w, g, c);
            for (int a = 0; a < 3; ++a)
                curl[a] = curl[a] + c[a];
        }
        o[0] = curl[0];
        o[1] = curl[1];
        o[2] = curl[2];
        o[3] = len3(curl);
    }
}

/* physics_fluid_apply_vorticity_confinement.cs.hlsl:16-94: eta = sum over same-material
 * neighbors of |curl_j| gradW(x_i - x_j); v_i += dt * confinement * ((eta / |eta|) x curl_i),
 * with the velocity's w written as zero. */
void cno_fluid_vorticity_apply(cno_scene *s, float dt, uint32_t cap, const uint32_t *counts,
                               const uint32_t *list, const float *vort)
{
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        if (s->kinds[i] != CN_PARTICLE_KIND_FLUID)
            continue;
        const cn_fluid_material *m = &s->fluid_mats[s->fluid_mat[i]];
        const float conf = maxf(m->vorticity_confinement, 0.0f);
        if (conf <= EPS)
            continue;
        const float *pi = s->pos + 4 * (size_t)i;
        if (pi[3] <= EPS)
            continue;
        const float *ci = vort + 4 * (size_t)i;
        if (ci[3] <= EPS)
            continue;
        const float h = maxf(m->smoothing_radius, 1.0e-4f);
        float eta[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t k = 0; k < counts[i]; ++k) {
            const uint32_t j = list[(uint64_t)i * cap + k];
            if (s->fluid_mat[j] != s->fluid_mat[i])
                continue;
            float d[3], g[3];
            sub3(pi, s->pos + 4 * (size_t)j, d);
            const float dist = len3(d);
            if (dist <= EPS || dist >= h)
                continue;
            spiky_grad(d, dist, h, g);
            const float mj = vort[4 * (size_t)j + 3];
            for (int a = 0; a < 3; ++a)
                eta[a] = eta[a] + mj * g[a];
        }
        const float gl = len3(eta);
        if (gl <= EPS)
            continue;
        float nrm[3], f[3];
        for (int a = 0; a < 3; ++a)
            nrm[a] = eta[a] / gl;
        cross3(nrm, ci, f);
        float *v = s->vel + 4 * (size_t)i;
        for (int a = 0; a < 3; ++a)
            v[a] = v[a] + dt * (conf * f[a]);
        v[3] = 0.0f;
    }
}

/* FluidAnisotropyJacobiRotate, physics_fluid_anisotropy_common.hlsli:28-74 (rsqrt read as
 * 1/sqrt) */
static void jacobi_rotate(float A[3][3], float R[3][3], int p, int q)
{
    if (fabsf(A[p][q]) <= 1.0e-20f)
        return;
    const float d = (A[p][p] - A[q][q]) / (2.0f * A[p][q]);
    float t = 1.0f / (fabsf(d) + sqrtf(d * d + 1.0f));
    if (d < 0.0f)
        t = -t;
    const float c = 1.0f / sqrtf(t * t + 1.0f);
    const float sn = t * c;
    A[p][p] = A[p][p] + t * A[p][q];
    A[q][q] = A[q][q] - t * A[p][q];
    A[p][q] = 0.0f;
    A[q][p] = 0.0f;
    for (int k = 0; k < 3; ++k) {
        if (k == p || k == q)
            continue;
        const float akp = c * A[k][p] + sn * A[k][q];
        const float akq = -sn * A[k][p] + c * A[k][q];
        A[k][p] = akp;
        A[p][k] = akp;
        A[k][q] = akq;
        A[q][k] = akq;
    }
    for (int k = 0; k < 3; ++k) {
        const float rkp = c * R[k][p] + sn * R[k][q];
        const float rkq = -sn * R[k][p] + c * R[k][q];
        R[k][p] = rkp;
        R[k][q] = rkq;
    }
}

/* FluidAnisotropyEigenDecomposition, physics_fluid_anisotropy_common.hlsli:76-122: four
 * largest-pivot Jacobi rotations */
static void jacobi_eigen(float A[3][3], float R[3][3], float ev[3])
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            R[r][c] = r == c ? 1.0f : 0.0f;
    for (int it = 0; it < 4; ++it) {
        int pair = 0;
        float mx = fabsf(A[0][1]);
        const float a02 = fabsf(A[0][2]);
        if (a02 > mx) {
            pair = 1;
            mx = a02;
        }
        const float a12 = fabsf(A[1][2]);
        if (a12 > mx) {
            pair = 2;
            mx = a12;
        }
        if (mx < 1.0e-15f)
            break;
        if (pair == 0)
            jacobi_rotate(A, R, 0, 1);
        else if (pair == 1)
            jacobi_rotate(A, R, 0, 2);
        else
            jacobi_rotate(A, R, 1, 2);
    }
    for (int k = 0; k < 3; ++k)
        ev[k] = maxf(A[k][k], 0.0f);
}

/* SafeNormalize of physics_fluid_build_render_anisotropy.cs.hlsl:21-25 (rsqrt read as 1/sqrt) */
static void aniso_normalize(const float *v, int axis, float *o)
{
    const float l2 = dot3(v, v);
    if (l2 > 1.0e-8f) {
        const float r = 1.0f / sqrtf(l2);
        for (int a = 0; a < 3; ++a)
            o[a] = v[a] * r;
    } else {
        for (int a = 0; a < 3; ++a)
            o[a] = a == axis ? 1.0f : 0.0f;
    }
}

static void write_aniso(float *out, size_t n, size_t i, const float ax[3][3], const float *scale)
{
    for (int k = 0; k < 3; ++k) {
        float *o = out + 4 * (k * n + i);
        o[0] = ax[k][0];
        o[1] = ax[k][1];
        o[2] = ax[k][2];
        o[3] = scale[k];
    }
}

/* physics_fluid_build_render_anisotropy.cs.hlsl:34-171: per fluid particle, the weighted
 * covariance of its listed fluid neighbors (weight 1 - (r / support)^3), its eigenvectors by
 * four Jacobi rotations, sorted by decreasing eigenvalue, scales 1.1 sqrt(eigenvalue) clamped
 * to [radius, 2 radius]; other particles and particles with no weight get the isotropic frame.
 * A pinned fluid particle (inverse mass <= epsilon) has no post-update list: the reference reads
 * a count its neighbor build never wrote there; here its list is empty. */
void cno_fluid_render_anisotropy(const cno_scene *s, uint32_t cap, const uint32_t *counts,
                                 const uint32_t *list, float *out)
{
    const size_t n = s->n;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const float radius = maxf(s->radii[i], 1.0e-4f);
        float iso_ax[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
        if (s->kinds[i] != CN_PARTICLE_KIND_FLUID) {
            const float sc[3] = {radius, radius, radius};
            write_aniso(out, n, i, iso_ax, sc);
            continue;
        }
        const cn_fluid_material *m = &s->fluid_mats[s->fluid_mat[i]];
        const float support = maxf(m->smoothing_radius, radius);
        const float inv = 1.0f / support;
        const float iso = maxf(1.0f * radius, radius);
        const float *pi = s->pos + 4 * (size_t)i;
        const uint32_t cnt = (pi[3] <= EPS || cap == 0) ? 0u : counts[i];
        float mean[3] = {0.0f, 0.0f, 0.0f}, wsum = 0.0f;
        for (uint32_t k = 0; k < cnt; ++k) {
            const uint32_t j = list[(uint64_t)i * cap + k];
            if (s->kinds[j] != CN_PARTICLE_KIND_FLUID)
                continue;
            const float *pj = s->pos + 4 * (size_t)j;
            float d[3];
            sub3(pi, pj, d);
            const float d2 = dot3(d, d);
            if (d2 <= 0.0f)
                continue;
            const float dist = sqrtf(d2);
            if (dist >= support)
                continue;
            const float x = dist * inv;
            const float w = 1.0f - x * x * x;
            wsum = wsum + w;
            for (int a = 0; a < 3; ++a)
                mean[a] = mean[a] + pj[a] * w;
        }
        if (wsum <= 0.0f) {
            const float sc[3] = {iso, iso, iso};
            write_aniso(out, n, i, iso_ax, sc);
            continue;
        }
        for (int a = 0; a < 3; ++a)
            mean[a] = mean[a] / wsum;
        float C[3][3] = {{0.0f}};
        for (uint32_t k = 0; k < cnt; ++k) {
            const uint32_t j = list[(uint64_t)i * cap + k];
            if (s->kinds[j] != CN_PARTICLE_KIND_FLUID)
                continue;
            const float *pj = s->pos + 4 * (size_t)j;
            float d[3];
            sub3(pi, pj, d);
            const float d2 = dot3(d, d);
            if (d2 <= 0.0f)
                continue;
            const float dist = sqrtf(d2);
            if (dist >= support)
                continue;
            const float x = dist * inv;
            const float w = 1.0f - x * x * x;
            float cc[3];
            sub3(pj, mean, cc);
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    C[r][c] = C[r][c] + w * cc[r] * cc[c];
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                C[r][c] = C[r][c] / wsum;
        C[0][1] = C[1][0] = 0.5f * (C[0][1] + C[1][0]);
        C[0][2] = C[2][0] = 0.5f * (C[0][2] + C[2][0]);
        C[1][2] = C[2][1] = 0.5f * (C[1][2] + C[2][1]);
        float R[3][3], ev[3], ax[3][3];
        jacobi_eigen(C, R, ev);
        for (int k = 0; k < 3; ++k) {
            const float col[3] = {R[0][k], R[1][k], R[2][k]};
            aniso_normalize(col, k, ax[k]);
        }
        /* FluidAnisotropySortDescending, physics_fluid_anisotropy_common.hlsli:138-156 */
        static const int order[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (int o = 0; o < 3; ++o) {
            const int a = order[o][0], b = order[o][1];
            if (ev[a] < ev[b]) {
                const float t = ev[a];
                ev[a] = ev[b];
                ev[b] = t;
                for (int c = 0; c < 3; ++c) {
                    const float u = ax[a][c];
                    ax[a][c] = ax[b][c];
                    ax[b][c] = u;
                }
            }
        }
        float sc[3];
        for (int k = 0; k < 3; ++k) {
            const float l = sqrtf(ev[k]) * 1.1f;
            sc[k] = minf(maxf(l, 1.0f * radius), 2.0f * radius);
        }
        write_aniso(out, n, i, ax, sc);
    }
}

/* Particle -> incident constraint lists (physics_scene_gpu_state.cpp:3273-3360): per particle the
 * references in ascending constraint index; ref = (constraint << 2) | slot.  Bend corrections are
 * indexed bends first, then dihedrals at softBendCount + dihedralIndex. */
typedef struct csr {
    uint32_t *start;
    uint32_t *ref;
} csr;

static void csr_build(uint32_t n, uint32_t nc, uint32_t arity, const uint32_t *(*get)(const void *, uint32_t),
                      const void *arr, uint32_t index_offset, csr *out, uint32_t *count)
{
    for (uint32_t c = 0; c < nc; ++c) {
        const uint32_t *p = get(arr, c);
        for (uint32_t k = 0; k < arity; ++k)
            if (p[k] < n)
                ++count[p[k]];
    }
    (void)index_offset;
    (void)out;
}

static const uint32_t *edge_parts(const void *a, uint32_t c) { return &((const cn_soft_edge *)a)[c].a; }
static const uint32_t *bend_parts(const void *a, uint32_t c) { return &((const cn_soft_bend *)a)[c].p0; }
static const uint32_t *dih_parts(const void *a, uint32_t c) { return &((const cn_cloth_dihedral *)a)[c].e0; }
static const uint32_t *tet_parts(const void *a, uint32_t c) { return ((const cn_soft_tet *)a)[c].p; }

static void csr_fill(uint32_t n, uint32_t nc, uint32_t arity, const uint32_t *(*get)(const void *, uint32_t),
                     const void *arr, uint32_t index_offset, csr *out, uint32_t *cursor)
{
    for (uint32_t c = 0; c < nc; ++c) {
        const uint32_t *p = get(arr, c);
        for (uint32_t k = 0; k < arity; ++k)
            if (p[k] < n)
                out->ref[cursor[p[k]]++] = ((c + index_offset) << 2) | k;
    }
}

static void csr_make(uint32_t n, csr *out, uint32_t groups, const uint32_t *ncs, const uint32_t *arities,
                     const uint32_t *(**gets)(const void *, uint32_t), const void **arrs, const uint32_t *offsets)
{
    uint32_t *count = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
    for (uint32_t g = 0; g < groups; ++g)
        csr_build(n, ncs[g], arities[g], gets[g], arrs[g], offsets[g], out, count);
    out->start = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    uint32_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        out->start[i] = acc;
        acc += count[i];
    }
    out->start[n] = acc;
    out->ref = (uint32_t *)malloc(sizeof(uint32_t) * (acc ? acc : 1));
    for (uint32_t i = 0; i < n; ++i)
        count[i] = out->start[i];
    for (uint32_t g = 0; g < groups; ++g)
        csr_fill(n, ncs[g], arities[g], gets[g], arrs[g], offsets[g], out, count);
    free(count);
}

static void csr_free(csr *c)
{
    free(c->start);
    free(c->ref);
}


/* physics_soft_solve_edge_constraints.cs.hlsl:10-74 */
static void solve_edges(cno_scene *s, const float *cfg, float dt, float *lam)
{
    const float relax = cfg[CFG_SOFT_RELAX];
#pragma omp parallel for schedule(static)
    for (int64_t ee = 0; ee < (int64_t)s->n_edges; ++ee) {
        const uint32_t e = (uint32_t)ee;
        const cn_soft_edge *E = &s->edges[e];
        float *out = s->edge_corr + 8 * e;
        memset(out, 0, 8 * sizeof(float));
        if (E->enabled == 0u)
            continue;
        const float *pa = s->pos + 4 * E->a, *pb = s->pos + 4 * E->b;
        const float wa = pa[3], wb = pb[3], wsum = wa + wb;
        if (wsum <= EPS)
            continue;
        float d[3];
        sub3(pb, pa, d);
        const float l2 = dot3(d, d);
        if (l2 <= EPS)
            continue;
        const float len = sqrtf(l2);
        float ga[3], gb[3];
        for (int k = 0; k < 3; ++k) {
            ga[k] = -d[k] / len;
            gb[k] = -ga[k];
        }
        const float comp = maxf(E->compliance, 0.0f);
        const float alpha = comp / maxf(dt * dt, EPS);
        const float C = len - E->rest_length;
        const float lambda = lam[e];
        const float den = wsum + alpha;
        if (den <= EPS)
            continue;
        const float dl = -(C + alpha * lambda) / den;
        lam[e] = lambda + dl;
        for (int k = 0; k < 3; ++k) {
            out[k] = wa * dl * ga[k] * relax;
            out[4 + k] = wb * dl * gb[k] * relax;
        }
    }
}

static void store_zero16(float *o) { memset(o, 0, 16 * sizeof(float)); }

/* physics_soft_solve_bend_constraints.cs.hlsl:20-86 */
static void solve_bends(cno_scene *s, const float *cfg, float dt, float *lam)
{
    const float relax = cfg[CFG_SOFT_RELAX];
#pragma omp parallel for schedule(static)
    for (int64_t bb = 0; bb < (int64_t)s->n_bends; ++bb) {
        const uint32_t b = (uint32_t)bb;
        const cn_soft_bend *B = &s->bends[b];
        float *out = s->bend_corr + 16 * b;
        const float *p0 = s->pos + 4 * B->p0, *p1 = s->pos + 4 * B->p1, *p2 = s->pos + 4 * B->p2;
        const float w0 = p0[3], w1 = p1[3], w2 = p2[3];
        if (w0 + w1 + w2 <= EPS) {
            store_zero16(out);
            continue;
        }
        float e0[3], e1[3];
        sub3(p0, p1, e0);
        sub3(p2, p1, e1);
        const float l0s = dot3(e0, e0), l1s = dot3(e1, e1);
        if (l0s <= EPS || l1s <= EPS) {
            store_zero16(out);
            continue;
        }
        const float l0 = sqrtf(l0s), l1 = sqrtf(l1s);
        float d0[3], d1[3];
        for (int k = 0; k < 3; ++k) {
            d0[k] = e0[k] / l0;
            d1[k] = e1[k] / l1;
        }
        const float ct = clampf(dot3(d0, d1), -1.0f + 1.0e-4f, 1.0f - 1.0e-4f);
        const float st2 = maxf(1.0f - ct * ct, 1.0e-8f);
        const float st = sqrtf(st2);
        const float theta = cno_acosf(ct);
        const float q0 = maxf(l0 * st, EPS), q2 = maxf(l1 * st, EPS);
        float g0[3], g1[3], g2[3];
        for (int k = 0; k < 3; ++k) {
            g0[k] = -(d1[k] - ct * d0[k]) / q0;
            g2[k] = -(d0[k] - ct * d1[k]) / q2;
            g1[k] = -(g0[k] + g2[k]);
        }
        const float den = w0 * dot3(g0, g0) + w1 * dot3(g1, g1) + w2 * dot3(g2, g2);
        const float alpha = maxf(B->compliance, 0.0f) / maxf(dt * dt, EPS);
        if (den + alpha <= EPS) {
            store_zero16(out);
            continue;
        }
        const float C = theta - B->rest_angle;
        const float lambda = lam[b];
        const float dl = -(C + alpha * lambda) / (den + alpha);
        lam[b] = lambda + dl;
        for (int k = 0; k < 3; ++k) {
            out[k] = w0 * dl * g0[k] * relax;
            out[4 + k] = w1 * dl * g1[k] * relax;
            out[8 + k] = w2 * dl * g2[k] * relax;
            out[12 + k] = 0.0f;
        }
        out[3] = out[7] = out[11] = out[15] = 0.0f;
    }
}

/* physics_cloth_solve_dihedral_constraints.cs.hlsl:20-95 */
static void solve_dihedrals(cno_scene *s, const float *cfg, float dt, float *lam)
{
    const float relax = cfg[CFG_SOFT_RELAX];
    const float two_pi = 6.283185307179586f, pi = 3.141592653589793f;
#pragma omp parallel for schedule(static)
    for (int64_t dd = 0; dd < (int64_t)s->n_dihedrals; ++dd) {
        const uint32_t d = (uint32_t)dd;
        const uint32_t ci = s->n_bends + d;
        const cn_cloth_dihedral *D = &s->dihedrals[d];
        float *out = s->bend_corr + 16 * ci;
        const float *e0m = s->pos + 4 * D->e0, *e1m = s->pos + 4 * D->e1;
        const float *o0m = s->pos + 4 * D->o0, *o1m = s->pos + 4 * D->o1;
        const float *p0 = o0m, *p1 = o1m, *p2 = e0m, *p3 = e1m;
        float edge[3], a[3], b[3], rn0[3], rn1[3];
        sub3(p3, p2, edge);
        const float el2 = dot3(edge, edge);
        sub3(p2, p0, a);
        sub3(p3, p0, b);
        cross3(a, b, rn0);
        sub3(p3, p1, a);
        sub3(p2, p1, b);
        cross3(a, b, rn1);
        const float n0s = dot3(rn0, rn0), n1s = dot3(rn1, rn1);
        if (el2 <= EPS || n0s <= EPS || n1s <= EPS) {
            store_zero16(out);
            continue;
        }
        const float el = sqrtf(el2);
        const float iel = 1.0f / el;
        float sn0[3], sn1[3], dop0[3], dop1[3], de0[3], de1[3];
        for (int k = 0; k < 3; ++k) {
            sn0[k] = rn0[k] / n0s;
            sn1[k] = rn1[k] / n1s;
            dop0[k] = el * sn0[k];
            dop1[k] = el * sn1[k];
        }
        float t0[3], t1[3];
        sub3(p0, p3, t0);
        sub3(p1, p3, t1);
        const float f00 = dot3(t0, edge) * iel, f01 = dot3(t1, edge) * iel;
        sub3(p2, p0, t0);
        sub3(p2, p1, t1);
        const float f10 = dot3(t0, edge) * iel, f11 = dot3(t1, edge) * iel;
        for (int k = 0; k < 3; ++k) {
            de0[k] = f00 * sn0[k] + f01 * sn1[k];
            de1[k] = f10 * sn0[k] + f11 * sn1[k];
        }
        float g0[3], g1[3], g2[3], g3[3];
        for (int k = 0; k < 3; ++k) {
            g0[k] = -de0[k];
            g1[k] = -de1[k];
            g2[k] = -dop0[k];
            g3[k] = -dop1[k];
        }
        const float r0 = sqrtf(n0s), r1 = sqrtf(n1s);
        float n0[3], n1[3], eu[3], cr[3];
        for (int k = 0; k < 3; ++k) {
            n0[k] = rn0[k] / r0;
            n1[k] = rn1[k] / r1;
            eu[k] = edge[k] * iel;
        }
        cross3(n0, n1, cr);
        const float theta = cno_atan2f(dot3(eu, cr), clampf(dot3(n0, n1), -1.0f, 1.0f));
        float C = theta - D->rest_angle;
        C = C - two_pi * floorf((C + pi) / two_pi);
        const float w0 = e0m[3], w1 = e1m[3], w2 = o0m[3], w3 = o1m[3];
        const float den = w0 * dot3(g0, g0) + w1 * dot3(g1, g1) + w2 * dot3(g2, g2) + w3 * dot3(g3, g3);
        const float alpha = maxf(D->compliance, 0.0f) / maxf(dt * dt, EPS);
        if (den + alpha <= EPS) {
            store_zero16(out);
            continue;
        }
        const float lambda = lam[ci];
        const float dl = -(C + alpha * lambda) / (den + alpha);
        lam[ci] = lambda + dl;
        for (int k = 0; k < 3; ++k) {
            out[k] = w0 * dl * g0[k] * relax;
            out[4 + k] = w1 * dl * g1[k] * relax;
            out[8 + k] = w2 * dl * g2[k] * relax;
            out[12 + k] = w3 * dl * g3[k] * relax;
        }
        out[3] = out[7] = out[11] = out[15] = 0.0f;
    }
}

/* physics_soft_solve_tet_constraints.cs.hlsl:15-87; an early return leaves the record as it was */
static void solve_tets(cno_scene *s, const float *cfg, float dt, float *lam)
{
    const float relax = cfg[CFG_SOFT_RELAX];
#pragma omp parallel for schedule(static)
    for (int64_t tt = 0; tt < (int64_t)s->n_tets; ++tt) {
        const uint32_t t = (uint32_t)tt;
        const cn_soft_tet *T = &s->tets[t];
        const float *p0 = s->pos + 4 * T->p[0], *p1 = s->pos + 4 * T->p[1];
        const float *p2 = s->pos + 4 * T->p[2], *p3 = s->pos + 4 * T->p[3];
        const float w0 = p0[3], w1 = p1[3], w2 = p2[3], w3 = p3[3];
        if ((w0 + w1 + w2 + w3) <= EPS)
            continue;
        float a[3], b[3], c[3], cr[3];
        sub3(p1, p0, a);
        sub3(p2, p0, b);
        sub3(p3, p0, c);
        cross3(a, b, cr);
        const float sv = dot3(cr, c) / 6.0f;
        const float sgn = sv >= 0.0f ? 1.0f : -1.0f;
        const float vol = fabsf(sv);
        const float rest = maxf(T->rest_volume, 0.0f);
        const float C = vol - rest;
        const float f = sgn / 6.0f;
        float g0[3], g1[3], g2[3], g3[3], u[3], v[3];
        sub3(p3, p1, u);
        sub3(p2, p1, v);
        cross3(u, v, g0);
        sub3(p2, p0, u);
        sub3(p3, p0, v);
        cross3(u, v, g1);
        sub3(p3, p0, u);
        sub3(p1, p0, v);
        cross3(u, v, g2);
        sub3(p1, p0, u);
        sub3(p2, p0, v);
        cross3(u, v, g3);
        for (int k = 0; k < 3; ++k) {
            g0[k] = g0[k] * f;
            g1[k] = g1[k] * f;
            g2[k] = g2[k] * f;
            g3[k] = g3[k] * f;
        }
        const float sg = w0 * dot3(g0, g0) + w1 * dot3(g1, g1) + w2 * dot3(g2, g2) + w3 * dot3(g3, g3);
        const float comp = maxf(T->compliance, 0.0f);
        const float alpha = comp / maxf(dt * dt, EPS);
        const float lambda = lam[t];
        const float den = sg + alpha;
        if (den <= EPS)
            continue;
        const float dl = -(C + alpha * lambda) / den;
        lam[t] = lambda + dl;
        float *out = s->tet_corr + 16 * t;
        for (int k = 0; k < 3; ++k) {
            out[k] = w0 * dl * g0[k] * relax;
            out[4 + k] = w1 * dl * g1[k] * relax;
            out[8 + k] = w2 * dl * g2[k] * relax;
            out[12 + k] = w3 * dl * g3[k] * relax;
        }
        out[3] = out[7] = out[11] = out[15] = 0.0f;
    }
}

/* physics_soft_apply_{edge,bend,tet}_corrections.cs.hlsl: per particle, the incident records in
 * list order summed onto the position (records of `stride` floats, slot k at 4 k). */
static void apply_corrections(cno_scene *s, const csr *c, const float *rec, uint32_t stride)
{
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        float *p = s->pos + 4 * i;
        if (p[3] <= EPS)
            continue;
        float t[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t k = c->start[i]; k < c->start[i + 1]; ++k) {
            const uint32_t ref = c->ref[k];
            const float *r = rec + (size_t)(ref >> 2) * stride + 4 * (ref & 3u);
            t[0] += r[0];
            t[1] += r[1];
            t[2] += r[2];
        }
        p[0] = p[0] + t[0];
        p[1] = p[1] + t[1];
        p[2] = p[2] + t[2];
    }
}

/* ------------------------------------------------------------------------------------------
 * PhysicsSolver::step for the particle path, src/physics/physics_solver.cpp:336-1401
 * ------------------------------------------------------------------------------------------ */
static uint32_t resolve_iters(uint32_t v, uint32_t d) { return v > 0 ? v : d; }

/* ---------------------------------------------------------------------------------------------
 * Particle-particle contacts.
 * Candidates once per substep from the predicted positions (buildParticleParticleCandidatePairs,
 * physics_particle_{count,emit}_particle_particle_candidates.cs.hlsl with
 * IsValidParticleParticleCandidate, physics_particle_broad_phase_query.hlsli:65-117); each contact
 * iteration generates the contacts at the current positions
 * (physics_particle_generate_explicit_contacts.cs.hlsl:80-189), solves them
 * (physics_particle_solve_explicit_contacts.cs.hlsl:25-246) and folds the corrections into the
 * positions (physics_particle_apply_position_corrections.cs.hlsl:8-30); after the velocity update
 * the velocity iterations solve the last contacts
 * (physics_particle_solve_contact_velocities.cs.hlsl:26-170,
 * physics_particle_apply_contact_velocities.cs.hlsl).  The reference accumulates corrections with
 * float CAS atomics in no fixed order; here each particle sums its contacts in its candidate walk
 * order (27 cells dz, dy, dx; ascending particle index in a cell).  SafeNormalize's rsqrt is
 * evaluated as 1 / sqrt.
 * ------------------------------------------------------------------------------------------- */
#define PHASE_GROUP_MASK 0x7fffffffu
#define PHASE_SELF_COLLIDE 0x80000000u
#define CFG_SOFT_CONTACT_RELAX 2
#define CFG_SOFT_MAX_CORRECTION 3
#define CFG_RESTITUTION_VELOCITY 4
#define CFG_RESTITUTION_PENETRATION 5

typedef struct contact_rec {
    float n[3];
    float pen;
    float mat[3];   /* friction, restitution, static friction */
    uint32_t active;
} contact_rec;

/* IsAdjacentParticle, physics_particle_broad_phase_query.hlsli:50-62: candidate among the
 * particle's adjacency list, scanned in order */
static int is_adjacent(const cno_scene *s, uint32_t particle, uint32_t candidate)
{
    if (!s->adj_offsets)
        return 0;
    const uint32_t off = s->adj_offsets[particle], cnt = s->adj_counts[particle];
    for (uint32_t k = 0; k < cnt; ++k)
        if (s->adj_indices[off + k] == candidate)
            return 1;
    return 0;
}

/* IsValidParticleParticleCandidate (physics_particle_broad_phase_query.hlsli:65-117) for the pair
 * lo < hi, lo being the reference's particleIndex; the test is symmetric in the pair because the
 * adjacency lists are */
static int pp_valid(const cno_scene *s, uint32_t lo, uint32_t hi)
{
    const uint32_t *ma = s->meta + 4 * (size_t)lo, *mb = s->meta + 4 * (size_t)hi;
    if (ma[0] != mb[0])
        return 0;
    if ((ma[3] & mb[2]) == 0u || (mb[3] & ma[2]) == 0u)
        return 0;
    const uint32_t ka = s->kinds[lo], kb = s->kinds[hi];
    if (ka == CN_PARTICLE_KIND_FLUID && kb == CN_PARTICLE_KIND_FLUID)
        return 0;
    if (ka == CN_PARTICLE_KIND_SOLID && kb == CN_PARTICLE_KIND_SOLID &&
        (ma[1] & PHASE_GROUP_MASK) == (mb[1] & PHASE_GROUP_MASK)) {
        /* one group: only when both have self collision on and they share no constraint */
        if (!(ma[1] & PHASE_SELF_COLLIDE) || !(mb[1] & PHASE_SELF_COLLIDE) || is_adjacent(s, lo, hi))
            return 0;
    }
    float d[3];
    for (int a = 0; a < 3; ++a)
        d[a] = s->pos[4 * (size_t)hi + a] - s->pos[4 * (size_t)lo + a];
    const float cr = s->radii[lo] + s->radii[hi];
    return dot3(d, d) <= cr * cr;
}

/* Candidate partners of every particle, in its 27-cell walk order; returns the overflow count. */
uint32_t cno_contact_partners(const cno_scene *s, const int32_t *cells, const uint32_t *si,
                              const uint32_t *tbl, uint32_t capacity, uint32_t pcap, uint32_t *pcount,
                              uint32_t *plist)
{
    uint32_t overflow = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : overflow)
    for (int64_t ii = 0; ii < (int64_t)s->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const int32_t *ci = cells + 3 * (size_t)i;
        uint32_t cnt = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int32_t tx = ci[0] + dx, ty = ci[1] + dy, tz = ci[2] + dz;
                    uint32_t start, end;
                    if (!cno_find_cell_range(tbl, capacity, cell_hash(tx, ty, tz), &start, &end))
                        continue;
                    for (uint32_t q = start; q < end; ++q) {
                        const uint32_t j = si[q];
                        const int32_t *cj = cells + 3 * (size_t)j;
                        if (j == i || cj[0] != tx || cj[1] != ty || cj[2] != tz)
                            continue;
                        if (!pp_valid(s, i < j ? i : j, i < j ? j : i))
                            continue;
                        if (cnt < pcap)
                            plist[(size_t)i * pcap + cnt] = j;
                        ++cnt;
                    }
                }
        if (cnt > pcap) {
            ++overflow;
            cnt = pcap;
        }
        pcount[i] = cnt;
    }
    return overflow;
}

/* CombineContactMaterial, physics_rigid_solver_shared.hlsli:37-43 */
static void combine_material(const cno_scene *s, uint32_t a, uint32_t b, float *m)
{
    const float *pa = s->part_mats + 4 * (size_t)s->part_mat[a], *pb = s->part_mats + 4 * (size_t)s->part_mat[b];
    m[0] = sqrtf(maxf(0.0f, pa[0]) * maxf(0.0f, pb[0]));
    m[1] = maxf(maxf(0.0f, pa[1]), maxf(0.0f, pb[1]));
    m[2] = sqrtf(maxf(0.0f, pa[3]) * maxf(0.0f, pb[3]));
}

/* the contact of the pair lo < hi at the current positions (generate_explicit_contacts :140-188) */
static void make_contact(const cno_scene *s, uint32_t lo, uint32_t hi, contact_rec *c)
{
    memset(c, 0, sizeof(*c));
    combine_material(s, lo, hi, c->mat);
    float d[3];
    for (int a = 0; a < 3; ++a)
        d[a] = s->pos[4 * (size_t)hi + a] - s->pos[4 * (size_t)lo + a];
    const float d2 = dot3(d, d);
    const float cr = s->radii[lo] + s->radii[hi];
    if (d2 <= EPS) {
        c->active = 1;
        c->n[1] = 1.0f;
        c->pen = cr;
        return;
    }
    const float dist = sqrtf(d2);
    const float pen = cr - dist;
    if (pen > 0.0f) {
        c->active = 1;
        for (int a = 0; a < 3; ++a)
            c->n[a] = d[a] / dist;
        c->pen = pen;
    }
}

/* SafeNormalize, physics_math.hlsli:6-12, with rsqrt as 1 / sqrt */
static void safe_normalize(const float *v, float *o)
{
    const float l2 = dot3(v, v);
    if (l2 <= EPS) {
        o[0] = 0.0f;
        o[1] = 1.0f;
        o[2] = 0.0f;
        return;
    }
    const float r = 1.0f / sqrtf(l2);
    for (int a = 0; a < 3; ++a)
        o[a] = v[a] * r;
}

/* solve_explicit_contacts :103-246 for a non-proxy pair; out = the correction of side hi (side_hi)
 * or lo, zero when that side takes none */
static void contact_position(const cno_scene *s, const float *cfg, uint32_t lo, uint32_t hi, const contact_rec *c,
                             int side_hi, float *out)
{
    out[0] = out[1] = out[2] = 0.0f;
    const float *pa = s->pos + 4 * (size_t)lo, *pb = s->pos + 4 * (size_t)hi;
    const float wa = pa[3], wb = pb[3];
    if (wa <= EPS && wb <= EPS)
        return;
    const float pen = minf(maxf(c->pen - cfg[CFG_CONTACT_SLOP], 0.0f), cfg[CFG_SOFT_MAX_CORRECTION]);
    if (pen <= 0.0f)
        return;
    float nrm[3], rel[3], tan[3], fd[3];
    safe_normalize(c->n, nrm);
    const float *va = s->prev + 4 * (size_t)lo, *vb = s->prev + 4 * (size_t)hi;
    for (int a = 0; a < 3; ++a)
        rel[a] = (pb[a] - vb[a]) - (pa[a] - va[a]);
    const float denom = wa + wb;
    if (denom <= EPS)
        return;
    const float relax = cfg[CFG_SOFT_CONTACT_RELAX];
    const float lambda = (pen / denom) * relax;
    float ca[3], cb[3];
    for (int a = 0; a < 3; ++a) {
        ca[a] = -nrm[a] * (wa * lambda);
        cb[a] = nrm[a] * (wb * lambda);
    }
    const float rn = dot3(rel, nrm);
    for (int a = 0; a < 3; ++a)
        tan[a] = rel[a] - nrm[a] * rn;
    friction_delta(cfg, tan, pen, c->mat[0], c->mat[2], fd);
    const float fdist = len3(fd);
    if (fdist > 0.0f) {
        const float tden = wa + wb;
        if (tden > EPS) {
            const float tl = (fdist / tden) * relax;
            for (int a = 0; a < 3; ++a) {
                const float t = fd[a] / fdist;
                ca[a] = ca[a] + t * (wa * tl);
                cb[a] = cb[a] - t * (wb * tl);
            }
        }
    }
    if (side_hi) {
        if (wb > EPS)
            memcpy(out, cb, sizeof(cb));
    } else if (wa > EPS) {
        memcpy(out, ca, sizeof(ca));
    }
}

/* solve_contact_velocities :100-170 for a non-proxy pair; out = the velocity correction of a side */
static void contact_velocity(const cno_scene *s, const float *cfg, uint32_t lo, uint32_t hi, const contact_rec *c,
                             int side_hi, float *out)
{
    out[0] = out[1] = out[2] = 0.0f;
    const float wa = s->pos[4 * (size_t)lo + 3], wb = s->pos[4 * (size_t)hi + 3];
    if (wa <= EPS && wb <= EPS)
        return;
    float nrm[3], rv[3];
    safe_normalize(c->n, nrm);
    const float *va = s->vel + 4 * (size_t)lo, *vb = s->vel + 4 * (size_t)hi;
    for (int a = 0; a < 3; ++a)
        rv[a] = vb[a] - va[a];
    const float nv = dot3(rv, nrm);
    if (nv >= 0.0f)
        return;
    const float denom = wa + wb;
    if (denom <= EPS)
        return;
    const int rest_on = (-nv > cfg[CFG_RESTITUTION_VELOCITY]) &&
                        (c->pen <= cfg[CFG_RESTITUTION_PENETRATION] * cfg[CFG_CONTACT_SLOP]);
    const float restitution = rest_on ? saturatef(c->mat[1]) : 0.0f;
    const float desired = rest_on ? (-restitution * nv) : 0.0f;
    const float imp = maxf(0.0f, (desired - nv) / denom);
    for (int a = 0; a < 3; ++a) {
        const float t = nrm[a] * imp;
        if (side_hi) {
            if (wb > EPS)
                out[a] = t * wb;
        } else if (wa > EPS) {
            out[a] = -t * wa;
        }
    }
}

/* ---------------------------------------------------------------------------------------------
 * Particle-rigid contacts of solid particles against the static colliders, every collider being a
 * static body of its own.  Candidates once per substep from the predicted positions
 * (buildParticleRigidCandidatePairs, physics_particle_count_particle_rigid_candidates.cs.hlsl
 * :112-151: the particle's radius box against the static LBVH, solid particles only for a static
 * body, the environment / layer / mask filters), in LBVH visit order.  Each contact iteration
 * generates every candidate's contact at the current position
 * (physics_particle_rigid_generate_contacts.cs.hlsl:30-190) and solves it
 * (physics_particle_rigid_solve_contacts.cs.hlsl:26-140) into the correction the
 * particle-particle contacts of the iteration sum into, after them.  After the particle-particle
 * velocity iterations, the rigid velocity iterations solve the contacts of the last contact
 * iteration (physics_particle_rigid_solve_contact_velocities.cs.hlsl:27-130), each followed by the
 * velocity apply.  A static body has zero inverse mass and inertia and does not move, so its
 * effective mass (ComputeContactEffectiveMass) is zero and its contact point's displacement and
 * velocity are exactly zero.
 * ------------------------------------------------------------------------------------------- */

/* candidate colliders of particle i: their LBVH visit positions, ascending, into cand (at most
 * nvis); returns the count */
static uint32_t rigid_candidates(const cno_scene *s, const ocol *oc, const uint32_t *order, uint32_t nvis, uint32_t i,
                                 uint32_t *cand)
{
    if (s->kinds[i] != CN_PARTICLE_KIND_SOLID)
        return 0u;
    const float *p = s->pos + 4 * (size_t)i;
    const uint32_t *md = s->meta + 4 * (size_t)i;
    const float r = s->radii[i];
    float qmin[3], qmax[3];
    for (int a = 0; a < 3; ++a) {
        qmin[a] = p[a] - r;
        qmax[a] = p[a] + r;
    }
    uint32_t n = 0;
    for (uint32_t k = 0; k < nvis; ++k) {
        const ocol *c = &oc[order[k]];
        const int overlap = qmin[0] <= c->mx[0] && qmax[0] >= c->mn[0] && qmin[1] <= c->mx[1] &&
                            qmax[1] >= c->mn[1] && qmin[2] <= c->mx[2] && qmax[2] >= c->mn[2];
        if (overlap && c->enabled && c->env == md[0] && (md[3] & c->layer) != 0u && (c->mask & md[2]) != 0u)
            cand[n++] = k;
    }
    return n;
}

/* generate_contacts :82-176 for one collider: the particle position p in the collider frame, the
 * shape's signed distance and local normal, penetration = radius - distance; returns 1 with the
 * world normal (SafeNormalize of the rotated local normal) when the penetration is positive */
static int rigid_contact(const ocol *c, const float *p, float radius, float *nrm, float *pen)
{
    if (!c->enabled)
        return 0;
    const float qc[4] = {-c->q[0], -c->q[1], -c->q[2], c->q[3]};
    float d[3], u[3], nl[3] = {0.0f, 1.0f, 0.0f};
    float sd = 0.0f;
    for (int a = 0; a < 3; ++a)
        d[a] = p[a] - c->c[a];
    quat_rotate(qc, d, u); /* QuaternionInverseRotate */
    if (c->shape == CN_SHAPE_SPHERE) {
        const float len = sqrtf(dot3(u, u));
        if (len > EPS)
            for (int a = 0; a < 3; ++a)
                nl[a] = u[a] / len;
        sd = len - c->he[0];
    } else if (c->shape == CN_SHAPE_BOX) {
        float e[3];
        for (int a = 0; a < 3; ++a)
            e[a] = u[a] - clampf(u[a], -c->he[a], c->he[a]);
        const float el = sqrtf(dot3(e, e));
        if (el > EPS) {
            for (int a = 0; a < 3; ++a)
                nl[a] = e[a] / el;
            sd = el;
        } else {
            /* ClosestFaceNormalLocal, physics_shape_common.hlsli:100-108 */
            const float dx = c->he[0] - fabsf(u[0]), dy = c->he[1] - fabsf(u[1]), dz = c->he[2] - fabsf(u[2]);
            nl[1] = 0.0f;
            if (dx <= dy && dx <= dz)
                nl[0] = u[0] >= 0.0f ? 1.0f : -1.0f;
            else if (dy <= dz)
                nl[1] = u[1] >= 0.0f ? 1.0f : -1.0f;
            else
                nl[2] = u[2] >= 0.0f ? 1.0f : -1.0f;
            sd = -minf(dx, minf(dy, dz));
        }
    } else {
        /* capsule: the closest point of the world segment, the world normal (fallback +x) taken
         * into the collider frame */
        const float g[3] = {c->sb[0] - c->sa[0], c->sb[1] - c->sa[1], c->sb[2] - c->sa[2]};
        const float lsq = dot3(g, g);
        float t = 0.0f;
        if (lsq > EPS) {
            const float w[3] = {p[0] - c->sa[0], p[1] - c->sa[1], p[2] - c->sa[2]};
            t = saturatef(dot3(w, g) / lsq);
        }
        float e[3], wn[3] = {1.0f, 0.0f, 0.0f};
        for (int a = 0; a < 3; ++a)
            e[a] = p[a] - (c->sa[a] + t * g[a]);
        const float len = sqrtf(dot3(e, e));
        if (len > EPS)
            for (int a = 0; a < 3; ++a)
                wn[a] = e[a] / len;
        sd = len - c->he[0];
        quat_rotate(qc, wn, nl);
    }
    *pen = radius - sd;
    if (!(*pen > 0.0f))
        return 0;
    float wn[3];
    quat_rotate(c->q, nl, wn);
    safe_normalize(wn, nrm);
    return 1;
}

int cno_rigid_contact(const cn_collider *g, const float *p, float radius, float *nrm, float *pen)
{
    ocol o;
    ocol_shape(g, &o);
    *pen = 0.0f;
    return rigid_contact(&o, p, radius, nrm, pen);
}

/* CombineContactMaterial of particle i's material and the collider's */
static void rigid_material(const cno_scene *s, uint32_t i, const float *cm, float *m)
{
    const float *pa = s->part_mats + 4 * (size_t)s->part_mat[i];
    m[0] = sqrtf(maxf(0.0f, pa[0]) * maxf(0.0f, cm[0]));
    m[1] = maxf(maxf(0.0f, pa[1]), maxf(0.0f, cm[1]));
    m[2] = sqrtf(maxf(0.0f, pa[3]) * maxf(0.0f, cm[3]));
}

/* solve_contacts :45-140 against a static body: particle i's position correction */
static void rigid_position(const cno_scene *s, const float *cfg, uint32_t i, const ocol *c, const float *n,
                           float pen0, float *out)
{
    out[0] = out[1] = out[2] = 0.0f;
    const float *p = s->pos + 4 * (size_t)i, *pv = s->prev + 4 * (size_t)i;
    const float w = p[3];
    if (w <= EPS)
        return;
    const float pen = minf(maxf(pen0 - cfg[CFG_CONTACT_SLOP], 0.0f), cfg[CFG_SOFT_MAX_CORRECTION]);
    if (pen <= 0.0f)
        return;
    float nrm[3], mat[3], rel[3], tan[3], fd[3], corr[3];
    safe_normalize(n, nrm);
    rigid_material(s, i, c->material, mat);
    const float denom = w + 0.0f;
    if (denom <= EPS)
        return;
    const float relax = cfg[CFG_SOFT_CONTACT_RELAX];
    const float lambda = (pen / denom) * relax;
    for (int a = 0; a < 3; ++a) {
        corr[a] = nrm[a] * (w * lambda);
        rel[a] = 0.0f - (p[a] - pv[a]);
    }
    const float rn = dot3(rel, nrm);
    for (int a = 0; a < 3; ++a)
        tan[a] = rel[a] - nrm[a] * rn;
    friction_delta(cfg, tan, pen, mat[0], mat[2], fd);
    const float fdist = len3(fd);
    if (fdist > 0.0f) {
        const float tden = w + 0.0f;
        if (tden > EPS) {
            const float tl = fdist / tden;
            for (int a = 0; a < 3; ++a)
                corr[a] = corr[a] + (fd[a] / fdist) * ((tl * w) * relax);
        }
    }
    memcpy(out, corr, sizeof(corr));
}

/* solve_contact_velocities :44-128 against a static body: particle i's velocity correction */
static void rigid_velocity(const cno_scene *s, const float *cfg, uint32_t i, const ocol *c, const float *n,
                           float pen0, float *out)
{
    out[0] = out[1] = out[2] = 0.0f;
    const float w = s->pos[4 * (size_t)i + 3];
    if (w <= EPS)
        return;
    float nrm[3], rv[3], mat[3];
    safe_normalize(n, nrm);
    const float *v = s->vel + 4 * (size_t)i;
    for (int a = 0; a < 3; ++a)
        rv[a] = v[a] - 0.0f;
    const float nv = dot3(rv, nrm);
    if (nv >= 0.0f)
        return;
    const float denom = w + 0.0f;
    if (denom <= EPS)
        return;
    rigid_material(s, i, c->material, mat);
    const int rest_on = (-nv > cfg[CFG_RESTITUTION_VELOCITY]) &&
                        (pen0 <= cfg[CFG_RESTITUTION_PENETRATION] * cfg[CFG_CONTACT_SLOP]);
    const float restitution = rest_on ? saturatef(mat[1]) : 0.0f;
    const float desired = rest_on ? (-restitution * nv) : 0.0f;
    const float imp = maxf(0.0f, (desired - nv) / denom);
    for (int a = 0; a < 3; ++a)
        out[a] = (nrm[a] * imp) * w;
}

/* the rigid contacts' corrections of particle i at the current position, added to acc after its
 * particle-particle ones; gen keeps the position they were generated at */
static void rigid_iteration_particle(const cno_scene *s, const float *cfg, const ocol *oc, const uint32_t *order,
                                     uint32_t i, uint32_t ncand, const uint32_t *cand, float *gen, float *acc)
{
    const float *p = s->pos + 4 * (size_t)i;
    memcpy(gen + 4 * (size_t)i, p, 4 * sizeof(float));
    for (uint32_t q = 0; q < ncand; ++q) {
        const ocol *c = &oc[order[cand[q]]];
        float nrm[3], pen, d[3];
        if (!rigid_contact(c, p, s->radii[i], nrm, &pen))
            continue;
        rigid_position(s, cfg, i, c, nrm, pen, d);
        for (int a = 0; a < 3; ++a)
            acc[a] = acc[a] + d[a];
    }
}

/* one rigid velocity iteration over the contacts generated at gen, then the apply */
static void rigid_velocity_iteration(cno_scene *s, const float *cfg, const ocol *oc, const uint32_t *order,
                                     const uint32_t *ncand, const uint32_t *cand, uint32_t stride, const float *gen,
                                     float *corr)
{
    const uint32_t n = s->n;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        float acc[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t q = 0; q < ncand[i]; ++q) {
            const ocol *c = &oc[order[cand[(size_t)i * stride + q]]];
            float nrm[3], pen, d[3];
            if (!rigid_contact(c, gen + 4 * (size_t)i, s->radii[i], nrm, &pen))
                continue;
            rigid_velocity(s, cfg, i, c, nrm, pen, d);
            for (int a = 0; a < 3; ++a)
                acc[a] = acc[a] + d[a];
        }
        memcpy(corr + 4 * (size_t)i, acc, sizeof(acc));
    }
    for (uint32_t i = 0; i < n; ++i) {
        float *v = s->vel + 4 * (size_t)i;
        if (s->pos[4 * (size_t)i + 3] <= EPS)
            continue;
        for (int a = 0; a < 3; ++a)
            v[a] = v[a] + corr[4 * (size_t)i + a];
    }
}

typedef struct rigid_state {
    const ocol *oc;        /* the scene's colliders (index order) */
    const uint32_t *order; /* LBVH visit order */
    uint32_t *ncand;       /* per particle: candidate count */
    uint32_t *cand;        /* per particle: candidate visit positions, stride entries */
    uint32_t stride;
    float *gen;            /* per particle: position of the last generation (float4) */
} rigid_state;

/* one contact iteration: generate every particle's contacts, sum its corrections (the
 * particle-particle ones in walk order when pcount is given, then the rigid ones in visit order
 * when rs is), then fold them into the positions */
static void contact_iteration(cno_scene *s, const float *cfg, uint32_t pcap, const uint32_t *pcount,
                              const uint32_t *plist, contact_rec *recs, const rigid_state *rs, float *corr)
{
    const uint32_t n = s->n;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        float acc[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t k = 0; pcount && k < pcount[i]; ++k) {
            const uint32_t j = plist[(size_t)i * pcap + k];
            const uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
            contact_rec *c = recs + (size_t)i * pcap + k;
            make_contact(s, lo, hi, c);
            if (!c->active)
                continue;
            float d[3];
            contact_position(s, cfg, lo, hi, c, i == hi, d);
            for (int a = 0; a < 3; ++a)
                acc[a] = acc[a] + d[a];
        }
        if (rs && rs->ncand[i])
            rigid_iteration_particle(s, cfg, rs->oc, rs->order, i, rs->ncand[i], rs->cand + (size_t)i * rs->stride,
                                     rs->gen, acc);
        memcpy(corr + 4 * (size_t)i, acc, sizeof(acc));
    }
    for (uint32_t i = 0; i < n; ++i) {
        float *p = s->pos + 4 * (size_t)i;
        if (p[3] <= EPS)
            continue;
        for (int a = 0; a < 3; ++a)
            p[a] = p[a] + corr[4 * (size_t)i + a];
    }
}

static void contact_velocity_iteration(cno_scene *s, const float *cfg, uint32_t pcap, const uint32_t *pcount,
                                       const uint32_t *plist, const contact_rec *recs, float *corr)
{
    const uint32_t n = s->n;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        float acc[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t k = 0; k < pcount[i]; ++k) {
            const contact_rec *c = recs + (size_t)i * pcap + k;
            if (!c->active)
                continue;
            const uint32_t j = plist[(size_t)i * pcap + k];
            const uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
            float d[3];
            contact_velocity(s, cfg, lo, hi, c, i == hi, d);
            for (int a = 0; a < 3; ++a)
                acc[a] = acc[a] + d[a];
        }
        memcpy(corr + 4 * (size_t)i, acc, sizeof(acc));
    }
    for (uint32_t i = 0; i < n; ++i) {
        float *v = s->vel + 4 * (size_t)i;
        if (s->pos[4 * (size_t)i + 3] <= EPS)
            continue;
        for (int a = 0; a < 3; ++a)
            v[a] = v[a] + corr[4 * (size_t)i + a];
    }
}

int cno_step(cno_scene *s, const cn_step_params *prm, cno_stats *stats)
{
    const uint32_t n = s->n;
    const uint32_t substeps = prm->substeps > 0 ? prm->substeps : 1u;
    const uint32_t def = prm->default_iterations > 0 ? prm->default_iterations : 1u;
    const uint32_t fluid_iters = resolve_iters(prm->fluid_iterations, def);
    const uint32_t soft_iters = resolve_iters(prm->soft_internal_iterations, def);
    const uint32_t contact_iters = resolve_iters(prm->soft_contact_iterations, def);
    uint32_t max_iters = fluid_iters;
    if (soft_iters > max_iters)
        max_iters = soft_iters;
    if (contact_iters > max_iters)
        max_iters = contact_iters;
    if (def > max_iters)
        max_iters = def; /* rigid joint / rigid contact iterations resolve to the default */
    const float dt = prm->dt / (float)substeps;
    const float *cfg = prm->solver_config;
    const uint32_t cap = prm->max_fluid_neighbors;
    float cell = prm->cell_size;
    if (cell <= 0.0f)
        cell = cno_grid_cell_size(n, s->radii, s->fluid_mat, s->n_fluid_mats, s->fluid_mats);

    uint32_t fluid_count = 0;
    for (uint32_t i = 0; i < n; ++i)
        fluid_count += s->kinds[i] == CN_PARTICLE_KIND_FLUID;
    const int has_fluid = fluid_count > 0 && n > 0;
    const int has_bnd = has_fluid && n_cols(s) > 0;

    const uint32_t capacity = cno_next_pow2(n * 2u > 1u ? n * 2u : 1u);
    uint32_t *keys = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    int32_t *cells = (int32_t *)malloc(sizeof(int32_t) * 3 * (n ? n : 1));
    uint32_t *tbl = (uint32_t *)malloc(sizeof(uint32_t) * 4 * capacity);
    uint32_t *counts = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    uint32_t *req = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    uint32_t *list = (uint32_t *)malloc(sizeof(uint32_t) * ((uint64_t)(n ? n : 1) * (cap ? cap : 1)));
    float *dacc = (float *)calloc(4 * (size_t)(n ? n : 1), sizeof(float));
    float *dtmp = (float *)calloc(4 * (size_t)(n ? n : 1), sizeof(float));
    float *nc = (float *)calloc(4 * (size_t)(n ? n : 1), sizeof(float));
    uint32_t *br = (uint32_t *)calloc(2 * (size_t)(n ? n : 1), sizeof(uint32_t));
    const uint32_t bcap = 32u * n;  /* fluidBoundaryCandidatePairCapacity, physics_scene_gpu_state.cpp:728-729 */
    uint32_t *bp = (uint32_t *)calloc(bcap ? bcap : 1, sizeof(uint32_t));
    uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * (n_cols(s) ? n_cols(s) : 1));
    if (!keys || !idx || !cells || !tbl || !counts || !req || !list || !dacc || !dtmp || !nc ||
        !br || !bp || !order)
        return -1;
    scene_collider_order(s, order);

    const int has_soft = n > 0 && (s->n_edges || s->n_bends || s->n_dihedrals || s->n_tets);
    /* particle-particle contacts: a pair can exist only between a solid and a fluid particle,
     * solids of two phase groups, or solids of a self-colliding group; without one the contact
     * passes would only add zero corrections (turning -0.0 coordinates into +0.0), so they are
     * skipped */
    uint32_t solid_count = 0, fluid_count_c = 0, group0 = 0xffffffffu;
    int two_groups = 0, self_group = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (s->kinds[i] == CN_PARTICLE_KIND_SOLID) {
            ++solid_count;
            self_group |= (s->meta[4 * (size_t)i + 1] & PHASE_SELF_COLLIDE) != 0u;
            const uint32_t g = s->meta[4 * (size_t)i + 1] & PHASE_GROUP_MASK;
            if (group0 == 0xffffffffu)
                group0 = g;
            two_groups |= g != group0;
        } else {
            ++fluid_count_c;
        }
    const int has_contacts =
        contact_iters > 0 && n > 1 && solid_count > 0 && (fluid_count_c > 0 || two_groups || self_group);
    /* particle-rigid contacts: a solid particle and a static collider (hasSoftRigidContactWork,
     * physics_solver.cpp:486-488; without a solid particle no candidate exists, and the passes
     * would again only add zero corrections) */
    const int has_rigid = contact_iters > 0 && solid_count > 0 && n_cols(s) > 0;
    const uint32_t pcap = CNO_CONTACT_PARTNERS;
    uint32_t *pcount = NULL, *plist = NULL;
    contact_rec *recs = NULL;
    float *ccorr = NULL;
    if (has_contacts) {
        pcount = (uint32_t *)calloc(n, sizeof(uint32_t));
        plist = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n * pcap);
        recs = (contact_rec *)calloc((size_t)n * pcap, sizeof(contact_rec));
        if (!pcount || !plist || !recs)
            return -1;
    }
    rigid_state rs = {NULL, order, NULL, NULL, 0, NULL};
    ocol *roc = NULL;
    uint32_t rnvis = 0;
    if (has_rigid) {
        const uint32_t nc = n_cols(s);
        roc = (ocol *)malloc(sizeof(ocol) * nc);
        rs.stride = nc;
        rs.ncand = (uint32_t *)calloc(n, sizeof(uint32_t));
        rs.cand = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n * nc);
        rs.gen = (float *)calloc(4 * (size_t)n, sizeof(float));
        if (!roc || !rs.ncand || !rs.cand || !rs.gen)
            return -1;
        for (uint32_t k = 0; k < nc; ++k)
            ocol_get(s, k, &roc[k]);
        rs.oc = roc;
        while (rnvis < nc && order[rnvis] != INVALID)
            ++rnvis;
    }
    if (has_contacts || has_rigid) {
        ccorr = (float *)calloc(4 * (size_t)n, sizeof(float));
        if (!ccorr)
            return -1;
    }
    csr inc_e = {0, 0}, inc_b = {0, 0}, inc_t = {0, 0};
    float *lam_e = NULL, *lam_b = NULL, *lam_t = NULL;
    if (has_soft) {
        const uint32_t *(*ge[1])(const void *, uint32_t) = {edge_parts};
        const void *ae[1] = {s->edges};
        const uint32_t ne[1] = {s->n_edges}, ar_e[1] = {2}, of_e[1] = {0};
        csr_make(n, &inc_e, 1, ne, ar_e, ge, ae, of_e);
        const uint32_t *(*gb[2])(const void *, uint32_t) = {bend_parts, dih_parts};
        const void *ab[2] = {s->bends, s->dihedrals};
        const uint32_t nb[2] = {s->n_bends, s->n_dihedrals}, ar_b[2] = {3, 4}, of_b[2] = {0, s->n_bends};
        csr_make(n, &inc_b, 2, nb, ar_b, gb, ab, of_b);
        const uint32_t *(*gt[1])(const void *, uint32_t) = {tet_parts};
        const void *at[1] = {s->tets};
        const uint32_t nt[1] = {s->n_tets}, ar_t[1] = {4}, of_t[1] = {0};
        csr_make(n, &inc_t, 1, nt, ar_t, gt, at, of_t);
        lam_e = (float *)calloc(s->n_edges + 1, sizeof(float));
        lam_b = (float *)calloc(s->n_bends + s->n_dihedrals + 1, sizeof(float));
        lam_t = (float *)calloc(s->n_tets + 1, sizeof(float));
    }

    cno_stats st;
    memset(&st, 0, sizeof(st));
    st.fluid_count = fluid_count;
    const int has_bp_work = (contact_iters > 0 && n > 1) || has_fluid || n_cols(s) > 0;

    for (uint32_t sub = 0; sub < substeps; ++sub) {
        cno_predict(s, dt, prm->gravity);
        if (has_bp_work) {
            cno_broadphase_entries(n, s->pos, cell, keys, cells);
            for (uint32_t i = 0; i < n; ++i)
                idx[i] = i;
            cno_radix_sort_1bit(n, keys, idx);
            cno_cell_ranges(n, keys, capacity, tbl);
        }
        if (has_bnd)
            cno_fluid_boundary_candidates(s, order, br, bp, bcap);
        if (has_contacts)
            st.contact_overflows += cno_contact_partners(s, cells, idx, tbl, capacity, pcap, pcount, plist);
        if (has_rigid)
            for (uint32_t i = 0; i < n; ++i)
                rs.ncand[i] = rigid_candidates(s, roc, order, rnvis, i, rs.cand + (size_t)i * rs.stride);
        if (has_soft) {
            /* clearSoftConstraintState, physics_clear_soft_constraint_state.cs.hlsl:14-54 */
            memset(lam_e, 0, sizeof(float) * (s->n_edges + 1));
            memset(lam_b, 0, sizeof(float) * (s->n_bends + s->n_dihedrals + 1));
            memset(lam_t, 0, sizeof(float) * (s->n_tets + 1));
        }
        for (uint32_t it = 0; it < max_iters; ++it) {
            if (has_fluid && it < fluid_iters) {
                cno_fluid_neighbors(s, cell, cells, keys, idx, tbl, capacity, dacc, it, cap,
                                    counts, list, req);
                if (it == 0) {
                    uint64_t tot = 0;
                    for (uint32_t i = 0; i < n; ++i)
                        if (s->kinds[i] == CN_PARTICLE_KIND_FLUID && s->pos[4 * i + 3] > EPS)
                            tot += counts[i];
                    st.neighbor_total_first_iter = tot;
                }
                for (uint32_t i = 0; i < n; ++i)
                    if (s->kinds[i] == CN_PARTICLE_KIND_FLUID && s->pos[4 * i + 3] > EPS) {
                        if (req[i] > st.max_required_neighbors)
                            st.max_required_neighbors = req[i];
                        if (req[i] > cap)
                            ++st.truncations;
                    }
                cno_fluid_density(s, cfg, dacc, it, cap, counts, list, br, bp, nc);
                cno_fluid_delta(s, cfg, dt, dacc, it, cap, counts, list, br, bp, nc, dtmp);
                cno_fluid_apply(s, cfg, it, dtmp, dacc);
                cno_fluid_clamp(s, dacc, br, bp, dacc);
            }
            if (has_soft && it < soft_iters) {
                /* physics_solver.cpp:915-1019: edges, then bends + dihedrals, then tets */
                solve_edges(s, cfg, dt, lam_e);
                apply_corrections(s, &inc_e, s->edge_corr, 8);
                solve_bends(s, cfg, dt, lam_b);
                solve_dihedrals(s, cfg, dt, lam_b);
                apply_corrections(s, &inc_b, s->bend_corr, 16);
                solve_tets(s, cfg, dt, lam_t);
                apply_corrections(s, &inc_t, s->tet_corr, 16);
            }
            if ((has_contacts || has_rigid) && it < contact_iters)
                contact_iteration(s, cfg, pcap, has_contacts ? pcount : NULL, plist, recs, has_rigid ? &rs : NULL,
                                  ccorr);
        }
        cno_update_velocities(s, dt, dacc);
        if (has_fluid) {
            cno_project_boundary_velocities(s, cfg, dt, br, bp);
            /* post-update neighbor rebuild (physics_solver.cpp:1262); it feeds vorticity
             * confinement and render anisotropy */
            cno_fluid_neighbors(s, cell, cells, keys, idx, tbl, capacity, dacc,
                                max_iters > 0 ? max_iters - 1 : 0, cap, counts, list, NULL);
            /* physics_solver.cpp:1269-1283 */
            cno_fluid_vorticity(s, cap, counts, list, dtmp);
            cno_fluid_vorticity_apply(s, dt, cap, counts, list, dtmp);
            /* physics_solver.cpp:1285-1290; each substep overwrites the last one's output */
            if (s->render_aniso && sub + 1 == substeps)
                cno_fluid_render_anisotropy(s, cap, counts, list, s->render_aniso);
        }
        if (has_contacts)
            for (uint32_t v = 0; v < contact_iters; ++v)
                contact_velocity_iteration(s, cfg, pcap, pcount, plist, recs, ccorr);
        /* physics_solver.cpp:1301-1309, after the particle-particle velocity iterations */
        if (has_rigid)
            for (uint32_t v = 0; v < contact_iters; ++v)
                rigid_velocity_iteration(s, cfg, roc, order, rs.ncand, rs.cand, rs.stride, rs.gen, ccorr);
    }
    /* density-constraint residual of the last density pass of the step */
    if (has_fluid) {
        double mx = 0.0, sum = 0.0;
        uint32_t cntf = 0;
        for (uint32_t i = 0; i < n; ++i)
            if (s->kinds[i] == CN_PARTICLE_KIND_FLUID && s->pos[4 * i + 3] > EPS) {
                const double c = nc[4 * i + 3];
                if (fabs(c) > mx)
                    mx = fabs(c);
                sum += c;
                ++cntf;
            }
        st.max_abs_constraint_last = mx;
        st.mean_constraint_last = cntf ? sum / cntf : 0.0;
    }
    if (stats)
        *stats = st;
    free(keys);
    free(idx);
    free(cells);
    free(tbl);
    free(counts);
    free(req);
    free(list);
    free(dacc);
    free(dtmp);
    free(nc);
    free(br);
    free(bp);
    free(order);
    free(pcount);
    free(plist);
    free(recs);
    free(ccorr);
    free(roc);
    free(rs.ncand);
    free(rs.cand);
    free(rs.gen);
    if (has_soft) {
        csr_free(&inc_e);
        csr_free(&inc_b);
        csr_free(&inc_t);
        free(lam_e);
        free(lam_b);
        free(lam_t);
    }
    return 0;
}
