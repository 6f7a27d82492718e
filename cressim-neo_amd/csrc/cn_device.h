// cn_device.h -- device-side types and per-element math shared by the gfx950 kernels.
//
// Every formula here restates one HLSL helper of the reference and keeps its operation order
// (the library is built with -ffp-contract=off, and divisions / square roots are the correctly
// rounded IEEE operations: '/' lowers to the div_scale/div_fmas/div_fixup sequence and
// __builtin_sqrtf to the corrected v_sqrt sequence, whereas sqrtf()/__fsqrt_rn lower to the
// 1-ulp v_sqrt_f32 on gfx950), so a kernel and the CPU oracle produce the same float32 bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cn_hip.h"

namespace cn {

constexpr float kEps = 1.0e-6f;              // kEpsilon, physics_base.hlsli:7
constexpr uint32_t kInvalid = 0xffffffffu;   // kInvalidIndex, physics_base.hlsli:6
constexpr float kKernelPi = 3.14159265359f;  // physics_fluid_common.hlsli:26
constexpr float kBroadPhaseMargin = 0.05f;   // physics_shape_common.hlsli:10
constexpr uint32_t kMaskColliders = 32;      // a scene of at most this many colliders: candidate bitmasks
constexpr uint32_t kMaxColliders = 65535;    // colliders of one scene (16-bit candidate entries)
constexpr uint32_t kBndList = 0x80000000u;   // candidate word: the low bits count list entries
constexpr uint32_t kBndListCap = 32;         // candidate colliders a particle holds in a list
constexpr uint32_t kColGroup = 32;           // colliders per group of the two-level candidate search
constexpr int kBlock = 256;                  // threads per workgroup for per-particle kernels

// GpuPhysicsSolverConfig float indices (physics_solver_config.hlsli:19-51)
constexpr int kCfgContactSlop = 0;
constexpr int kCfgPosFrictionMin = 7;
constexpr int kCfgSoftRelax = 24;
constexpr int kCfgFluidSolve = 32;
constexpr int kCfgFluidRelax = 33;
constexpr int kCfgFluidBndDensity = 34;
constexpr int kCfgFluidBndDelta = 35;
constexpr int kCfgFluidMaxUnder = 36;

// Static per-particle attributes, deduplicated on upload; each particle stores a class index.
struct ParticleClass {
    float radius;
    uint32_t kind;
    uint32_t fluid_mat;
    uint32_t part_mat;
    uint32_t env;
    uint32_t phase;
    uint32_t layer;
    uint32_t mask;
    uint32_t scene;  // scene of the pool the particle belongs to
};

// One scene of the pool: its cell size, its key range and its static colliders.  Scenes never
// interact: the scene index occupies the key bits above the Morton code, so no cell, run or
// block holds particles of two scenes.
struct SceneInfo {
    float inv_cell;       // 1 / max(cell size, 1e-4) of the scene
    uint32_t key_prefix;  // scene index << (3 * morton_bits)
    uint32_t col0;        // first collider of the scene in the visit-ordered collider array
    uint32_t ncol;        // enabled colliders of the scene (at most kMaxColliders)
    uint32_t grp0;        // first collider group of the scene (scenes of more than kMaskColliders)
    uint32_t fluid_only;  // 1: no solid particle and no constraint, so fluid positions change
                          // within a substep only by their deltas
    uint32_t has_fluid;   // 1: some fluid particle (the fluid passes have work in the scene)
    uint32_t pairs;       // 1: a particle-particle contact pair can exist (a solid particle and a
                          // fluid one, solid particles of two phase groups, or a self-colliding
                          // phase group)
    uint32_t rigid;       // 1: particle-rigid contacts can exist (a solid particle and a static
                          // collider slot), solved in the slot-order contact pass
    uint32_t rigid_pid;   // 1: the same, for a scene without particle-particle pairs whose every
                          // particle is constrained: solved in particle-index order in the soft phase
};

// kColGroup consecutive colliders of a scene's visit order and the union of their broad-phase
// AABBs: a scene of more than kMaskColliders colliders tests a group before its members.
struct ColGroup {
    float mn[3], mx[3];
    uint32_t first, count;  // colliders first .. first + count - 1 of the scene's visit order
};

// Calls f(k) for every candidate collider of a slot, k its index in the scene's visit order, in
// that order: `word` is a bitmask over the scene's colliders, or kBndList | n with the candidates in
// the slot's kBndListCap-entry list.
template <class F>
__device__ __forceinline__ void each_candidate(uint32_t word, const uint16_t *__restrict__ list, uint32_t slot, F &&f)
{
    if (!(word & kBndList)) {
        for (uint32_t m = word; m; m &= m - 1u) f((uint32_t)__builtin_ctz(m));
    } else {
        const uint32_t n = min(word & 0xffu, kBndListCap);
        const uint16_t *l = list + (size_t)slot * kBndListCap;
        for (uint32_t q = 0; q < n; ++q) f((uint32_t)l[q]);
    }
}

// A static collider in LBVH visit order, its world pose resolved on the host once (the body pose
// composed with the collider's local pose, quaternions normalized) and its broad-phase AABB
// precomputed.
struct Collider {
    float c[3];          // world position of the collider
    float he[3];         // box: half extents; sphere and capsule: radius in he[0]
    float aabb_min[3];
    float aabb_max[3];
    float material[4];
    uint32_t env, layer, mask, enabled;
    uint32_t index;      // collider index in the caller's tables (boxes, then shape colliders)
    uint32_t shape;      // CN_SHAPE_*
    uint32_t rotated;    // 0: identity orientation (the box path rotates nothing)
    uint32_t pad;
    float q[4];          // world orientation (x, y, z, w), normalized
    float sa[4], sb[4];  // capsule segment ends (xyz)
};

// Solver-wide scalars of one substep.
struct SubstepConsts {
    float dt;
    float inv_cell;       // > 0: 1 / max(cell, 1e-4) for every scene; 0: each scene's own
    float gravity[3];
    uint32_t n;           // slots stepped (the pool, or the listed scenes of a compact tick)
    uint32_t npid;        // particles of the pool: the bound of particle-index arrays
    uint32_t cap;         // neighbor slots per particle
    uint32_t key_mode;
    uint32_t morton_bits; // bits per axis of the Morton key
    uint32_t bucket_shift;
    uint32_t uniform_class;  // 1: every particle has class 0
    uint32_t fluid_uniform;  // 1: every scene with fluid has one particle class (a cell block's class
                             // is its first particle's)
    uint32_t tension;        // 1: some fluid material has surface tension
    uint32_t tile_max;       // halo particles a fluid-pass workgroup may stage in LDS (upper bound)
    uint32_t tile_dens;      // halo capacity of the density pass
    uint32_t tile_delta;     // halo capacity of the delta pass from the second iteration on
    uint32_t tile_wide;      // halo capacity of the wide delta pass (previous delta read from HBM)
    uint32_t blk_cap;        // cell blocks with a stored halo descriptor
    uint32_t grid_max;       // workgroups of a persistent fluid launch (upper bound)
    uint32_t superset;       // 1: LDS blocks of fluid-only scenes reuse iteration 0's candidate lists
    uint32_t fluid_iters;    // fluid iterations per substep
    uint32_t fluid_scenes;   // scenes with fluid particles: they hold key prefixes [0, fluid_scenes)
    float cfl_guard;         // relative margin of the CFL-term skip test
    float cfg[CN_SOLVER_CONFIG_FLOATS];
};

__device__ __forceinline__ float fdiv(float a, float b) { return a / b; }
__device__ __forceinline__ float fsqrt(float a) { return __builtin_sqrtf(a); }

// a / b for b > 0 from y = 1 / b, the correctly rounded reciprocal computed once per divisor:
// q0 = a * y lies within one ulp of a / b, the remainder a - b * q0 is exact in an fma, and
// q0 + remainder * y rounds to the correctly rounded quotient (Markstein's theorem).  Exact while
// neither the remainder nor the quotient underflows: a == 0, or |a| >= 2^-96 with
// 2^-100 <= |a / b| < 2^100.  The last step is taken negated so that a zero dividend gives the
// signed zero a / b gives.  Three VALU operations against the ten of the scaled IEEE sequence.
__device__ __forceinline__ float fdiv_rcp_neg(float a, float b, float y)
{
    const float q0 = __fmul_rn(a, y);
    const float r = __builtin_fmaf(-q0, b, a);
    return __builtin_fmaf(-r, y, -q0);  // -(a / b)
}
__device__ __forceinline__ float fdiv_rcp(float a, float b, float y) { return -fdiv_rcp_neg(a, b, y); }

// 1 / b, correctly rounded, for a normal b > 0: the hardware estimate (within one ulp) refined by
// two Newton steps in fma, which reach the correctly rounded reciprocal from any estimate within
// two ulps except for the significand 1.11...1 (checked exhaustively over significands); that one
// takes the IEEE division.
__device__ __forceinline__ float frcp_exact(float b)
{
    float y = __builtin_amdgcn_rcpf(b);
    y = __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
    y = __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
    if ((__float_as_uint(b) & 0x7fffffu) == 0x7fffffu) {
        __asm__ volatile("");  // a side effect keeps this rare path a branch rather than a select
        y = fdiv(1.0f, b);
    }
    return y;
}

// 1 / b for b >= 1, +inf or NaN: frcp_exact while 1 / b stays normal (b < 2^126), else the IEEE
// division.
__device__ __forceinline__ float frcp_exact_ge1(float b)
{
    float y = __builtin_amdgcn_rcpf(b);
    y = __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
    y = __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
    if ((__float_as_uint(b) & 0x7fffffu) == 0x7fffffu || !(b < 0x1p126f)) {
        __asm__ volatile("");
        y = fdiv(1.0f, b);
    }
    return y;
}

// Square root by the hardware estimate (within one ulp) and the rounding test on its two
// neighbours: s - ulp is the root when x - (s - ulp) * s <= 0, s + ulp when x - (s + ulp) * s > 0.
// Equal to fsqrt for x == 0 and for 2^-96 <= x < inf, where fsqrt applies no scaling; for
// 0 < x < 2^-96 the result is within a few ulps of the root and below 2^-47, which the fluid
// terms consume only through dist <= eps tests and through d / h < 2^-25 (1 - d / h rounds to 1).
__device__ __forceinline__ float fsqrt_nearzero_free(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u);
    const float sp = __uint_as_float(__float_as_uint(s) + 1u);
    const float rm = __builtin_fmaf(-sm, s, x);
    const float rp = __builtin_fmaf(-sp, s, x);
    const float t = rm <= 0.0f ? sm : s;
    return rp > 0.0f ? sp : t;
}
__device__ __forceinline__ float fmaxd(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fmind(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fsat(float a) { return fmind(fmaxd(a, 0.0f), 1.0f); }
__device__ __forceinline__ float fclamp(float x, float lo, float hi) { return fmind(fmaxd(x, lo), hi); }

// dot(a, a) then sqrt: HLSL length(), evaluated as ((x*x + y*y) + z*z).
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    return __fadd_rn(__fadd_rn(__fmul_rn(ax, bx), __fmul_rn(ay, by)), __fmul_rn(az, bz));
}
__device__ __forceinline__ float len3(float x, float y, float z) { return fsqrt(dot3(x, y, z, x, y, z)); }

// ComputeParticleGridCellKey, physics_particle_grid.hlsli:4-10
__device__ __forceinline__ uint32_t cell_hash(int gx, int gy, int gz)
{
    uint32_t s = (uint32_t)gx * 73856093u;
    s ^= (uint32_t)gy * 19349663u;
    s ^= (uint32_t)gz * 83492791u;
    return s;
}

// 3-D Morton interleave of the low `bits` bits of each wrapped cell coordinate.
__device__ __forceinline__ uint32_t spread3(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t cell_morton(int gx, int gy, int gz, uint32_t bits)
{
    const uint32_t m = (1u << bits) - 1u;
    return (spread3((uint32_t)gx & m) << 2) | (spread3((uint32_t)gy & m) << 1) | spread3((uint32_t)gz & m);
}
// Key of a cell of the scene whose key range starts at prefix (0 in hash mode, which holds one scene).
__device__ __forceinline__ uint32_t cell_key(int gx, int gy, int gz, uint32_t prefix, uint32_t mode, uint32_t bits)
{
    return mode == CN_KEY_HASH ? cell_hash(gx, gy, gz) : (prefix | cell_morton(gx, gy, gz, bits));
}

// FluidSpikyKernel / FluidSpikyKernelDerivative / FluidCohesionKernel,
// physics_fluid_common.hlsli:17-65
// The kernels' normalisation 15 / (pi h^3) depends on h only: callers that evaluate many
// neighbors of one particle compute it once (spiky_coef) and pass it to the _k forms.
__device__ __forceinline__ float spiky_coef(float h)
{
    const float h3 = __fmul_rn(__fmul_rn(h, h), h);
    return fdiv(15.0f, __fmul_rn(kKernelPi, h3));
}
__device__ __forceinline__ float spiky_w_k(float d, float h, float k)
{
    if (d >= h) return 0.0f;
    const float omq = __fsub_rn(1.0f, fdiv(d, h));
    return __fmul_rn(__fmul_rn(k, omq), omq);
}
__device__ __forceinline__ float spiky_dw_k(float d, float h, float k)
{
    if (d <= kEps || d >= h) return 0.0f;
    const float omq = __fsub_rn(1.0f, fdiv(d, h));
    return fdiv(__fmul_rn(__fmul_rn(-2.0f, k), omq), h);
}
__device__ __forceinline__ float spiky_w(float d, float h) { return spiky_w_k(d, h, spiky_coef(h)); }
__device__ __forceinline__ float spiky_dw(float d, float h) { return spiky_dw_k(d, h, spiky_coef(h)); }
__device__ __forceinline__ float cohesion_kernel(float d, float h, float c1, float c2)
{
    if (d <= kEps || d >= h) return 0.0f;
    const float q = fdiv(d, h);
    const float a = __fmul_rn(__fmul_rn(__fmul_rn(c1, q), q), q);
    const float b = __fmul_rn(__fmul_rn(c2, q), q);
    return __fsub_rn(__fadd_rn(a, b), 1.0f);
}

// The same kernels from q = d / h supplied by the caller (the per-neighbor fluid terms share
// one quotient between the kernels of a pair).
__device__ __forceinline__ float spiky_w_om(float d, float h, float omq, float k)  // omq = 1 - d / h
{
    if (d >= h) return 0.0f;
    return __fmul_rn(__fmul_rn(k, omq), omq);
}
// The same for a pair known to lie within h (a listed neighbor: its squared distance is below the
// threshold whose rounded root is h), where the range test is always false.
__device__ __forceinline__ float spiky_w_om_in(float omq, float k) { return __fmul_rn(__fmul_rn(k, omq), omq); }
__device__ __forceinline__ float cohesion_q(float d, float h, float q, float c1, float c2)
{
    if (d <= kEps || d >= h) return 0.0f;
    const float a = __fmul_rn(__fmul_rn(__fmul_rn(c1, q), q), q);
    const float b = __fmul_rn(__fmul_rn(c2, q), q);
    return __fsub_rn(__fadd_rn(a, b), 1.0f);
}

// QuaternionRotate (physics_math.hlsli:33-37): v + w t + q x t with t = 2 (q x v), left to right.
__device__ __forceinline__ void quat_rotate(float qx, float qy, float qz, float qw, float vx, float vy, float vz,
                                            float &ox, float &oy, float &oz)
{
    const float tx = 2.0f * __fsub_rn(__fmul_rn(qy, vz), __fmul_rn(qz, vy));
    const float ty = 2.0f * __fsub_rn(__fmul_rn(qz, vx), __fmul_rn(qx, vz));
    const float tz = 2.0f * __fsub_rn(__fmul_rn(qx, vy), __fmul_rn(qy, vx));
    ox = __fadd_rn(__fadd_rn(vx, __fmul_rn(qw, tx)), __fsub_rn(__fmul_rn(qy, tz), __fmul_rn(qz, ty)));
    oy = __fadd_rn(__fadd_rn(vy, __fmul_rn(qw, ty)), __fsub_rn(__fmul_rn(qz, tx), __fmul_rn(qx, tz)));
    oz = __fadd_rn(__fadd_rn(vz, __fmul_rn(qw, tz)), __fsub_rn(__fmul_rn(qx, ty), __fmul_rn(qy, tx)));
}

// The box branch of ComputeFluidBoundaryContactInfo in the box's frame (local position l):
// closest point, outside distance and normal, or the inside depth and ClosestFaceNormalLocal
// (physics_shape_common.hlsli:100-108).
__device__ __forceinline__ void box_local_contact(const float *he, float lx, float ly, float lz, float &sd,
                                                  float &nx, float &ny, float &nz)
{
    const float dx = __fsub_rn(lx, fclamp(lx, -he[0], he[0]));
    const float dy = __fsub_rn(ly, fclamp(ly, -he[1], he[1]));
    const float dz = __fsub_rn(lz, fclamp(lz, -he[2], he[2]));
    const float dl = len3(dx, dy, dz);
    if (dl > kEps) {
        nx = fdiv(dx, dl); ny = fdiv(dy, dl); nz = fdiv(dz, dl);
        sd = dl;
    } else {
        const float ex = __fsub_rn(he[0], fabsf(lx));
        const float ey = __fsub_rn(he[1], fabsf(ly));
        const float ez = __fsub_rn(he[2], fabsf(lz));
        nx = 0.0f; ny = 0.0f; nz = 0.0f;
        if (ex <= ey && ex <= ez) nx = lx >= 0.0f ? 1.0f : -1.0f;
        else if (ey <= ez) ny = ly >= 0.0f ? 1.0f : -1.0f;
        else nz = lz >= 0.0f ? 1.0f : -1.0f;
        sd = -fmind(ex, fmind(ey, ez));
    }
}

// ComputeFluidBoundaryContactInfo, physics_fluid_boundary_common.hlsli:10-84: signed distance and
// outward normal of point p against a sphere, a box or a capsule.  A box with identity orientation
// takes its frame without a rotation (the quaternion rotation by the identity changes at most the
// sign of a zero).
__device__ __forceinline__ bool collider_contact(const Collider &c, float px, float py, float pz, float &sd,
                                                 float &nx, float &ny, float &nz)
{
    nx = 0.0f; ny = 1.0f; nz = 0.0f; sd = 0.0f;
    if (!c.enabled) return false;
    if (c.shape == CN_SHAPE_BOX) {
        const float lx = __fsub_rn(px, c.c[0]), ly = __fsub_rn(py, c.c[1]), lz = __fsub_rn(pz, c.c[2]);
        if (!c.rotated) {
            box_local_contact(c.he, lx, ly, lz, sd, nx, ny, nz);
            return true;
        }
        float ux, uy, uz, mx, my, mz;
        quat_rotate(-c.q[0], -c.q[1], -c.q[2], c.q[3], lx, ly, lz, ux, uy, uz);  // QuaternionInverseRotate
        box_local_contact(c.he, ux, uy, uz, sd, mx, my, mz);
        quat_rotate(c.q[0], c.q[1], c.q[2], c.q[3], mx, my, mz, nx, ny, nz);
        return true;
    }
    float dx, dy, dz;
    if (c.shape == CN_SHAPE_SPHERE) {
        dx = __fsub_rn(px, c.c[0]); dy = __fsub_rn(py, c.c[1]); dz = __fsub_rn(pz, c.c[2]);
    } else {  // capsule: closest point of the segment, t = saturate(dot(p - a, b - a) / |b - a|^2)
        const float gx = __fsub_rn(c.sb[0], c.sa[0]), gy = __fsub_rn(c.sb[1], c.sa[1]), gz = __fsub_rn(c.sb[2], c.sa[2]);
        const float lsq = dot3(gx, gy, gz, gx, gy, gz);
        float t = 0.0f;
        if (lsq > kEps) {
            const float wx = __fsub_rn(px, c.sa[0]), wy = __fsub_rn(py, c.sa[1]), wz = __fsub_rn(pz, c.sa[2]);
            t = fsat(fdiv(dot3(wx, wy, wz, gx, gy, gz), lsq));
        }
        // lerp(a, b, t) = a + t (b - a)
        dx = __fsub_rn(px, __fadd_rn(c.sa[0], __fmul_rn(t, gx)));
        dy = __fsub_rn(py, __fadd_rn(c.sa[1], __fmul_rn(t, gy)));
        dz = __fsub_rn(pz, __fadd_rn(c.sa[2], __fmul_rn(t, gz)));
    }
    const float dist = len3(dx, dy, dz);
    if (dist > kEps) {
        nx = fdiv(dx, dist); ny = fdiv(dy, dist); nz = fdiv(dz, dist);
    }
    sd = __fsub_rn(dist, c.he[0]);
    return true;
}

// ComputeFluidBoundaryGhostDelta, physics_fluid_boundary_common.hlsli:90-110
__device__ __forceinline__ bool ghost_delta(const Collider &c, float px, float py, float pz, float h,
                                            float &gx, float &gy, float &gz)
{
    float sd, nx, ny, nz;
    gx = gy = gz = 0.0f;
    if (!collider_contact(c, px, py, pz, sd, nx, ny, nz)) return false;
    if (sd < 0.0f || sd >= h) return false;
    const float t = __fmul_rn(2.0f, sd);
    gx = __fmul_rn(nx, t); gy = __fmul_rn(ny, t); gz = __fmul_rn(nz, t);
    return true;
}

// Single-precision atan / atan2 / asin / acos after the Cephes Math Library (S. L. Moshier): the
// same range reductions and polynomial coefficients as its atanf.c / asinf.c, one rounding per
// operation.  Used by the bend (acos) and dihedral (atan2) constraints so the kernels and the
// CPU oracle agree bit for bit.
constexpr float kCephesPi = 3.141592653589793238f;
constexpr float kCephesPio2 = 1.5707963267948966192f;
constexpr float kCephesPio4 = 0.7853981633974483096f;

__device__ __forceinline__ float cephes_atanf(float xx)
{
    float x = xx, y;
    const bool neg = x < 0.0f;
    if (neg) x = -x;
    if (x > 2.414213562373095f) {
        y = kCephesPio2;
        x = -(1.0f / x);
    } else if (x > 0.4142135623730950f) {
        y = kCephesPio4;
        x = (x - 1.0f) / (x + 1.0f);
    } else {
        y = 0.0f;
    }
    const float z = x * x;
    const float p = (((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) *
                        z * x + x;
    y = y + p;
    return neg ? -y : y;
}

__device__ __forceinline__ float cephes_atan2f(float y, float x)
{
    int code = x < 0.0f ? 2 : 0;
    if (y < 0.0f) code |= 1;
    if (x == 0.0f) {
        if (code & 1) return -kCephesPio2;
        if (y == 0.0f) return 0.0f;
        return kCephesPio2;
    }
    if (y == 0.0f) return (code & 2) ? kCephesPi : 0.0f;
    const float w = code == 2 ? kCephesPi : (code == 3 ? -kCephesPi : 0.0f);
    return w + cephes_atanf(y / x);
}

__device__ __forceinline__ float cephes_asinf(float xx)
{
    const float a = fabsf(xx);
    if (a > 1.0f) return __builtin_nanf("");
    if (a < 1.0e-4f) return xx;
    float x, z;
    bool flag = false;
    if (a > 0.5f) {
        z = 0.5f * (1.0f - a);
        x = fsqrt(z);
        flag = true;
    } else {
        x = a;
        z = x * x;
    }
    z = ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z +
         1.6666752422e-1f) * z * x + x;
    if (flag) {
        z = z + z;
        z = kCephesPio2 - z;
    }
    return xx < 0.0f ? -z : z;
}

__device__ __forceinline__ float cephes_acosf(float x)
{
    if (x < -1.0f || x > 1.0f) return __builtin_nanf("");
    if (x < -0.5f) return kCephesPi - 2.0f * cephes_asinf(fsqrt(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * cephes_asinf(fsqrt(0.5f * (1.0f - x)));
    return kCephesPio2 - cephes_asinf(x);
}

// Exclusive scan over a kBlock-thread workgroup (sh holds kBlock words); total receives the sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *sh, uint32_t &total)
{
    const uint32_t t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < (uint32_t)kBlock; off <<= 1) {
        const uint32_t add = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    total = sh[kBlock - 1];
    const uint32_t ex = sh[t] - v;
    __syncthreads();
    return ex;
}

// XCD-aware tile index: consecutive logical tiles land on the same XCD (and its L2), since the
// dispatcher deals workgroups to the 8 XCDs round-robin.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t bid, uint32_t nblocks)
{
    const uint32_t per = nblocks / 8u;
    if (per == 0u) return bid;
    const uint32_t full = per * 8u;
    if (bid >= full) return bid;
    return (bid % 8u) * per + bid / 8u;
}

}  // namespace cn
