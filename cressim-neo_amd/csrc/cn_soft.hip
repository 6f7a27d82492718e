// cn_soft.hip -- XPBD soft-body and cloth constraints on gfx950.
//
// Reference passes (assets/shaders/physics/soft/solver/): physics_soft_solve_edge_constraints.cs.hlsl
// :10-74, physics_soft_solve_bend_constraints.cs.hlsl:20-86,
// physics_cloth_solve_dihedral_constraints.cs.hlsl:20-95, physics_soft_solve_tet_constraints.cs.hlsl
// :15-87, physics_soft_apply_{edge,bend,tet}_corrections.cs.hlsl and
// physics_clear_soft_constraint_state.cs.hlsl:14-54.  Each solve is Jacobi over its constraint type:
// one thread per constraint reads the current positions, updates its lambda and writes one
// correction record per vertex; the apply pass then sums, per particle, the records of its incident
// constraints in ascending constraint order (the incidence lists of
// physics_scene_gpu_state.cpp:3273-3360) and adds the sum to the position.  Constraints refer to
// particle indices; positions live in slot order, reached through slot_of[particle index].
#include <hip/hip_runtime.h>

#include "cn_internal.h"

namespace cn {

// Workgroup size of the soft passes.  They run beside the persistent fluid workgroups (eight waves
// each); measured on a pool of fluid, tet and cloth scenes, 512 threads step it in 49.3 ms against
// 50.6 at 256, 51.5 at 128 and 55.7 at 1024.
constexpr uint32_t kSoftBlock = 512;

static inline uint32_t grid_n(uint32_t n) { return (n + kSoftBlock - 1) / kSoftBlock; }

// A constraint's correction records, written where its vertices' incidence lists read them.
struct RecOut {
    CorrRec *rec;
    const uint32_t *map;
    __device__ CorrRec &operator[](uint32_t k) const { return rec[map[k]]; }
};

__device__ __forceinline__ CorrRec rec3(float x, float y, float z) { return CorrRec{x, y, z}; }

__device__ __forceinline__ float4 ldp(const SoftArgs &a, uint32_t particle) { return a.posi[particle]; }

// Constraints and particles of scenes that this step leaves untouched are skipped (records kept);
// a step of every scene skips the two dependent loads of the check.
__device__ __forceinline__ bool live(const SoftArgs &a, uint32_t particle)
{
    return a.all_live || a.active[a.pid_scene[particle]] != 0u;
}

__device__ __forceinline__ void cross3d(float ax, float ay, float az, float bx, float by, float bz, float &ox, float &oy,
                                        float &oz)
{
    ox = ay * bz - az * by;
    oy = az * bx - ax * bz;
    oz = ax * by - ay * bx;
}

__global__ __launch_bounds__(kSoftBlock) void k_solve_edges(SoftArgs a)
{
    const uint32_t e = xcd_tile(blockIdx.x, gridDim.x) * kSoftBlock + threadIdx.x;
    if (e >= a.ne) return;
    const EdgeDev E = a.edges[e];
    if (!live(a, E.a)) return;
    const RecOut out{a.corr_e, a.map_e + 2 * (size_t)e};
    const CorrRec z4 = rec3(0.0f, 0.0f, 0.0f);
    if (E.compliance != E.compliance) {  // disabled
        out[0] = z4;
        out[1] = z4;
        if (a.first) a.lam_e[e] = 0.0f;
        return;
    }
    const float4 pa = ldp(a, E.a), pb = ldp(a, E.b);
    const float wa = pa.w, wb = pb.w, wsum = wa + wb;
    if (wsum <= kEps) {
        out[0] = z4;
        out[1] = z4;
        if (a.first) a.lam_e[e] = 0.0f;
        return;
    }
    const float dx = pb.x - pa.x, dy = pb.y - pa.y, dz = pb.z - pa.z;
    const float l2 = dot3(dx, dy, dz, dx, dy, dz);
    if (l2 <= kEps) {
        out[0] = z4;
        out[1] = z4;
        if (a.first) a.lam_e[e] = 0.0f;
        return;
    }
    const float len = fsqrt(l2);
    const float gax = fdiv(-dx, len), gay = fdiv(-dy, len), gaz = fdiv(-dz, len);
    const float gbx = -gax, gby = -gay, gbz = -gaz;
    const float alpha = fdiv(E.compliance, fmaxd(a.dt * a.dt, kEps));
    const float C = len - E.rest_length;
    const float lambda = a.first ? 0.0f : a.lam_e[e];
    const float den = wsum + alpha;
    if (den <= kEps) {
        out[0] = z4;
        out[1] = z4;
        if (a.first) a.lam_e[e] = 0.0f;
        return;
    }
    const float dl = fdiv(-(C + alpha * lambda), den);
    a.lam_e[e] = lambda + dl;
    const float fa = wa * dl, fb = wb * dl;
    out[0] = rec3(fa * gax * a.relax, fa * gay * a.relax, fa * gaz * a.relax);
    out[1] = rec3(fb * gbx * a.relax, fb * gby * a.relax, fb * gbz * a.relax);
}

__global__ __launch_bounds__(kSoftBlock) void k_solve_bends(SoftArgs a)
{
    const uint32_t b = xcd_tile(blockIdx.x, gridDim.x) * kSoftBlock + threadIdx.x;
    if (b >= a.nb) return;
    const BendDev B = a.bends[b];
    if (!live(a, B.p0)) return;
    const RecOut out{a.corr_b, a.map_b + 4 * (size_t)b};
    const CorrRec z4 = rec3(0.0f, 0.0f, 0.0f);
    const float4 p0 = ldp(a, B.p0), p1 = ldp(a, B.p1), p2 = ldp(a, B.p2);
    const float w0 = p0.w, w1 = p1.w, w2 = p2.w;
    bool zero = w0 + w1 + w2 <= kEps;
    float g0x = 0, g0y = 0, g0z = 0, g1x = 0, g1y = 0, g1z = 0, g2x = 0, g2y = 0, g2z = 0, dl = 0;
    if (!zero) {
        const float e0x = p0.x - p1.x, e0y = p0.y - p1.y, e0z = p0.z - p1.z;
        const float e1x = p2.x - p1.x, e1y = p2.y - p1.y, e1z = p2.z - p1.z;
        const float l0s = dot3(e0x, e0y, e0z, e0x, e0y, e0z), l1s = dot3(e1x, e1y, e1z, e1x, e1y, e1z);
        zero = l0s <= kEps || l1s <= kEps;
        if (!zero) {
            const float l0 = fsqrt(l0s), l1 = fsqrt(l1s);
            const float d0x = fdiv(e0x, l0), d0y = fdiv(e0y, l0), d0z = fdiv(e0z, l0);
            const float d1x = fdiv(e1x, l1), d1y = fdiv(e1y, l1), d1z = fdiv(e1z, l1);
            const float ct = fclamp(dot3(d0x, d0y, d0z, d1x, d1y, d1z), -1.0f + 1.0e-4f, 1.0f - 1.0e-4f);
            const float st = fsqrt(fmaxd(1.0f - ct * ct, 1.0e-8f));
            const float theta = cephes_acosf(ct);
            const float q0 = fmaxd(l0 * st, kEps), q2 = fmaxd(l1 * st, kEps);
            g0x = fdiv(-(d1x - ct * d0x), q0);
            g0y = fdiv(-(d1y - ct * d0y), q0);
            g0z = fdiv(-(d1z - ct * d0z), q0);
            g2x = fdiv(-(d0x - ct * d1x), q2);
            g2y = fdiv(-(d0y - ct * d1y), q2);
            g2z = fdiv(-(d0z - ct * d1z), q2);
            g1x = -(g0x + g2x);
            g1y = -(g0y + g2y);
            g1z = -(g0z + g2z);
            const float den = w0 * dot3(g0x, g0y, g0z, g0x, g0y, g0z) + w1 * dot3(g1x, g1y, g1z, g1x, g1y, g1z) +
                              w2 * dot3(g2x, g2y, g2z, g2x, g2y, g2z);
            const float alpha = fdiv(B.compliance, fmaxd(a.dt * a.dt, kEps));
            zero = den + alpha <= kEps;
            if (!zero) {
                const float C = theta - B.rest_angle;
                const float lambda = a.first ? 0.0f : a.lam_b[b];
                dl = fdiv(-(C + alpha * lambda), den + alpha);
                a.lam_b[b] = lambda + dl;
            }
        }
    }
    if (zero) {
        out[0] = out[1] = out[2] = out[3] = z4;
        if (a.first) a.lam_b[b] = 0.0f;
        return;
    }
    const float f0 = w0 * dl, f1 = w1 * dl, f2 = w2 * dl, r = a.relax;
    out[0] = rec3(f0 * g0x * r, f0 * g0y * r, f0 * g0z * r);
    out[1] = rec3(f1 * g1x * r, f1 * g1y * r, f1 * g1z * r);
    out[2] = rec3(f2 * g2x * r, f2 * g2y * r, f2 * g2z * r);
    out[3] = z4;
}

__global__ __launch_bounds__(kSoftBlock) void k_solve_dihedrals(SoftArgs a)
{
    const uint32_t d = xcd_tile(blockIdx.x, gridDim.x) * kSoftBlock + threadIdx.x;
    if (d >= a.nd) return;
    const uint32_t ci = a.nb + d;
    const DihedralDev D = a.dih[d];
    if (!live(a, D.e0)) return;
    const RecOut out{a.corr_b, a.map_b + 4 * (size_t)ci};
    const CorrRec z4 = rec3(0.0f, 0.0f, 0.0f);
    const float4 e0m = ldp(a, D.e0), e1m = ldp(a, D.e1), o0m = ldp(a, D.o0), o1m = ldp(a, D.o1);
    // p0, p1 opposite vertices; p2, p3 edge vertices
    const float ex = e1m.x - e0m.x, ey = e1m.y - e0m.y, ez = e1m.z - e0m.z;
    const float el2 = dot3(ex, ey, ez, ex, ey, ez);
    float n0x, n0y, n0z, n1x, n1y, n1z;
    cross3d(e0m.x - o0m.x, e0m.y - o0m.y, e0m.z - o0m.z, e1m.x - o0m.x, e1m.y - o0m.y, e1m.z - o0m.z, n0x, n0y, n0z);
    cross3d(e1m.x - o1m.x, e1m.y - o1m.y, e1m.z - o1m.z, e0m.x - o1m.x, e0m.y - o1m.y, e0m.z - o1m.z, n1x, n1y, n1z);
    const float n0s = dot3(n0x, n0y, n0z, n0x, n0y, n0z), n1s = dot3(n1x, n1y, n1z, n1x, n1y, n1z);
    if (el2 <= kEps || n0s <= kEps || n1s <= kEps) {
        out[0] = out[1] = out[2] = out[3] = z4;
        if (a.first) a.lam_b[ci] = 0.0f;
        return;
    }
    const float el = fsqrt(el2);
    const float iel = fdiv(1.0f, el);
    const float s0x = fdiv(n0x, n0s), s0y = fdiv(n0y, n0s), s0z = fdiv(n0z, n0s);
    const float s1x = fdiv(n1x, n1s), s1y = fdiv(n1y, n1s), s1z = fdiv(n1z, n1s);
    const float f00 = dot3(o0m.x - e1m.x, o0m.y - e1m.y, o0m.z - e1m.z, ex, ey, ez) * iel;
    const float f01 = dot3(o1m.x - e1m.x, o1m.y - e1m.y, o1m.z - e1m.z, ex, ey, ez) * iel;
    const float f10 = dot3(e0m.x - o0m.x, e0m.y - o0m.y, e0m.z - o0m.z, ex, ey, ez) * iel;
    const float f11 = dot3(e0m.x - o1m.x, e0m.y - o1m.y, e0m.z - o1m.z, ex, ey, ez) * iel;
    const float g0x = -(f00 * s0x + f01 * s1x), g0y = -(f00 * s0y + f01 * s1y), g0z = -(f00 * s0z + f01 * s1z);
    const float g1x = -(f10 * s0x + f11 * s1x), g1y = -(f10 * s0y + f11 * s1y), g1z = -(f10 * s0z + f11 * s1z);
    const float g2x = -(el * s0x), g2y = -(el * s0y), g2z = -(el * s0z);
    const float g3x = -(el * s1x), g3y = -(el * s1y), g3z = -(el * s1z);
    const float r0 = fsqrt(n0s), r1 = fsqrt(n1s);
    const float u0x = fdiv(n0x, r0), u0y = fdiv(n0y, r0), u0z = fdiv(n0z, r0);
    const float u1x = fdiv(n1x, r1), u1y = fdiv(n1y, r1), u1z = fdiv(n1z, r1);
    float cx, cy, cz;
    cross3d(u0x, u0y, u0z, u1x, u1y, u1z, cx, cy, cz);
    const float theta = cephes_atan2f(dot3(ex * iel, ey * iel, ez * iel, cx, cy, cz),
                                      fclamp(dot3(u0x, u0y, u0z, u1x, u1y, u1z), -1.0f, 1.0f));
    const float two_pi = 6.283185307179586f, pi = 3.141592653589793f;
    float C = theta - D.rest_angle;
    C = C - two_pi * floorf(fdiv(C + pi, two_pi));
    const float w0 = e0m.w, w1 = e1m.w, w2 = o0m.w, w3 = o1m.w;
    const float den = w0 * dot3(g0x, g0y, g0z, g0x, g0y, g0z) + w1 * dot3(g1x, g1y, g1z, g1x, g1y, g1z) +
                      w2 * dot3(g2x, g2y, g2z, g2x, g2y, g2z) + w3 * dot3(g3x, g3y, g3z, g3x, g3y, g3z);
    const float alpha = fdiv(D.compliance, fmaxd(a.dt * a.dt, kEps));
    if (den + alpha <= kEps) {
        out[0] = out[1] = out[2] = out[3] = z4;
        if (a.first) a.lam_b[ci] = 0.0f;
        return;
    }
    const float lambda = a.first ? 0.0f : a.lam_b[ci];
    const float dl = fdiv(-(C + alpha * lambda), den + alpha);
    a.lam_b[ci] = lambda + dl;
    const float f0 = w0 * dl, f1 = w1 * dl, f2 = w2 * dl, f3 = w3 * dl, r = a.relax;
    out[0] = rec3(f0 * g0x * r, f0 * g0y * r, f0 * g0z * r);
    out[1] = rec3(f1 * g1x * r, f1 * g1y * r, f1 * g1z * r);
    out[2] = rec3(f2 * g2x * r, f2 * g2y * r, f2 * g2z * r);
    out[3] = rec3(f3 * g3x * r, f3 * g3y * r, f3 * g3z * r);
}

__global__ __launch_bounds__(kSoftBlock) void k_solve_tets(SoftArgs a)
{
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x) * kSoftBlock + threadIdx.x;
    if (t >= a.nt) return;
    const TetDev T = a.tets[t];
    if (!live(a, T.p[0])) return;
    const float4 p0 = ldp(a, T.p[0]), p1 = ldp(a, T.p[1]), p2 = ldp(a, T.p[2]), p3 = ldp(a, T.p[3]);
    const float w0 = p0.w, w1 = p1.w, w2 = p2.w, w3 = p3.w;
    if ((w0 + w1 + w2 + w3) <= kEps) {  // the record keeps its previous value
        if (a.first) a.lam_t[t] = 0.0f;
        return;
    }
    float cx, cy, cz;
    cross3d(p1.x - p0.x, p1.y - p0.y, p1.z - p0.z, p2.x - p0.x, p2.y - p0.y, p2.z - p0.z, cx, cy, cz);
    const float sv = fdiv(dot3(cx, cy, cz, p3.x - p0.x, p3.y - p0.y, p3.z - p0.z), 6.0f);
    const float sgn = sv >= 0.0f ? 1.0f : -1.0f;
    const float C = fabsf(sv) - T.rest_volume;
    const float f = fdiv(sgn, 6.0f);
    float g0x, g0y, g0z, g1x, g1y, g1z, g2x, g2y, g2z, g3x, g3y, g3z;
    cross3d(p3.x - p1.x, p3.y - p1.y, p3.z - p1.z, p2.x - p1.x, p2.y - p1.y, p2.z - p1.z, g0x, g0y, g0z);
    cross3d(p2.x - p0.x, p2.y - p0.y, p2.z - p0.z, p3.x - p0.x, p3.y - p0.y, p3.z - p0.z, g1x, g1y, g1z);
    cross3d(p3.x - p0.x, p3.y - p0.y, p3.z - p0.z, p1.x - p0.x, p1.y - p0.y, p1.z - p0.z, g2x, g2y, g2z);
    cross3d(p1.x - p0.x, p1.y - p0.y, p1.z - p0.z, p2.x - p0.x, p2.y - p0.y, p2.z - p0.z, g3x, g3y, g3z);
    g0x = g0x * f; g0y = g0y * f; g0z = g0z * f;
    g1x = g1x * f; g1y = g1y * f; g1z = g1z * f;
    g2x = g2x * f; g2y = g2y * f; g2z = g2z * f;
    g3x = g3x * f; g3y = g3y * f; g3z = g3z * f;
    const float sg = w0 * dot3(g0x, g0y, g0z, g0x, g0y, g0z) + w1 * dot3(g1x, g1y, g1z, g1x, g1y, g1z) +
                     w2 * dot3(g2x, g2y, g2z, g2x, g2y, g2z) + w3 * dot3(g3x, g3y, g3z, g3x, g3y, g3z);
    const float alpha = fdiv(T.compliance, fmaxd(a.dt * a.dt, kEps));
    const float lambda = a.first ? 0.0f : a.lam_t[t];
    const float den = sg + alpha;
    if (den <= kEps) {
        if (a.first) a.lam_t[t] = 0.0f;
        return;
    }
    const float dl = fdiv(-(C + alpha * lambda), den);
    a.lam_t[t] = lambda + dl;
    const float f0 = w0 * dl, f1 = w1 * dl, f2 = w2 * dl, f3 = w3 * dl, r = a.relax;
    const RecOut out{a.corr_t, a.map_t + 4 * (size_t)t};
    out[0] = rec3(f0 * g0x * r, f0 * g0y * r, f0 * g0z * r);
    out[1] = rec3(f1 * g1x * r, f1 * g1y * r, f1 * g1z * r);
    out[2] = rec3(f2 * g2x * r, f2 * g2y * r, f2 * g2z * r);
    out[3] = rec3(f3 * g3x * r, f3 * g3y * r, f3 * g3z * r);
}

// Per constrained particle, in particle-index order: its incident records, stored contiguously in incidence
// order, summed in that order, then added to its position.
__global__ __launch_bounds__(kSoftBlock) void k_apply_corrections(SoftArgs a, const uint32_t *__restrict__ start,
                                                              const CorrRec *__restrict__ rec)
{
    const uint32_t j = xcd_tile(blockIdx.x, gridDim.x) * kSoftBlock + threadIdx.x;
    if (j >= a.nsoft) return;
    const uint32_t pid = a.soft_ids[j];
    const uint32_t k0 = start[pid], k1 = start[pid + 1];
    if (k0 == k1 || !live(a, pid)) return;
    float4 p = a.posi[pid];
    if (p.w <= kEps) return;
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;
    for (uint32_t k = k0; k < k1; k += 4) {
        CorrRec cr[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) cr[u] = rec[min(k + u, k1 - 1)];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u)
            if (k + u < k1) {
                tx = tx + cr[u].x;
                ty = ty + cr[u].y;
                tz = tz + cr[u].z;
            }
    }
    p.x = p.x + tx;
    p.y = p.y + ty;
    p.z = p.z + tz;
    a.posi[pid] = p;
}

// The soft phase works on the constrained particles in particle-index order, where the
// constraints' vertices are neighbors in memory: gathered from slot order before it, put back after.
__global__ __launch_bounds__(kSoftBlock) void k_soft_gather(SoftArgs a)
{
    const uint32_t j = xcd_tile(blockIdx.x, gridDim.x) * kSoftBlock + threadIdx.x;
    if (j >= a.nsoft) return;
    const uint32_t pid = a.soft_ids[j];
    if (!live(a, pid)) return;  // a scene left out of the step may not be in the slots stepped
    a.posi[pid] = a.pos[a.slot_of[pid]];
}

__global__ __launch_bounds__(kSoftBlock) void k_soft_scatter(SoftArgs a)
{
    const uint32_t j = xcd_tile(blockIdx.x, gridDim.x) * kSoftBlock + threadIdx.x;
    if (j >= a.nsoft) return;
    const uint32_t pid = a.soft_ids[j];
    if (!live(a, pid)) return;
    a.pos[a.slot_of[pid]] = a.posi[pid];
}

void launch_soft_iteration(const SoftArgs &a, bool gather, bool scatter, bool solve, hipStream_t st)
{
    // physics_solver.cpp:915-1019: edges solve/apply, bends + dihedrals solve/apply, tets solve/apply
    if (a.nsoft && gather) hipLaunchKernelGGL(k_soft_gather, dim3(grid_n(a.nsoft)), dim3(kSoftBlock), 0, st, a);
    if (!solve) {
        if (a.nsoft && scatter) hipLaunchKernelGGL(k_soft_scatter, dim3(grid_n(a.nsoft)), dim3(kSoftBlock), 0, st, a);
        return;
    }
    // the applies run over the constrained particles; a group without constraints has nothing to apply
    const uint32_t g = grid_n(a.nsoft);
    if (a.ne) {
        hipLaunchKernelGGL(k_solve_edges, dim3(grid_n(a.ne)), dim3(kSoftBlock), 0, st, a);
        hipLaunchKernelGGL(k_apply_corrections, dim3(g), dim3(kSoftBlock), 0, st, a, a.inc_e_start, a.corr_e);
    }
    if (a.nb) hipLaunchKernelGGL(k_solve_bends, dim3(grid_n(a.nb)), dim3(kSoftBlock), 0, st, a);
    if (a.nd) hipLaunchKernelGGL(k_solve_dihedrals, dim3(grid_n(a.nd)), dim3(kSoftBlock), 0, st, a);
    if (a.nb + a.nd)
        hipLaunchKernelGGL(k_apply_corrections, dim3(g), dim3(kSoftBlock), 0, st, a, a.inc_b_start, a.corr_b);
    if (a.nt) {
        hipLaunchKernelGGL(k_solve_tets, dim3(grid_n(a.nt)), dim3(kSoftBlock), 0, st, a);
        hipLaunchKernelGGL(k_apply_corrections, dim3(g), dim3(kSoftBlock), 0, st, a, a.inc_t_start, a.corr_t);
    }
    if (a.nsoft && scatter) hipLaunchKernelGGL(k_soft_scatter, dim3(grid_n(a.nsoft)), dim3(kSoftBlock), 0, st, a);
}

}  // namespace cn
