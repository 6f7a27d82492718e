// cn_contact.hip -- particle-particle contacts on gfx950.
//
// Reference passes: buildParticleParticleCandidatePairs (physics_pass_dispatcher.cpp:1102-1197,
// physics_particle_{count,emit}_particle_particle_candidates.cs.hlsl with
// IsValidParticleParticleCandidate, physics_particle_broad_phase_query.hlsli:65-117) once per
// substep from the predicted positions; per contact iteration
// physics_particle_generate_explicit_contacts.cs.hlsl:80-189, the compaction of active contacts,
// physics_particle_solve_explicit_contacts.cs.hlsl:25-246 and
// physics_particle_apply_position_corrections.cs.hlsl:8-30; after the velocity update
// physics_particle_solve_contact_velocities.cs.hlsl:26-170 and
// physics_particle_apply_contact_velocities.cs.hlsl per velocity iteration.
//
// The reference emits each pair once (lower particle index first) and accumulates both sides'
// corrections with float CAS atomics, in no fixed order.  Here every particle keeps its candidate
// partners in its own 27-cell walk order, evaluates each pair's contact itself (the arithmetic of
// the pair is the same from either side: lower index first) and sums its side's corrections in
// that order, so results are deterministic and equal to the oracle's.  Contacts are Jacobi: a
// solve pass writes every particle's summed correction, an apply pass adds them.  Particle-
// particle contacts are those of free particles (no rigid proxies, no suturing).  Solids of one
// phase group pair up only when the group collides with itself, minus constraint-adjacent pairs.
//
// Particle-rigid contacts: solid particles against the scene's static colliders, every collider a
// static body of its own (buildParticleRigidCandidatePairs with
// physics_particle_count_particle_rigid_candidates.cs.hlsl:112-151; per contact iteration
// physics_particle_rigid_generate_contacts.cs.hlsl:30-190 and
// physics_particle_rigid_solve_contacts.cs.hlsl:26-140 into the same correction, after the
// particle-particle terms; after the particle-particle velocity iterations
// physics_particle_rigid_solve_contact_velocities.cs.hlsl:27-130 with its apply per iteration).
// The candidates are the particle's collider bitmask of the reorder pass, in LBVH visit order.  A
// static body has zero inverse mass and inertia and does not move: its effective mass is zero and
// its contact point's displacement and velocity are exactly zero.  The velocity pass regenerates
// each contact from the position of the last contact iteration, which is kept per particle.
#include <hip/hip_runtime.h>

#include "cn_internal.h"
#include "cn_walk.h"

namespace cn {

constexpr uint32_t kPhaseGroupMask = 0x7fffffffu;
constexpr uint32_t kPhaseSelfCollide = 0x80000000u;
constexpr int kCfgSoftContactRelax = 2;
constexpr int kCfgSoftMaxCorrection = 3;
constexpr int kCfgRestitutionVelocity = 4;
constexpr int kCfgRestitutionPenetration = 5;

struct ContactArgs {
    float4 *pos;              // slot order
    float4 *vel;
    const float4 *prev;
    float4 *cur;              // fluid solver positions x* + delta
    const float4 *dacc;       // current fluid delta
    const uint32_t *cls;
    const uint32_t *id;
    const ParticleClass *classes;
    const float4 *part_mats;
    const uint32_t *active;   // per scene
    const SceneInfo *scenes;
    RunTable rt;
    uint32_t *plist;          // partner slots, k-major: plist[k * n + t]
    uint32_t *pcount;
    float4 *crec;             // contact of each partner: (normal, penetration), k-major
    uint32_t *cbits;          // bit k: partner k's contact is active
    float4 *corr;             // summed correction of the pass
    uint32_t *meta;
    const uint32_t *adj_off;  // constraint adjacency by particle index (null: none)
    const uint32_t *adj_cnt;
    const uint32_t *adj_idx;
    const Collider *col;      // static colliders in visit order (SceneInfo.col0 per scene)
    const uint32_t *bnd;      // per slot: candidate colliders (a solid's rigid candidates)
    const uint16_t *bndl;     // their lists (scenes of many colliders)
    float4 *rgen;             // per slot: position of the last contact generation
    uint32_t n, pcap;
};

__device__ __forceinline__ const ParticleClass &pclass(const ContactArgs &a, uint32_t slot)
{
    return a.classes[a.cls[slot]];
}

// IsAdjacentParticle, physics_particle_broad_phase_query.hlsli:50-62: particle index `other` among
// the adjacency list of particle index `self`.
__device__ __forceinline__ bool adjacent(const ContactArgs &a, uint32_t self, uint32_t other)
{
    if (!a.adj_off) return false;
    const uint32_t off = a.adj_off[self], cnt = a.adj_cnt[self];
    for (uint32_t k = 0; k < cnt; ++k)
        if (a.adj_idx[off + k] == other) return true;
    return false;
}

// IsValidParticleParticleCandidate for the pair (lo, hi) = slots ordered by particle index (lo is
// the reference's particleIndex, the lower one).
__device__ __forceinline__ bool pair_valid(const ContactArgs &a, uint32_t lo, uint32_t hi)
{
    const ParticleClass &A = pclass(a, lo), &B = pclass(a, hi);
    if (A.env != B.env) return false;
    if ((A.mask & B.layer) == 0u || (B.mask & A.layer) == 0u) return false;
    if (A.kind == CN_PARTICLE_KIND_FLUID && B.kind == CN_PARTICLE_KIND_FLUID) return false;
    // one phase group: only when both have self collision on and they share no constraint
    if (A.kind == CN_PARTICLE_KIND_SOLID && B.kind == CN_PARTICLE_KIND_SOLID &&
        (A.phase & kPhaseGroupMask) == (B.phase & kPhaseGroupMask)) {
        if (!(A.phase & kPhaseSelfCollide) || !(B.phase & kPhaseSelfCollide) || adjacent(a, a.id[lo], a.id[hi]))
            return false;
    }
    const float4 pa = a.pos[lo], pb = a.pos[hi];
    const float dx = pb.x - pa.x, dy = pb.y - pa.y, dz = pb.z - pa.z;
    const float cr = A.radius + B.radius;
    return dot3(dx, dy, dz, dx, dy, dz) <= cr * cr;
}

__device__ __forceinline__ void order_pair(const ContactArgs &a, uint32_t t, uint32_t q, uint32_t &lo, uint32_t &hi,
                                           bool &t_is_hi)
{
    t_is_hi = a.id[q] < a.id[t];
    lo = t_is_hi ? q : t;
    hi = t_is_hi ? t : q;
}

// CombineContactMaterial, physics_rigid_solver_shared.hlsli:37-43: (friction, restitution, static)
__device__ __forceinline__ float3 combine_material(const ContactArgs &a, uint32_t lo, uint32_t hi)
{
    const float4 ma = a.part_mats[pclass(a, lo).part_mat], mb = a.part_mats[pclass(a, hi).part_mat];
    return make_float3(fsqrt(fmaxd(0.0f, ma.x) * fmaxd(0.0f, mb.x)), fmaxd(fmaxd(0.0f, ma.y), fmaxd(0.0f, mb.y)),
                       fsqrt(fmaxd(0.0f, ma.w) * fmaxd(0.0f, mb.w)));
}

// SafeNormalize, physics_math.hlsli:6-12, with rsqrt as 1 / sqrt
__device__ __forceinline__ float3 safe_normalize(float x, float y, float z)
{
    const float l2 = dot3(x, y, z, x, y, z);
    if (l2 <= kEps) return make_float3(0.0f, 1.0f, 0.0f);
    const float r = fdiv(1.0f, fsqrt(l2));
    return make_float3(x * r, y * r, z * r);
}

// ComputePositionFrictionDelta, physics_rigid_solver_shared.hlsli:50-67
__device__ __forceinline__ float3 friction_delta(const float *cfg, float ex, float ey, float ez, float pen, float kin,
                                                 float stat)
{
    const float td = len3(ex, ey, ez);
    if (td <= cfg[kCfgPosFrictionMin] || pen <= kEps) return make_float3(0.0f, 0.0f, 0.0f);
    const float sl = fsat(stat) * pen;
    if (td <= sl) return make_float3(ex, ey, ez);
    const float ks = fmind(fdiv(fsat(kin) * pen, td), 1.0f);
    return make_float3(ex * ks, ey * ks, ez * ks);
}

// physics_particle_rigid_generate_contacts.cs.hlsl:82-176 for one static collider: the position
// in the collider frame, the shape's signed distance and local normal, penetration = radius -
// distance.  Returns true with (world normal, penetration) when the penetration is positive; the
// world normal is SafeNormalize of the rotated local normal.
__device__ __forceinline__ bool rigid_contact(const Collider &c, float px, float py, float pz, float radius,
                                              float4 &rec)
{
    if (!c.enabled) return false;
    float ux, uy, uz;
    quat_rotate(-c.q[0], -c.q[1], -c.q[2], c.q[3], __fsub_rn(px, c.c[0]), __fsub_rn(py, c.c[1]),
                __fsub_rn(pz, c.c[2]), ux, uy, uz);  // QuaternionInverseRotate
    float nx = 0.0f, ny = 1.0f, nz = 0.0f, sd = 0.0f;
    if (c.shape == CN_SHAPE_SPHERE) {
        const float len = len3(ux, uy, uz);
        if (len > kEps) {
            nx = fdiv(ux, len); ny = fdiv(uy, len); nz = fdiv(uz, len);
        }
        sd = __fsub_rn(len, c.he[0]);
    } else if (c.shape == CN_SHAPE_BOX) {
        box_local_contact(c.he, ux, uy, uz, sd, nx, ny, nz);
    } else {
        // capsule: the closest point of the world segment; the world normal (fallback +x) is taken
        // into the collider frame
        const float gx = __fsub_rn(c.sb[0], c.sa[0]), gy = __fsub_rn(c.sb[1], c.sa[1]), gz = __fsub_rn(c.sb[2], c.sa[2]);
        const float lsq = dot3(gx, gy, gz, gx, gy, gz);
        float t = 0.0f;
        if (lsq > kEps) {
            const float wx = __fsub_rn(px, c.sa[0]), wy = __fsub_rn(py, c.sa[1]), wz = __fsub_rn(pz, c.sa[2]);
            t = fsat(fdiv(dot3(wx, wy, wz, gx, gy, gz), lsq));
        }
        const float ex = __fsub_rn(px, __fadd_rn(c.sa[0], __fmul_rn(t, gx)));
        const float ey = __fsub_rn(py, __fadd_rn(c.sa[1], __fmul_rn(t, gy)));
        const float ez = __fsub_rn(pz, __fadd_rn(c.sa[2], __fmul_rn(t, gz)));
        const float len = len3(ex, ey, ez);
        float wx = 1.0f, wy = 0.0f, wz = 0.0f;
        if (len > kEps) {
            wx = fdiv(ex, len); wy = fdiv(ey, len); wz = fdiv(ez, len);
        }
        sd = __fsub_rn(len, c.he[0]);
        quat_rotate(-c.q[0], -c.q[1], -c.q[2], c.q[3], wx, wy, wz, nx, ny, nz);
    }
    const float pen = __fsub_rn(radius, sd);
    if (!(pen > 0.0f)) return false;
    float wx, wy, wz;
    quat_rotate(c.q[0], c.q[1], c.q[2], c.q[3], nx, ny, nz, wx, wy, wz);
    const float3 n = safe_normalize(wx, wy, wz);
    rec = make_float4(n.x, n.y, n.z, pen);
    return true;
}

// CombineContactMaterial of a particle material and a collider's
__device__ __forceinline__ float3 rigid_material(const float4 pm, const float *cm)
{
    return make_float3(fsqrt(fmaxd(0.0f, pm.x) * fmaxd(0.0f, cm[0])), fmaxd(fmaxd(0.0f, pm.y), fmaxd(0.0f, cm[1])),
                       fsqrt(fmaxd(0.0f, pm.w) * fmaxd(0.0f, cm[3])));
}

// rigid solve_contacts :45-140 against a static body: the position correction of a particle at p
// (inverse mass p.w; pv its position at the start of the substep) with contact rec = (normal,
// penetration) against collider cc; false when the contact adds nothing.  The displacement of a
// static body's contact point is +0, its effective mass 0.
__device__ __forceinline__ bool rigid_position(const float4 p, const float4 pv, const float4 rec, const float4 pm,
                                               const Collider &cc, const SubstepConsts &c, float3 &d)
{
    const float w = p.w;
    if (w <= kEps) return false;
    const float pn = fmind(fmaxd(rec.w - c.cfg[kCfgContactSlop], 0.0f), c.cfg[kCfgSoftMaxCorrection]);
    if (pn <= 0.0f) return false;
    const float3 nr = safe_normalize(rec.x, rec.y, rec.z);
    const float3 mat = rigid_material(pm, cc.material);
    const float denom = w + 0.0f;
    if (denom <= kEps) return false;
    const float relax = c.cfg[kCfgSoftContactRelax];
    const float lambda = fdiv(pn, denom) * relax;
    float cx = nr.x * (w * lambda), cy = nr.y * (w * lambda), cz = nr.z * (w * lambda);
    // relative displacement = (contact point - its previous position) - (x - x_prev), the first
    // term +0 for a static body
    const float rx = 0.0f - (p.x - pv.x), ry = 0.0f - (p.y - pv.y), rz = 0.0f - (p.z - pv.z);
    const float rn = dot3(rx, ry, rz, nr.x, nr.y, nr.z);
    const float3 fd = friction_delta(c.cfg, rx - nr.x * rn, ry - nr.y * rn, rz - nr.z * rn, pn, mat.x, mat.z);
    const float fdist = len3(fd.x, fd.y, fd.z);
    if (fdist > 0.0f) {
        const float tden = w + 0.0f;
        if (tden > kEps) {
            const float tl = fdiv(fdist, tden);
            const float k = (tl * w) * relax;
            cx = cx + fdiv(fd.x, fdist) * k;
            cy = cy + fdiv(fd.y, fdist) * k;
            cz = cz + fdiv(fd.z, fdist) * k;
        }
    }
    d = make_float3(cx, cy, cz);
    return true;
}


// ---------------------------------------------------------------------------------------------
// candidates
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_contact_partners(ContactArgs a, SubstepConsts c)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= a.n) return;
    const uint32_t scene = pclass(a, t).scene;
    if (!a.active[scene]) return;
    uint32_t cnt = 0;
    if (a.scenes[scene].pairs) walk_runs(a.rt.cellc[t], a.rt, c, [&](uint32_t q) {
        if (q == t) return true;
        uint32_t lo, hi;
        bool t_hi;
        order_pair(a, t, q, lo, hi, t_hi);
        if (!pair_valid(a, lo, hi)) return true;
        if (cnt < a.pcap) a.plist[(size_t)cnt * a.n + t] = q;
        ++cnt;
        return true;
    });
    if (cnt > a.pcap) {
        atomicAdd(&a.meta[4], 1u);
        cnt = a.pcap;
    }
    a.pcount[t] = cnt;
}

// ---------------------------------------------------------------------------------------------
// position iteration: generate + solve (sum in partner order), then apply
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_contact_solve(ContactArgs a, SubstepConsts c)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= a.n) return;
    const ParticleClass &pc = pclass(a, t);
    if (!a.active[pc.scene]) return;
    const SceneInfo sc = a.scenes[pc.scene];
    if (!(sc.pairs | sc.rigid)) return;
    const uint32_t cnt = a.pcount[t];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    uint32_t bits = 0;
    const float relax = c.cfg[kCfgSoftContactRelax];
    for (uint32_t k = 0; k < cnt; ++k) {
        const uint32_t q = a.plist[(size_t)k * a.n + t];
        uint32_t lo, hi;
        bool t_hi;
        order_pair(a, t, q, lo, hi, t_hi);
        // generate_explicit_contacts :140-188
        const float4 pa = a.pos[lo], pb = a.pos[hi];
        const float dx = pb.x - pa.x, dy = pb.y - pa.y, dz = pb.z - pa.z;
        const float d2 = dot3(dx, dy, dz, dx, dy, dz);
        const float cr = pclass(a, lo).radius + pclass(a, hi).radius;
        float nx = 0.0f, ny = 0.0f, nz = 0.0f, pen = 0.0f;
        bool on = false;
        if (d2 <= kEps) {
            on = true;
            ny = 1.0f;
            pen = cr;
        } else {
            const float dist = fsqrt(d2);
            const float p0 = cr - dist;
            if (p0 > 0.0f) {
                on = true;
                nx = fdiv(dx, dist);
                ny = fdiv(dy, dist);
                nz = fdiv(dz, dist);
                pen = p0;
            }
        }
        a.crec[(size_t)k * a.n + t] = make_float4(nx, ny, nz, pen);
        if (!on) continue;
        bits |= 1u << k;
        // solve_explicit_contacts :103-246 for a pair of free particles
        const float wa = pa.w, wb = pb.w;
        if (wa <= kEps && wb <= kEps) continue;
        const float pn = fmind(fmaxd(pen - c.cfg[kCfgContactSlop], 0.0f), c.cfg[kCfgSoftMaxCorrection]);
        if (pn <= 0.0f) continue;
        const float3 nr = safe_normalize(nx, ny, nz);
        const float4 va = a.prev[lo], vb = a.prev[hi];
        const float rx = (pb.x - vb.x) - (pa.x - va.x);
        const float ry = (pb.y - vb.y) - (pa.y - va.y);
        const float rz = (pb.z - vb.z) - (pa.z - va.z);
        const float denom = wa + wb;
        if (denom <= kEps) continue;
        const float lambda = fdiv(pn, denom) * relax;
        float cax = -nr.x * (wa * lambda), cay = -nr.y * (wa * lambda), caz = -nr.z * (wa * lambda);
        float cbx = nr.x * (wb * lambda), cby = nr.y * (wb * lambda), cbz = nr.z * (wb * lambda);
        const float rn = dot3(rx, ry, rz, nr.x, nr.y, nr.z);
        const float3 mat = combine_material(a, lo, hi);
        const float3 fd = friction_delta(c.cfg, rx - nr.x * rn, ry - nr.y * rn, rz - nr.z * rn, pn, mat.x, mat.z);
        const float fdist = len3(fd.x, fd.y, fd.z);
        if (fdist > 0.0f) {
            const float tden = wa + wb;
            if (tden > kEps) {
                const float tl = fdiv(fdist, tden) * relax;
                const float tx = fdiv(fd.x, fdist), ty = fdiv(fd.y, fdist), tz = fdiv(fd.z, fdist);
                cax = cax + tx * (wa * tl);
                cay = cay + ty * (wa * tl);
                caz = caz + tz * (wa * tl);
                cbx = cbx - tx * (wb * tl);
                cby = cby - ty * (wb * tl);
                cbz = cbz - tz * (wb * tl);
            }
        }
        if (t_hi) {
            if (wb > kEps) {
                sx = sx + cbx;
                sy = sy + cby;
                sz = sz + cbz;
            }
        } else if (wa > kEps) {
            sx = sx + cax;
            sy = sy + cay;
            sz = sz + caz;
        }
    }
    a.cbits[t] = bits;
    // particle-rigid contacts of a solid particle (rigid solve_contacts :45-140 with a static body)
    const uint32_t rm = pc.kind == CN_PARTICLE_KIND_SOLID ? a.bnd[t] : 0u;
    if (rm) {
        const float4 p = a.pos[t];
        a.rgen[t] = p;
        const float4 pv = a.prev[t];
        const float4 pm = a.part_mats[pc.part_mat];
        each_candidate(rm, a.bndl, t, [&](uint32_t k) {
            const Collider &cc = a.col[sc.col0 + k];
            float4 rec;
            float3 d;
            if (!rigid_contact(cc, p.x, p.y, p.z, pc.radius, rec)) return;
            if (!rigid_position(p, pv, rec, pm, cc, c, d)) return;
            sx = sx + d.x;
            sy = sy + d.y;
            sz = sz + d.z;
        });
    }
    a.corr[t] = make_float4(sx, sy, sz, 0.0f);
}

// apply_position_corrections :8-30 (every free particle; those without contacts add zero); a fluid
// particle's solver position follows its new x*
__global__ __launch_bounds__(kBlock) void k_contact_apply(ContactArgs a)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= a.n) return;
    const ParticleClass &pc = pclass(a, t);
    if (!a.active[pc.scene]) return;
    const SceneInfo &sc = a.scenes[pc.scene];
    if (!(sc.pairs | sc.rigid)) return;
    float4 p = a.pos[t];
    if (p.w <= kEps) return;
    const float4 d = a.corr[t];
    p.x = p.x + d.x;
    p.y = p.y + d.y;
    p.z = p.z + d.z;
    a.pos[t] = p;
    if (pc.kind == CN_PARTICLE_KIND_FLUID) {
        const float4 dl = a.dacc[t];
        a.cur[t] = make_float4(p.x + dl.x, p.y + dl.y, p.z + dl.z, p.w);
    }
}

// ---------------------------------------------------------------------------------------------
// velocity iteration over the last generated contacts
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_contact_velocity(ContactArgs a, SubstepConsts c)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= a.n) return;
    const uint32_t scene = pclass(a, t).scene;
    if (!a.active[scene] || !a.scenes[scene].pairs) return;
    const uint32_t cnt = a.pcount[t], bits = a.cbits[t];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (uint32_t k = 0; k < cnt; ++k) {
        if (!((bits >> k) & 1u)) continue;
        const uint32_t q = a.plist[(size_t)k * a.n + t];
        uint32_t lo, hi;
        bool t_hi;
        order_pair(a, t, q, lo, hi, t_hi);
        const float wa = a.pos[lo].w, wb = a.pos[hi].w;
        if (wa <= kEps && wb <= kEps) continue;
        const float4 rec = a.crec[(size_t)k * a.n + t];
        const float3 nr = safe_normalize(rec.x, rec.y, rec.z);
        const float4 va = a.vel[lo], vb = a.vel[hi];
        const float nv = dot3(vb.x - va.x, vb.y - va.y, vb.z - va.z, nr.x, nr.y, nr.z);
        if (nv >= 0.0f) continue;
        const float denom = wa + wb;
        if (denom <= kEps) continue;
        const bool rest_on = (-nv > c.cfg[kCfgRestitutionVelocity]) &&
                             (rec.w <= c.cfg[kCfgRestitutionPenetration] * c.cfg[kCfgContactSlop]);
        const float restitution = rest_on ? fsat(combine_material(a, lo, hi).y) : 0.0f;
        const float desired = rest_on ? (-restitution * nv) : 0.0f;
        const float imp = fmaxd(0.0f, fdiv(desired - nv, denom));
        const float tx = nr.x * imp, ty = nr.y * imp, tz = nr.z * imp;
        if (t_hi) {
            if (wb > kEps) {
                sx = sx + tx * wb;
                sy = sy + ty * wb;
                sz = sz + tz * wb;
            }
        } else if (wa > kEps) {
            sx = sx + -tx * wa;
            sy = sy + -ty * wa;
            sz = sz + -tz * wa;
        }
    }
    a.corr[t] = make_float4(sx, sy, sz, 0.0f);
}

// apply_contact_velocities: v += correction for free particles
__global__ __launch_bounds__(kBlock) void k_contact_velocity_apply(ContactArgs a)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= a.n) return;
    const uint32_t scene = pclass(a, t).scene;
    if (!a.active[scene] || !a.scenes[scene].pairs) return;
    if (a.pos[t].w <= kEps) return;
    float4 v = a.vel[t];
    const float4 d = a.corr[t];
    v.x = v.x + d.x;
    v.y = v.y + d.y;
    v.z = v.z + d.z;
    a.vel[t] = v;
}

// The particle-rigid velocity iterations (rigid solve_contact_velocities :44-128 with a static
// body, each followed by apply_contact_velocities): a particle's correction reads only its own
// velocity and the static colliders, so all iterations of a particle run in one thread, each
// adding its summed correction (zero without contacts) to a free particle's velocity.
__global__ __launch_bounds__(kBlock) void k_rigid_velocity(ContactArgs a, SubstepConsts c, uint32_t iters)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= a.n) return;
    const ParticleClass &pc = pclass(a, t);
    if (!a.active[pc.scene]) return;
    const SceneInfo sc = a.scenes[pc.scene];
    if (!(sc.rigid | sc.rigid_pid)) return;
    const float w = a.pos[t].w;
    if (w <= kEps) return;
    const uint32_t rm = pc.kind == CN_PARTICLE_KIND_SOLID ? a.bnd[t] : 0u;
    const float4 g = rm ? a.rgen[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 pm = a.part_mats[pc.part_mat];
    float4 v = a.vel[t];
    for (uint32_t it = 0; it < iters; ++it) {
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        each_candidate(rm, a.bndl, t, [&](uint32_t k) {
            const Collider &cc = a.col[sc.col0 + k];
            float4 rec;
            if (!rigid_contact(cc, g.x, g.y, g.z, pc.radius, rec)) return;
            const float3 nr = safe_normalize(rec.x, rec.y, rec.z);
            // the contact point's velocity is +0, so the relative velocity is v - 0 = v
            const float nv = dot3(v.x, v.y, v.z, nr.x, nr.y, nr.z);
            if (nv >= 0.0f) return;
            const float denom = w + 0.0f;
            if (denom <= kEps) return;
            const bool rest_on = (-nv > c.cfg[kCfgRestitutionVelocity]) &&
                                 (rec.w <= c.cfg[kCfgRestitutionPenetration] * c.cfg[kCfgContactSlop]);
            const float restitution = rest_on ? fsat(rigid_material(pm, cc.material).y) : 0.0f;
            const float desired = rest_on ? (-restitution * nv) : 0.0f;
            const float imp = fmaxd(0.0f, fdiv(desired - nv, denom));
            sx = sx + (nr.x * imp) * w;
            sy = sy + (nr.y * imp) * w;
            sz = sz + (nr.z * imp) * w;
        });
        v.x = v.x + sx;
        v.y = v.y + sy;
        v.z = v.z + sz;
    }
    a.vel[t] = v;
}

// The particle-rigid contacts of a scene whose every particle is constrained and which holds no
// particle-particle pair, in particle-index order on the soft phase's positions right after its
// last apply of the iteration (the reference's contact pass of the iteration, for such a scene):
// each candidate's contact at the current position, the static-body solve, their sum (zero without
// contacts) added to a free particle.  The scene then needs no slot-order contact pass, and its
// constrained positions stay in particle-index order across the iteration.
struct RigidPidArgs {
    float4 *posi;
    const uint32_t *soft_ids;
    uint32_t nsoft;
    const uint32_t *slot_of;
    const uint32_t *pid_scene;
    const uint32_t *active;
    const SceneInfo *scenes;
    const Collider *col;
    const uint32_t *bnd;      // slot order
    const uint16_t *bndl;
    const float4 *prev;       // slot order
    const uint32_t *cls;      // slot order
    const ParticleClass *classes;
    const float4 *part_mats;
    float4 *rgen;             // slot order
    uint32_t *pick;           // per constrained particle: kPickSkip, kPickNone or its slot
    uint32_t npid;
};

// What the iterations of a substep do for a constrained particle, fixed for the substep (slot,
// class, collider candidates): nothing (its scene is not stepped or not solved here), add a zero
// sum (no candidates), or solve the candidates of its slot.
constexpr uint32_t kPickSkip = 0xffffffffu, kPickNone = 0xfffffffeu;

__global__ __launch_bounds__(kBlock) void k_rigid_pick(RigidPidArgs a, SubstepConsts c)
{
    const uint32_t j = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (j >= a.nsoft) return;
    const uint32_t pid = a.soft_ids[j];
    const uint32_t scene = a.pid_scene[pid];
    uint32_t v = kPickSkip;
    if (a.active[scene] && a.scenes[scene].rigid_pid) {
        const uint32_t t = a.slot_of[min(pid, a.npid - 1u)];
        const ParticleClass &pc = a.classes[c.uniform_class ? 0u : a.cls[t]];
        v = pc.kind == CN_PARTICLE_KIND_SOLID && a.bnd[t] != 0u ? t : kPickNone;
    }
    a.pick[j] = v;
}

__global__ __launch_bounds__(kBlock) void k_rigid_pid(RigidPidArgs a, SubstepConsts c)
{
    const uint32_t j = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (j >= a.nsoft) return;
    const uint32_t v = a.pick[j];
    if (v == kPickSkip) return;
    const uint32_t pid = a.soft_ids[j];
    float4 p = a.posi[pid];
    if (p.w <= kEps) return;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    if (v != kPickNone) {
        const uint32_t t = v, rm = a.bnd[t];
        const ParticleClass &pc = a.classes[c.uniform_class ? 0u : a.cls[t]];
        const SceneInfo &sc = a.scenes[a.pid_scene[pid]];
        a.rgen[t] = p;
        const float4 pv = a.prev[t];
        const float4 pm = a.part_mats[pc.part_mat];
        each_candidate(rm, a.bndl, t, [&](uint32_t k) {
            const Collider &cc = a.col[sc.col0 + k];
            float4 rec;
            float3 d;
            if (!rigid_contact(cc, p.x, p.y, p.z, pc.radius, rec)) return;
            if (!rigid_position(p, pv, rec, pm, cc, c, d)) return;
            sx = sx + d.x;
            sy = sy + d.y;
            sz = sz + d.z;
        });
    }
    p.x = p.x + sx;
    p.y = p.y + sy;
    p.z = p.z + sz;
    a.posi[pid] = p;
}

static ContactArgs contact_args(const SceneBuffers &s, const SubstepConsts &c, int din)
{
    ContactArgs a;
    const int cur = s.cur;
    a.pos = s.pos[cur];
    a.vel = s.vel[cur];
    a.prev = s.prev[cur];
    a.cur = s.curpos;
    a.dacc = s.dacc[din];
    a.cls = s.cls[cur];
    a.id = s.id[cur];
    a.classes = s.classes;
    a.part_mats = s.part_mats;
    a.active = s.scene_active;
    a.scenes = s.scenes;
    a.rt = RunTable{s.run_key, s.run_start, s.run_info, s.bucket, c.bucket_shift, s.cellc};
    a.plist = s.ct_list;
    a.pcount = s.ct_count;
    a.crec = s.ct_rec;
    a.cbits = s.ct_bits;
    a.corr = s.ct_corr;
    a.meta = s.meta;
    a.adj_off = s.adj_off;
    a.adj_cnt = s.adj_cnt;
    a.adj_idx = s.adj_idx;
    a.col = s.colliders;
    a.bnd = s.bnd;
    a.bndl = s.bndl;
    a.rgen = s.ct_gen;
    a.n = c.n;
    a.pcap = kContactPartners;
    return a;
}

static uint32_t grid_c(uint32_t n) { return (n + kBlock - 1) / kBlock; }

void launch_contact_partners(const SceneBuffers &s, const SubstepConsts &c, hipStream_t st)
{
    const ContactArgs a = contact_args(s, c, 0);
    hipLaunchKernelGGL(k_contact_partners, dim3(grid_c(c.n)), dim3(kBlock), 0, st, a, c);
}

void launch_contact_iteration(const SceneBuffers &s, const SubstepConsts &c, int din, hipStream_t st)
{
    const ContactArgs a = contact_args(s, c, din);
    hipLaunchKernelGGL(k_contact_solve, dim3(grid_c(c.n)), dim3(kBlock), 0, st, a, c);
    hipLaunchKernelGGL(k_contact_apply, dim3(grid_c(c.n)), dim3(kBlock), 0, st, a);
}

void launch_contact_velocity(const SceneBuffers &s, const SubstepConsts &c, hipStream_t st)
{
    const ContactArgs a = contact_args(s, c, 0);
    hipLaunchKernelGGL(k_contact_velocity, dim3(grid_c(c.n)), dim3(kBlock), 0, st, a, c);
    hipLaunchKernelGGL(k_contact_velocity_apply, dim3(grid_c(c.n)), dim3(kBlock), 0, st, a);
}

void launch_rigid_pid(const SceneBuffers &s, const SubstepConsts &c, bool pick, hipStream_t st)
{
    if (!s.nsoft) return;
    RigidPidArgs a;
    a.posi = s.posi;
    a.soft_ids = s.soft_ids;
    a.nsoft = s.nsoft;
    a.slot_of = s.slot_of;
    a.pid_scene = s.pid_scene;
    a.active = s.scene_active;
    a.scenes = s.scenes;
    a.col = s.colliders;
    a.bnd = s.bnd;
    a.bndl = s.bndl;
    a.prev = s.prev[s.cur];
    a.cls = s.cls[s.cur];
    a.classes = s.classes;
    a.part_mats = s.part_mats;
    a.rgen = s.ct_gen;
    a.pick = s.rigid_pick;
    a.npid = s.npid;
    if (pick) hipLaunchKernelGGL(k_rigid_pick, dim3(grid_c(s.nsoft)), dim3(kBlock), 0, st, a, c);
    hipLaunchKernelGGL(k_rigid_pid, dim3(grid_c(s.nsoft)), dim3(kBlock), 0, st, a, c);
}

void launch_rigid_velocity(const SceneBuffers &s, const SubstepConsts &c, uint32_t iters, hipStream_t st)
{
    const ContactArgs a = contact_args(s, c, 0);
    hipLaunchKernelGGL(k_rigid_velocity, dim3(grid_c(c.n)), dim3(kBlock), 0, st, a, c, iters);
}

}  // namespace cn
