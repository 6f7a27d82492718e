// cn_physics.cpp -- the C++ host side of the particle path (see cn_physics.hpp).
#include "cn_physics.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <unordered_set>

namespace cressim::neo::hip {

namespace {

void logError(const std::string &msg) { std::fprintf(stderr, "[cn] error: %s\n", msg.c_str()); }

// phases: one group per authored part (packParticlePhase, include/physics/particle_phase.h:16-27);
// fluids carry the self-collide flag, soft bodies do not
constexpr std::uint32_t kPhaseSelfCollide = 0x80000000u;

float dot3(const Float3 &a, const Float3 &b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
Float3 sub3(const Float3 &a, const Float3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Float3 cross3(const Float3 &a, const Float3 &b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

}  // namespace

PhysicsSolverDesc::PhysicsSolverDesc()
{
    cn_step_params p;
    cn_default_step_params(&p);
    substeps = p.substeps;
    defaultIterations = p.default_iterations;
    gravity = {p.gravity[0], p.gravity[1], p.gravity[2]};
    maxFluidNeighbors = p.max_fluid_neighbors;
    std::memcpy(solverConfig, p.solver_config, sizeof(solverConfig));
}

std::uint32_t PhysicsWorld::appendParticles(std::size_t count, std::uint32_t kind, float invMass, float radius,
                                            std::uint32_t fluidMaterial, std::uint32_t env, std::uint32_t phase)
{
    const std::uint32_t first = particleCount();
    const std::size_t n = first + count;
    positionsInvMass.resize(n);
    previousPositions.resize(n);
    velocities.resize(n);
    radii.resize(n, radius);
    kinds.resize(n, kind);
    fluidMaterialIndices.resize(n, fluidMaterial);
    materialIndices.resize(n, 0u);
    environmentIndices.resize(n, env);
    phases.resize(n, phase);
    collisionLayers.resize(n, 1u);
    collisionMasks.resize(n, 0xFFFFFFFFu);
    for (std::size_t i = first; i < n; ++i) positionsInvMass[i].w = invMass;
    if (particleContactMaterials.empty()) particleContactMaterials.push_back({0.0f, 0.0f, 0.0f, -1.0f});
    ++revision_;
    return first;
}

// prepareFluidStateForInsert (src/physics/physics_world.cpp:6857-6892) for an untransformed grid
// placed at desc.position, and toFluidSolverMaterial through the library
std::uint32_t PhysicsWorld::addFluid(const FluidDesc &desc)
{
    const float spacing = desc.targetParticleSpacing;
    if (!(spacing > 0.0f) || !(desc.particleMass > 0.0f) || !(desc.particleRadius > 0.0f)) {
        logError("fluid: spacing, mass and radius must be positive");
        return UINT32_MAX;
    }
    cn_fluid_material m;
    if (cn_fluid_material_from_desc(&desc.material, desc.particleRadius, &m) != CN_OK) {
        logError("fluid: material derivation failed");
        return UINT32_MAX;
    }
    const auto cells = [&](float extent) {
        return std::max<std::uint32_t>(1u, static_cast<std::uint32_t>(std::floor(extent / spacing + 0.5f)));
    };
    const std::uint32_t nx = cells(desc.size.x), ny = cells(desc.size.y), nz = cells(desc.size.z);
    fluidMaterials.push_back(m);
    const std::uint32_t first =
        appendParticles(static_cast<std::size_t>(nx) * ny * nz, CN_PARTICLE_KIND_FLUID, 1.0f / desc.particleMass,
                        desc.particleRadius, static_cast<std::uint32_t>(fluidMaterials.size() - 1),
                        desc.environment, nextGroup_++ | kPhaseSelfCollide);
    std::uint32_t i = first;
    for (std::uint32_t z = 0; z < nz; ++z)
        for (std::uint32_t y = 0; y < ny; ++y)
            for (std::uint32_t x = 0; x < nx; ++x, ++i) {
                Float4 &p = positionsInvMass[i];
                p.x = desc.position.x + (-desc.size.x * 0.5f + (static_cast<float>(x) + 0.5f) * spacing);
                p.y = desc.position.y + (-desc.size.y * 0.5f + (static_cast<float>(y) + 0.5f) * spacing);
                p.z = desc.position.z + (-desc.size.z * 0.5f + (static_cast<float>(z) + 0.5f) * spacing);
                previousPositions[i] = p;
            }
    return first;
}

std::uint32_t PhysicsWorld::addSoftBlock(const SoftBlockDesc &desc)
{
    const std::uint32_t rx = desc.resolution[0], ry = desc.resolution[1], rz = desc.resolution[2];
    if (rx < 2 || ry < 2 || rz < 2 || !(desc.particleMass > 0.0f)) {
        logError("soft block: resolution below 2 or non-positive mass");
        return UINT32_MAX;
    }
    const float sx = std::max(desc.size.x, 1.0e-4f), sy = std::max(desc.size.y, 1.0e-4f),
                sz = std::max(desc.size.z, 1.0e-4f);
    const std::uint32_t first =
        appendParticles(static_cast<std::size_t>(rx) * ry * rz, CN_PARTICLE_KIND_SOLID, 1.0f / desc.particleMass,
                        desc.particleRadius, 0xFFFFFFFFu, desc.environment, nextGroup_++);
    const auto flat = [&](std::uint32_t x, std::uint32_t y, std::uint32_t z) { return (x * ry + y) * rz + z; };
    for (std::uint32_t x = 0; x < rx; ++x)
        for (std::uint32_t y = 0; y < ry; ++y)
            for (std::uint32_t z = 0; z < rz; ++z) {
                Float4 &p = positionsInvMass[first + flat(x, y, z)];
                p.x = (static_cast<float>(x) / static_cast<float>(rx - 1) - 0.5f) * sx + desc.position.x;
                p.y = (static_cast<float>(y) / static_cast<float>(ry - 1) - 0.5f) * sy + desc.position.y;
                p.z = (static_cast<float>(z) / static_cast<float>(rz - 1) - 0.5f) * sz + desc.position.z;
                previousPositions[first + flat(x, y, z)] = p;
            }
    // five tets per cube, the split alternating with the cube's (x + y + z) parity
    std::vector<std::uint32_t> tets;
    for (std::uint32_t x = 0; x + 1 < rx; ++x)
        for (std::uint32_t y = 0; y + 1 < ry; ++y)
            for (std::uint32_t z = 0; z + 1 < rz; ++z) {
                const std::uint32_t p0 = flat(x, y, z), p1 = flat(x, y, z + 1), p3 = flat(x + 1, y, z),
                                    p2 = flat(x + 1, y, z + 1), p7 = flat(x + 1, y + 1, z),
                                    p6 = flat(x + 1, y + 1, z + 1), p4 = flat(x, y + 1, z),
                                    p5 = flat(x, y + 1, z + 1);
                const std::uint32_t odd[5][4] = {{p2, p1, p6, p3}, {p6, p3, p4, p7}, {p4, p1, p6, p5},
                                                 {p3, p1, p4, p0}, {p6, p1, p4, p3}};
                const std::uint32_t even[5][4] = {{p0, p2, p5, p1}, {p7, p2, p0, p3}, {p5, p2, p7, p6},
                                                  {p7, p0, p5, p4}, {p0, p2, p7, p5}};
                const auto &t = ((x + y + z) & 1u) ? odd : even;
                for (int k = 0; k < 5; ++k) tets.insert(tets.end(), t[k], t[k] + 4);
            }
    const auto pos3 = [&](std::uint32_t local) {
        const Float4 &p = positionsInvMass[first + local];
        return Float3{p.x, p.y, p.z};
    };
    // edges in first-seen order over each tet's (0,1),(0,2),(0,3),(1,2),(1,3),(2,3) pairs
    static const int kPairs[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    std::unordered_set<std::uint64_t> seen;
    for (std::size_t t = 0; t < tets.size(); t += 4) {
        for (const auto &pr : kPairs) {
            const std::uint32_t a = tets[t + pr[0]], b = tets[t + pr[1]];
            const std::uint32_t lo = std::min(a, b), hi = std::max(a, b);
            if (!seen.insert((static_cast<std::uint64_t>(lo) << 32) | hi).second) continue;
            const Float3 d = sub3(pos3(hi), pos3(lo));
            cn_soft_edge e{};
            e.a = first + lo;
            e.b = first + hi;
            e.rest_length = std::sqrt(dot3(d, d));
            e.compliance = desc.edgeCompliance;
            e.enabled = 1u;
            softEdges.push_back(e);
        }
        const Float3 q0 = pos3(tets[t]), q1 = pos3(tets[t + 1]), q2 = pos3(tets[t + 2]), q3 = pos3(tets[t + 3]);
        cn_soft_tet tt{};
        for (int k = 0; k < 4; ++k) tt.p[k] = first + tets[t + k];
        tt.rest_volume = std::fabs(dot3(cross3(sub3(q1, q0), sub3(q2, q0)), sub3(q3, q0))) / 6.0f;
        tt.compliance = desc.volumeCompliance;
        softTets.push_back(tt);
    }
    return first;
}

void PhysicsWorld::addStaticBox(const StaticBoxDesc &desc)
{
    cn_box_collider c{};
    c.center[0] = desc.center.x;
    c.center[1] = desc.center.y;
    c.center[2] = desc.center.z;
    c.half_extents[0] = desc.halfExtents.x;
    c.half_extents[1] = desc.halfExtents.y;
    c.half_extents[2] = desc.halfExtents.z;
    c.material[0] = desc.material.x;
    c.material[1] = desc.material.y;
    c.material[2] = desc.material.z;
    c.material[3] = desc.material.w;
    c.environment = desc.environment;
    c.layer = 1u;
    c.mask = 0xFFFFFFFFu;
    c.enabled = 1u;
    staticBoxes.push_back(c);
    ++revision_;
}

bool PhysicsSolver::fail(const std::string &what)
{
    error_ = what;
    if (ctx_) {
        const char *e = cn_last_error(ctx_);
        if (e && *e) error_ += std::string(": ") + e;
    }
    logError(error_);
    return false;
}

bool PhysicsSolver::initialize()
{
    if (ctx_) return true;
    cn_device_desc d{device_, 0u};
    if (cn_ctx_create(&d, &ctx_) != CN_OK) {
        ctx_ = nullptr;
        return fail("no HIP device " + std::to_string(device_));
    }
    return true;
}

void PhysicsSolver::shutdown()
{
    if (ctx_) cn_ctx_destroy(ctx_);
    ctx_ = nullptr;
    uploadedRevision_ = ~0ull;
}

bool PhysicsSolver::upload(PhysicsWorld &w)
{
    const std::uint32_t n = w.particleCount();
    cn_particle_soa_view v{};
    v.count = n;
    v.positions_inv_mass = &w.positionsInvMass[0].x;
    v.velocities = &w.velocities[0].x;
    v.radii = w.radii.data();
    v.kinds = w.kinds.data();
    v.fluid_material_indices = w.fluidMaterialIndices.data();
    v.particle_material_indices = w.materialIndices.data();
    v.environment_indices = w.environmentIndices.data();
    v.phases = w.phases.data();
    v.collision_layers = w.collisionLayers.data();
    v.collision_masks = w.collisionMasks.data();
    cn_scene_tables t{};
    t.fluid_material_count = static_cast<std::uint32_t>(w.fluidMaterials.size());
    t.fluid_materials = w.fluidMaterials.data();
    t.particle_material_count = static_cast<std::uint32_t>(w.particleContactMaterials.size());
    t.particle_materials = &w.particleContactMaterials[0].x;
    t.collider_count = static_cast<std::uint32_t>(w.staticBoxes.size());
    t.colliders = w.staticBoxes.data();
    cn_constraint_view c{};
    c.edge_count = static_cast<std::uint32_t>(w.softEdges.size());
    c.edges = w.softEdges.data();
    c.tet_count = static_cast<std::uint32_t>(w.softTets.size());
    c.tets = w.softTets.data();
    if (cn_scene_reserve(ctx_, n, std::max<std::uint32_t>(desc_.maxFluidNeighbors, 64u)) != CN_OK)
        return fail("reserve failed");
    if (cn_scene_upload(ctx_, &v, &t, &c) != CN_OK) return fail("upload failed");
    uploadedRevision_ = w.revision();
    uploadedColliders_ = w.colliderRevision();
    uploadedMaterials_ = w.materialRevision();
    uploadedConstraints_ = w.constraintRevision();
    uploadedCount_ = n;
    w.clearDirtyParticles();
    lastUpload_ = 1;
    return true;
}

// The categories whose revisions moved, and the dirty particles as contiguous runs
// (uploadContiguousRuns, physics_scene_gpu_state.cpp:255-284).
bool PhysicsSolver::uploadChanges(PhysicsWorld &w)
{
    lastUpload_ = 0;
    lastRuns_ = 0;
    if (w.colliderRevision() != uploadedColliders_) {
        if (cn_scene_set_colliders(ctx_, 0, w.staticBoxes.data(), static_cast<std::uint32_t>(w.staticBoxes.size())) !=
            CN_OK)
            return fail("collider update failed");
        uploadedColliders_ = w.colliderRevision();
        lastUpload_ = 2;
    }
    if (w.materialRevision() != uploadedMaterials_) {
        if (cn_scene_set_materials(ctx_, 0, w.fluidMaterials.data(), static_cast<std::uint32_t>(w.fluidMaterials.size()),
                                   &w.particleContactMaterials[0].x,
                                   static_cast<std::uint32_t>(w.particleContactMaterials.size())) != CN_OK)
            return fail("material update failed");
        uploadedMaterials_ = w.materialRevision();
        lastUpload_ = 2;
    }
    if (w.constraintRevision() != uploadedConstraints_) {
        cn_constraint_view c{};
        c.edge_count = static_cast<std::uint32_t>(w.softEdges.size());
        c.edges = w.softEdges.data();
        c.tet_count = static_cast<std::uint32_t>(w.softTets.size());
        c.tets = w.softTets.data();
        if (cn_scene_set_constraints(ctx_, 0, &c) != CN_OK) return fail("constraint update failed");
        uploadedConstraints_ = w.constraintRevision();
        lastUpload_ = 2;
    }
    std::vector<std::uint32_t> idx = w.dirtyParticles();
    if (!idx.empty()) {
        std::sort(idx.begin(), idx.end());
        idx.erase(std::unique(idx.begin(), idx.end()), idx.end());
        if (idx.back() >= uploadedCount_) return fail("dirty particle index out of range");
        std::size_t r0 = 0;
        for (std::size_t k = 1; k <= idx.size(); ++k) {
            if (k < idx.size() && idx[k] == idx[k - 1] + 1u) continue;
            const std::uint32_t first = idx[r0], count = static_cast<std::uint32_t>(k - r0);
            if (cn_particles_update_range(ctx_, first, count, &w.positionsInvMass[first].x, &w.velocities[first].x, 0u) !=
                CN_OK)
                return fail("particle update failed");
            ++lastRuns_;
            r0 = k;
        }
        w.clearDirtyParticles();
        lastUpload_ = 2;
    }
    return true;
}

bool PhysicsSolver::step(const FrameContext &frame, PhysicsWorld &world)
{
    if (!ctx_) return fail("step before initialize");
    if (world.particleCount() == 0) return true;  // nothing on the particle path
    if (world.revision() != uploadedRevision_) {
        if (!upload(world)) return false;
    } else if (!uploadChanges(world)) {
        return false;
    }
    cn_step_params p;
    cn_default_step_params(&p);
    p.dt = frame.deltaSeconds;
    p.substeps = desc_.substeps;
    p.default_iterations = desc_.defaultIterations;
    p.fluid_iterations = desc_.fluidIterations;
    p.soft_internal_iterations = desc_.softInternalIterations;
    p.soft_contact_iterations = desc_.softContactIterations;
    p.gravity[0] = desc_.gravity.x;
    p.gravity[1] = desc_.gravity.y;
    p.gravity[2] = desc_.gravity.z;
    std::memcpy(p.solver_config, desc_.solverConfig, sizeof(p.solver_config));
    p.max_fluid_neighbors = desc_.maxFluidNeighbors;
    if (cn_step(ctx_, &p) != CN_OK) return fail("step failed");
    if (cn_readback(ctx_, &world.positionsInvMass[0].x, &world.previousPositions[0].x, &world.velocities[0].x,
                    uploadedCount_) != CN_OK)
        return fail("readback failed");
    return true;
}

bool PhysicsSolver::validateGpuMetaBlocking()
{
    if (!ctx_) return fail("validate before initialize");
    const cn_status st = cn_validate_meta(ctx_, &meta_);
    if (st == CN_ERR_CAPACITY) return fail("capacity overflow");
    if (st != CN_OK) return fail("validate failed");
    return true;
}

}  // namespace cressim::neo::hip
