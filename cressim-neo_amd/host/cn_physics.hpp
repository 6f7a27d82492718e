// cn_physics.hpp -- C++ host side of the particle path over libcn_hip.so, with the shape of the
// reference's engine-facing classes: a PhysicsWorld that owns the authored unified particle pool
// (include/physics/physics_world.h, ParticleSoAHost of include/physics/physics_types.h:858-882)
// and a PhysicsSolver with initialize / step(frame, world) / validateGpuMetaBlocking / shutdown
// (include/physics/physics_solver.h:173, src/physics/physics_solver.cpp:336-1492).  Calls return
// bool and log the reason, as the reference does; lastError() keeps the last message.
//
// Only the particle path exists here: fluids from a regular grid source, tet-grid soft blocks,
// static boxes.  Rigid bodies, joints, strands and rendering are not part of this library.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "cn_hip.h"

namespace cressim::neo::hip {

struct Float3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
struct Float4 {
    float x = 0.0f, y = 0.0f, z = 0.0f, w = 0.0f;
};

// common::FrameContext (include/common/frame_context.h)
struct FrameContext {
    std::uint64_t frameIndex = 0;
    double timeSeconds = 0.0;
    float deltaSeconds = 1.0f / 60.0f;
};

// PhysicsSolverDesc (include/physics/physics_solver.h:21-143), particle-path fields
struct PhysicsSolverDesc {
    std::uint32_t substeps = 1;
    std::uint32_t defaultIterations = 20;
    std::uint32_t fluidIterations = 0;         // 0: defaultIterations
    std::uint32_t softInternalIterations = 0;  // 0: defaultIterations
    std::uint32_t softContactIterations = 0;   // 0: defaultIterations
    Float3 gravity{0.0f, -9.81f, 0.0f};
    std::uint32_t maxFluidNeighbors = 128;
    float solverConfig[CN_SOLVER_CONFIG_FLOATS] = {};  // GpuPhysicsSolverConfig order
    PhysicsSolverDesc();
};

// FluidComponent with a RegularGrid source (include/engine/components.h, physics_world.cpp:6857-6892)
struct FluidDesc {
    Float3 position;                    // translation of the grid's centre
    Float3 size{0.2f, 0.2f, 0.2f};      // regularGrid.size
    float targetParticleSpacing = 0.1f; // regularGrid.targetParticleSpacing
    float particleMass = 1.0f;
    float particleRadius = 0.05f;
    cn_fluid_material_desc material{0.01f, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f};
    std::uint32_t environment = 0;
};

// A static box collider (RigidBodyType::Static with a Box shape: half extents = params * |scale|)
struct StaticBoxDesc {
    Float3 center;
    Float3 halfExtents{0.5f, 0.5f, 0.5f};
    Float4 material{0.0f, 0.0f, 0.0f, -1.0f};  // friction, restitution, damping, static friction
    std::uint32_t environment = 0;
};

// SoftBodyComponent with a RegularGrid source: resolution^3 vertices spanning size, five tets per
// cube (src/physics/soft_body_authoring.cpp:81-177), edges and volume constraints
struct SoftBlockDesc {
    Float3 position;
    Float3 size{0.2f, 0.2f, 0.2f};
    std::uint32_t resolution[3] = {4, 4, 4};
    float particleMass = 0.1f;
    float particleRadius = 0.025f;
    float edgeCompliance = 1.0e-6f;
    float volumeCompliance = 1.0e-6f;
    std::uint32_t environment = 0;
};

// The unified particle pool and the tables uploaded beside it.
class PhysicsWorld {
public:
    // ParticleSoAHost
    std::vector<Float4> positionsInvMass, previousPositions, velocities;
    std::vector<float> radii;
    std::vector<std::uint32_t> kinds, fluidMaterialIndices, materialIndices, environmentIndices, phases,
        collisionLayers, collisionMasks;
    // tables
    std::vector<cn_fluid_material> fluidMaterials;
    std::vector<Float4> particleContactMaterials;
    std::vector<cn_box_collider> staticBoxes;
    std::vector<cn_soft_edge> softEdges;
    std::vector<cn_soft_tet> softTets;

    // Returns the first particle index of the new part, or UINT32_MAX on failure.
    std::uint32_t addFluid(const FluidDesc &desc);
    std::uint32_t addSoftBlock(const SoftBlockDesc &desc);
    void addStaticBox(const StaticBoxDesc &desc);

    std::uint32_t particleCount() const { return static_cast<std::uint32_t>(positionsInvMass.size()); }
    // Layout revision: particle count, per-particle attributes, topology.  A change re-uploads the
    // world on the next step.
    std::uint64_t revision() const { return revision_; }
    // Call after editing the arrays directly (the next step uploads the world again).
    void markDirty() { ++revision_; }

    // Narrower edits, uploaded without a re-upload on the next step (the reference's per-category
    // revisions and dirty-index runs, physics_scene_gpu_state.cpp:255-284, 2348-2420):
    // positionsInvMass / velocities of these particles (sent as contiguous runs)
    void markParticleDirty(std::uint32_t index) { dirtyParticles_.push_back(index); }
    void markParticleRangeDirty(std::uint32_t first, std::uint32_t count)
    {
        for (std::uint32_t i = 0; i < count; ++i) dirtyParticles_.push_back(first + i);
    }
    // staticBoxes moved, resized, re-materialled, enabled or disabled (same count)
    void markCollidersDirty() { ++colliderRevision_; }
    // fluidMaterials (same smoothing radii) or particleContactMaterials (same counts)
    void markMaterialsDirty() { ++materialRevision_; }
    // rest lengths, rest volumes, compliances, enable flags of softEdges / softTets (same topology)
    void markConstraintsDirty() { ++constraintRevision_; }
    std::uint64_t colliderRevision() const { return colliderRevision_; }
    std::uint64_t materialRevision() const { return materialRevision_; }
    std::uint64_t constraintRevision() const { return constraintRevision_; }
    const std::vector<std::uint32_t> &dirtyParticles() const { return dirtyParticles_; }
    void clearDirtyParticles() { dirtyParticles_.clear(); }

private:
    std::uint32_t appendParticles(std::size_t count, std::uint32_t kind, float invMass, float radius,
                                  std::uint32_t fluidMaterial, std::uint32_t env, std::uint32_t phase);
    std::uint32_t nextGroup_ = 1;
    std::uint64_t revision_ = 0;
    std::uint64_t colliderRevision_ = 0, materialRevision_ = 0, constraintRevision_ = 0;
    std::vector<std::uint32_t> dirtyParticles_;
};

class PhysicsSolver {
public:
    explicit PhysicsSolver(int deviceOrdinal = 0) : device_(deviceOrdinal) {}
    ~PhysicsSolver() { shutdown(); }
    PhysicsSolver(const PhysicsSolver &) = delete;
    PhysicsSolver &operator=(const PhysicsSolver &) = delete;

    bool initialize();
    void shutdown();
    PhysicsSolverDesc &desc() { return desc_; }

    // Uploads the world when its layout changed since the last step, otherwise only what its
    // narrower revisions and dirty particles name; runs one frame (desc().substeps substeps) and
    // reads the particle state back into the world (blocking, like the reference's
    // readbackParticles after a step).
    bool step(const FrameContext &frame, PhysicsWorld &world);
    // what the last step sent: 0 nothing, 1 a full upload, 2 partial updates
    int lastUploadKind() const { return lastUpload_; }
    std::uint32_t lastUploadRuns() const { return lastRuns_; }
    // false when an overflow counter of the last step is set (fluid-neighbor truncation, contact
    // candidates beyond a particle's list); the counters stay in meta()
    bool validateGpuMetaBlocking();
    const cn_meta &meta() const { return meta_; }
    const std::string &lastError() const { return error_; }
    cn_ctx *context() const { return ctx_; }

private:
    bool fail(const std::string &what);
    bool upload(PhysicsWorld &world);
    bool uploadChanges(PhysicsWorld &world);
    int device_;
    std::uint64_t uploadedColliders_ = 0, uploadedMaterials_ = 0, uploadedConstraints_ = 0;
    int lastUpload_ = 0;
    std::uint32_t lastRuns_ = 0;
    cn_ctx *ctx_ = nullptr;
    PhysicsSolverDesc desc_;
    std::uint64_t uploadedRevision_ = ~0ull;
    std::uint32_t uploadedCount_ = 0;
    cn_meta meta_{};
    std::string error_;
};

}  // namespace cressim::neo::hip
