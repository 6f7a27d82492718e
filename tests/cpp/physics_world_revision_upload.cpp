// Revision-tracked uploads: after a few frames, edits of single particles, a run of particles, a
// static box, the materials and the soft constraints are sent as partial updates (dirty particle
// runs and per-category revisions, physics_scene_gpu_state.cpp:255-284, 2348-2420) instead of a
// re-upload, and the next frame equals, bit for bit, a fresh solver that uploads the whole edited
// world.  A layout change (a new part) still re-uploads.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "cn_physics.hpp"

using namespace cressim::neo::hip;

static PhysicsWorld makeWorld()
{
    PhysicsWorld world;
    world.addStaticBox({{0.0f, -0.05f, 0.0f}, {1.0f, 0.05f, 1.0f}});
    world.addStaticBox({{-0.45f, 0.5f, 0.0f}, {0.05f, 0.5f, 1.0f}});
    FluidDesc f;
    f.position = {0.0f, 0.25f, 0.0f};
    f.size = {0.3f, 0.3f, 0.3f};
    f.targetParticleSpacing = 0.05f;
    f.particleRadius = 0.025f;
    f.particleMass = 0.125f;
    world.addFluid(f);
    SoftBlockDesc b;
    b.position = {0.2f, 0.7f, 0.0f};
    b.size = {0.15f, 0.15f, 0.15f};
    b.resolution[0] = b.resolution[1] = b.resolution[2] = 4;
    b.particleRadius = 0.02f;
    world.addSoftBlock(b);
    return world;
}

static void edit(PhysicsWorld &w, bool track)
{
    w.positionsInvMass[5].y += 0.01f;
    w.velocities[5].x = 0.3f;
    for (std::uint32_t i = 40; i < 52; ++i) w.velocities[i].y = -0.5f;
    w.staticBoxes[1].center[0] = -0.4f;
    w.fluidMaterials[0].viscosity = 0.05f;
    for (cn_soft_edge &e : w.softEdges) e.compliance = 1.0e-4f;
    if (track) {
        w.markParticleDirty(5);
        w.markParticleRangeDirty(40, 12);
        w.markCollidersDirty();
        w.markMaterialsDirty();
        w.markConstraintsDirty();
    } else {
        w.markDirty();
    }
}

static bool setup(PhysicsSolver &s)
{
    if (!s.initialize()) return false;
    s.desc().substeps = 2;
    s.desc().defaultIterations = 4;
    s.desc().maxFluidNeighbors = 64;
    return true;
}

int main()
{
    PhysicsSolver a(0), b(0);
    if (!setup(a) || !setup(b)) return 1;
    PhysicsWorld wa = makeWorld();
    FrameContext frame;
    for (int i = 0; i < 3; ++i)
        if (!a.step(frame, wa)) return 1;
    if (a.lastUploadKind() != 0) {
        std::fprintf(stderr, "An unchanged world sent %d.\n", a.lastUploadKind());
        return 1;
    }
    PhysicsWorld wb = wa;  // the state after three frames, as the readback left it
    edit(wa, true);
    edit(wb, false);
    if (!a.step(frame, wa) || !b.step(frame, wb)) return 1;
    if (a.lastUploadKind() != 2 || a.lastUploadRuns() != 2 || b.lastUploadKind() != 1) {
        std::fprintf(stderr, "Upload kinds %d (%u runs) / %d.\n", a.lastUploadKind(), a.lastUploadRuns(),
                     b.lastUploadKind());
        return 1;
    }
    for (std::uint32_t i = 0; i < wa.particleCount(); ++i)
        if (std::memcmp(&wa.positionsInvMass[i], &wb.positionsInvMass[i], sizeof(Float4)) != 0 ||
            std::memcmp(&wa.velocities[i], &wb.velocities[i], sizeof(Float4)) != 0) {
            std::fprintf(stderr, "Particle %u differs after the partial update.\n", i);
            return 1;
        }
    // a new part changes the layout: a full upload again
    FluidDesc g;
    g.position = {0.0f, 0.9f, 0.0f};
    g.size = {0.1f, 0.1f, 0.1f};
    g.targetParticleSpacing = 0.05f;
    g.particleRadius = 0.025f;
    wa.addFluid(g);
    if (!a.step(frame, wa) || a.lastUploadKind() != 1) {
        std::fprintf(stderr, "A new part did not re-upload (%d).\n", a.lastUploadKind());
        return 1;
    }
    std::printf("Physics revision upload checks passed (%u particles).\n", wa.particleCount());
    return 0;
}
