// The reference's tests/physics/physics_fluid_compute_smoke.cpp on the HIP particle path: one fluid
// particle from a 0.2 m regular grid above a static floor (box half extents 1 x 0.1 x 1) falls for
// 30 frames of 1/60 s and stays above the floor.
#include <cstdio>

#include "cn_physics.hpp"

int main()
{
    using namespace cressim::neo::hip;
    PhysicsSolver solver(0);
    if (!solver.initialize()) {
        std::fprintf(stderr, "Physics solver initialization failed.\n");
        return 1;
    }
    PhysicsWorld world;
    world.addStaticBox({{0.0f, 0.0f, 0.0f}, {1.0f, 0.1f, 1.0f}});
    FluidDesc fluid;
    fluid.position = {0.0f, 0.45f, 0.0f};
    fluid.size = {0.2f, 0.2f, 0.2f};
    fluid.targetParticleSpacing = 0.2f;
    fluid.particleMass = 0.5f;
    fluid.particleRadius = 0.1f;
    fluid.material.viscosity = 0.05f;
    const std::uint32_t first = world.addFluid(fluid);
    if (first == UINT32_MAX || world.particleCount() == 0) {
        std::fprintf(stderr, "Fluid authoring did not create any particles.\n");
        return 1;
    }
    const float initialY = world.positionsInvMass[first].y;
    FrameContext frame;
    frame.deltaSeconds = 1.0f / 60.0f;
    for (std::uint32_t i = 0; i < 30; ++i) {
        frame.frameIndex = i;
        frame.timeSeconds = static_cast<double>(i) * frame.deltaSeconds;
        if (!solver.step(frame, world)) {
            std::fprintf(stderr, "Physics step failed: %s\n", solver.lastError().c_str());
            return 1;
        }
    }
    const float finalY = world.positionsInvMass[first].y;
    if (!(finalY < initialY)) {
        std::fprintf(stderr, "Fluid particle did not advance under gravity.\n");
        return 1;
    }
    if (finalY < -0.05f) {
        std::fprintf(stderr, "Fluid particle escaped through the authored static floor.\n");
        return 1;
    }
    if (!solver.validateGpuMetaBlocking()) return 1;
    std::printf("Physics fluid compute smoke checks passed (y %.4f -> %.4f).\n", initialY, finalY);
    return 0;
}
