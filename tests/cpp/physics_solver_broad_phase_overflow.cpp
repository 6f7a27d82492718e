// The reference's tests/physics/physics_solver_broad_phase_overflow.cpp on the particle path: the
// step succeeds, and validateGpuMetaBlocking reports the overflow -- here 128 fluid particles
// within one smoothing radius of each other against 64 neighbor slots (the reference's test
// overflows its rigid broad phase the same way with 128 overlapping spheres).
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
    solver.desc().maxFluidNeighbors = 64;
    PhysicsWorld world;
    FluidDesc fluid;
    fluid.position = {0.0f, 0.0f, 0.0f};
    fluid.size = {0.08f, 0.08f, 0.04f};  // 8 x 8 x 4 = 256 particles at 0.01 m
    fluid.targetParticleSpacing = 0.01f;
    fluid.particleRadius = 0.03f;        // h = 0.1 m covers the whole block
    world.addFluid(fluid);
    FrameContext frame;
    frame.deltaSeconds = 1.0f / 60.0f;
    if (!solver.step(frame, world)) {
        std::fprintf(stderr, "Broad-phase overflow test failed during step.\n");
        return 1;
    }
    if (solver.validateGpuMetaBlocking()) {
        std::fprintf(stderr, "Broad-phase overflow test unexpectedly succeeded.\n");
        return 1;
    }
    if (solver.meta().fluid_neighbor_truncations == 0 || solver.meta().max_fluid_neighbor_count <= 64) {
        std::fprintf(stderr, "Overflow counters not set.\n");
        return 1;
    }
    std::printf("Physics broad-phase overflow checks passed (%u truncated lists, largest %u neighbors).\n",
                solver.meta().fluid_neighbor_truncations, solver.meta().max_fluid_neighbor_count);
    return 0;
}
