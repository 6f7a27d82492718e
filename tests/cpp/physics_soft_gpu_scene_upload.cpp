// A soft block (tet grid, 5 tets per cube) uploaded with its edges and volume constraints falls
// freely for 30 frames: every particle stays finite, the centre of mass follows g t^2 / 2 and the
// block keeps its volume within 1 % (after the reference's
// tests/physics/physics_soft_gpu_scene_upload.cpp and physics_soft_body_generation_regular_tet_grid.cpp).
// Static boxes bound fluids only on this path, so there is no floor.
#include <cmath>
#include <cstdio>

#include "cn_physics.hpp"

using namespace cressim::neo::hip;

static double blockVolume(const PhysicsWorld &w)
{
    double v = 0.0;
    for (const cn_soft_tet &t : w.softTets) {
        const Float4 &a = w.positionsInvMass[t.p[0]], &b = w.positionsInvMass[t.p[1]],
                     &c = w.positionsInvMass[t.p[2]], &d = w.positionsInvMass[t.p[3]];
        const double ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
        const double vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
        const double wx = d.x - a.x, wy = d.y - a.y, wz = d.z - a.z;
        v += std::fabs((uy * vz - uz * vy) * wx + (uz * vx - ux * vz) * wy + (ux * vy - uy * vx) * wz) / 6.0;
    }
    return v;
}

int main()
{
    PhysicsSolver solver(0);
    if (!solver.initialize()) return 1;
    solver.desc().substeps = 4;
    solver.desc().defaultIterations = 8;
    PhysicsWorld world;
    SoftBlockDesc block;
    block.position = {0.0f, 0.3f, 0.0f};
    block.size = {0.2f, 0.2f, 0.2f};
    block.resolution[0] = block.resolution[1] = block.resolution[2] = 6;
    block.particleRadius = 0.02f;
    block.edgeCompliance = block.volumeCompliance = 0.0f;
    world.addSoftBlock(block);
    if (world.particleCount() != 216 || world.softTets.size() != 5u * 125u) {
        std::fprintf(stderr, "Tet grid authoring: %u particles, %zu tets.\n", world.particleCount(),
                     world.softTets.size());
        return 1;
    }
    const double v0 = blockVolume(world);
    double y0 = 0.0;
    for (const Float4 &p : world.positionsInvMass) y0 += p.y;
    y0 /= world.particleCount();
    FrameContext frame;
    frame.deltaSeconds = 1.0f / 60.0f;
    const std::uint32_t frames = 30;
    for (std::uint32_t i = 0; i < frames; ++i) {
        frame.frameIndex = i;
        if (!solver.step(frame, world)) return 1;
    }
    double y1 = 0.0;
    for (const Float4 &p : world.positionsInvMass) {
        if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z)) {
            std::fprintf(stderr, "Soft particle lost: (%f, %f, %f).\n", p.x, p.y, p.z);
            return 1;
        }
        y1 += p.y;
    }
    y1 /= world.particleCount();
    const double t = frames / 60.0, fall = 0.5 * 9.81 * t * t;
    if (std::fabs((y0 - y1) - fall) > 0.02 * fall) {
        std::fprintf(stderr, "Centre of mass fell %g m, free fall %g m.\n", y0 - y1, fall);
        return 1;
    }
    const double v1 = blockVolume(world);
    if (std::fabs(v1 - v0) > 0.01 * v0) {
        std::fprintf(stderr, "Soft block volume drifted: %g -> %g.\n", v0, v1);
        return 1;
    }
    std::printf("Physics soft upload checks passed (volume %.6g -> %.6g).\n", v0, v1);
    return 0;
}
