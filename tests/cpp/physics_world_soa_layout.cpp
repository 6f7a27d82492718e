// The layout half of the reference's tests/physics/physics_world_soa_layout.cpp for the C ABI: the
// records the engine uploads unchanged keep the reference's GPU layouts (GpuFluidMaterial 48 B,
// GpuSoftEdge / GpuSoftTet 32 B, float4 particle arrays), and authoring fills every SoA array to
// the particle count.  No device is touched.
#include <cstddef>
#include <cstdio>

#include "cn_physics.hpp"

using namespace cressim::neo::hip;

static_assert(sizeof(cn_fluid_material) == 48, "GpuFluidMaterial is 48 bytes");
static_assert(sizeof(cn_soft_edge) == 32 && sizeof(cn_soft_bend) == 32, "soft records are 32 bytes");
static_assert(sizeof(cn_cloth_dihedral) == 32 && sizeof(cn_soft_tet) == 32, "soft records are 32 bytes");
static_assert(sizeof(cn_box_collider) == 64, "collider record is 64 bytes");
static_assert(sizeof(Float4) == 16, "particle arrays are float4 records");
static_assert(offsetof(cn_step_params, solver_config) == 40, "solver config after the gravity float4");
static_assert(sizeof(cn_step_params) == 4 * (6 + 4 + CN_SOLVER_CONFIG_FLOATS + 4), "step params");

int main()
{
    PhysicsWorld world;
    world.addStaticBox({});
    FluidDesc fluid;
    fluid.size = {0.4f, 0.2f, 0.2f};
    fluid.targetParticleSpacing = 0.1f;
    const std::uint32_t f0 = world.addFluid(fluid);
    SoftBlockDesc block;
    block.resolution[0] = 3;
    block.resolution[1] = 3;
    block.resolution[2] = 3;
    const std::uint32_t s0 = world.addSoftBlock(block);
    const std::uint32_t n = world.particleCount();
    // 4 x 2 x 2 fluid grid (floor(extent / spacing + 0.5) per axis), then a 3^3 soft block
    if (f0 != 0 || s0 != 16 || n != 16 + 27) {
        std::fprintf(stderr, "unexpected part offsets %u %u %u\n", f0, s0, n);
        return 1;
    }
    const std::size_t sizes[] = {world.previousPositions.size(), world.velocities.size(), world.radii.size(),
                                 world.kinds.size(), world.fluidMaterialIndices.size(), world.materialIndices.size(),
                                 world.environmentIndices.size(), world.phases.size(), world.collisionLayers.size(),
                                 world.collisionMasks.size()};
    for (std::size_t s : sizes)
        if (s != n) {
            std::fprintf(stderr, "SoA array of %zu entries for %u particles\n", s, n);
            return 1;
        }
    // the reference grid places particles at -size/2 + (i + 0.5) spacing
    const float x0 = -0.4f * 0.5f + 0.5f * 0.1f, x3 = -0.4f * 0.5f + 3.5f * 0.1f;
    if (world.positionsInvMass[0].x != x0 || world.positionsInvMass[3].x != x3) return 1;
    // two 2x2x2-cube soft block: 8 cubes x 5 tets; every constraint refers to the soft particles
    if (world.softTets.size() != 40) return 1;
    for (const cn_soft_edge &e : world.softEdges)
        if (e.a < s0 || e.b >= n || e.a >= e.b) return 1;
    if (world.kinds[0] != CN_PARTICLE_KIND_FLUID || world.kinds[s0] != CN_PARTICLE_KIND_SOLID) return 1;
    if ((world.phases[0] & 0x80000000u) == 0u || (world.phases[s0] & 0x80000000u) != 0u) return 1;
    std::printf("Physics world SoA layout checks passed.\n");
    return 0;
}
