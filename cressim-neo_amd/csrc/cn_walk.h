// cn_walk.h -- the reference's 27-cell candidate walk over the run table of the sorted cell keys.
#pragma once

#include "cn_device.h"

namespace cn {

// The run table of one substep: runs of equal keys in sorted order, a bucket index over the top
// key bits, and the cell coordinates (w: the scene's key prefix) of every slot.
struct RunTable {
    const uint32_t *run_key;
    const uint32_t *run_start;
    const int4 *run_info;   // (cx, cy, cz, pure)
    const uint32_t *bucket;
    uint32_t bucket_shift;
    const int4 *cellc;
};

__device__ __forceinline__ int find_run(uint32_t key, const RunTable &t)
{
    const uint32_t b = key >> t.bucket_shift;
    uint32_t r = t.bucket[b];
    const uint32_t e = t.bucket[b + 1];
    while (r < e && t.run_key[r] < key) ++r;
    return (r < e && t.run_key[r] == key) ? (int)r : -1;
}

// f(slot) for the particles of the 27 cells around cell ci in dz, dy, dx order, each cell's
// particles in ascending particle index (its slots), until f returns false.  A cell whose hash
// equals the empty-table marker is never found (physics_fluid_build_neighbors.cs.hlsl:81-85 with
// FindParticleCellRange); a run shared by several cells (hash collision, Morton wrap) is filtered
// by cell coordinates (:94-98).
template <class F>
__device__ __forceinline__ void walk_runs(int4 ci, const RunTable &t, const SubstepConsts &c, F &&f)
{
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int tx = ci.x + dx, ty = ci.y + dy, tz = ci.z + dz;
                if (cell_hash(tx, ty, tz) == kInvalid) continue;
                const int r = find_run(cell_key(tx, ty, tz, (uint32_t)ci.w, c.key_mode, c.morton_bits), t);
                if (r < 0) continue;
                const int4 info = t.run_info[r];
                const bool pure = info.w != 0;
                if (pure && (info.x != tx || info.y != ty || info.z != tz)) continue;
                const uint32_t qe = t.run_start[r + 1];
                for (uint32_t q = t.run_start[r]; q < qe; ++q) {
                    if (!pure) {
                        const int4 cq = t.cellc[q];
                        if (cq.x != tx || cq.y != ty || cq.z != tz) continue;
                    }
                    if (!f(q)) return;
                }
            }
}

}  // namespace cn
