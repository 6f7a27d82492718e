// cn_passes.hip -- single reference passes in the reference's own data order, for parity tests of
// the broadphase stages (keys, sort, cell-range table).
#include <hip/hip_runtime.h>

#include "cn_internal.h"

namespace cn {

// physics_particle_build_broad_phase_entries.cs.hlsl:9-35 + ..._keys.cs.hlsl:8-23
__global__ __launch_bounds__(kBlock) void k_ref_keys(const float4 *__restrict__ pos, uint32_t n, float inv_cell,
                                                     uint32_t *__restrict__ keys, int32_t *__restrict__ cells)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    const int gx = (int)floorf(p.x * inv_cell), gy = (int)floorf(p.y * inv_cell), gz = (int)floorf(p.z * inv_cell);
    keys[i] = cell_hash(gx, gy, gz);
    cells[3 * i + 0] = gx;
    cells[3 * i + 1] = gy;
    cells[3 * i + 2] = gz;
}

void launch_ref_keys(const float4 *pos, uint32_t n, float inv_cell, uint32_t *keys, int32_t *cells, hipStream_t st)
{
    hipLaunchKernelGGL(k_ref_keys, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, pos, n, inv_cell, keys,
                       cells);
}

// clearParticleCellRanges + mark + scan + build (physics_particle_clear_cell_ranges.cs.hlsl:6-21,
// ..._mark_cell_range_starts.cs.hlsl:7-28, ..._build_cell_ranges.cs.hlsl:10-58): the table slot of a
// run is the number of run starts before it; each run start writes {key, start, end, 0}.
__global__ __launch_bounds__(kBlock) void k_ref_clear_ranges(uint4 *__restrict__ ranges, uint32_t capacity)
{
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r < capacity) ranges[r] = make_uint4(kInvalid, 0u, 0u, 0u);
}

__global__ __launch_bounds__(kBlock) void k_ref_build_ranges(const uint32_t *__restrict__ sk, uint32_t n,
                                                             const uint32_t *__restrict__ run_index,
                                                             uint32_t capacity, uint4 *__restrict__ ranges)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t key = sk[i];
    if (i != 0 && sk[i - 1] == key) return;
    uint32_t end = i + 1;
    while (end < n && sk[end] == key) ++end;
    const uint32_t r = run_index[i];
    if (r >= capacity) return;
    ranges[r] = make_uint4(key, i, end, 0u);
}

// run_index[i] = number of run starts in [0, i)
__global__ __launch_bounds__(1024) void k_ref_run_index(const uint32_t *__restrict__ sk, uint32_t n,
                                                        uint32_t *__restrict__ run_index)
{
    __shared__ uint32_t sh[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b0 = t * per;
    uint32_t local = 0;
    for (uint32_t k = 0; k < per; ++k) {
        const uint32_t i = b0 + k;
        if (i < n && (i == 0 || sk[i - 1] != sk[i])) ++local;
    }
    sh[t] = local;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t add = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    uint32_t run = sh[t] - local;
    for (uint32_t k = 0; k < per; ++k) {
        const uint32_t i = b0 + k;
        if (i >= n) break;
        run_index[i] = run;
        if (i == 0 || sk[i - 1] != sk[i]) ++run;
    }
}

void launch_ref_cell_ranges(const uint32_t *sorted_keys, uint32_t n, uint32_t capacity, uint32_t *ranges,
                            hipStream_t st)
{
    // ranges buffer holds capacity uint4 records followed by n uint32 of run index scratch
    uint4 *r4 = reinterpret_cast<uint4 *>(ranges);
    uint32_t *run_index = ranges + 4ull * capacity;
    hipLaunchKernelGGL(k_ref_clear_ranges, dim3((capacity + kBlock - 1) / kBlock), dim3(kBlock), 0, st, r4,
                       capacity);
    if (n == 0) return;
    hipLaunchKernelGGL(k_ref_run_index, dim3(1), dim3(1024), 0, st, sorted_keys, n, run_index);
    hipLaunchKernelGGL(k_ref_build_ranges, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, sorted_keys, n,
                       run_index, capacity, r4);
}

}  // namespace cn
