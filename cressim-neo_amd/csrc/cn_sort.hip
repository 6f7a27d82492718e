// cn_sort.hip -- stable LSD radix sort of (key, value) uint32 pairs for gfx950.
//
// Replaces the reference's 32-pass one-bit split sort (physics_pass_dispatcher.cpp:3963-4063,
// physics_radix_{classify,finalize,scatter}.cs.hlsl) with 8-bit digits: per digit pass one
// histogram kernel (LDS atomics per 4096-element tile), one digit-row scan kernel and one scatter
// kernel that ranks the tile with wave64 ballots.  Any stable LSD sort yields the same order as the
// reference's stable one-bit sort, so keys and permutation are bit-identical to it.
#include <hip/hip_runtime.h>

#include "cn_internal.h"

namespace cn {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortThreads * kSortRounds;  // 4096 elements per workgroup

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint2 *__restrict__ kv, uint32_t n,
                                                             uint32_t shift, uint32_t nb,
                                                             uint32_t *__restrict__ counts)
{
    __shared__ uint32_t hist[256];
    const uint32_t t = threadIdx.x;
    const uint32_t b = blockIdx.x;
    hist[t] = 0;
    __syncthreads();
    const uint32_t base = b * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const uint32_t e = base + r * kSortThreads + t;
        if (e < n) atomicAdd(&hist[(kv[e].x >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    counts[t * nb + b] = hist[t];
}

// One workgroup per digit: exclusive scan of the digit's row of per-tile counts, row total out.
__global__ __launch_bounds__(kSortThreads) void k_radix_rowscan(uint32_t *__restrict__ counts, uint32_t nb,
                                                                uint32_t *__restrict__ totals)
{
    __shared__ uint32_t sh[kSortThreads];
    const uint32_t d = blockIdx.x;
    const uint32_t t = threadIdx.x;
    uint32_t *row = counts + (size_t)d * nb;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += kSortThreads * 4) {
        uint32_t v[4];
        uint32_t local = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t idx = c0 + t * 4 + k;
            v[k] = idx < nb ? row[idx] : 0u;
            local += v[k];
        }
        sh[t] = local;
        __syncthreads();
        for (uint32_t off = 1; off < kSortThreads; off <<= 1) {
            const uint32_t add = t >= off ? sh[t - off] : 0u;
            __syncthreads();
            sh[t] += add;
            __syncthreads();
        }
        uint32_t run = carry + sh[t] - local;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t idx = c0 + t * 4 + k;
            if (idx < nb) row[idx] = run;
            run += v[k];
        }
        carry += sh[kSortThreads - 1];
        __syncthreads();
    }
    if (t == 0) totals[d] = carry;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint2 *__restrict__ in, uint2 *__restrict__ out,
                                                                uint32_t n, uint32_t shift, uint32_t nb,
                                                                const uint32_t *__restrict__ counts,
                                                                const uint32_t *__restrict__ totals)
{
    __shared__ uint32_t base_sh[256];
    __shared__ uint32_t running[256];
    __shared__ uint32_t wc[kSortThreads / 64][256];
    const uint32_t t = threadIdx.x;
    const uint32_t b = blockIdx.x;
    const uint32_t lane = t & 63u;
    const uint32_t wave = t >> 6;

    // digit bases: exclusive scan of the digit totals + this tile's row prefix
    base_sh[t] = totals[t];
    running[t] = 0;
#pragma unroll
    for (int w = 0; w < kSortThreads / 64; ++w) wc[w][t] = 0;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        const uint32_t add = t >= off ? base_sh[t - off] : 0u;
        __syncthreads();
        base_sh[t] += add;
        __syncthreads();
    }
    const uint32_t my_excl = base_sh[t] - totals[t];
    __syncthreads();
    base_sh[t] = my_excl + counts[t * nb + b];

    const uint32_t tile = b * kSortTile;
    uint2 items[kSortRounds];
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const uint32_t e = tile + r * kSortThreads + t;
        items[r] = e < n ? in[e] : make_uint2(0u, 0u);
    }
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64u - lane));
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const uint32_t e = tile + r * kSortThreads + t;
        const bool valid = e < n;
        const uint32_t d = (items[r].x >> shift) & 0xffu;
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (d >> bit) & 1u;
            const uint64_t bal = __ballot(on);
            match &= on ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(match & lt_mask);
        if (valid && rank == 0) wc[wave][d] = (uint32_t)__popcll(match);
        __syncthreads();
        if (valid) {
            uint32_t pos = running[d] + rank;
            for (uint32_t w = 0; w < wave; ++w) pos += wc[w][d];
            out[base_sh[d] + pos] = items[r];
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kSortThreads / 64; ++w) {
            add += wc[w][t];
            wc[w][t] = 0;
        }
        running[t] += add;
        __syncthreads();
    }
}

// (key, value) pairs from two separate arrays, and back: the layout of the pair entry points that
// take the caller's device arrays.
__global__ __launch_bounds__(kSortThreads) void k_pack_pairs(const uint32_t *__restrict__ keys,
                                                             const uint32_t *__restrict__ vals, uint32_t n,
                                                             uint2 *__restrict__ out)
{
    const uint32_t i = blockIdx.x * kSortThreads + threadIdx.x;
    if (i < n) out[i] = make_uint2(keys[i], vals[i]);
}

__global__ __launch_bounds__(kSortThreads) void k_unpack_pairs(const uint2 *__restrict__ in, uint32_t n,
                                                               uint32_t *__restrict__ keys,
                                                               uint32_t *__restrict__ vals)
{
    const uint32_t i = blockIdx.x * kSortThreads + threadIdx.x;
    if (i < n) {
        const uint2 v = in[i];
        keys[i] = v.x;
        vals[i] = v.y;
    }
}

void launch_pack_pairs(const uint32_t *keys, const uint32_t *vals, uint32_t n, uint2 *out, hipStream_t stream)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_pack_pairs, dim3((n + kSortThreads - 1) / kSortThreads), dim3(kSortThreads), 0, stream,
                       keys, vals, n, out);
}

void launch_unpack_pairs(const uint2 *in, uint32_t n, uint32_t *keys, uint32_t *vals, hipStream_t stream)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_unpack_pairs, dim3((n + kSortThreads - 1) / kSortThreads), dim3(kSortThreads), 0, stream,
                       in, n, keys, vals);
}

// Sorts n pairs by the low key_bits of .x, stably.  buf_a holds the input; the result ends in
// the returned buffer (buf_a or buf_b).  scratch needs sort_scratch_words(n) uint32 words.
size_t sort_scratch_words(uint32_t n)
{
    const uint32_t nb = (n + kSortTile - 1) / kSortTile;
    return (size_t)256 * (nb ? nb : 1) + 256;
}

uint2 *radix_sort_pairs(uint2 *buf_a, uint2 *buf_b, uint32_t n, uint32_t key_bits, uint32_t *scratch,
                        hipStream_t stream)
{
    if (n <= 1) return buf_a;
    const uint32_t nb = (n + kSortTile - 1) / kSortTile;
    uint32_t *counts = scratch;
    uint32_t *totals = scratch + (size_t)256 * nb;
    uint2 *src = buf_a, *dst = buf_b;
    for (uint32_t shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kSortThreads), 0, stream, src, n, shift, nb, counts);
        hipLaunchKernelGGL(k_radix_rowscan, dim3(256), dim3(kSortThreads), 0, stream, counts, nb, totals);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kSortThreads), 0, stream, src, dst, n, shift, nb,
                           counts, totals);
        uint2 *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

}  // namespace cn
