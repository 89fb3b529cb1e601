// Workgroup-wide ranks and prefix sums for 256-thread workgroups (four
// wave64 wavefronts): the building blocks of the stream compactions that
// keep their output in input order (selection's pick list, sun flips).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vgi {

// Workgroup-wide exclusive ranks of up to 8 flags (ballots per wavefront,
// wavefront totals through LDS): rank[k] for this thread, total[k] for the
// workgroup. blockDim.x == 256.
template <int K>
__device__ __forceinline__ void block_ranks(const bool (&flag)[K], uint32_t (&rank)[K], uint32_t (&total)[K],
                                            uint32_t (*wave_totals)[4]) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t local[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        unsigned long long b = __ballot(flag[k]);
        local[k] = (uint32_t)__popcll(b & below);
        if (lane == 0u) wave_totals[k][wave] = (uint32_t)__popcll(b);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        uint32_t before = 0u, sum = 0u;
        for (uint32_t w = 0; w < 4u; ++w) {
            uint32_t n = wave_totals[k][w];
            before += w < wave ? n : 0u;
            sum += n;
        }
        rank[k] = before + local[k];
        total[k] = sum;
    }
    __syncthreads();
}

// Exclusive prefix sums of part[k][0..255] for k < K in one workgroup of 256
// (Hillis-Steele in LDS); returns the totals.
template <int K>
__device__ __forceinline__ void block_exclusive_scan(uint32_t (*part)[256], uint32_t (&total)[K]) {
    const uint32_t t = threadIdx.x;
    uint32_t mine[K];
#pragma unroll
    for (int k = 0; k < K; ++k) mine[k] = part[k][t];
    for (uint32_t step = 1u; step < 256u; step <<= 1) {
        uint32_t add[K];
#pragma unroll
        for (int k = 0; k < K; ++k) add[k] = t >= step ? part[k][t - step] : 0u;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) part[k][t] += add[k];
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        total[k] = part[k][255];
        mine[k] = part[k][t] - mine[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) part[k][t] = mine[k];
    __syncthreads();
}

}  // namespace vgi
