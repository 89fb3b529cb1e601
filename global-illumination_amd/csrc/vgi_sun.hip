// Sun corner words of the bound faces and their flips (voxel_path.wgsl:245-266
// sky_corner_visibility; voxel_view_runtime.wgsl:168-200 the per-face words,
// 222-268 report_sun_flips), for the world's faces instead of a camera's.
//
// A corner word is one byte per face corner: bit 7 "known", bit i "sees
// directional light i" (lights with visible colour, at most four). A face
// whose corner bits changed while both words were known reports the light its
// corners gained or lost as a flip; flips become Light retention events on the
// host. The reference appends flips with one atomic per face, in whatever
// order its threads arrive; here they are compacted in bound-face order (a
// count pass, one scan workgroup, an emit pass), so the records, the overflow
// split at 2,048 and the host's per-region sums are reproducible.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "vgi_blockscan.h"
#include "vgi_kernels.h"
#include "vgi_trace.h"

namespace vgi {

namespace {

// voxel_path.wgsl:245-266: which of the first four coloured directional
// lights a lattice point sees (bit 31 known; 0 when a ray ran out of steps).
__device__ __forceinline__ uint32_t sky_corner_visibility(const SceneView &scene, const Optics &optics, f3 position) {
    uint32_t word = 0x80000000u, bit = 0u;
    NoTally tally;
    for (uint32_t i = 0u; i < optics.directional && bit < 4u; ++i) {
        float4 pr = optics.lights[i * 3u], en = optics.lights[i * 3u + 1u], ex = optics.lights[i * 3u + 2u];
        if (fmx(en.x - ex.y, 0.0f) == 0.0f && fmx(en.y - ex.z, 0.0f) == 0.0f && fmx(en.z - ex.w, 0.0f) == 0.0f) continue;
        f3 direction = normalize3(mk(pr.x, pr.y, pr.z));
        f3 origin = position + direction * 0.002f;
        uint32_t visibility = trace_ray<NoTally, true>(scene, origin, direction, optics.max_distance, tally).status;
        if (visibility == 2u) {
            word = 0u;
            break;
        }
        if (visibility == 0u) word |= 1u << bit;
        bit += 1u;
    }
    return word;
}

// The light a face's flipped corners gained or lost, a quarter face per
// corner, times its albedo (report_sun_flips' flux).
__device__ __forceinline__ float flip_flux(const Optics &optics, f3 normal, uint32_t material, uint32_t before, uint32_t after) {
    const f3 w = mk(0.2126f, 0.7152f, 0.0722f);
    float flux = 0.0f;
    uint32_t bit = 0u;
    for (uint32_t i = 0u; i < optics.directional && bit < 4u; ++i) {
        float4 pr = optics.lights[i * 3u], en = optics.lights[i * 3u + 1u], ex = optics.lights[i * 3u + 2u];
        f3 color = mk(fmx(en.x - ex.y, 0.0f), fmx(en.y - ex.z, 0.0f), fmx(en.z - ex.w, 0.0f));
        if (color.x == 0.0f && color.y == 0.0f && color.z == 0.0f) continue;
        float cosine = fmx(dot3(normal, normalize3(mk(pr.x, pr.y, pr.z))), 0.0f);
        for (uint32_t corner = 0u; corner < 4u; ++corner) {
            uint32_t a = (before >> (corner * 8u)) & 255u, b = (after >> (corner * 8u)) & 255u;
            if ((a & b & 128u) != 0u && ((a ^ b) & (1u << bit)) != 0u) flux = flux + 0.25f * dot3(color, w) * cosine;
        }
        bit += 1u;
    }
    if (flux <= 0.0f) return 0.0f;
    float4 md = optics.materials[(material & 255u) * 2u];
    return flux * dot3(mk(md.x, md.y, md.z), w);
}

// One thread per bound face: faces that keep their word (carried between
// refreshes) or face away from every sun (known, all corners unlit) are
// settled here; the faces that trace are marked (flip flux -1, every other
// face 0) and then compacted into a queue in bound-face order, so tracing
// wavefronts are dense whatever the refresh stagger and neighbouring faces
// stay together. A face without a known word traces at once.
__global__ __launch_bounds__(256) void sun_classify_kernel(Optics optics, const uint4 *__restrict__ tasks,
                                                           const uint32_t *__restrict__ slots, uint32_t count,
                                                           uint32_t *__restrict__ corners, float *__restrict__ flips, uint32_t frame,
                                                           uint32_t refresh) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool trace = false;
    if (i < count) {
        uint4 t = tasks[i * 2u];
        uint32_t slot = slots[i];
        if (t.w < 6u && slot != NONE) {
            f3 normal = face_normal(t.w);
            bool lit_face = false;
            for (uint32_t l = 0u; l < optics.directional; ++l) {
                float4 pr = optics.lights[l * 3u];
                lit_face = lit_face || dot3(normal, mk(pr.x, pr.y, pr.z)) > 0.0f;
            }
            uint32_t old = corners[slot];
            bool carried = old != 0u && refresh > 1u && (frame + i) % refresh != 0u;
            if (!carried && !lit_face) corners[slot] = 0x80808080u;
            trace = lit_face && !carried;
        }
        flips[i] = trace ? -1.0f : 0.0f;
    }
}

// The queued faces' corner words and flip flux (grid-stride over the queue,
// whose length only the device knows).
__global__ __launch_bounds__(256) void sun_trace_kernel(SceneView scene, Optics optics, const uint4 *__restrict__ tasks,
                                                        const uint32_t *__restrict__ slots, const uint32_t *__restrict__ meta,
                                                        uint32_t *__restrict__ corners, float *__restrict__ flips,
                                                        const uint32_t *__restrict__ queue) {
    const uint32_t n = queue[0];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        uint32_t i = queue[1u + j];
        uint4 t = tasks[i * 2u];
        uint32_t slot = slots[i];
        f3 normal = face_normal(t.w);
        // Corner lattice points: u, v span the face; the face lies on the
        // voxel's positive side along its axis for +x/+y/+z faces.
        f3 u = mk(0.0f, 1.0f, 0.0f), v = mk(0.0f, 0.0f, 1.0f);
        if (t.w / 2u == 1u) u = mk(1.0f, 0.0f, 0.0f);
        else if (t.w / 2u == 2u) { u = mk(1.0f, 0.0f, 0.0f); v = mk(0.0f, 1.0f, 0.0f); }
        f3 base = mk((float)(int)t.x, (float)(int)t.y, (float)(int)t.z) +
                  mk(fmx(normal.x, 0.0f), fmx(normal.y, 0.0f), fmx(normal.z, 0.0f));
        uint32_t packed = 0u;
        for (uint32_t corner = 0u; corner < 4u; ++corner) {
            f3 position = base + u * (float)(corner & 1u) + v * (float)(corner >> 1u);
            uint32_t w = sky_corner_visibility(scene, optics, position);
            packed |= ((w & 15u) | ((w >> 24u) & 128u)) << (corner * 8u);
        }
        uint32_t old = corners[slot];
        flips[i] = old != 0u ? flip_flux(optics, normal, meta[(size_t)slot * 12u + 3u] >> 16, old, packed) : 0.0f;
        corners[slot] = packed;
    }
}

// Marked faces (flux < 0: to trace) or flipped faces (flux > 0) per range of
// `per_block` bound faces.
__device__ __forceinline__ bool selected(float flux, bool queued) { return queued ? flux < 0.0f : flux > 0.0f; }

__global__ __launch_bounds__(256) void sun_count_kernel(const float *__restrict__ flips, uint32_t count, uint32_t per_block,
                                                        bool queued, uint32_t *__restrict__ counts) {
    __shared__ uint32_t wave_sums[4];
    uint32_t begin = blockIdx.x * per_block, end = min(begin + per_block, count);
    uint32_t n = 0u;
    for (uint32_t i = begin + threadIdx.x; i < end; i += 256u) n += selected(flips[i], queued) ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) n += (uint32_t)__shfl_xor((int)n, off, 64);
    if ((threadIdx.x & 63u) == 0u) wave_sums[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0u) counts[blockIdx.x] = wave_sums[0] + wave_sums[1] + wave_sums[2] + wave_sums[3];
}

// One workgroup: each range's first output index; the total into *total.
__global__ __launch_bounds__(256) void sun_offsets_kernel(const uint32_t *__restrict__ counts, uint32_t ranges,
                                                          uint32_t *__restrict__ offsets, uint32_t *__restrict__ total_out) {
    __shared__ uint32_t part[1][256];
    const uint32_t t = threadIdx.x, per = (ranges + 255u) / 256u;
    uint32_t first = t * per, last = min(first + per, ranges), mine = 0u;
    for (uint32_t r = first; r < last; ++r) mine += counts[r];
    part[0][t] = mine;
    __syncthreads();
    uint32_t total[1];
    block_exclusive_scan<1>(part, total);
    uint32_t before = part[0][t];
    for (uint32_t r = first; r < last; ++r) {
        offsets[r] = before;
        before += counts[r];
    }
    if (t == 0u) *total_out = total[0];
}

// The marked faces' indices in bound-face order after the count word.
__global__ __launch_bounds__(256) void sun_queue_kernel(const float *__restrict__ flips, uint32_t count, uint32_t per_block,
                                                        const uint32_t *__restrict__ offsets, uint32_t *__restrict__ queue) {
    __shared__ uint32_t wave_totals[1][4];
    uint32_t begin = blockIdx.x * per_block, end = min(begin + per_block, count);
    uint32_t base_index = offsets[blockIdx.x];
    for (uint32_t base = begin; base < end; base += 256u) {
        uint32_t i = base + threadIdx.x;
        bool flag[1] = {i < end && flips[i] < 0.0f};
        uint32_t rank[1], total[1];
        block_ranks<1>(flag, rank, total, wave_totals);
        if (flag[0]) queue[1u + base_index + rank[0]] = i;
        base_index += total[0];
    }
}

// Records in bound-face order: the first SUN_FLIP_RECORDS as (voxel, flux);
// the rest widen the overflow box (biased maxima) and add flux x 64. Each
// range reduces its overflow part in registers, across wavefronts and then
// into `partial` (8 words per range); sun_overflow_kernel combines the
// ranges. Maxima and an integer sum do not depend on order, and no word is
// contended however many faces flip.
__global__ __launch_bounds__(256) void sun_emit_kernel(const float *__restrict__ flips, const uint4 *__restrict__ tasks,
                                                       uint32_t count, uint32_t per_block, const uint32_t *__restrict__ offsets,
                                                       uint32_t *__restrict__ region, uint32_t *__restrict__ partial) {
    __shared__ uint32_t wave_totals[1][4];
    __shared__ uint32_t wave_over[4][7];
    uint32_t begin = blockIdx.x * per_block, end = min(begin + per_block, count);
    uint32_t base_record = offsets[blockIdx.x];
    uint32_t over[7] = {0u, 0u, 0u, 0u, 0u, 0u, 0u};
    for (uint32_t base = begin; base < end; base += 256u) {
        uint32_t i = base + threadIdx.x;
        float flux = i < end ? flips[i] : 0.0f;
        bool flag[1] = {flux > 0.0f};
        uint32_t rank[1], total[1];
        block_ranks<1>(flag, rank, total, wave_totals);
        uint32_t record = base_record + rank[0];
        base_record += total[0];
        if (!flag[0]) continue;
        uint4 t = tasks[i * 2u];
        int x = (int)t.x, y = (int)t.y, z = (int)t.z;
        if (record < SUN_FLIP_RECORDS) {
            uint4 *r = reinterpret_cast<uint4 *>(region) + 1u + record;
            *r = make_uint4((uint32_t)x, (uint32_t)y, (uint32_t)z, __float_as_uint(flux));
            continue;
        }
        over[0] = max(over[0], (uint32_t)(SUN_FLIP_BIAS - x));
        over[1] = max(over[1], (uint32_t)(SUN_FLIP_BIAS - y));
        over[2] = max(over[2], (uint32_t)(SUN_FLIP_BIAS - z));
        over[3] = max(over[3], (uint32_t)(x + SUN_FLIP_BIAS));
        over[4] = max(over[4], (uint32_t)(y + SUN_FLIP_BIAS));
        over[5] = max(over[5], (uint32_t)(z + SUN_FLIP_BIAS));
        over[6] += (uint32_t)(flux * 64.0f + 0.5f);
    }
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 6; ++k) over[k] = max(over[k], (uint32_t)__shfl_xor((int)over[k], off, 64));
        over[6] += (uint32_t)__shfl_xor((int)over[6], off, 64);
    }
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0u)
        for (int k = 0; k < 7; ++k) wave_over[wave][k] = over[k];
    __syncthreads();
    if (threadIdx.x < 7u) {
        uint32_t k = threadIdx.x, v = wave_over[0][k];
        for (uint32_t w = 1u; w < 4u; ++w) v = k == 6u ? v + wave_over[w][k] : max(v, wave_over[w][k]);
        partial[blockIdx.x * 8u + k] = v;
    }
}

// One workgroup: the ranges' overflow parts into the region's overflow words.
__global__ __launch_bounds__(256) void sun_overflow_kernel(const uint32_t *__restrict__ partial, uint32_t ranges,
                                                           uint32_t *__restrict__ region) {
    __shared__ uint32_t wave_over[4][7];
    uint32_t over[7] = {0u, 0u, 0u, 0u, 0u, 0u, 0u};
    for (uint32_t r = threadIdx.x; r < ranges; r += 256u) {
#pragma unroll
        for (int k = 0; k < 6; ++k) over[k] = max(over[k], partial[r * 8u + k]);
        over[6] += partial[r * 8u + 6u];
    }
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 6; ++k) over[k] = max(over[k], (uint32_t)__shfl_xor((int)over[k], off, 64));
        over[6] += (uint32_t)__shfl_xor((int)over[6], off, 64);
    }
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0u)
        for (int k = 0; k < 7; ++k) wave_over[wave][k] = over[k];
    __syncthreads();
    if (threadIdx.x < 7u) {
        uint32_t k = threadIdx.x, v = wave_over[0][k];
        for (uint32_t w = 1u; w < 4u; ++w) v = k == 6u ? v + wave_over[w][k] : max(v, wave_over[w][k]);
        region[SUN_FLIP_OVERFLOW + k] = v;
    }
}

__global__ void sun_clear_kernel(uint32_t *__restrict__ region) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < SUN_FLIP_WORDS) region[i] = 0u;
}

inline uint32_t blocks_of(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

}  // namespace

void launch_sun_corners(const SceneView &scene, const Optics &optics, const uint4 *tasks, const uint32_t *slots, uint32_t count,
                        const uint32_t *meta, uint32_t *corners, float *flips, uint32_t *queue, uint32_t trace_blocks, uint32_t frame,
                        uint32_t refresh, uint32_t *ranges, uint32_t *region, hipStream_t stream) {
    hipLaunchKernelGGL(sun_clear_kernel, dim3(blocks_of(SUN_FLIP_WORDS, 256)), dim3(256), 0, stream, region);
    if (count == 0u) return;
    uint32_t per_block = (blocks_of(count, SUN_RANGES) + 255u) / 256u * 256u;
    uint32_t n_ranges = blocks_of(count, per_block);
    uint32_t *counts = ranges, *offsets = ranges + SUN_RANGES;
    hipLaunchKernelGGL(sun_classify_kernel, dim3(blocks_of(count, 256)), dim3(256), 0, stream, optics, tasks, slots, count, corners,
                       flips, frame, refresh);
    hipLaunchKernelGGL(sun_count_kernel, dim3(n_ranges), dim3(256), 0, stream, flips, count, per_block, true, counts);
    hipLaunchKernelGGL(sun_offsets_kernel, dim3(1), dim3(256), 0, stream, counts, n_ranges, offsets, queue);
    hipLaunchKernelGGL(sun_queue_kernel, dim3(n_ranges), dim3(256), 0, stream, flips, count, per_block, offsets, queue);
    hipLaunchKernelGGL(sun_trace_kernel, dim3(std::min(trace_blocks, blocks_of(count, 256))), dim3(256), 0, stream, scene, optics, tasks,
                       slots, meta, corners, flips, queue);
    hipLaunchKernelGGL(sun_count_kernel, dim3(n_ranges), dim3(256), 0, stream, flips, count, per_block, false, counts);
    hipLaunchKernelGGL(sun_offsets_kernel, dim3(1), dim3(256), 0, stream, counts, n_ranges, offsets, region);
    uint32_t *partial = ranges + 2u * SUN_RANGES;
    hipLaunchKernelGGL(sun_emit_kernel, dim3(n_ranges), dim3(256), 0, stream, flips, tasks, count, per_block, offsets, region, partial);
    hipLaunchKernelGGL(sun_overflow_kernel, dim3(1), dim3(256), 0, stream, partial, n_ranges, region);
}

}  // namespace vgi
