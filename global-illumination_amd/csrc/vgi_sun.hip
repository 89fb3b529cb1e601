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

// One thread per bound face: its corner word (refreshed every
// `refresh` calls per face, staggered by bound index; a face without a known
// word traces at once) and its flip's flux (0 when none).
__global__ __launch_bounds__(256) void sun_corners_kernel(SceneView scene, Optics optics, const uint4 *__restrict__ tasks,
                                                          const uint32_t *__restrict__ slots, uint32_t count,
                                                          const uint32_t *__restrict__ meta, uint32_t *__restrict__ corners,
                                                          float *__restrict__ flips, uint32_t frame, uint32_t refresh) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float flux = 0.0f;
    uint4 t = tasks[i * 2u];
    uint32_t slot = slots[i];
    if (t.w < 6u && slot != NONE) {
        f3 normal = face_normal(t.w);
        bool lit_face = false;
        for (uint32_t l = 0u; l < optics.directional; ++l) {
            float4 pr = optics.lights[l * 3u];
            lit_face = lit_face || dot3(normal, mk(pr.x, pr.y, pr.z)) > 0.0f;
        }
        uint32_t old = corners[slot], word = 0x80808080u;
        bool carried = old != 0u && refresh > 1u && (frame + i) % refresh != 0u;
        if (carried) word = old;
        if (lit_face && !carried) {
            // Corner lattice points: u, v span the face; the face lies on
            // the voxel's positive side along its axis for +x/+y/+z faces.
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
            word = packed;
            if (old != 0u) flux = flip_flux(optics, normal, meta[(size_t)slot * 12u + 3u] >> 16, old, packed);
        }
        corners[slot] = word;
    }
    flips[i] = flux;
}

// Flipped faces per range of `per_block` bound faces.
__global__ __launch_bounds__(256) void sun_count_kernel(const float *__restrict__ flips, uint32_t count, uint32_t per_block,
                                                        uint32_t *__restrict__ counts) {
    __shared__ uint32_t wave_sums[4];
    uint32_t begin = blockIdx.x * per_block, end = min(begin + per_block, count);
    uint32_t n = 0u;
    for (uint32_t i = begin + threadIdx.x; i < end; i += 256u) n += flips[i] > 0.0f ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) n += (uint32_t)__shfl_xor((int)n, off, 64);
    if ((threadIdx.x & 63u) == 0u) wave_sums[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0u) counts[blockIdx.x] = wave_sums[0] + wave_sums[1] + wave_sums[2] + wave_sums[3];
}

// One workgroup: each range's first record index; the flip count.
__global__ __launch_bounds__(256) void sun_offsets_kernel(const uint32_t *__restrict__ counts, uint32_t ranges,
                                                          uint32_t *__restrict__ offsets, uint32_t *__restrict__ region) {
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
    if (t == 0u) region[0] = total[0];
}

// Records in bound-face order: the first SUN_FLIP_RECORDS as (voxel, flux);
// the rest widen the overflow box (biased maxima) and add flux x 64.
__global__ __launch_bounds__(256) void sun_emit_kernel(const float *__restrict__ flips, const uint4 *__restrict__ tasks,
                                                       uint32_t count, uint32_t per_block, const uint32_t *__restrict__ offsets,
                                                       uint32_t *__restrict__ region) {
    __shared__ uint32_t wave_totals[1][4];
    uint32_t begin = blockIdx.x * per_block, end = min(begin + per_block, count);
    uint32_t base_record = offsets[blockIdx.x];
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
        uint32_t *o = region + SUN_FLIP_OVERFLOW;
        atomicMax(&o[0], (uint32_t)(SUN_FLIP_BIAS - x));
        atomicMax(&o[1], (uint32_t)(SUN_FLIP_BIAS - y));
        atomicMax(&o[2], (uint32_t)(SUN_FLIP_BIAS - z));
        atomicMax(&o[3], (uint32_t)(x + SUN_FLIP_BIAS));
        atomicMax(&o[4], (uint32_t)(y + SUN_FLIP_BIAS));
        atomicMax(&o[5], (uint32_t)(z + SUN_FLIP_BIAS));
        atomicAdd(&o[6], (uint32_t)(flux * 64.0f + 0.5f));
    }
}

__global__ void sun_clear_kernel(uint32_t *__restrict__ region) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < SUN_FLIP_WORDS) region[i] = 0u;
}

inline uint32_t blocks_of(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

}  // namespace

void launch_sun_corners(const SceneView &scene, const Optics &optics, const uint4 *tasks, const uint32_t *slots, uint32_t count,
                        const uint32_t *meta, uint32_t *corners, float *flips, uint32_t frame, uint32_t refresh, uint32_t *ranges,
                        uint32_t *region, hipStream_t stream) {
    hipLaunchKernelGGL(sun_clear_kernel, dim3(blocks_of(SUN_FLIP_WORDS, 256)), dim3(256), 0, stream, region);
    if (count == 0u) return;
    hipLaunchKernelGGL(sun_corners_kernel, dim3(blocks_of(count, 256)), dim3(256), 0, stream, scene, optics, tasks, slots, count, meta,
                       corners, flips, frame, refresh);
    uint32_t per_block = (blocks_of(count, SUN_RANGES) + 255u) / 256u * 256u;
    uint32_t n_ranges = blocks_of(count, per_block);
    uint32_t *counts = ranges, *offsets = ranges + SUN_RANGES;
    hipLaunchKernelGGL(sun_count_kernel, dim3(n_ranges), dim3(256), 0, stream, flips, count, per_block, counts);
    hipLaunchKernelGGL(sun_offsets_kernel, dim3(1), dim3(256), 0, stream, counts, n_ranges, offsets, region);
    hipLaunchKernelGGL(sun_emit_kernel, dim3(n_ranges), dim3(256), 0, stream, flips, tasks, count, per_block, offsets, region);
}

}  // namespace vgi
