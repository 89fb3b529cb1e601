// GPU work selection over the resident faces (voxel_accumulate.wgsl:448-695):
// which faces a batch traces when the GPU picks them itself.
//
//   classify   one thread per slot of this batch's rotating quarter: the
//              slot's class and key bucket (persisting until its quarter comes
//              round again), and statistics-mirror records of faces whose
//              host mirror is stale (one atomic per wavefront).
//   histogram  per workgroup, a 4 x 256-bin LDS histogram of the keys over a
//              grid-stride range; only touched bins reach the global one.
//   threshold  one workgroup: an inclusive scan per class, the class shares
//              of the pick budget, and each class's threshold bucket.
//   emit       one thread per slot: keys below the threshold are picked,
//              ties at it up to the class's remainder; ties and picks are
//              reserved with one atomic per wavefront (ballot + prefix).
//
// All four are HBM-bound integer work: classify reads 48 B of metadata and
// writes a 4 B key per classified slot, histogram and emit read 4 B per slot.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vgi_kernels.h"
#include "vgi_math.h"

namespace vgi {

namespace {

constexpr uint32_t BOOTSTRAP = 8u;  // samples below which a face wants coverage
constexpr uint32_t CLASSES = 4u;    // coverage, refinement, sweep, urgent
constexpr uint32_t NO_CLASS = 4u;

// log2 of a nonnegative float (-inf at 0): the exponent plus ln(m)/ln(2) of
// the mantissa m in [sqrt(1/2), sqrt(2)) from the atanh series
// ln(m) = 2 (z + z^3/3 + z^5/5 + ...), z = (m - 1) / (m + 1), |z| < 0.172,
// to z^11 (truncation below 4e-10). The selection buckets only need it to a
// few 1e-7; the host oracle evaluates the same operations in the same order.
__device__ __forceinline__ float log2_own(float x) {
    if (!(x > 0.0f)) return -__builtin_inff();
    uint32_t u = __float_as_uint(x);
    int e = (int)((u >> 23) & 255u);
    if (e == 0) {  // subnormal: scale into the normal range
        x = x * 8388608.0f;
        u = __float_as_uint(x);
        e = (int)((u >> 23) & 255u) - 23;
    }
    e -= 127;
    float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    float z = (m - 1.0f) / (m + 1.0f);
    float z2 = z * z;
    float series = 1.0f / 11.0f;
    series = series * z2 + 1.0f / 9.0f;
    series = series * z2 + 1.0f / 7.0f;
    series = series * z2 + 1.0f / 5.0f;
    series = series * z2 + 1.0f / 3.0f;
    series = series * z2 + 1.0f;
    float ln_m = 2.0f * z * series;
    return (float)e + ln_m * 1.44269504f;
}

__device__ __forceinline__ int page_of(int v) { return (v - (v < 0 ? 31 : 0)) / 32; }

// voxel_accumulate.wgsl:474-516 (candidate): the slot's (class << 8 | bucket),
// or NONE when it is not a candidate this batch.
__device__ __forceinline__ uint32_t candidate_key(const uint32_t *__restrict__ m, const SelectParams &p) {
    uint32_t identity = m[3];
    if ((identity & 8u) == 0u) return NONE;
    uint32_t schedule = m[5];
    // A host job's face stays the host's for its readback window (at most
    // three batches in flight): never raced by a pick.
    uint32_t last_cpu = schedule >> 4;
    if (last_cpu != 0u && ((p.batch - last_cpu) & 0x0fffffffu) < 4u) return NONE;
    uint32_t rank_w = m[11];
    uint32_t count = rank_w & 0x1fffffffu;
    bool fresh = (schedule & 8u) != 0u && m[7] == p.revision;
    uint32_t since = p.now_ms - m[9];
    uint32_t age_ms = rank_w < 0x80000000u ? 0x7fffffffu : (since > 0x7fffffffu ? 0u : since);
    uint32_t required = __uint_as_float(m[10]) < 0.02f ? max(p.stable_samples, 256u) : p.stable_samples;
    bool due = !fresh || count < required || (schedule & 7u) < 4u || age_ms >= p.audit_ms || (rank_w & RANK_URGENT) != 0u;
    if (!due) return NONE;
    int px = (int)m[0], py = (int)m[1], pz = (int)m[2];
    if ((rank_w & RANK_URGENT) != 0u || count < BOOTSTRAP) {
        float dx = (float)(px - p.fx), dy = (float)(py - p.fy), dz = (float)(pz - p.fz);
        uint32_t bucket = min((uint32_t)(log2_own(dx * dx + dy * dy + dz * dz + 1.0f) * 8.0f), 255u);
        return (((rank_w & RANK_URGENT) != 0u ? 3u : 0u) << 8) | bucket;
    }
    int gx = abs(page_of(px) - page_of(p.fx)), gy = abs(page_of(py) - page_of(p.fy)), gz = abs(page_of(pz) - page_of(p.fz));
    int page_distance = max(gx, max(gy, gz));
    if (page_distance <= 2) {
        float age = fmn((float)age_ms * 0.001f, 10.0f);
        float error = count < 2u ? 1.0f : fmn(__uint_as_float(m[8]) / fsqrt((float)count), 8.0f);
        float priority = error + age * 0.1f + (page_distance <= 1 ? 0.25f : 0.0f);
        // Log buckets: about 4% relative resolution from 0.004 to 50.
        uint32_t level = min((uint32_t)(fmx(log2_own(priority) + 8.0f, 0.0f) * 16.0f), 255u);
        return (1u << 8) | (255u - level);
    }
    uint32_t level = min((uint32_t)(log2_own((float)age_ms + 1.0f) * 8.0f), 255u);
    return (2u << 8) | (255u - level);
}

// One wavefront-wide reservation: lanes with `want` get consecutive indices
// from `*counter` (one atomic per wavefront); returns this lane's index.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t *__restrict__ counter, bool want) {
    const uint32_t lane = threadIdx.x & 63u;
    unsigned long long b = __ballot(want);
    uint32_t base = 0u;
    if (b != 0ull) {
        uint32_t leader = (uint32_t)__ffsll((long long)b) - 1u;
        if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(b));
        base = (uint32_t)__shfl((int)base, (int)leader, 64);
    }
    return base + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
}

__global__ void select_clear_kernel(uint32_t *__restrict__ selection, uint4 *__restrict__ pick_tasks,
                                    uint32_t *__restrict__ pick_slots, uint32_t picks) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < SEL_WORDS) selection[i] = 0u;
    if (i < picks) {
        pick_tasks[i * 2u] = make_uint4(0u, 0u, 0u, NONE);
        pick_tasks[i * 2u + 1u] = make_uint4(0u, 0u, 0u, 0u);
        pick_slots[i] = NONE;
    }
}

// voxel_accumulate.wgsl:518-551 (select_classify)
__global__ __launch_bounds__(256) void select_classify_kernel(uint32_t *__restrict__ meta, const uint32_t *__restrict__ state,
                                                              uint32_t *__restrict__ keys, uint32_t *__restrict__ selection,
                                                              uint4 *__restrict__ mirror, SelectParams p) {
    uint32_t slot = p.first + blockIdx.x * blockDim.x + threadIdx.x;
    bool stale = false;
    uint32_t *m = meta + (size_t)slot * 12u;
    if (slot < p.end) {
        keys[slot] = candidate_key(m, p);
        stale = (m[11] & MIRROR_STALE) != 0u && (m[3] & 8u) != 0u;
    }
    uint32_t record = wave_reserve(&selection[SEL_MIRRORS], stale);
    if (!stale || record >= p.mirror_budget) return;
    // Emitted: the mirror is current once the host merges this record.
    m[11] &= ~MIRROR_STALE;
    const uint4 *st = reinterpret_cast<const uint4 *>(state + (size_t)slot * 36u);
    uint4 *out = mirror + (size_t)record * 7u;
    out[0] = make_uint4(slot, st[7].x, m[6], 0u);
#pragma unroll
    for (int w = 0; w < 6; ++w) out[1 + w] = st[w];
    if ((m[5] & 15u) == 12u) out[2].y |= 0x80000000u;  // settled_flag
}

// voxel_accumulate.wgsl:553-575 (select_histogram), grid-stride.
__global__ __launch_bounds__(256) void select_histogram_kernel(const uint32_t *__restrict__ keys, uint32_t slots,
                                                               uint32_t *__restrict__ selection) {
    __shared__ uint32_t histogram[CLASSES * 256u];
    for (uint32_t b = threadIdx.x; b < CLASSES * 256u; b += blockDim.x) histogram[b] = 0u;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += gridDim.x * blockDim.x) {
        uint32_t key = keys[i];
        if (key != NONE) atomicAdd(&histogram[min(key >> 8, 3u) * 256u + (key & 255u)], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < CLASSES * 256u; b += blockDim.x) {
        uint32_t n = histogram[b];
        if (n != 0u) atomicAdd(&selection[b], n);
    }
}

// voxel_accumulate.wgsl:577-633 (select_threshold): one workgroup of 256.
__global__ __launch_bounds__(256) void select_threshold_kernel(uint32_t *__restrict__ selection, uint32_t budget) {
    __shared__ uint32_t scan[CLASSES][256];
    __shared__ uint32_t taken[CLASSES];
    const uint32_t t = threadIdx.x;
    for (uint32_t c = 0; c < CLASSES; ++c) scan[c][t] = selection[c * 256u + t];
    __syncthreads();
    for (uint32_t step = 1u; step < 256u; step <<= 1) {
        uint32_t add[CLASSES];
        for (uint32_t c = 0; c < CLASSES; ++c) add[c] = t >= step ? scan[c][t - step] : 0u;
        __syncthreads();
        for (uint32_t c = 0; c < CLASSES; ++c) scan[c][t] += add[c];
        __syncthreads();
    }
    if (t == 0u) {
        // Urgent faces first, up to three quarters of the budget; the rest by
        // quota (coverage 1/4, refinement 1/2, sweep 1/4), the unused share
        // lent to refinement, coverage, sweep, urgent in that order.
        uint32_t take[CLASSES];
        take[3] = min(scan[3][255], budget - budget / 4u);
        uint32_t rest = budget - take[3];
        take[0] = min(rest / 4u, scan[0][255]);
        take[1] = min(rest / 2u, scan[1][255]);
        take[2] = min(rest - rest / 4u - rest / 2u, scan[2][255]);
        uint32_t spare = budget - take[0] - take[1] - take[2] - take[3];
        const uint32_t lend[CLASSES] = {1u, 0u, 2u, 3u};
        for (uint32_t o = 0; o < CLASSES; ++o) {
            uint32_t c = lend[o];
            uint32_t extra = min(spare, scan[c][255] - take[c]);
            take[c] += extra;
            spare -= extra;
        }
        for (uint32_t c = 0; c < CLASSES; ++c) {
            taken[c] = take[c];
            selection[SEL_THRESHOLD + c] = 256u;
            selection[SEL_REMAINDER + c] = 0u;
        }
    }
    __syncthreads();
    for (uint32_t c = 0; c < CLASSES; ++c) {
        uint32_t below = t > 0u ? scan[c][t - 1u] : 0u;
        if (taken[c] > 0u && below < taken[c] && scan[c][t] >= taken[c]) {
            selection[SEL_THRESHOLD + c] = t;
            selection[SEL_REMAINDER + c] = taken[c] - below;
        }
    }
}

// voxel_accumulate.wgsl:635-695 (select_emit). Picks fill the pick list in
// reservation order; a pick's task takes the slot's current sequence.
__global__ __launch_bounds__(256) void select_emit_kernel(const uint32_t *__restrict__ meta, const uint32_t *__restrict__ state,
                                                          uint32_t *__restrict__ keys, uint32_t *__restrict__ selection,
                                                          uint4 *__restrict__ pick_tasks, uint32_t *__restrict__ pick_slots,
                                                          uint32_t slots, uint32_t budget) {
    uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t key = slot < slots ? keys[slot] : NONE;
    uint32_t group = min(key >> 8, 3u), bucket = key & 255u;
    uint32_t threshold = selection[SEL_THRESHOLD + group];
    bool candidate = key != NONE && bucket <= threshold;
    bool tie = candidate && bucket == threshold;
    uint32_t tie_index = 0u;
    for (uint32_t c = 0; c < CLASSES; ++c) {
        uint32_t index = wave_reserve(&selection[SEL_TIES + c], tie && group == c);
        if (group == c) tie_index = index;
    }
    bool accept = candidate && (!tie || tie_index < selection[SEL_REMAINDER + group]);
    uint32_t pick = wave_reserve(&selection[SEL_TOTAL], accept);
    if (!accept || pick >= budget) return;
    // Re-ranked when its quarter comes round; until then not picked again.
    keys[slot] = NONE;
    const uint32_t *m = meta + (size_t)slot * 12u;
    // A stale key of a retired slot: a hole in the pick list.
    if ((m[3] & 8u) == 0u) return;
    pick_tasks[pick * 2u] = make_uint4(m[0], m[1], m[2], m[3] & 7u);
    pick_tasks[pick * 2u + 1u] = make_uint4(state[(size_t)slot * 36u + 6u], 0u, 0u, 0u);
    pick_slots[pick] = slot;
}

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

}  // namespace

void launch_select(const SelectParams &p, uint32_t *meta, const uint32_t *state, uint32_t *keys, uint32_t *selection, uint4 *mirror,
                   uint4 *pick_tasks, uint32_t *pick_slots, uint32_t histogram_blocks, hipStream_t stream) {
    uint32_t clear = p.budget > SEL_WORDS ? p.budget : SEL_WORDS;
    hipLaunchKernelGGL(select_clear_kernel, dim3(blocks_for(clear, 256)), dim3(256), 0, stream, selection, pick_tasks, pick_slots,
                       p.budget);
    if (p.end > p.first)
        hipLaunchKernelGGL(select_classify_kernel, dim3(blocks_for(p.end - p.first, 256)), dim3(256), 0, stream, meta, state, keys,
                           selection, mirror, p);
    if (p.slots == 0u) return;
    uint32_t hb = histogram_blocks ? histogram_blocks : 1u;
    hb = hb < blocks_for(p.slots, 256) ? hb : blocks_for(p.slots, 256);
    hipLaunchKernelGGL(select_histogram_kernel, dim3(hb), dim3(256), 0, stream, keys, p.slots, selection);
    hipLaunchKernelGGL(select_threshold_kernel, dim3(1), dim3(256), 0, stream, selection, p.budget);
    hipLaunchKernelGGL(select_emit_kernel, dim3(blocks_for(p.slots, 256)), dim3(256), 0, stream, meta, state, keys, selection,
                       pick_tasks, pick_slots, p.slots, p.budget);
}

}  // namespace vgi
