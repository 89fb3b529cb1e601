// GPU work selection over the resident faces (voxel_accumulate.wgsl:448-695):
// which faces a batch traces when the GPU picks them itself.
//
//   classify   one thread per slot of this batch's rotating quarter: the
//              slot's class and key bucket (persisting until its quarter comes
//              round again), and statistics-mirror records of faces whose
//              host mirror is stale (one atomic per workgroup).
//   histogram  per workgroup, a 4 x 256-bin LDS histogram of the keys over a
//              grid-stride range; only touched bins reach the global one.
//   threshold  one workgroup: an inclusive scan per class, the class shares
//              of the pick budget, and each class's threshold bucket.
//   emit       keys below the threshold are picked, ties at it up to the
//              class's remainder: a stream compaction over contiguous slot
//              ranges (count, one-workgroup offset scan, write), ranks from
//              wavefront ballots.
//
// All four are HBM-bound integer work: classify reads 48 B of metadata and
// writes a 4 B key per classified slot, histogram and emit read 4 B per slot.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vgi_blockscan.h"
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
__device__ __forceinline__ uint32_t candidate_key(const uint32_t (&m)[12], const SelectParams &p) {
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
        const uint4 *m4 = reinterpret_cast<const uint4 *>(m);
        uint4 q0 = m4[0], q1 = m4[1], q2 = m4[2];
        const uint32_t words[12] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
        keys[slot] = candidate_key(words, p);
        stale = (q2.w & MIRROR_STALE) != 0u && (q0.w & 8u) != 0u;
    }
    // Mirror records reserved once per workgroup.
    __shared__ uint32_t wave_totals[1][4];
    __shared__ uint32_t group_base;
    bool flag[1] = {stale};
    uint32_t rank[1], total[1];
    block_ranks<1>(flag, rank, total, wave_totals);
    if (threadIdx.x == 0u) group_base = total[0] ? atomicAdd(&selection[SEL_MIRRORS], total[0]) : 0u;
    __syncthreads();
    uint32_t record = group_base + rank[0];
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
    // Lanes of a wavefront with equal keys add once (one lane per distinct
    // key): faces merged together share a bucket, and 64 lanes on one LDS
    // word would serialise.
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (slots + stride - 1u) / stride;
    for (uint32_t r = 0; r < rounds; ++r) {
        uint32_t i = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        uint32_t key = i < slots ? keys[i] : NONE;
        unsigned long long pending = __ballot(key != NONE);
        while (pending) {
            uint32_t leader = (uint32_t)__ffsll((long long)pending) - 1u;
            uint32_t k = (uint32_t)__shfl((int)key, (int)leader, 64);
            unsigned long long same = __ballot(key == k) & pending;
            if (lane == leader) atomicAdd(&histogram[min(k >> 8, 3u) * 256u + (k & 255u)], (uint32_t)__popcll(same));
            pending &= ~same;
        }
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

// select_emit (voxel_accumulate.wgsl:635-695) as a stream compaction in three
// passes instead of reservations on five global counters (those serialise
// ~10^5 atomics on one address per batch): each workgroup owns a contiguous
// slot range; count its accepted keys and its ties per class; one workgroup
// scans the counts into per-range offsets; then each range writes its picks
// at its offset. The rule is the reference's: keys below a class's threshold
// are accepted, and of the ties at it as many as the class's remainder --
// here the first ones in slot order, so the pick list is deterministic.
struct EmitFlags {
    bool strict;      // accepted below the threshold (or a class without one)
    bool tie;         // at the threshold bucket
    uint32_t group;
};
__device__ __forceinline__ EmitFlags emit_flags(uint32_t key, const uint32_t *__restrict__ selection) {
    EmitFlags f;
    f.group = min(key >> 8, 3u);
    uint32_t bucket = key & 255u, threshold = selection[SEL_THRESHOLD + f.group];
    bool candidate = key != NONE && bucket <= threshold;
    f.tie = candidate && bucket == threshold;
    f.strict = candidate && !f.tie;
    return f;
}

__global__ __launch_bounds__(256) void select_count_kernel(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ selection,
                                                           uint32_t slots, uint32_t per_block, uint32_t *__restrict__ counts) {
    __shared__ uint32_t wave_sums[5][4];
    uint32_t begin = blockIdx.x * per_block, end = min(begin + per_block, slots);
    uint32_t sum[5] = {0u, 0u, 0u, 0u, 0u};
    for (uint32_t slot = begin + threadIdx.x; slot < end; slot += 256u) {
        EmitFlags f = emit_flags(keys[slot], selection);
        sum[0] += f.strict ? 1u : 0u;
        if (f.tie) sum[1u + f.group] += 1u;
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        uint32_t v = sum[k];
        for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
        if (lane == 0u) wave_sums[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 5u)
        counts[blockIdx.x * 8u + threadIdx.x] = wave_sums[threadIdx.x][0] + wave_sums[threadIdx.x][1] + wave_sums[threadIdx.x][2] +
                                                wave_sums[threadIdx.x][3];
}

// One workgroup of 256: per range, the ties before it per class, the ties it
// accepts (up to each class's remainder), and the picks before it; the
// selection's tie and pick totals.
__global__ __launch_bounds__(256) void select_offsets_kernel(const uint32_t *__restrict__ counts, uint32_t ranges,
                                                             uint32_t *__restrict__ selection, uint32_t *__restrict__ offsets) {
    __shared__ uint32_t part[5][256];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (ranges + 255u) / 256u;
    uint32_t first = t * per, last = min(first + per, ranges);
    // This thread's ties per class, then their exclusive prefix over threads.
    uint32_t ties[4] = {0u, 0u, 0u, 0u};
    for (uint32_t r = first; r < last; ++r)
        for (int c = 0; c < 4; ++c) ties[c] += counts[r * 8u + 1u + c];
    for (int c = 0; c < 4; ++c) part[c][t] = ties[c];
    uint32_t tie_total[4];
    block_exclusive_scan<4>(part, tie_total);
    if (t < 4u) selection[SEL_TIES + t] = tie_total[t];
    uint32_t remainder[4], tie_before[4];
    for (int c = 0; c < 4; ++c) {
        remainder[c] = selection[SEL_REMAINDER + c];
        tie_before[c] = part[c][t];
    }
    // Picks of this thread's ranges: strict keys plus the ties within the remainder.
    uint32_t picks = 0u;
    for (uint32_t r = first; r < last; ++r) {
        uint32_t accepted = counts[r * 8u];
        for (int c = 0; c < 4; ++c) {
            uint32_t n = counts[r * 8u + 1u + c];
            uint32_t room = remainder[c] > tie_before[c] ? remainder[c] - tie_before[c] : 0u;
            accepted += min(room, n);
            tie_before[c] += n;
        }
        picks += accepted;
    }
    part[4][t] = picks;
    __syncthreads();
    uint32_t pick_total[1];
    block_exclusive_scan<1>(part + 4, pick_total);
    if (t == 0u) selection[SEL_TOTAL] = pick_total[0];
    // Each range's offsets.
    for (int c = 0; c < 4; ++c) tie_before[c] = part[c][t];
    uint32_t pick_before = part[4][t];
    for (uint32_t r = first; r < last; ++r) {
        offsets[r * 8u] = pick_before;
        uint32_t accepted = counts[r * 8u];
        for (int c = 0; c < 4; ++c) {
            offsets[r * 8u + 1u + c] = tie_before[c];
            uint32_t n = counts[r * 8u + 1u + c];
            uint32_t room = remainder[c] > tie_before[c] ? remainder[c] - tie_before[c] : 0u;
            accepted += min(room, n);
            tie_before[c] += n;
        }
        pick_before += accepted;
    }
}

// Picks fill the pick list in slot order; a pick's task takes the slot's
// current sequence (begin_trace).
__global__ __launch_bounds__(256) void select_emit_kernel(const uint32_t *__restrict__ meta, const uint32_t *__restrict__ state,
                                                          uint32_t *__restrict__ keys, const uint32_t *__restrict__ selection,
                                                          const uint32_t *__restrict__ offsets, uint32_t slots, uint32_t per_block,
                                                          uint32_t budget, uint4 *__restrict__ pick_tasks,
                                                          uint32_t *__restrict__ pick_slots) {
    __shared__ uint32_t wave_totals[5][4];
    uint32_t begin = blockIdx.x * per_block, end = min(begin + per_block, slots);
    uint32_t pick_base = offsets[blockIdx.x * 8u];
    uint32_t tie_base[4], remainder[4];
    for (int c = 0; c < 4; ++c) {
        tie_base[c] = offsets[blockIdx.x * 8u + 1u + c];
        remainder[c] = selection[SEL_REMAINDER + c];
    }
    for (uint32_t base = begin; base < end; base += 256u) {
        uint32_t slot = base + threadIdx.x;
        EmitFlags f = emit_flags(slot < end ? keys[slot] : NONE, selection);
        bool tflag[4] = {f.tie && f.group == 0u, f.tie && f.group == 1u, f.tie && f.group == 2u, f.tie && f.group == 3u};
        uint32_t trank[4], ttotal[4];
        block_ranks<4>(tflag, trank, ttotal, wave_totals);
        uint32_t tie_index = 0u, base_of_group = 0u, rank_of_group = 0u;
        for (int c = 0; c < 4; ++c)
            if (f.group == (uint32_t)c) { base_of_group = tie_base[c]; rank_of_group = trank[c]; }
        tie_index = base_of_group + rank_of_group;
        bool accept = f.strict || (f.tie && tie_index < remainder[f.group]);
        bool aflag[1] = {accept};
        uint32_t arank[1], atotal[1];
        block_ranks<1>(aflag, arank, atotal, wave_totals);
        uint32_t pick = pick_base + arank[0];
        pick_base += atotal[0];
        for (int c = 0; c < 4; ++c) tie_base[c] += ttotal[c];
        if (!accept || pick >= budget) continue;
        // Re-ranked when its quarter comes round; until then not picked again.
        keys[slot] = NONE;
        const uint32_t *m = meta + (size_t)slot * 12u;
        // A stale key of a retired slot: a hole in the pick list.
        if ((m[3] & 8u) == 0u) continue;
        pick_tasks[pick * 2u] = make_uint4(m[0], m[1], m[2], m[3] & 7u);
        pick_tasks[pick * 2u + 1u] = make_uint4(state[(size_t)slot * 36u + 6u], 0u, 0u, 0u);
        pick_slots[pick] = slot;
    }
}

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

}  // namespace

void launch_select(const SelectParams &p, uint32_t *meta, const uint32_t *state, uint32_t *keys, uint32_t *selection, uint4 *mirror,
                   uint4 *pick_tasks, uint32_t *pick_slots, uint32_t *ranges, uint32_t histogram_blocks, hipStream_t stream) {
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
    // Emit over contiguous slot ranges of whole 256-slot chunks.
    uint32_t n_ranges = blocks_for(p.slots, 256) < SELECT_RANGES ? blocks_for(p.slots, 256) : SELECT_RANGES;
    uint32_t per_range = blocks_for(blocks_for(p.slots, n_ranges), 256) * 256u;
    n_ranges = blocks_for(p.slots, per_range);
    uint32_t *counts = ranges, *offsets = ranges + SELECT_RANGES * 8u;
    hipLaunchKernelGGL(select_count_kernel, dim3(n_ranges), dim3(256), 0, stream, keys, selection, p.slots, per_range, counts);
    hipLaunchKernelGGL(select_offsets_kernel, dim3(1), dim3(256), 0, stream, counts, n_ranges, selection, offsets);
    hipLaunchKernelGGL(select_emit_kernel, dim3(n_ranges), dim3(256), 0, stream, meta, state, keys, selection, offsets, p.slots,
                       per_range, p.budget, pick_tasks, pick_slots);
}

}  // namespace vgi
