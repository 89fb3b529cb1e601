// HIP kernels for gfx950 (CDNA4, wave64): face tracing, GPU NAADF build and
// packing, and the resident world cache passes.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

#include <type_traits>

#include "vgi_blockscan.h"
#include "vgi_kernels.h"
#include "vgi_mlp.h"
#include "vgi_trace.h"

namespace vgi {

// ============================================================== face tracing
// voxel_trace.wgsl:44-92 (trace_moments + trace_faces_parallel): `lanes` lanes
// share a face; lane l traces samples l, l+lanes, ...; lane 0 adds the lanes'
// moments in lane order (zero-initialised, as the reference's workgroup sum)
// and writes the 96-byte record after one more analytic direct estimate
// (write_result, voxel_trace.wgsl:25-37). lanes divides 64, so a face's lanes
// share a wavefront and the sum uses cross-lane reads instead of LDS.
// Occupancy: the path estimator is one long dependent gather chain, so the
// kernel is latency bound; it is compiled for VGI_TRACE_WAVES waves per SIMD
// (the register budget that allows), trading a little spill for latency hiding.
// One face's samples on `lanes` consecutive threads of a wavefront (every
// thread of the wavefront calls this; out-of-range faces pass valid = false).
// kLearned: paths stop at their first hit into the learned cache; each
// sample's partial energy, prediction factor and cache query go to
// `learned` (2 float4 per sample) and `queries`, and the combine pass forms the
// face moments once the network has answered.
template <class Tally, bool kHits, bool kLearned>
__device__ __forceinline__ void trace_face(SceneView scene, Optics optics, const uint4 *__restrict__ tasks,
                                           uint32_t face_index, bool valid, uint32_t lane, uint32_t samples, uint32_t lanes,
                                           uint32_t sequence_base, float4 *__restrict__ results, uint32_t *__restrict__ hits,
                                           float4 *__restrict__ learned, MlpQuery *__restrict__ queries, Tally &tally) {
    uint4 t0 = make_uint4(0u, 0u, 0u, NONE), t1 = make_uint4(0u, 0u, 0u, 0u);
    if (valid) {
        t0 = tasks[face_index * 2u];
        t1 = tasks[face_index * 2u + 1u];
        valid = t0.w != NONE;
    }
    f3 mean = mk(0.0f, 0.0f, 0.0f), second = mk(0.0f, 0.0f, 0.0f);
    uint32_t known = 0u;
    int vx = (int)t0.x, vy = (int)t0.y, vz = (int)t0.z;
    uint32_t face = t0.w;
    f3 normal = mk(0.0f, 0.0f, 0.0f);
    if (valid) {
        normal = face_normal(face);
        f3 surface = mk((float)vx, (float)vy, (float)vz) + mk(0.5f, 0.5f, 0.5f) + normal * 0.5f;
        uint32_t stream = hash_voxel(vx, vy, vz) ^ pcg(face + 1u);
        for (uint32_t i = lane; i < samples; i += lanes) {
            bool k;
            FirstHit first;
            uint32_t rays_before = tally.ray_count();
            CacheQuery cq;
            cq.vx = cq.vy = cq.vz = 0;
            cq.face_material = NONE;
            cq.factor = mk(0.0f, 0.0f, 0.0f);
            f3 value = trace_receiver<Tally, kLearned>(scene, optics, surface, normal, stream, t1.x + sequence_base + i, k, first,
                                                       tally, &cq);
            if (kHits) {
                uint4 *h = reinterpret_cast<uint4 *>(hits + ((size_t)face_index * samples + i) * 8u);
                h[0] = make_uint4((uint32_t)first.vx, (uint32_t)first.vy, (uint32_t)first.vz, first.face | (first.status << 8));
                h[1] = make_uint4(first.material, first.slot, tally.ray_count() - rays_before, k ? 1u : 0u);
            }
            if constexpr (kLearned) {
                size_t at = (size_t)face_index * samples + i;
                learned[at * 2u] = make_float4(value.x, value.y, value.z, k ? 1.0f : 0.0f);
                learned[at * 2u + 1u] = make_float4(cq.factor.x, cq.factor.y, cq.factor.z, 0.0f);
                queries[at] = MlpQuery{cq.vx, cq.vy, cq.vz, k ? cq.face_material : NONE};
                continue;
            }
            if (!k) continue;
            mean = mean + value;
            second = second + value * value;
            known += 1u;
        }
    }
    // Lane 0 gathers the face's lanes in order.
    uint32_t base = (threadIdx.x & 63u) - lane;
    f3 tmean = mk(0.0f, 0.0f, 0.0f), tsecond = mk(0.0f, 0.0f, 0.0f);
    uint32_t tknown = 0u;
    for (uint32_t l = 0u; l < lanes; ++l) {
        int src = (int)(base + l);
        f3 m = mk(__shfl(mean.x, src, 64), __shfl(mean.y, src, 64), __shfl(mean.z, src, 64));
        f3 q = mk(__shfl(second.x, src, 64), __shfl(second.y, src, 64), __shfl(second.z, src, 64));
        uint32_t kn = (uint32_t)__shfl((int)known, src, 64);
        tmean = tmean + m;
        tsecond = tsecond + q;
        tknown += kn;
    }
    if (valid && lane == 0u) {
        float inverse = 1.0f / (float)(tknown > 1u ? tknown : 1u);
        float4 *r = results + (size_t)face_index * 6u;
        if constexpr (!kLearned) {
            r[0] = make_float4(tmean.x * inverse, tmean.y * inverse, tmean.z * inverse, 0.0f);
            r[1] = make_float4(0.0f, (float)tknown, 0.0f, 0.0f);
            r[2] = make_float4(tsecond.x * inverse, tsecond.y * inverse, tsecond.z * inverse, 0.0f);
        }
        r[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (!optics.direct_reuse) {
            f3 dv, dc;
            direct_at(scene, optics, mk((float)vx, (float)vy, (float)vz) + mk(0.5f, 0.5f, 0.5f) + normal * 0.5f, normal, dv, dc, tally);
            r[4] = make_float4(dv.x, dv.y, dv.z, 0.0f);
            r[5] = make_float4(0.0f, dc.x, dc.y, dc.z);
        }
    }
}

template <class Tally, bool kHits, bool kLearned>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VGI_TRACE_WAVES))) void trace_faces_kernel(
    SceneView scene, Optics optics, const uint4 *__restrict__ tasks, uint32_t count, uint32_t samples, uint32_t lanes,
    uint32_t sequence_base, float4 *__restrict__ results, uint32_t *__restrict__ hits, unsigned long long *__restrict__ counters,
    float4 *__restrict__ learned, MlpQuery *__restrict__ queries, const uint32_t *__restrict__ order) {
    uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t slot = tid / lanes;
    uint32_t lane = tid % lanes;
    bool valid = slot < count;
    uint32_t face_index = (order && valid) ? order[slot] : slot;
    Tally tally;
    trace_face<Tally, kHits, kLearned>(scene, optics, tasks, face_index, valid, lane, samples, lanes, sequence_base, results, hits,
                                       learned, queries, tally);
    if (counters == nullptr) return;
    if constexpr (std::is_same<Tally, FullTally>::value) {
        unsigned long long c[5] = {tally.rays, tally.steps, tally.headers, tally.bricks, tally.cells};
#pragma unroll
        for (int k = 0; k < 5; ++k)
            for (int off = 32; off > 0; off >>= 1) c[k] += __shfl_xor(c[k], off, 64);
        if ((threadIdx.x & 63u) == 0u)
            for (int k = 0; k < 5; ++k) atomicAdd(&counters[k], c[k]);
    } else if constexpr (std::is_same<Tally, RayTally>::value) {
        unsigned long long r = tally.rays;
        for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off, 64);
        if ((threadIdx.x & 63u) == 0u) atomicAdd(&counters[0], r);
    }
}

// ============================================================== batched frames
// Many scenes (or uneven chunks) in one persistent launch: workgroups take
// queue tiles (scene, face range) from a device counter until the queue is
// empty, and trace each tile's faces exactly as trace_faces_kernel does, so
// every scene's records equal its own launch's bit for bit. The last tile of
// a scene stamps the scene's finish time (per-scene latency).
__global__ void batch_reset_kernel(BatchControl *__restrict__ control, uint32_t *__restrict__ done,
                                   unsigned long long *__restrict__ finish, uint32_t scenes) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0u) {
        control->head = 0u;
        control->frame_seq = control->next_seq;
        control->next_seq = control->next_seq + control->samples;
        control->rays = 0ull;
        control->start = wall_clock64();
    }
    for (uint32_t k = i; k < scenes; k += gridDim.x * blockDim.x) {
        done[k] = 0u;
        finish[k] = 0ull;
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VGI_TRACE_WAVES))) void batch_trace_kernel(
    const BatchScene *__restrict__ scenes, const BatchTile *__restrict__ tiles, uint32_t n_tiles, uint32_t lanes,
    BatchControl *__restrict__ control, uint32_t *__restrict__ done, unsigned long long *__restrict__ finish) {
    __shared__ uint32_t s_tile;
    const uint32_t samples = control->samples;
    const uint32_t sequence = control->frame_seq;
    const uint32_t per_pass = blockDim.x / lanes;
    const uint32_t local = threadIdx.x / lanes, lane = threadIdx.x % lanes;
    RayTally tally;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0u) s_tile = atomicAdd(&control->head, 1u);
        __syncthreads();
        uint32_t ti = __builtin_amdgcn_readfirstlane(s_tile);
        if (ti >= n_tiles) break;
        BatchTile tile = tiles[ti];
        const BatchScene &sc = scenes[tile.scene];
        for (uint32_t f = 0u; f < tile.faces; f += per_pass) {
            uint32_t at = f + local;
            bool valid = at < tile.faces;
            uint32_t face = tile.first + at;
            if (sc.order && valid) face = sc.order[face];
            trace_face<RayTally, false, false>(sc.scene, sc.optics, sc.tasks, face, valid, lane, samples, lanes, sequence, sc.results,
                                               nullptr, nullptr, nullptr, tally);
        }
        __syncthreads();
        if (threadIdx.x == 0u) {
            __threadfence();
            uint32_t d = atomicAdd(&done[tile.scene], 1u);
            if (d + 1u == sc.tiles) finish[tile.scene] = wall_clock64();
        }
    }
    unsigned long long r = tally.rays;
    for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off, 64);
    if ((threadIdx.x & 63u) == 0u) atomicAdd(&control->rays, r);
}

// ============================================================== NAADF build
// One wavefront computes the NAADF of a 4^3 grid, one cell per lane (lane =
// x + 4y + 16z): 3 rounds of x, y, z Jacobi sweeps (naadf.rs:65-112, rounds =
// max(size) - 1). A cell's six u8 extents travel as two u32 (bytes 0-3, 4-5);
// neighbours along x, y, z are lanes +-1, +-4, +-16 of the same wavefront, read
// with cross-lane shuffles. The sweep from zero reaches the closed forms of
// naadf.rs:42-63 for uniform grids (naadf/construction_tests.rs:110-145).
__device__ __forceinline__ uint32_t ext_get(uint32_t lo, uint32_t hi, int d) {
    return d < 4 ? (lo >> (d * 8)) & 255u : (hi >> ((d - 4) * 8)) & 255u;
}
__device__ __forceinline__ void aadf_4cubed(bool blocked, uint32_t lane, uint32_t &lo, uint32_t &hi) {
    lo = 0u;
    hi = 0u;
    int x = (int)(lane & 3u), y = (int)((lane >> 2) & 3u), z = (int)(lane >> 4);
    for (int round = 0; round < 3; ++round) {
        for (int axis = 0; axis < 3; ++axis) {
            int stride = axis == 0 ? 1 : (axis == 1 ? 4 : 16);
            int coord = axis == 0 ? x : (axis == 1 ? y : z);
            // Neighbour data (source = previous sweep), fetched by every lane.
            int nlane = (int)lane - stride, plane = (int)lane + stride;
            uint32_t nlo = (uint32_t)__shfl((int)lo, nlane < 0 ? 0 : nlane, 64);
            uint32_t nhi = (uint32_t)__shfl((int)hi, nlane < 0 ? 0 : nlane, 64);
            uint32_t nbl = (uint32_t)__shfl((int)blocked, nlane < 0 ? 0 : nlane, 64);
            uint32_t plo = (uint32_t)__shfl((int)lo, plane > 63 ? 63 : plane, 64);
            uint32_t phi = (uint32_t)__shfl((int)hi, plane > 63 ? 63 : plane, 64);
            uint32_t pbl = (uint32_t)__shfl((int)blocked, plane > 63 ? 63 : plane, 64);
            uint32_t cur[6];
            for (int d = 0; d < 6; ++d) cur[d] = ext_get(lo, hi, d);
            if (!blocked) {
                // try_grow (naadf.rs:117-148) toward -axis, then +axis, reading
                // the neighbour's previous-sweep extents.
                for (int side = 0; side < 2; ++side) {
                    int direction = axis * 2 + side;
                    bool inside = side == 0 ? coord > 0 : coord < 3;
                    uint32_t ol = side == 0 ? nlo : plo, oh = side == 0 ? nhi : phi;
                    bool ob = side == 0 ? nbl != 0u : pbl != 0u;
                    if (!inside || cur[direction] == 255u || ob) continue;
                    int opposite = direction ^ 1;
                    bool grow = true;
                    for (int d = 0; d < 6; ++d)
                        if (d != opposite && ext_get(ol, oh, d) < cur[d]) grow = false;
                    if (grow) cur[direction] += 1u;
                }
                lo = cur[0] | (cur[1] << 8) | (cur[2] << 16) | (cur[3] << 24);
                hi = cur[4] | (cur[5] << 8);
            } else {
                lo = 0u;
                hi = 0u;
            }
        }
    }
}
// gpu_scene.rs:462-467: six 2-bit extents, NEG_X in bits 0-1 ... POS_Z in 10-11.
__device__ __forceinline__ uint32_t pack_extents2(uint32_t lo, uint32_t hi) {
    uint32_t bits = 0u;
    for (int d = 0; d < 6; ++d) bits |= ext_get(lo, hi, d) << (d * 2);
    return bits;
}

// world.rs:243-319 (NaadfRegion::build) per region, one 256-thread block per
// listed region: storage class, per-brick uniformity, brick materials and
// occupancy. Voxels come from the dense u16 grid (x fastest).
__global__ __launch_bounds__(256) void classify_regions_kernel(const uint16_t *__restrict__ grid, int gx, int gy, int gz,
                                                               int vminx, int vminy, int vminz, const int4 *__restrict__ list,
                                                               uint32_t list_count, RegionInfo *__restrict__ info) {
    __shared__ uint16_t mats[4096];
    __shared__ uint32_t flags[4];
    uint32_t r = blockIdx.x;
    if (r >= list_count) return;
    int4 reg = list[r];
    int ox = reg.x * 16 - vminx, oy = reg.y * 16 - vminy, oz = reg.z * 16 - vminz;
    // 256 rows of 16 voxels: row = y + 16 z.
    {
        int row = (int)threadIdx.x;
        int y = row & 15, z = row >> 4;
        const uint16_t *src = grid + (size_t)(ox) + (size_t)gx * ((size_t)(oy + y) + (size_t)gy * (size_t)(oz + z));
        for (int x = 0; x < 16; ++x) mats[row * 16 + x] = src[x];
    }
    if (threadIdx.x < 4) flags[threadIdx.x] = 0u;
    __syncthreads();
    uint16_t first = mats[0];
    bool any_solid_local = false, all_same_local = true;
    for (int i = (int)threadIdx.x; i < 4096; i += 256) {
        any_solid_local |= mats[i] != AIR;
        all_same_local &= mats[i] == first;
    }
    if (any_solid_local) atomicOr(&flags[0], 1u);
    if (!all_same_local) atomicOr(&flags[1], 1u);
    // Bricks: thread b < 64 scans brick b (x fastest brick order).
    uint32_t brick_uniform = 1u, brick_solid = 0u;
    uint16_t brick_first = AIR;
    if (threadIdx.x < 64) {
        int b = (int)threadIdx.x;
        int bx = b & 3, by = (b >> 2) & 3, bz = b >> 4;
        brick_first = mats[(bz * 4) * 256 + (by * 4) * 16 + bx * 4];
        for (int i = 0; i < 64; ++i) {
            int lx = i & 3, ly = (i >> 2) & 3, lz = i >> 4;
            uint16_t m = mats[(bz * 4 + lz) * 256 + (by * 4 + ly) * 16 + bx * 4 + lx];
            if (m != brick_first) brick_uniform = 0u;
            if (m != AIR) brick_solid = 1u;
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        // wave 0 holds the 64 bricks: ballots give the masks.
        uint64_t uniform_mask = __ballot(brick_uniform != 0u);
        uint64_t solid_mask = __ballot(brick_solid != 0u);
        if (threadIdx.x < 64) info[r].brick_material[threadIdx.x] = brick_first;
        if (threadIdx.x == 0) {
            RegionInfo &ri = info[r];
            bool any_solid = flags[0] != 0u;
            bool all_same = flags[1] == 0u;
            uint32_t storage;
            if (!any_solid) storage = 0u;                 // Empty
            else if (all_same) storage = 1u;              // Uniform
            else if (uniform_mask == ~0ull) storage = 3u; // Bricks
            else storage = 2u;                            // Dense
            ri.storage = storage;
            ri.uniform = first;
            uint64_t mixed = storage == 2u ? ~uniform_mask : 0ull;
            ri.mixed_lo = (uint32_t)mixed;
            ri.mixed_hi = (uint32_t)(mixed >> 32);
            // any_solid (world.rs:308): Dense from brick occupancy, Bricks from
            // any non-air brick; both equal "some voxel is solid" here.
            ri.any_solid = any_solid ? 1u : 0u;
            ri.occupied_lo = (uint32_t)solid_mask;
            ri.occupied_hi = (uint32_t)(solid_mask >> 32);
            ri.mixed_count = (uint32_t)__popcll(mixed);
            ri.table = storage >= 2u ? 1u : 0u;
        }
    }
}

// Exclusive scans of table and mixed-brick counts over the listed regions, in
// list order, one block of 1024 threads looping over chunks. Offsets start at
// `cells_base` (words): tables first, then bricks.
__global__ __launch_bounds__(1024) void scan_regions_kernel(RegionInfo *__restrict__ info, uint32_t count, uint32_t cells_base,
                                                            uint32_t *__restrict__ totals) {
    __shared__ uint32_t sums_t[1024], sums_b[1024];
    __shared__ uint32_t carry_t, carry_b;
    if (threadIdx.x == 0) { carry_t = 0u; carry_b = 0u; }
    __syncthreads();
    // Pass 1: tables.
    uint32_t total_tables = 0u;
    for (uint32_t start = 0; start < count; start += 1024u) {
        uint32_t i = start + threadIdx.x;
        uint32_t vt = i < count ? info[i].table : 0u;
        uint32_t vb = i < count ? info[i].mixed_count : 0u;
        sums_t[threadIdx.x] = vt;
        sums_b[threadIdx.x] = vb;
        __syncthreads();
        for (uint32_t off = 1; off < 1024u; off <<= 1) {
            uint32_t at = threadIdx.x >= off ? sums_t[threadIdx.x - off] : 0u;
            uint32_t ab = threadIdx.x >= off ? sums_b[threadIdx.x - off] : 0u;
            __syncthreads();
            sums_t[threadIdx.x] += at;
            sums_b[threadIdx.x] += ab;
            __syncthreads();
        }
        if (i < count) {
            info[i].table_rank = carry_t + sums_t[threadIdx.x] - vt;
            info[i].brick_rank = carry_b + sums_b[threadIdx.x] - vb;
        }
        __syncthreads();
        if (threadIdx.x == 1023u) { carry_t += sums_t[1023]; carry_b += sums_b[1023]; }
        __syncthreads();
    }
    total_tables = carry_t;
    if (threadIdx.x == 0) {
        totals[0] = carry_t;
        totals[1] = carry_b;
    }
    // Offsets in words.
    for (uint32_t i = threadIdx.x; i < count; i += 1024u) {
        info[i].table_offset = cells_base + info[i].table_rank * 64u;
        info[i].brick_offset = cells_base + total_tables * 64u + info[i].brick_rank * 32u;
    }
}

// gpu_scene.rs:186-239 (brick table and mixed-brick payloads), one 256-thread
// block per listed region with a table: wavefront w handles bricks w, w+4, ...
// Brick bounds (world.rs:372-384) come from the occupancy mask, voxel bounds
// (world.rs:352-370) from each mixed brick's own voxels.
__global__ __launch_bounds__(256) void pack_regions_kernel(const uint16_t *__restrict__ grid, int gx, int gy, int gz, int vminx,
                                                           int vminy, int vminz, const int4 *__restrict__ list,
                                                           uint32_t list_count, const RegionInfo *__restrict__ info,
                                                           uint32_t *__restrict__ cells) {
    __shared__ uint16_t mats[4096];
    __shared__ uint32_t brick_extents[64];
    uint32_t r = blockIdx.x;
    if (r >= list_count) return;
    const RegionInfo &ri = info[r];
    if (ri.table == 0u) return;
    int4 reg = list[r];
    int ox = reg.x * 16 - vminx, oy = reg.y * 16 - vminy, oz = reg.z * 16 - vminz;
    {
        int row = (int)threadIdx.x;
        int y = row & 15, z = row >> 4;
        const uint16_t *src = grid + (size_t)(ox) + (size_t)gx * ((size_t)(oy + y) + (size_t)gy * (size_t)(oz + z));
        for (int x = 0; x < 16; ++x) mats[row * 16 + x] = src[x];
    }
    uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    uint64_t occupied = ((uint64_t)ri.occupied_hi << 32) | ri.occupied_lo;
    uint64_t mixed = ((uint64_t)ri.mixed_hi << 32) | ri.mixed_lo;
    if (wave == 0) {
        uint32_t lo, hi;
        aadf_4cubed(((occupied >> lane) & 1ull) != 0ull, lane, lo, hi);
        brick_extents[lane] = pack_extents2(lo, hi);
    }
    __syncthreads();
    for (uint32_t b = wave; b < 64u; b += 4u) {
        uint32_t word;
        uint16_t bm = ri.brick_material[b];
        bool is_mixed = ((mixed >> b) & 1ull) != 0ull;
        if (!is_mixed) {
            word = bm == AIR ? (brick_extents[b] << 2) : (((uint32_t)bm << 2) | 1u);
        } else {
            uint32_t bx = b & 3u, by = (b >> 2) & 3u, bz = b >> 4;
            uint32_t lx = lane & 3u, ly = (lane >> 2) & 3u, lz = lane >> 4;
            uint16_t m = mats[(bz * 4u + lz) * 256u + (by * 4u + ly) * 16u + bx * 4u + lx];
            bool blocked = m != AIR;
            uint32_t lo, hi;
            aadf_4cubed(blocked, lane, lo, hi);
            uint32_t cell = blocked ? (((uint32_t)m << 1) | 1u) : (pack_extents2(lo, hi) << 1);
            uint32_t other = (uint32_t)__shfl((int)cell, (int)(lane | 1u), 64);
            uint32_t rank = (uint32_t)__popcll(mixed & ((1ull << b) - 1ull));
            uint32_t pointer = ri.brick_offset + rank * 32u;
            if ((lane & 1u) == 0u) cells[pointer + (lane >> 1)] = cell | (other << 16);
            word = (pointer << 2) | 2u;
        }
        if (lane == 0u) cells[ri.table_offset + b] = word;
    }
}

// Region headers (gpu_scene.rs:174-195): kind 1 empty / 2 uniform / 3 table,
// with the top-level extents supplied by the host (2 words per listed region).
__global__ void write_headers_kernel(const int4 *__restrict__ list, uint32_t list_count, const RegionInfo *__restrict__ info,
                                     const uint32_t *__restrict__ top, uint4 *__restrict__ headers, int rminx, int rminy, int rminz,
                                     int rsx, int rsy) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= list_count) return;
    int4 reg = list[i];
    const RegionInfo &ri = info[i];
    uint4 h;
    if (ri.storage == 0u) h = make_uint4(1u, AIR, top[i * 2], top[i * 2 + 1]);
    else if (ri.storage == 1u) h = make_uint4(2u, ri.uniform, top[i * 2], top[i * 2 + 1]);
    else h = make_uint4(3u, ri.table_offset, top[i * 2], top[i * 2 + 1]);
    size_t at = (size_t)(reg.x - rminx) + (size_t)rsx * ((size_t)(reg.y - rminy) + (size_t)rsy * (size_t)(reg.z - rminz));
    headers[at] = h;
}

// Top-level extents of every loaded region (header words z, w).
__global__ void write_top_kernel(const uint32_t *__restrict__ top, uint32_t cells, uint4 *__restrict__ headers) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    uint4 h = headers[i];
    if (h.x == 0u) return;
    h.z = top[i * 2];
    h.w = top[i * 2 + 1];
    headers[i] = h;
}

// Voxel edits into the dense grid.
__global__ void scatter_voxels_kernel(uint16_t *__restrict__ grid, const uint32_t *__restrict__ index,
                                      const uint16_t *__restrict__ values, uint32_t count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) grid[index[i]] = values[i];
}

// ============================================================== resident cache
// prepare (voxel_accumulate.wgsl:83-133) for listed faces: a traced face takes
// its slot's sequence (begin_trace); then mark_changed (:432-446).
__device__ __forceinline__ f3 load3(const uint32_t *w) { return mk(__uint_as_float(w[0]), __uint_as_float(w[1]), __uint_as_float(w[2])); }
__device__ __forceinline__ bool binds(const uint32_t *st, float change, float physical) {
    f3 w = mk(0.2126f, 0.7152f, 0.0722f);
    f3 mean = load3(st);
    f3 sec = load3(st + 8);
    f3 var = mk(fmx(sec.x - mean.x * mean.x, 0.0f), fmx(sec.y - mean.y * mean.y, 0.0f), fmx(sec.z - mean.z * mean.z, 0.0f));
    bool lit = dot3(w, mean) >= 0.05f * change;
    float keep = physical * dot3(w, var) / fmx(change * change, 1e-30f);
    return !lit || keep < (float)(st[5] & 0x7fffffffu);
}

// prepare (voxel_accumulate.wgsl:83-133) for the faces a batch traces:
// begin_trace (the task takes the slot's sequence) and, for host-requested
// faces (batch != 0), mirror_count and the CPU job's claim on the face (GPU
// selection leaves it for four batches); then mark_changed
// (voxel_accumulate.wgsl:432-446).
__global__ void prepare_kernel(uint4 *__restrict__ tasks, const uint32_t *__restrict__ slots, uint32_t count,
                               uint32_t *__restrict__ state, uint32_t *__restrict__ meta, float physical, uint32_t batch,
                               uint32_t *__restrict__ keys) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t s = slots[i];
    if (s == NONE) return;
    uint32_t *st = state + (size_t)s * 36u;
    tasks[i * 2u + 1u].x = st[6];
    uint32_t *m = meta + (size_t)s * 12u;
    if (batch != 0u) {
        m[11] = (m[11] & 0xe0000000u) | (st[5] & 0x1fffffffu);
        m[5] = (m[5] & 0xfu) | (batch << 4);
        if (keys) keys[s] = NONE;
    }
    uint32_t *pw = m + 4u;
    float pending = __uint_as_float(*pw);
    if (pending > 0.0f) {
        if (binds(st, pending, physical)) st[5] |= 0x80000000u;
        else *pw = 0u;
    }
}

// Worst per-channel relative standard deviation of one sample, RGB with a
// 0.002 floor (voxel_accumulate.wgsl:132-137 sample_noise).
__device__ __forceinline__ float sample_noise(f3 mean, f3 second) {
    return max3(mk(fsqrt(fmx(second.x - mean.x * mean.x, 0.0f)) / (0.002f + fabsf(mean.x)),
                   fsqrt(fmx(second.y - mean.y * mean.y, 0.0f)) / (0.002f + fabsf(mean.y)),
                   fsqrt(fmx(second.z - mean.z * mean.z, 0.0f)) / (0.002f + fabsf(mean.z))));
}
// voxel_accumulate.wgsl:145-151 refinement_error: the same over sqrt(count), at most 8.
__device__ __forceinline__ float refinement_error(f3 mean, f3 second, uint32_t count) {
    if (count < 2u) return 1.0f;
    float c = (float)count;
    return fmn(max3(mk(fsqrt(fmx(second.x - mean.x * mean.x, 0.0f) / c) / (0.002f + fabsf(mean.x)),
                       fsqrt(fmx(second.y - mean.y * mean.y, 0.0f) / c) / (0.002f + fabsf(mean.y)),
                       fsqrt(fmx(second.z - mean.z * mean.z, 0.0f) / c) / (0.002f + fabsf(mean.z)))),
               8.0f);
}
// voxel_accumulate.wgsl:138-144: dark faces need more evidence before resting.
__device__ __forceinline__ uint32_t required_samples(f3 mean, uint32_t stable_samples) {
    return max3(mean) < 0.02f ? max(stable_samples, 256u) : stable_samples;
}

// record_stability (voxel_accumulate.wgsl:99-130, scheduling.rs
// record_face_batch): a face's stable-batch count advances while it stays in
// the same transport revision, its display has settled on its mean, it has
// its required samples and both its relative error and this batch's drift
// are small; then its ranking record (noise, merge time, brightest channel,
// measured | mirror-stale | count) for GPU selection.
__device__ __forceinline__ void record_stability(uint32_t *__restrict__ m, const uint32_t *__restrict__ st, float4 before,
                                                 const AccumulateParams &p) {
    uint32_t schedule = m[5];
    bool same_revision = (schedule & 8u) != 0u && m[7] == p.revision;
    uint32_t count = st[5] & 0x7fffffffu;
    f3 after = load3(st);
    f3 second = load3(st + 8);
    float tolerance = p.stable_error;
    f3 b = mk(before.x, before.y, before.z);
    float drift = max3(mk(fabsf(b.x - after.x) / (0.002f + fmx(fabsf(b.x), fabsf(after.x))),
                          fabsf(b.y - after.y) / (0.002f + fmx(fabsf(b.y), fabsf(after.y))),
                          fabsf(b.z - after.z) / (0.002f + fmx(fabsf(b.z), fabsf(after.z)))));
    f3 display = load3(st + 24);
    bool settled = max3(mk(fabsf(display.x - after.x), fabsf(display.y - after.y), fabsf(display.z - after.z))) <=
                   0.001f + max3(after) * tolerance * 0.5f;
    uint32_t stable = (same_revision && settled && count >= required_samples(after, p.stable_samples) &&
                       refinement_error(after, second, count) <= tolerance && drift <= tolerance * 0.5f)
                          ? min((schedule & 7u) + 1u, 4u)
                          : 0u;
    m[5] = (schedule & 0xfffffff0u) | 8u | stable;
    m[7] = p.revision;
    uint32_t stale = p.picked ? MIRROR_STALE : (m[11] & MIRROR_STALE);
    m[8] = __float_as_uint(sample_noise(after, second));
    m[9] = p.clock_ms;
    m[10] = __float_as_uint(max3(after));
    m[11] = 0x80000000u | count | stale;
}

// accumulate (voxel_accumulate.wgsl:135-227): merge the batch's raw moments
// into the slot with the history cap, recency and measured-retention keep,
// then the display EMA; record_stability; the settled flag on the record.
__global__ void accumulate_kernel(const uint32_t *__restrict__ slots, uint32_t count, float4 *__restrict__ results,
                                  uint32_t *__restrict__ state, uint32_t *__restrict__ checkpoint, uint32_t *__restrict__ meta,
                                  AccumulateParams p) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t s = slots[i];
    if (s == NONE) return;
    // The slot's state and metadata rows and the batch record, held in
    // registers: whole rows move as 16-byte loads and stores (the rows are
    // 144, 48 and 96 bytes) instead of word by word.
    uint4 *srow = reinterpret_cast<uint4 *>(state + (size_t)s * 36u);
    uint4 *crow = reinterpret_cast<uint4 *>(checkpoint + (size_t)s * 12u);
    uint4 *mrow = reinterpret_cast<uint4 *>(meta + (size_t)s * 12u);
    uint4 *orow = reinterpret_cast<uint4 *>(results + (size_t)i * 6u);
    uint32_t st[36], cp[12], mw[12], out[24];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        uint4 v = srow[k];
        st[4 * k] = v.x; st[4 * k + 1] = v.y; st[4 * k + 2] = v.z; st[4 * k + 3] = v.w;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        uint4 v = orow[k];
        out[4 * k] = v.x; out[4 * k + 1] = v.y; out[4 * k + 2] = v.z; out[4 * k + 3] = v.w;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint4 v = mrow[k];
        mw[4 * k] = v.x; mw[4 * k + 1] = v.y; mw[4 * k + 2] = v.z; mw[4 * k + 3] = v.w;
    }
    cp[0] = st[0]; cp[1] = st[1]; cp[2] = st[2]; cp[3] = st[3];
    cp[4] = st[8]; cp[5] = st[9]; cp[6] = st[10]; cp[7] = st[11];
    cp[8] = st[4]; cp[9] = st[5] & 0x7fffffffu; cp[10] = st[12]; cp[11] = 0u;
    uint32_t fresh_count = (uint32_t)__uint_as_float(out[5]);
    uint32_t *pw = mw + 4;
    float pending = __uint_as_float(*pw);
    if (fresh_count > 0u) *pw = 0u;
    uint32_t previous_count = st[5] & 0x7fffffffu;
    float physical = p.physical;
    float change = (st[5] & 0x80000000u) != 0u ? pending : 0.0f;
    if (physical > 0.0f && fresh_count > 0u && change > 0.0f) {
        f3 w = mk(0.2126f, 0.7152f, 0.0722f);
        f3 om = load3(st), os = load3(st + 8);
        f3 ov = mk(fmx(os.x - om.x * om.x, 0.0f), fmx(os.y - om.y * om.y, 0.0f), fmx(os.z - om.z * om.z, 0.0f));
        f3 fm = load3(out), fs = load3(out + 8);
        f3 fv = mk(fmx(fs.x - fm.x * fm.x, 0.0f), fmx(fs.y - fm.y * fm.y, 0.0f), fmx(fs.z - fm.z * fm.z, 0.0f));
        float variance = dot3(w, om) < 0.05f * change ? dot3(w, fv) : dot3(w, ov);
        float keep = physical * variance / (change * change);
        uint32_t kept = (uint32_t)fmn(keep, 65536.0f);
        previous_count = previous_count < kept ? previous_count : kept;
        cp[9] = cp[9] < previous_count ? cp[9] : previous_count;
    }
    const float4 before = make_float4(__uint_as_float(st[0]), __uint_as_float(st[1]), __uint_as_float(st[2]), __uint_as_float(st[3]));
    uint32_t total = previous_count + fresh_count;
    uint32_t cnt = total < p.history_limit ? total : p.history_limit;
    float alpha = fresh_count > 0u ? fmn(p.recency * (float)fresh_count / (float)(cnt > 1u ? cnt : 1u), 1.0f) : 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float pm = __uint_as_float(st[k]), m = __uint_as_float(out[k]);
        st[k] = __float_as_uint(pm + (m - pm) * alpha);
        float ps = __uint_as_float(st[8 + k]), q = __uint_as_float(out[8 + k]);
        st[8 + k] = __float_as_uint(ps + (q - ps) * alpha);
    }
    st[5] = cnt;
    st[6] = st[6] + p.samples;
#pragma unroll
    for (int k = 0; k < 8; ++k) st[16 + k] = out[16 + k];
    f3 radiance = load3(st), moment = load3(st + 8);
    float fc = fmx((float)cnt, 1.0f);
    f3 error = mk(fsqrt(fmx(moment.x - radiance.x * radiance.x, 0.0f) / fc), fsqrt(fmx(moment.y - radiance.y * radiance.y, 0.0f) / fc),
                  fsqrt(fmx(moment.z - radiance.z * radiance.z, 0.0f) / fc));
    float uncertainty = cnt < 2u ? 1.0f
                                 : fmn(8.0f, max3(mk(error.x / (0.02f + fabsf(radiance.x)), error.y / (0.02f + fabsf(radiance.y)),
                                                     error.z / (0.02f + fabsf(radiance.z)))));
    float noise = cnt < 2u ? 1.0f
                           : fmn(8.0f, max3(mk(error.x / (0.002f + fabsf(radiance.x)), error.y / (0.002f + fabsf(radiance.y)),
                                               error.z / (0.002f + fabsf(radiance.z)))));
    float elapsed = st[30] != 0u ? (float)(p.clock_ms - st[29]) * 0.001f : 1.0f / 60.0f;
    float time_alpha = 1.0f - exp_own(-elapsed / p.response);
    float evidence = (float)cnt / ((float)cnt + 32.0f);
    float blend = fmn(time_alpha, evidence / (1.0f + 16.0f * noise * noise));
    f3 prev = load3(st + 24);
    float prev_a = __uint_as_float(st[27]);
    float old_weight = (1.0f - blend) / (0.25f + max3(prev));
    float new_weight = blend / (0.25f + max3(radiance));
    float alpha_display = fmn(blend, new_weight / fmx(old_weight + new_weight, 0.0000001f));
    float confidence = fclamp((float)cnt / ((float)cnt + 8.0f) / (1.0f + uncertainty), 0.125f, 1.0f);
    f3 disp = prev * (1.0f - alpha_display) + radiance * alpha_display;
    float disp_a = prev_a * (1.0f - blend) + confidence * blend;
    st[24] = __float_as_uint(disp.x); st[25] = __float_as_uint(disp.y); st[26] = __float_as_uint(disp.z);
    st[27] = __float_as_uint(disp_a);
    st[29] = p.clock_ms;
    st[30] = 0xffffffffu;
    st[7] = st[24];
    st[13] = st[25]; st[14] = st[26]; st[15] = st[27];
#pragma unroll
    for (int k = 0; k < 24; ++k) out[k] = st[k];
    record_stability(mw, st, before, p);
    if ((mw[5] & 15u) == 12u) out[5] |= 0x80000000u;  // settled_flag
#pragma unroll
    for (int k = 0; k < 9; ++k) srow[k] = make_uint4(st[4 * k], st[4 * k + 1], st[4 * k + 2], st[4 * k + 3]);
#pragma unroll
    for (int k = 0; k < 3; ++k) crow[k] = make_uint4(cp[4 * k], cp[4 * k + 1], cp[4 * k + 2], cp[4 * k + 3]);
#pragma unroll
    for (int k = 0; k < 3; ++k) mrow[k] = make_uint4(mw[4 * k], mw[4 * k + 1], mw[4 * k + 2], mw[4 * k + 3]);
#pragma unroll
    for (int k = 0; k < 6; ++k) orow[k] = make_uint4(out[4 * k], out[4 * k + 1], out[4 * k + 2], out[4 * k + 3]);
}

// voxel_retention.wgsl:14-18
__device__ __forceinline__ uint32_t box_distance2(int lx, int ly, int lz, int hx, int hy, int hz, int px, int py, int pz) {
    int sx = max(max(lx - px, px - hx), 0), sy = max(max(ly - py, py - hy), 0), sz = max(max(lz - pz, pz - hz), 0);
    uint32_t ux = (uint32_t)min(sx, 4096), uy = (uint32_t)min(sy, 4096), uz = (uint32_t)min(sz, 4096);
    return ux * ux + uy * uy + uz * uz;
}
__device__ __forceinline__ uint32_t curve_limit(uint32_t d2) {
    if (d2 >= 36864u) return NONE;
    return (d2 * d2 + 10368u) / 20736u;
}

// apply_events (voxel_accumulate.wgsl:295-332) over every slot below the
// high-water mark, with resolve_retention (voxel_retention.wgsl:29-96).
// A workgroup of 256 slots stages each event's change boxes through LDS, 256
// at a time: each thread loads one box and keeps it only when it lies within
// the reach of the workgroup's slots' bounding box (a box no slot can reach
// changes no slot: resolve_retention skips it per face); the kept boxes are
// compacted in index order (ballot ranks), and every slot then runs the
// exact per-face rule over them from LDS -- the same boxes in the same order,
// so the sums are the per-face loop's.
__device__ __forceinline__ uint32_t box_box_distance2(int alx, int aly, int alz, int ahx, int ahy, int ahz, int blx, int bly, int blz, int bhx,
                                                      int bhy, int bhz) {
    uint32_t sx = (uint32_t)min(max(max(alx - bhx, blx - ahx), 0), 4096);
    uint32_t sy = (uint32_t)min(max(max(aly - bhy, bly - ahy), 0), 4096);
    uint32_t sz = (uint32_t)min(max(max(alz - bhz, blz - ahz), 0), 4096);
    return sx * sx + sy * sy + sz * sz;
}

__global__ __launch_bounds__(256) void apply_events_kernel(const uint4 *__restrict__ events, uint32_t event_count, uint32_t high_water,
                                                          uint32_t *__restrict__ state, uint32_t *__restrict__ checkpoint,
                                                          uint32_t *__restrict__ meta, float physical, uint32_t reach,
                                                          uint32_t urgent_measured, uint32_t urgent_curve) {
    __shared__ uint4 s_box[256][2];
    __shared__ uint32_t s_wave[1][4];
    __shared__ int s_bounds[4][6];
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    bool live = false;
    int px = 0, py = 0, pz = 0;
    if (slot < high_water) {
        const uint4 id = reinterpret_cast<const uint4 *>(meta)[(size_t)slot * 3u];
        live = (id.w & 8u) != 0u;
        px = (int)id.x; py = (int)id.y; pz = (int)id.z;
    }
    // The workgroup's bounding box of live slots (empty: every box culled).
    int b[6] = {live ? px : INT_MAX, live ? py : INT_MAX, live ? pz : INT_MAX, live ? px : INT_MIN, live ? py : INT_MIN, live ? pz : INT_MIN};
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            int o = __shfl_xor(b[k], off, 64);
            b[k] = k < 3 ? min(b[k], o) : max(b[k], o);
        }
    if (lane == 0u)
        for (int k = 0; k < 6; ++k) s_bounds[wave][k] = b[k];
    __syncthreads();
    int g[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        g[k] = s_bounds[0][k];
        for (int w = 1; w < 4; ++w) g[k] = k < 3 ? min(g[k], s_bounds[w][k]) : max(g[k], s_bounds[w][k]);
    }
    const bool any_live = g[0] <= g[3];
    bool measured_mode = physical > 0.0f;
    uint32_t cap = NONE, source_cap = NONE;
    float change = 0.0f;
    bool known = true, woken = false, urgent_if_binding = false, urgent = false;
    for (uint32_t e = 0u; e < event_count && any_live; ++e) {
        const uint4 h0 = events[e * 3u], h1 = events[e * 3u + 1u], h2 = events[e * 3u + 2u];
        if ((h0.w & 2u) != 0u) {
            uint32_t limit = h2.x;
            for (uint32_t k = 0u; k < h2.w; ++k) {
                uint4 sp = events[event_count * 3u + (h2.z + k) * 2u];
                f3 d = mk((float)px, (float)py, (float)pz) - mk(__uint_as_float(sp.x), __uint_as_float(sp.y), __uint_as_float(sp.z));
                if (dot3(d, d) <= __uint_as_float(sp.w)) limit = min(limit, 4u);
            }
            source_cap = min(source_cap, limit);
            woken = true;
            continue;
        }
        // No slot of the workgroup within reach of the event's union: no box
        // of it can change any of them (uniform across the workgroup).
        if (box_box_distance2((int)h0.x, (int)h0.y, (int)h0.z, (int)h1.x, (int)h1.y, (int)h1.z, g[0], g[1], g[2], g[3], g[4], g[5]) > reach)
            continue;
        const bool in_reach = live && box_distance2((int)h0.x, (int)h0.y, (int)h0.z, (int)h1.x, (int)h1.y, (int)h1.z, px, py, pz) <= reach;
        if (in_reach) woken = true;
        bool exact = known;
        uint32_t nearest = NONE;
        for (uint32_t base = 0u; base < h2.w; base += 256u) {
            uint32_t k = base + threadIdx.x;
            uint4 b0 = make_uint4(0u, 0u, 0u, 0u), b1 = b0;
            bool keep = false;
            if (k < h2.w) {
                uint32_t record = event_count * 3u + (h2.z + k) * 2u;
                b0 = events[record];
                b1 = events[record + 1u];
                keep = box_box_distance2((int)b0.x, (int)b0.y, (int)b0.z, (int)b1.x, (int)b1.y, (int)b1.z, g[0], g[1], g[2], g[3], g[4],
                                         g[5]) <= reach;
            }
            bool flag[1] = {keep};
            uint32_t rank[1], total[1];
            block_ranks<1>(flag, rank, total, s_wave);
            if (keep) {
                s_box[rank[0]][0] = b0;
                s_box[rank[0]][1] = b1;
            }
            __syncthreads();
            if (in_reach)
                for (uint32_t j = 0u; j < total[0]; ++j) {
                    const uint4 c0 = s_box[j][0], c1 = s_box[j][1];
                    uint32_t d = box_distance2((int)c0.x, (int)c0.y, (int)c0.z, (int)c1.x, (int)c1.y, (int)c1.z, px, py, pz);
                    if (d > reach) continue;
                    nearest = min(nearest, d);
                    uint32_t edit_kind = c1.w & 7u;
                    if ((c1.w & 8u) != 0u && (edit_kind == 0u || edit_kind == 3u)) {
                        if (edit_kind == 3u && measured_mode) urgent_if_binding = urgent_if_binding || d <= urgent_measured;
                        else urgent = urgent || d <= urgent_curve;
                    }
                    if (exact) {
                        uint32_t kind = c1.w & 7u;
                        if (kind == 0u) { known = false; exact = false; }
                        else if (kind != 1u) {
                            float area = 1.0f;
                            if (kind == 2u) area = (float)min(max(((int)c1.w >> 3) + 1 - py, 0), 16);
                            change = change + __uint_as_float(c0.w) * area / (3.14159265358979f * (float)max(d, 1u));
                        }
                    }
                }
            __syncthreads();
        }
        if (nearest != NONE) cap = min(cap, curve_limit(nearest));
    }
    if (!live) return;
    bool measured = measured_mode && known;
    uint32_t limit = min(source_cap, measured ? NONE : cap);
    uint32_t *st = state + (size_t)slot * 36u;
    if (limit != NONE) {
        st[5] = min(st[5], limit);
        uint32_t *c = checkpoint + (size_t)slot * 12u + 9u;
        *c = min(*c, limit);
    }
    uint32_t *pw = meta + (size_t)slot * 12u + 4u;
    if (measured && change > 0.0f) *pw = __float_as_uint(__uint_as_float(*pw) + change);
    if (urgent_if_binding) {
        float pending = __uint_as_float(*pw);
        urgent = urgent || !measured || (pending > 0.0f && binds(st, pending, physical));
    }
    if (woken || urgent) meta[(size_t)slot * 12u + 5u] &= 0xfffffff0u;
    if (urgent) meta[(size_t)slot * 12u + 11u] |= 0x40000000u;
}

// Learned mode, per face: each sample's value is its partial energy plus
// factor * max(prediction, 0); moments are summed per lane subsequence and
// then in lane order, as trace_faces_kernel does (voxel_trace.wgsl:72-92).
__global__ void learned_combine_kernel(const float4 *__restrict__ learned, const float *__restrict__ pred, uint32_t count,
                                       uint32_t samples, uint32_t lanes, const uint4 *__restrict__ tasks,
                                       float4 *__restrict__ results) {
    uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= count || tasks[f * 2u].w == NONE) return;
    f3 tmean = mk(0.0f, 0.0f, 0.0f), tsecond = mk(0.0f, 0.0f, 0.0f);
    uint32_t tknown = 0u;
    for (uint32_t lane = 0u; lane < lanes; ++lane) {
        f3 mean = mk(0.0f, 0.0f, 0.0f), second = mk(0.0f, 0.0f, 0.0f);
        uint32_t known = 0u;
        for (uint32_t i = lane; i < samples; i += lanes) {
            size_t at = (size_t)f * samples + i;
            float4 a = learned[at * 2u], b = learned[at * 2u + 1u];
            if (a.w == 0.0f) continue;
            const float *p = pred + at * 3u;
            f3 value = mk(a.x, a.y, a.z) + mk(b.x, b.y, b.z) * mk(fmx(p[0], 0.0f), fmx(p[1], 0.0f), fmx(p[2], 0.0f));
            mean = mean + value;
            second = second + value * value;
            known += 1u;
        }
        tmean = tmean + mean;
        tsecond = tsecond + second;
        tknown += known;
    }
    float inverse = 1.0f / (float)(tknown > 1u ? tknown : 1u);
    float4 *r = results + (size_t)f * 6u;
    r[0] = make_float4(tmean.x * inverse, tmean.y * inverse, tmean.z * inverse, 0.0f);
    r[1] = make_float4(0.0f, (float)tknown, 0.0f, 0.0f);
    r[2] = make_float4(tsecond.x * inverse, tsecond.y * inverse, tsecond.z * inverse, 0.0f);
}

// Training set of a frame: a window of a fixed permutation of the bound
// faces, as tasks (traced plain, one sample) and as cache queries.
__global__ void train_gather_kernel(const uint4 *__restrict__ bound, const uint32_t *__restrict__ perm,
                                    const uint32_t *__restrict__ face_material, uint32_t n_bound, uint32_t start, uint32_t count,
                                    uint4 *__restrict__ tasks, MlpQuery *__restrict__ queries) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t f = perm[(uint32_t)(((uint64_t)start + i) % n_bound)];
    uint4 a = bound[f * 2u], b = bound[f * 2u + 1u];
    tasks[i * 2u] = a;
    tasks[i * 2u + 1u] = b;
    queries[i] = MlpQuery{(int)a.x, (int)a.y, (int)a.z, a.w == NONE ? NONE : (a.w | (face_material[f] << 8))};
}

// Targets: the traced sample's incident radiance, weight 1 when known.
__global__ void train_targets_kernel(const float4 *__restrict__ results, uint32_t count, float4 *__restrict__ targets) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 m = results[(size_t)i * 6u];
    float known = results[(size_t)i * 6u + 1u].y;
    targets[i] = make_float4(m.x, m.y, m.z, known > 0.0f ? 1.0f : 0.0f);
}

// Long-horizon FP64 sums of per-launch face moments: the converged reference
// of the error harness. Per face: sum of mean*known (rgb), sum of
// second*known (rgb), sum of known, and the analytic direct term's luminance
// slot left for the host.
__global__ void fp64_accumulate_kernel(const float4 *__restrict__ results, uint32_t count, double *__restrict__ sums) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float4 *r = results + (size_t)i * 6u;
    float4 m = r[0], k4 = r[1], q = r[2];
    double k = (double)k4.y;
    double *s = sums + (size_t)i * 8u;
    s[0] += (double)m.x * k;
    s[1] += (double)m.y * k;
    s[2] += (double)m.z * k;
    s[3] += (double)q.x * k;
    s[4] += (double)q.y * k;
    s[5] += (double)q.z * k;
    s[6] += k;
}

// The face index (voxel_accumulate.wgsl:357-373 face_slot): a face's slot
// or NONE.
__device__ __forceinline__ uint32_t face_slot(const uint4 *__restrict__ index, uint32_t mask, int vx, int vy, int vz, uint32_t face) {
    uint32_t probe = (hash_voxel(vx, vy, vz) ^ pcg(face + 1u)) & mask;
    for (uint32_t attempt = 0u; attempt < 64u; ++attempt) {
        uint4 e = index[probe];
        if (e.w == NONE) return NONE;
        if ((int)e.x == vx && (int)e.y == vy && (int)e.z == vz && (e.w >> 29u) == face) return e.w & 0x1fffffffu;
        probe = (probe + 1u) & mask;
    }
    return NONE;
}

// measure_edits (voxel_accumulate.wgsl:375-407): one scan position per
// thread; the brightest outgoing luminance (albedo luminance times cached
// incident plus emission luminance) of any resident face of that voxel,
// max-reduced per edit. Entries hold ascending, contiguous position ranges,
// so a binary search finds the same entry as the reference's linear scan.
__global__ void measure_edits_kernel(const uint4 *__restrict__ measures, uint32_t entries, uint32_t positions,
                                     const uint4 *__restrict__ index, uint32_t mask, const uint32_t *__restrict__ state,
                                     const uint32_t *__restrict__ meta, const float2 *__restrict__ material_light,
                                     uint32_t *__restrict__ measure_max) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= positions) return;
    uint32_t lo = 0u, hi = entries;  // the last entry whose first position <= t
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (measures[mid * 3u].y <= t) lo = mid;
        else hi = mid;
    }
    uint32_t e = lo;
    uint4 a = measures[e * 3u], b = measures[e * 3u + 1u], c = measures[e * 3u + 2u];
    uint32_t local = t - a.y;
    int px = (int)b.x + (int)(local % c.x), py = (int)b.y + (int)((local / c.x) % c.y), pz = (int)b.z + (int)(local / (c.x * c.y));
    const f3 w = mk(0.2126f, 0.7152f, 0.0722f);
    float brightest = 0.0f;
    bool observed = false;
    for (uint32_t face = 0u; face < 6u; ++face) {
        uint32_t slot = face_slot(index, mask, px, py, pz, face);
        if (slot == NONE) continue;
        uint32_t identity = meta[(size_t)slot * 12u + 3u];
        if ((identity & 8u) == 0u) continue;
        const uint32_t *st = state + (size_t)slot * 36u;
        if ((st[5] & 0x7fffffffu) == 0u) continue;
        observed = true;
        float incident = dot3(w, load3(st)) + dot3(w, load3(st + 16));
        float2 light = material_light[identity >> 16];
        brightest = fmx(brightest, light.x * incident + light.y);
    }
    if (observed) atomicMax(measure_max + e, __float_as_uint(brightest) | 0x80000000u);
}

// finish_edits (voxel_accumulate.wgsl:409-430): each measured edit's flux,
// base plus the brightest outgoing luminance times two faces per changed
// voxel; an edit no resident face observed falls back to unknown flux.
__global__ void finish_edits_kernel(const uint4 *__restrict__ measures, uint32_t entries, const uint32_t *__restrict__ measure_max,
                                    uint4 *__restrict__ events) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= entries) return;
    uint4 a = measures[e * 3u];
    float base_flux = __uint_as_float(measures[e * 3u + 1u].w);
    uint32_t observation = measure_max[e];
    if ((observation & 0x80000000u) == 0u) {
        events[a.x + 1u].w &= 0xfffffff8u;
        return;
    }
    float brightest = __uint_as_float(observation & 0x7fffffffu);
    events[a.x].w = __float_as_uint(base_flux + brightest * 2.0f * (float)a.w);
    events[a.x + 1u].w = (events[a.x + 1u].w & 0xfffffff8u) | 3u;
}

// Bound-face updates after edits: entries (task, slot) written at their
// indices; indices are distinct.
__global__ void scatter_bound_kernel(const uint32_t *__restrict__ index, const uint4 *__restrict__ tasks,
                                     const uint32_t *__restrict__ slots, uint32_t n, uint4 *__restrict__ bound_tasks,
                                     uint32_t *__restrict__ bound_slots) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t at = index[i];
    bound_tasks[at * 2u] = tasks[i * 2u];
    bound_tasks[at * 2u + 1u] = tasks[i * 2u + 1u];
    if (bound_slots) bound_slots[at] = slots[i];
}

// Single-word writes into the slot metadata (identity words after a face's
// material changed): pairs (word index, value), indices distinct.
__global__ void scatter_words_kernel(const uint2 *__restrict__ writes, uint32_t n, uint32_t *__restrict__ words) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint2 w = writes[i];
    words[w.x] = w.y;
}

// A plain streaming copy (16 B per lane), the known-size reference against
// which hardware byte counters are scaled.
__global__ void copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// ============================================================== change footprints
// voxel_trace.wgsl:94-202: sunlight a changed box admits or blocks, traced in
// the new scene. One task per changed box, one thread per tube of a square
// grid across the box's diagonal, per light: a tube whose sun side is clear
// beyond the box is traced back through it; where it lands, the landing
// voxel widens the task's footprint and the light reflected there adds
// flux. Tasks (int4): a count, then per task (min, header record), (max,
// bound record), (first thread, grid, 0, 0). Sums (8 words per task): the
// landing voxel's biased min xyz and max xyz, the flux x 64, spare.
constexpr int FOOTPRINT_BIAS = 1073741824;
constexpr float FOOTPRINT_FLUX_SCALE = 64.0f;

// The last task whose first thread is at most t (first threads ascend).
__device__ __forceinline__ uint32_t footprint_task(const int4 *__restrict__ tasks, uint32_t t) {
    uint32_t lo = 0u, hi = (uint32_t)tasks[0].x;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) / 2u;
        if ((uint32_t)tasks[1u + mid * 3u + 2u].x <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void footprint_trace_kernel(SceneView scene, Optics optics, const int4 *__restrict__ tasks,
                                                              uint32_t threads, uint32_t *__restrict__ sums) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= threads || tasks[0].x == 0) return;
    uint32_t e = footprint_task(tasks, id);
    int4 a = tasks[1u + e * 3u], b = tasks[1u + e * 3u + 1u], c = tasks[1u + e * 3u + 2u];
    uint32_t grid = (uint32_t)c.y, local = id - (uint32_t)c.x;
    if (local >= grid * grid) return;
    f3 lo = mk((float)a.x, (float)a.y, (float)a.z);
    f3 hi = mk((float)b.x, (float)b.y, (float)b.z) + mk(1.0f, 1.0f, 1.0f);
    f3 center = (lo + hi) * 0.5f;
    // A square grid across the box's diagonal covers its section from any
    // direction; tubes that miss the box are skipped below.
    float diagonal = len3(hi - lo);
    float spacing = diagonal / (float)grid;
    float ox = ((float)(local % grid) + 0.5f) * spacing - 0.5f * diagonal;
    float oy = ((float)(local / grid) + 0.5f) * spacing - 0.5f * diagonal;
    const f3 w = mk(0.2126f, 0.7152f, 0.0722f);
    const float reach = optics.max_distance;
    const uint32_t base = e * 8u;
    NoTally tally;
    float flux = 0.0f;
    for (uint32_t i = 0u; i < optics.directional; ++i) {
        float4 pr = optics.lights[i * 3u], en = optics.lights[i * 3u + 1u], ex = optics.lights[i * 3u + 2u];
        f3 color = mk(fmx(en.x - ex.y, 0.0f), fmx(en.y - ex.z, 0.0f), fmx(en.z - ex.w, 0.0f));
        if (color.x == 0.0f && color.y == 0.0f && color.z == 0.0f) continue;
        f3 direction = normalize3(mk(pr.x, pr.y, pr.z));
        f3 side = fabsf(direction.y) > 0.9f ? mk(1.0f, 0.0f, 0.0f) : mk(0.0f, 1.0f, 0.0f);
        f3 u = normalize3(cross3(direction, side));
        f3 v = cross3(direction, u);
        f3 point = center + u * ox + v * oy;
        // Exact axis directions would make 0 * inf; a tiny component is exact enough.
        f3 inverse = mk(1.0f / (fabsf(direction.x) < 1e-8f ? 1e-8f : direction.x), 1.0f / (fabsf(direction.y) < 1e-8f ? 1e-8f : direction.y),
                        1.0f / (fabsf(direction.z) < 1e-8f ? 1e-8f : direction.z));
        f3 nr = (lo - point) * inverse, fr = (hi - point) * inverse;
        float enter = fmx(fmx(fmn(nr.x, fr.x), fmn(nr.y, fr.y)), fmn(nr.z, fr.z));
        float exit = fmn(fmn(fmx(nr.x, fr.x), fmx(nr.y, fr.y)), fmx(nr.z, fr.z));
        if (enter > exit) continue;
        // Shadowed beyond the box either way: the change cannot matter.
        if (trace_ray<NoTally, true>(scene, point + direction * (exit + 0.01f), direction, reach, tally).status != 0u) continue;
        f3 start = point + direction * (enter - 0.01f);
        RayResult hit = trace_ray<NoTally, true>(scene, start, -direction, reach, tally);
        // A tube starting inside solid (the box rests on it) lands nowhere.
        if (hit.status != 1u || dot3(hit.normal, direction) <= 0.0f || len3(hit.position - start) < 0.02f) continue;
        // Reflected power through the tube: irradiance across its section
        // (perpendicular to the light, so no cosine) times the albedo.
        float4 md = optics.materials[hit.material * 2u];
        flux = flux + dot3(color, w) * dot3(mk(md.x, md.y, md.z), w) * spacing * spacing;
        uint32_t px = (uint32_t)(hit.vx + FOOTPRINT_BIAS), py = (uint32_t)(hit.vy + FOOTPRINT_BIAS), pz = (uint32_t)(hit.vz + FOOTPRINT_BIAS);
        atomicMin(&sums[base], px);
        atomicMin(&sums[base + 1u], py);
        atomicMin(&sums[base + 2u], pz);
        atomicMax(&sums[base + 3u], px);
        atomicMax(&sums[base + 4u], py);
        atomicMax(&sums[base + 5u], pz);
    }
    if (flux > 0.0f) atomicAdd(&sums[base + 6u], (uint32_t)(flux * FOOTPRINT_FLUX_SCALE + 0.5f));
}

// Each task's landing patch becomes its reserved Light event: the header's
// union and its one bound. A task that landed nowhere keeps its empty bound,
// out of every receiver's reach.
__global__ void footprint_finish_kernel(const int4 *__restrict__ tasks, const uint32_t *__restrict__ sums, uint4 *__restrict__ events) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint32_t)tasks[0].x) return;
    uint32_t base = e * 8u;
    uint32_t x = sums[base];
    if (x == NONE) return;
    int mnx = (int)x - FOOTPRINT_BIAS, mny = (int)sums[base + 1u] - FOOTPRINT_BIAS, mnz = (int)sums[base + 2u] - FOOTPRINT_BIAS;
    int mxx = (int)sums[base + 3u] - FOOTPRINT_BIAS, mxy = (int)sums[base + 4u] - FOOTPRINT_BIAS, mxz = (int)sums[base + 5u] - FOOTPRINT_BIAS;
    float flux = (float)sums[base + 6u] / FOOTPRINT_FLUX_SCALE;
    uint32_t header = (uint32_t)tasks[1u + e * 3u].w, bound = (uint32_t)tasks[1u + e * 3u + 1u].w;
    events[header] = make_uint4((uint32_t)mnx, (uint32_t)mny, (uint32_t)mnz, events[header].w);
    events[header + 1u] = make_uint4((uint32_t)mxx, (uint32_t)mxy, (uint32_t)mxz, events[header + 1u].w);
    events[bound] = make_uint4((uint32_t)mnx, (uint32_t)mny, (uint32_t)mnz, __float_as_uint(flux));
    events[bound + 1u] = make_uint4((uint32_t)mxx, (uint32_t)mxy, (uint32_t)mxz, 4u);
}

// ============================================================== launchers
static inline uint32_t blocks(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

// Kernel variant by what the caller asked for: hit records (with full
// counters), all counters, rays only, or nothing.
void launch_trace(const SceneView &scene, const Optics &optics, const uint4 *tasks, uint32_t count, uint32_t samples,
                  uint32_t lanes, uint32_t sequence_base, float4 *results, uint32_t *hits, unsigned long long *counters,
                  bool full_counters, hipStream_t stream, const uint32_t *order) {
    uint64_t threads = (uint64_t)count * lanes;
    if (threads == 0) return;
    // One wavefront per workgroup: a workgroup's slots free as soon as its
    // wavefront's longest path ends, not its slowest of four (512^3 cave,
    // plain frame: 256 threads 53.2 ms, 128: 49.0, 64: 47.4).
    dim3 grid(blocks(threads, 64)), block(64);
    float4 *no_l = nullptr;
    MlpQuery *no_q = nullptr;
    if (hits)
        hipLaunchKernelGGL((trace_faces_kernel<FullTally, true, false>), grid, block, 0, stream, scene, optics, tasks, count, samples,
                           lanes, sequence_base, results, hits, counters, no_l, no_q, order);
    else if (counters && full_counters)
        hipLaunchKernelGGL((trace_faces_kernel<FullTally, false, false>), grid, block, 0, stream, scene, optics, tasks, count, samples,
                           lanes, sequence_base, results, hits, counters, no_l, no_q, order);
    else
        // Frames that count nothing run the ray-counting variant too (with no
        // counters it adds nothing up): its build measured 47.3 ms a plain
        // frame against 51.2 ms for the variant without a tally.
        hipLaunchKernelGGL((trace_faces_kernel<RayTally, false, false>), grid, block, 0, stream, scene, optics, tasks, count, samples,
                           lanes, sequence_base, results, hits, counters, no_l, no_q, order);
}

void launch_trace_learned(const SceneView &scene, const Optics &optics, const uint4 *tasks, uint32_t count, uint32_t samples,
                          uint32_t lanes, uint32_t sequence_base, float4 *results, unsigned long long *counters, float4 *learned,
                          MlpQuery *queries, hipStream_t stream, const uint32_t *order) {
    uint64_t threads = (uint64_t)count * lanes;
    if (threads == 0) return;
    // As launch_trace: one wavefront per workgroup, the ray-counting variant
    // whether or not the frame counts.
    dim3 grid(blocks(threads, 64)), block(64);
    hipLaunchKernelGGL((trace_faces_kernel<RayTally, false, true>), grid, block, 0, stream, scene, optics, tasks, count, samples,
                       lanes, sequence_base, results, nullptr, counters, learned, queries, order);
}

void launch_learned_combine(const float4 *learned, const float *pred, uint32_t count, uint32_t samples, uint32_t lanes,
                            const uint4 *tasks, float4 *results, hipStream_t stream) {
    if (count == 0) return;
    hipLaunchKernelGGL(learned_combine_kernel, dim3(blocks(count, 256)), dim3(256), 0, stream, learned, pred, count, samples, lanes,
                       tasks, results);
}
void launch_train_gather(const uint4 *bound, const uint32_t *perm, const uint32_t *face_material, uint32_t n_bound, uint32_t start,
                         uint32_t count, uint4 *tasks, MlpQuery *queries, hipStream_t stream) {
    if (count == 0) return;
    hipLaunchKernelGGL(train_gather_kernel, dim3(blocks(count, 256)), dim3(256), 0, stream, bound, perm, face_material, n_bound, start,
                       count, tasks, queries);
}
void launch_train_targets(const float4 *results, uint32_t count, float4 *targets, hipStream_t stream) {
    if (count == 0) return;
    hipLaunchKernelGGL(train_targets_kernel, dim3(blocks(count, 256)), dim3(256), 0, stream, results, count, targets);
}
void launch_classify(const uint16_t *grid, int gx, int gy, int gz, int vx, int vy, int vz, const int4 *list, uint32_t n,
                     RegionInfo *info, hipStream_t stream) {
    if (n == 0) return;
    hipLaunchKernelGGL(classify_regions_kernel, dim3(n), dim3(256), 0, stream, grid, gx, gy, gz, vx, vy, vz, list, n, info);
}
void launch_scan(RegionInfo *info, uint32_t n, uint32_t cells_base, uint32_t *totals, hipStream_t stream) {
    hipLaunchKernelGGL(scan_regions_kernel, dim3(1), dim3(1024), 0, stream, info, n, cells_base, totals);
}
void launch_pack(const uint16_t *grid, int gx, int gy, int gz, int vx, int vy, int vz, const int4 *list, uint32_t n,
                 const RegionInfo *info, uint32_t *cells, hipStream_t stream) {
    if (n == 0) return;
    hipLaunchKernelGGL(pack_regions_kernel, dim3(n), dim3(256), 0, stream, grid, gx, gy, gz, vx, vy, vz, list, n, info, cells);
}
void launch_headers(const int4 *list, uint32_t n, const RegionInfo *info, const uint32_t *top, uint4 *headers, int rx, int ry,
                    int rz, int sx, int sy, hipStream_t stream) {
    if (n == 0) return;
    hipLaunchKernelGGL(write_headers_kernel, dim3(blocks(n, 256)), dim3(256), 0, stream, list, n, info, top, headers, rx, ry, rz, sx,
                       sy);
}
void launch_top(const uint32_t *top, uint32_t cells, uint4 *headers, hipStream_t stream) {
    if (cells == 0) return;
    hipLaunchKernelGGL(write_top_kernel, dim3(blocks(cells, 256)), dim3(256), 0, stream, top, cells, headers);
}
void launch_scatter_voxels(uint16_t *grid, const uint32_t *index, const uint16_t *values, uint32_t n, hipStream_t stream) {
    if (n == 0) return;
    hipLaunchKernelGGL(scatter_voxels_kernel, dim3(blocks(n, 256)), dim3(256), 0, stream, grid, index, values, n);
}
void launch_prepare(uint4 *tasks, const uint32_t *slots, uint32_t n, uint32_t *state, uint32_t *meta, float physical,
                    uint32_t batch, uint32_t *keys, hipStream_t stream) {
    if (n == 0) return;
    hipLaunchKernelGGL(prepare_kernel, dim3(blocks(n, 256)), dim3(256), 0, stream, tasks, slots, n, state, meta, physical, batch,
                       keys);
}
void launch_accumulate(const uint32_t *slots, uint32_t n, float4 *results, uint32_t *state, uint32_t *checkpoint, uint32_t *meta,
                       const AccumulateParams &p, hipStream_t stream) {
    if (n == 0) return;
    hipLaunchKernelGGL(accumulate_kernel, dim3(blocks(n, 256)), dim3(256), 0, stream, slots, n, results, state, checkpoint, meta, p);
}
void launch_apply_events(const uint4 *events, uint32_t event_count, uint32_t high_water, uint32_t *state, uint32_t *checkpoint,
                         uint32_t *meta, float physical, hipStream_t stream) {
    if (high_water == 0) return;
    hipLaunchKernelGGL(apply_events_kernel, dim3(blocks(high_water, 256)), dim3(256), 0, stream, events, event_count, high_water,
                       state, checkpoint, meta, physical, 192u * 192u, 64u * 64u, 48u * 48u);
}

void launch_measure_edits(const uint4 *measures, uint32_t entries, uint32_t positions, const uint4 *index, uint32_t mask,
                          const uint32_t *state, const uint32_t *meta, const float2 *material_light, uint32_t *measure_max,
                          uint4 *events, hipStream_t stream) {
    if (entries == 0) return;
    if (positions)
        hipLaunchKernelGGL(measure_edits_kernel, dim3(blocks(positions, 256)), dim3(256), 0, stream, measures, entries, positions, index,
                           mask, state, meta, material_light, measure_max);
    hipLaunchKernelGGL(finish_edits_kernel, dim3(blocks(entries, 256)), dim3(256), 0, stream, measures, entries, measure_max, events);
}

void launch_scatter_bound(const uint32_t *index, const uint4 *tasks, const uint32_t *slots, uint32_t n, uint4 *bound_tasks,
                          uint32_t *bound_slots, hipStream_t stream) {
    if (n == 0) return;
    hipLaunchKernelGGL(scatter_bound_kernel, dim3(blocks(n, 256)), dim3(256), 0, stream, index, tasks, slots, n, bound_tasks, bound_slots);
}
void launch_scatter_words(const uint2 *writes, uint32_t n, uint32_t *words, hipStream_t stream) {
    if (n == 0) return;
    hipLaunchKernelGGL(scatter_words_kernel, dim3(blocks(n, 256)), dim3(256), 0, stream, writes, n, words);
}

void launch_batch_frame(const BatchScene *scenes, uint32_t n_scenes, const BatchTile *tiles, uint32_t n_tiles, uint32_t lanes,
                        uint32_t block, uint32_t grid, BatchControl *control, uint32_t *done, unsigned long long *finish,
                        hipStream_t stream) {
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(256), 0, stream, control, done, finish, n_scenes);
    hipLaunchKernelGGL(batch_trace_kernel, dim3(grid), dim3(block), 0, stream, scenes, tiles, n_tiles, lanes, control, done, finish);
}

void launch_fp64_accumulate(const float4 *results, uint32_t count, double *sums, hipStream_t stream) {
    if (count == 0) return;
    hipLaunchKernelGGL(fp64_accumulate_kernel, dim3(blocks(count, 256)), dim3(256), 0, stream, results, count, sums);
}

void launch_footprints(const SceneView &scene, const Optics &optics, const int4 *tasks, uint32_t task_count, uint32_t threads,
                       uint32_t *sums, uint4 *events, hipStream_t stream) {
    if (task_count == 0u) return;
    if (threads) hipLaunchKernelGGL(footprint_trace_kernel, dim3(blocks(threads, 256)), dim3(256), 0, stream, scene, optics, tasks, threads, sums);
    hipLaunchKernelGGL(footprint_finish_kernel, dim3(blocks(task_count, 64)), dim3(64), 0, stream, tasks, sums, events);
}

// Fills `words` u32 with `value` (a kernel, so it may sit in captured graphs).
__global__ void fill_words_kernel(uint32_t *__restrict__ words, uint64_t n, uint32_t value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) words[i] = value;
}

void launch_fill_words(uint32_t *words, uint64_t n, uint32_t value, hipStream_t stream) {
    if (n == 0) return;
    uint64_t want = (n + 255) / 256;
    hipLaunchKernelGGL(fill_words_kernel, dim3((uint32_t)(want < 4096 ? want : 4096)), dim3(256), 0, stream, words, n, value);
}

void launch_copy(const void *src, void *dst, size_t bytes, hipStream_t stream) {
    size_t n = bytes / 16;
    if (n == 0) return;
    hipLaunchKernelGGL(copy_kernel, dim3(2048), dim3(256), 0, stream, (const uint4 *)src, (uint4 *)dst, n);
}

}  // namespace vgi
