// The view tier's passes (vgi_view.h). One thread per prepass pixel for
// discovery, indexing and the exchange; one thread per owner (a compacted
// queue) for tracing. Every pass is a function of the pixel's or owner's own
// data and the frame's read-only inputs, so the results do not depend on the
// order in which threads run; the only atomics pick the owner of a face (the
// smallest pixel index of its duplicates) and append owners to the queue.
#include "vgi_view.h"

namespace vgi {

// clip -> world: h = M (x, y, z, 1), columns summed in order, then divided.
__device__ __forceinline__ f3 view_unproject(const ViewParams &p, float px, float py, float z) {
    float ux = (px + 0.5f) / (float)p.width, uy = (py + 0.5f) / (float)p.height;
    float x = ux * 2.0f - 1.0f, y = 1.0f - uy * 2.0f;
    float h[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) h[i] = ((p.m[i] * x + p.m[4 + i] * y) + p.m[8 + i] * z) + p.m[12 + i];
    return mk(h[0] / h[3], h[1] / h[3], h[2] / h[3]);
}

__device__ __forceinline__ float view_distance(f3 a, f3 b) {
    f3 d = a - b;
    return fsqrt((d.x * d.x + d.y * d.y) + d.z * d.z);
}

// voxel_view_discovery.wgsl:4-11: an axis normal, or none.
__device__ __forceinline__ uint32_t prepass_face(f3 n) {
    if (fmx(fmx(fabsf(n.x), fabsf(n.y)), fabsf(n.z)) < 0.99f) return NONE;
    if (fabsf(n.x) > 0.99f) return n.x < 0.0f ? 1u : 0u;
    if (fabsf(n.y) > 0.99f) return n.y < 0.0f ? 3u : 2u;
    return n.z < 0.0f ? 5u : 4u;
}

__device__ __forceinline__ int axis_step(uint32_t face) { return (face & 1u) == 0u ? 1 : -1; }

// voxel_view_discovery.wgsl:214-226 (unit voxels): a solid voxel whose
// neighbour across the face is air (or not loaded).
__device__ __forceinline__ bool receiver_exists(const SceneView &s, int vx, int vy, int vz, uint32_t face) {
    if (face >= 6u) return false;
    if (voxel_material(s, vx, vy, vz) == AIR) return false;
    int o[3] = {vx, vy, vz};
    o[face / 2u] += axis_step(face);
    return voxel_material(s, o[0], o[1], o[2]) == AIR;
}

__device__ __forceinline__ f3 face_centre(int vx, int vy, int vz, f3 n) {
    return mk((float)vx, (float)vy, (float)vz) + mk(0.5f, 0.5f, 0.5f) + n * 0.5f;
}

__device__ __forceinline__ uint32_t receiver_hash(const ViewReceiver &r) { return hash_voxel(r.vx, r.vy, r.vz) ^ pcg(r.face + 1u); }

__device__ __forceinline__ bool same_receiver(const ViewReceiver &a, const ViewReceiver &b) {
    return a.face == b.face && a.vx == b.vx && a.vy == b.vy && a.vz == b.vz && a.sx == b.sx && a.sy == b.sy && a.sz == b.sz;
}

// prepass_voxel_receiver (voxel_view_discovery.wgsl:40-104) for unit voxels:
// the decoded point, its normal axis snapped to the face plane, the voxel
// behind it; a point within one pixel's world footprint of an in-plane voxel
// boundary takes the exposed neighbour instead. A pixel whose surface is not
// an exposed face of the loaded scene owns no receiver.
__global__ void view_discover_kernel(ViewParams p, SceneView s, const float *__restrict__ depth, const float *__restrict__ normals,
                                     ViewReceiver *__restrict__ receivers, ViewEstimate *__restrict__ estimates) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.pixels) return;
    ViewReceiver r = {0, 0, 0, NONE, 0.0f, 0.0f, 0.0f, 0u};
    float px = (float)(i % p.width), py = (float)(i / p.width);
    float z = depth[i];
    bool found = false;
    if (z > 0.0f) {
        f3 position = view_unproject(p, px, py, z);
        bool inner = p.cam[3] > 0.0f && view_distance(position, mk(p.cam[0], p.cam[1], p.cam[2])) < p.cam[3];
        uint32_t face = inner ? NONE : prepass_face(mk(normals[i * 3u], normals[i * 3u + 1u], normals[i * 3u + 2u]));
        if (face < 6u) {
            f3 n = face_normal(face);
            const uint32_t axis = face / 2u;
            float sv[3] = {position.x, position.y, position.z};
            sv[axis] = rintf(sv[axis]);
            f3 surface = mk(sv[0], sv[1], sv[2]);
            f3 inside = surface - n * 0.01f;
            int vx = (int)floorf(inside.x), vy = (int)floorf(inside.y), vz = (int)floorf(inside.z);
            if (receiver_exists(s, vx, vy, vz, face)) {
                r = ViewReceiver{vx, vy, vz, face, 0.0f, 0.0f, 0.0f, 0u};
                found = true;
            } else {
                float footprint = view_distance(view_unproject(p, px + 1.0f, py, z), position);
                const uint32_t u = (axis + 1u) % 3u, v = (axis + 2u) % 3u;
                float lu = sv[u] - floorf(sv[u]), lv = sv[v] - floorf(sv[v]);
                float gu = fmn(lu, 1.0f - lu), gv = fmn(lv, 1.0f - lv);
                int su = lu < 0.5f ? -1 : 1, sw = lv < 0.5f ? -1 : 1;
                float best_gap = 2.0f;
                for (uint32_t c = 1u; c < 4u; ++c) {
                    int off[3] = {0, 0, 0};
                    float needed = 0.0f;
                    if (c & 1u) { off[u] = su; needed = fmx(needed, gu); }
                    if (c & 2u) { off[v] = sw; needed = fmx(needed, gv); }
                    if (needed > footprint || needed >= best_gap) continue;
                    int nx = vx + off[0], ny = vy + off[1], nz = vz + off[2];
                    if (receiver_exists(s, nx, ny, nz, face)) {
                        r = ViewReceiver{nx, ny, nz, face, 0.0f, 0.0f, 0.0f, 0u};
                        best_gap = needed;
                        found = true;
                    }
                }
            }
            if (found) {
                f3 c = face_centre(r.vx, r.vy, r.vz, n);
                r.sx = c.x; r.sy = c.y; r.sz = c.z;
            }
        }
    }
    receivers[i] = r;
    ViewEstimate e;
    e.mean = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    e.moment = e.mean;
    e.direct = e.mean;
    e.identity = make_uint4(0u, found ? voxel_material(s, r.vx, r.vy, r.vz) : 0u, 0u, 0u);
    estimates[i] = e;
}

// index_receivers (voxel_view_runtime.wgsl:128-142): publish each face once.
// Every duplicate of a key reaches the same hash entry, claims it or lowers
// it to its own index: the owner is the smallest pixel index of the face,
// whatever order the threads run in.
__global__ void view_index_kernel(ViewParams p, const ViewReceiver *__restrict__ receivers, uint32_t *__restrict__ hash,
                                  uint32_t *__restrict__ queue) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.pixels) return;
    const ViewReceiver r = receivers[i];
    if (r.face >= 6u) return;
    uint32_t slot = receiver_hash(r) & p.mask;
    for (uint32_t step = 0u; step < 64u; ++step) {
        uint32_t old = atomicCAS(&hash[slot], 0u, i + 1u);
        if (old == 0u) return;
        if (same_receiver(receivers[old - 1u], r)) {
            atomicMin(&hash[slot], i + 1u);
            return;
        }
        slot = (slot + 1u) & p.mask;
    }
    atomicAdd(&queue[1], 1u);  // overflow: the face owns no receiver this frame
}

__device__ __forceinline__ uint32_t owner_of(const ViewParams &p, const ViewReceiver &r, const ViewReceiver *__restrict__ receivers,
                                             const uint32_t *__restrict__ hash) {
    uint32_t slot = receiver_hash(r) & p.mask;
    for (uint32_t step = 0u; step < 64u; ++step) {
        uint32_t value = hash[slot];
        if (value == 0u) return NONE;
        if (same_receiver(receivers[value - 1u], r)) return value - 1u;
        slot = (slot + 1u) & p.mask;
    }
    return NONE;
}

// voxel_view_runtime.wgsl:105-116
__device__ __forceinline__ uint32_t previous_owner(const ViewParams &p, const ViewReceiver &r, const ViewReceiver *__restrict__ receivers,
                                                   const uint32_t *__restrict__ hash) {
    if (p.prev_pixels == 0u) return NONE;
    uint32_t slot = receiver_hash(r) & p.prev_mask;
    for (uint32_t step = 0u; step < 64u; ++step) {
        uint32_t value = hash[slot];
        if (value == 0u) return NONE;
        if (value <= p.prev_pixels && same_receiver(receivers[value - 1u], r)) return value - 1u;
        slot = (slot + 1u) & p.prev_mask;
    }
    return NONE;
}

__global__ void view_owner_kernel(ViewParams p, const ViewReceiver *__restrict__ receivers, const uint32_t *__restrict__ hash,
                                  uint32_t *__restrict__ owners) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.pixels) return;
    const ViewReceiver r = receivers[i];
    owners[i] = r.face < 6u ? owner_of(p, r, receivers, hash) : NONE;
}

// reproject_paths (voxel_view_runtime.wgsl:148-217): an owner takes its
// face's estimate of the previous frame when the material is unchanged, else
// starts empty; owners join the trace queue.
__global__ void view_reproject_kernel(ViewParams p, ViewReceiver *__restrict__ receivers, ViewEstimate *__restrict__ estimates,
                                      const uint32_t *__restrict__ owners, const ViewReceiver *__restrict__ prev_receivers,
                                      const ViewEstimate *__restrict__ prev_estimates, const uint32_t *__restrict__ prev_hash,
                                      uint32_t *__restrict__ queue) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.pixels || owners[i] != i) return;
    const ViewReceiver r = receivers[i];
    const uint32_t material = estimates[i].identity.y;
    uint32_t previous = previous_owner(p, r, prev_receivers, prev_hash);
    ViewEstimate e;
    e.mean = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    e.moment = e.mean;
    e.direct = e.mean;
    e.identity = make_uint4(0u, 0u, 0u, 0u);
    if (previous != NONE) {
        ViewEstimate old = prev_estimates[previous];
        if (old.identity.y == material) e = old;
    }
    e.identity.y = material;
    e.identity.z = 0u;
    receivers[i].sequence = e.identity.x;
    estimates[i] = e;
    uint32_t entry = atomicAdd(&queue[0], 1u);
    queue[2u + entry] = i;
}

// shared_slot (voxel_view_cache.wgsl:6-23): the face's resident slot when
// the world cache holds it live with this material.
__device__ __forceinline__ uint32_t shared_slot(const SharedCache &c, const ViewReceiver &r, uint32_t material) {
    uint32_t slot = receiver_hash(r) & c.mask;
    for (uint32_t attempt = 0u; attempt < 64u; ++attempt) {
        uint4 e = c.index[slot];
        if (e.w == NONE) return NONE;
        if ((int)e.x == r.vx && (int)e.y == r.vy && (int)e.z == r.vz && (e.w >> 29u) == r.face) {
            uint32_t s = e.w & 0x1fffffffu;
            uint32_t key = c.meta[(size_t)s * 12u + 3u];
            if ((key & 8u) == 0u || (key >> 16u) != material) return NONE;
            return s;
        }
        slot = (slot + 1u) & c.mask;
    }
    return NONE;
}

// load_shared_face (voxel_view_cache.wgsl:25-99): an admitted face (its
// incarnation recorded) always loads the resident authority; a face entering
// the view loads it when the authority holds at least as many samples, else
// keeps its own history (promotion). The owner then reserves this frame's
// path sequences on the authority.
__global__ void view_load_kernel(ViewParams p, SharedCache c, ViewReceiver *__restrict__ receivers, ViewEstimate *__restrict__ estimates,
                                 const uint32_t *__restrict__ owners) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.pixels || owners[i] != i) return;
    const ViewReceiver r = receivers[i];
    ViewEstimate e = estimates[i];
    uint32_t slot = shared_slot(c, r, e.identity.y);
    if (slot == NONE) {
        estimates[i].identity.w = 0u;
        return;
    }
    uint32_t *st = c.state + (size_t)slot * 36u;
    const uint32_t incarnation = st[28];
    const uint32_t count = st[5] & 0x7fffffffu;
    const uint32_t next = st[6];
    if (e.identity.w == incarnation || e.mean.w <= (float)count) {
        e.mean = make_float4(__uint_as_float(st[0]), __uint_as_float(st[1]), __uint_as_float(st[2]), (float)count);
        e.moment = make_float4(__uint_as_float(st[8]), __uint_as_float(st[9]), __uint_as_float(st[10]),
                               __uint_as_float(c.meta[(size_t)slot * 12u + 4u]));
        e.identity.x = next;
    } else {
        e.identity.x = max(e.identity.x, next);
    }
    e.identity.w = incarnation;
    const uint32_t first = max(next, e.identity.x);
    st[6] = first + p.samples;
    e.identity.x = first;
    estimates[i] = e;
    receivers[i].sequence = first;
}

// trace_visible_paths (voxel_view_runtime.wgsl:273-324) with every path
// completed in its visit: `samples` paths per owner through the world tier's
// estimator; each known sample first applies a pending predicted change (the
// measured-retention keep rule) and then merges into the running mean with
// weight 1 / min(count + 1, 65536). Unknown samples advance the sequence
// only. The face's analytic direct term is traced once.
// An owner's samples run on `lanes` consecutive threads (a power of two up
// to 8, so a group never straddles a wavefront): lane l traces samples l,
// l + lanes, ...; after each round the group's first lane merges the round's
// values in sample order, so the running mean is the sequential one.
__global__ __launch_bounds__(256) void view_trace_kernel(ViewParams p, SceneView s, Optics o, ViewReceiver *__restrict__ receivers,
                                                         ViewEstimate *__restrict__ estimates, const uint32_t *__restrict__ queue,
                                                         uint32_t lanes) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k = tid / lanes, lane = tid % lanes;
    const bool valid = k < queue[0];
    const uint32_t i = valid ? queue[2u + k] : 0u;
    ViewReceiver r = {0, 0, 0, 0u, 0.0f, 0.0f, 0.0f, 0u};
    ViewEstimate e;
    e.mean = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    e.moment = e.mean;
    e.direct = e.mean;
    e.identity = make_uint4(0u, 0u, 0u, 0u);
    if (valid) {
        r = receivers[i];
        e = estimates[i];
    }
    const f3 normal = face_normal(r.face);
    const f3 surface = mk(r.sx, r.sy, r.sz);
    const uint32_t stream = hash_voxel(r.vx, r.vy, r.vz) ^ pcg(r.face + 1u);
    const uint32_t start = e.identity.x;
    const int group = (int)((threadIdx.x & 63u) - lane);
    NoTally tally;
    for (uint32_t base = 0u; base < p.samples; base += lanes) {
        f3 value = mk(0.0f, 0.0f, 0.0f);
        bool known = false;
        if (valid && base + lane < p.samples) {
            FirstHit first;
            value = trace_receiver<NoTally>(s, o, surface, normal, stream, start + base + lane, known, first, tally);
        }
        for (uint32_t l = 0u; l < lanes; ++l) {
            const int src = group + (int)l;
            const f3 v = mk(__shfl(value.x, src, 64), __shfl(value.y, src, 64), __shfl(value.z, src, 64));
            const bool kn = __shfl((int)known, src, 64) != 0;
            if (lane != 0u || !valid || base + l >= p.samples || !kn) continue;
            if (e.moment.w > 0.0f) {
                float change = e.moment.w;
                f3 variance = mk(fmx(e.moment.x - e.mean.x * e.mean.x, 0.0f), fmx(e.moment.y - e.mean.y * e.mean.y, 0.0f),
                                 fmx(e.moment.z - e.mean.z * e.mean.z, 0.0f));
                float keep = p.retention * ((variance.x * 0.2126f + variance.y * 0.7152f) + variance.z * 0.0722f) / (change * change);
                e.mean.w = fmn(e.mean.w, floorf(keep));
                e.moment.w = 0.0f;
            }
            float count = fmn(e.mean.w + 1.0f, 65536.0f);
            float t = 1.0f / count, u = 1.0f - t;
            e.mean = make_float4(e.mean.x * u + v.x * t, e.mean.y * u + v.y * t, e.mean.z * u + v.z * t, count);
            e.moment.x = e.moment.x * u + (v.x * v.x) * t;
            e.moment.y = e.moment.y * u + (v.y * v.y) * t;
            e.moment.z = e.moment.z * u + (v.z * v.z) * t;
        }
    }
    if (!valid || lane != 0u) return;
    e.identity.x = start + p.samples;
    f3 dv, dc;
    direct_at(s, o, surface, normal, dv, dc, tally);
    e.direct = make_float4(dv.x, dv.y, dv.z, 0.0f);
    estimates[i] = e;
    receivers[i].sequence = e.identity.x;
}

// store_shared_faces (voxel_view_cache.wgsl:214-258): an admitted owner
// writes its estimate back as the resident authority (raw moments, count and
// pending flag, next sequence, pending change) and refreshes the slot's
// ranking record when its count or mean changed.
__global__ void view_store_kernel(ViewParams p, SharedCache c, const ViewReceiver *__restrict__ receivers,
                                  const ViewEstimate *__restrict__ estimates, const uint32_t *__restrict__ owners) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.pixels || owners[i] != i) return;
    const ViewReceiver r = receivers[i];
    const ViewEstimate e = estimates[i];
    uint32_t slot = shared_slot(c, r, e.identity.y);
    if (slot == NONE) return;
    uint32_t *st = c.state + (size_t)slot * 36u;
    if (e.identity.w != st[28]) return;
    const uint32_t count = (uint32_t)e.mean.w;
    const bool changed_count = (st[5] & 0x7fffffffu) != count;
    const bool changed_mean = st[0] != __float_as_uint(e.mean.x) || st[1] != __float_as_uint(e.mean.y) || st[2] != __float_as_uint(e.mean.z);
    st[0] = __float_as_uint(e.mean.x); st[1] = __float_as_uint(e.mean.y); st[2] = __float_as_uint(e.mean.z); st[3] = 0u;
    st[8] = __float_as_uint(e.moment.x); st[9] = __float_as_uint(e.moment.y); st[10] = __float_as_uint(e.moment.z); st[11] = 0u;
    st[4] = 0u;
    st[12] = 0u;
    st[5] = count | (e.moment.w > 0.0f ? 0x80000000u : 0u);
    st[6] = e.identity.x;
    uint32_t *m = c.meta + (size_t)slot * 12u;
    m[4] = __float_as_uint(e.moment.w);
    if (changed_mean || changed_count) {
        float nx = fsqrt(fmx(e.moment.x - e.mean.x * e.mean.x, 0.0f)) / (0.002f + fabsf(e.mean.x));
        float ny = fsqrt(fmx(e.moment.y - e.mean.y * e.mean.y, 0.0f)) / (0.002f + fabsf(e.mean.y));
        float nz = fsqrt(fmx(e.moment.z - e.mean.z * e.mean.z, 0.0f)) / (0.002f + fabsf(e.mean.z));
        m[8] = __float_as_uint(fmx(nx, fmx(ny, nz)));
        m[9] = p.clock_ms;
        m[10] = __float_as_uint(fmx(e.mean.x, fmx(e.mean.y, e.mean.z)));
        m[11] = 0xa0000000u | count;
    }
}

void launch_view_frame(const ViewParams &p, const SceneView &s, const Optics &o, const SharedCache &cache, const float *depth,
                       const float *normals, ViewReceiver *receivers, ViewEstimate *estimates, uint32_t *hash, uint32_t *owners,
                       const ViewReceiver *prev_receivers, const ViewEstimate *prev_estimates, const uint32_t *prev_hash,
                       uint32_t *queue, hipStream_t stream, hipEvent_t before_trace, hipEvent_t after_trace) {
    if (p.pixels == 0u) return;
    const dim3 block(256), grid((p.pixels + 255u) / 256u);
    hipLaunchKernelGGL(view_discover_kernel, grid, block, 0, stream, p, s, depth, normals, receivers, estimates);
    hipLaunchKernelGGL(view_index_kernel, grid, block, 0, stream, p, receivers, hash, queue);
    hipLaunchKernelGGL(view_owner_kernel, grid, block, 0, stream, p, receivers, hash, owners);
    hipLaunchKernelGGL(view_reproject_kernel, grid, block, 0, stream, p, receivers, estimates, owners, prev_receivers, prev_estimates,
                       prev_hash, queue);
    if (p.exchange) hipLaunchKernelGGL(view_load_kernel, grid, block, 0, stream, p, cache, receivers, estimates, owners);
    if (before_trace) (void)hipEventRecord(before_trace, stream);
    uint32_t lanes = 1u;
    while (lanes < 8u && lanes < p.samples) lanes *= 2u;
    const dim3 trace_grid((uint32_t)(((uint64_t)p.pixels * lanes + 255u) / 256u));
    hipLaunchKernelGGL(view_trace_kernel, trace_grid, block, 0, stream, p, s, o, receivers, estimates, queue, lanes);
    if (after_trace) (void)hipEventRecord(after_trace, stream);
    if (p.exchange) hipLaunchKernelGGL(view_store_kernel, grid, block, 0, stream, p, cache, receivers, estimates, owners);
}

}  // namespace vgi
