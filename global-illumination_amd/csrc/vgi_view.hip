// The view tier's passes (vgi_view.h). One thread per prepass pixel for
// discovery, indexing, the exchange, presentation and publication; one
// thread per owner (a compacted queue) for tracing; one thread per far cell
// for the far tier's classification and per queued cell for its paths. Every
// pass is a function of the pixel's, owner's or cell's own data and the
// frame's read-only inputs, so the results do not depend on the order in
// which threads run; the atomics pick the owner of a face (the smallest
// pixel index of its duplicates), append owners and cells to their queues
// and add integer counters.
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

__device__ __forceinline__ uint32_t pixel_stamp(uint32_t frame) { return frame & 0x1fffffffu; }

// in_interest (voxel_view_runtime.wgsl:26-30): the voxel's region inside the
// host's geometry interest -- the regions of the loaded pages (pages.rs:
// 186-240), here the scene's loaded box. Outside it a rendered surface never
// gets GI geometry: no receiver waits for it and no far cell lights it.
__device__ __forceinline__ bool in_interest(const SceneView &s, int vx, int vy, int vz) {
    const int rx = vx >> 4, ry = vy >> 4, rz = vz >> 4;
    return rx >= s.minx && ry >= s.miny && rz >= s.minz && rx <= s.maxx && ry <= s.maxy && rz <= s.maxz;
}

// pixel_owned (voxel_view_pixels.wgsl:54-64): the far tier takes a face whose
// voxel, at its distance from the camera (not its view depth, so turning
// does not move the boundary), is under `threshold` pixels across (measure
// 0) or covers under `threshold` pixels (side squared times the cosine to
// the view ray; measures 1 and 2 agree for unit voxels).
__device__ __forceinline__ bool pixel_owned(const ViewParams &p, const FarParams &f, int vx, int vy, int vz, uint32_t face) {
    if ((f.flags & 1u) == 0u || face >= 6u) return false;
    const uint32_t mode = (f.flags >> 1u) & 3u;
    const f3 ray = mk((float)vx + 0.5f, (float)vy + 0.5f, (float)vz + 0.5f) - mk(p.cam[0], p.cam[1], p.cam[2]);
    if (dot3(ray, mk(f.axis[0], f.axis[1], f.axis[2])) <= 0.0f) return false;
    const float side = f.axis[3] / len3(ray);
    if (mode == 0u) return side < f.threshold;
    return side * side * fabsf(dot3(face_normal(face), normalize3(ray))) < f.threshold;
}

// prepass_voxel_receiver (voxel_view_discovery.wgsl:40-104) for unit voxels:
// the decoded point, its normal axis snapped to the face plane, the voxel
// behind it; a point within one pixel's world footprint of an in-plane voxel
// boundary takes the exposed neighbour instead. A pixel whose surface is not
// an exposed face of the loaded scene owns no receiver.
__global__ void view_discover_kernel(ViewParams p, FarParams f, SceneView s, const float *__restrict__ depth,
                                     const float *__restrict__ normals, ViewReceiver *__restrict__ receivers,
                                     ViewEstimate *__restrict__ estimates) {
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
            // Faces small enough for the far tier are lit per pixel and own
            // no receiver (voxel_view_discovery.wgsl:59-64).
            const bool per_pixel = pixel_owned(p, f, vx, vy, vz, face);
            if (!in_interest(s, vx, vy, vz) || per_pixel) {
                // outside the interest, or lit by the far tier
            } else if (receiver_exists(s, vx, vy, vz, face)) {
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
    e.shown = e.mean;
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
    e.shown = e.mean;
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

// previous_cell (voxel_view_pixels.wgsl:83-127): the far cell at this world
// point last frame, bilinear over the previous cells around its old screen
// position, keeping cells of the same face stamped by the previous frame
// within four voxels along the normal and two cells' world footprint across.
// Returns the weighted mean (rgb, count; count 0 if none) and the sequence of
// the best-weighted cell.
__device__ __forceinline__ float4 previous_cell(const ViewParams &p, const FarParams &f, const PixelHistory *__restrict__ previous,
                                                f3 surface, uint32_t face, uint32_t &sequence) {
    float4 found = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    sequence = 0u;
    if (f.previous == NONE || previous == nullptr) return found;
    float c[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] = ((f.clip[i] * surface.x + f.clip[4 + i] * surface.y) + f.clip[8 + i] * surface.z) + f.clip[12 + i];
    if (c[3] <= 0.0f) return found;
    const f3 normal = face_normal(face);
    const float ux = c[0] / c[3] * 0.5f + 0.5f, uy = c[1] / c[3] * -0.5f + 0.5f;
    const float qx = ux * (float)p.width / (float)f.size - 0.5f, qy = uy * (float)p.height / (float)f.size - 0.5f;
    const float bx = floorf(qx), by = floorf(qy);
    // Off screen (every neighbour outside the cell grid) or not a number: no
    // cell (and no float-to-int conversion of an out-of-range value).
    if (!(bx >= -1.0f && by >= -1.0f && bx < (float)f.cells_x && by < (float)f.cells_y)) return found;
    const float fx = qx - bx, fy = qy - by;
    const float reach = 2.0f * (float)f.size * len3(surface - mk(p.cam[0], p.cam[1], p.cam[2])) / f.axis[3];
    const float tolerance = 4.0f * 1.0f + 0.01f;
    const uint32_t expected = face | (pixel_stamp(f.previous) << 3u);
    float weight = 0.0f, samples = 0.0f, best = 0.0f;
    f3 sum = mk(0.0f, 0.0f, 0.0f);
    for (int y = 0; y <= 1; ++y) {
        for (int x = 0; x <= 1; ++x) {
            const int cx = (int)bx + x, cy = (int)by + y;
            if (cx < 0 || cy < 0 || cx >= (int)f.cells_x || cy >= (int)f.cells_y) continue;
            const PixelHistory old = previous[(uint32_t)cy * f.cells_x + (uint32_t)cx];
            if (old.face != expected || old.mean.w == 0.0f) continue;
            const f3 delta = mk(old.sx, old.sy, old.sz) - surface;
            if (fabsf(dot3(delta, normal)) > tolerance || dot3(delta, delta) > reach * reach) continue;
            const float w = (x == 1 ? fx : 1.0f - fx) * (y == 1 ? fy : 1.0f - fy) + 1e-4f;
            sum = sum + mk(old.mean.x, old.mean.y, old.mean.z) * w;
            samples = samples + old.mean.w * w;
            weight = weight + w;
            if (w > best) {
                best = w;
                sequence = old.sequence;
            }
        }
    }
    if (weight > 0.0f) found = make_float4(sum.x / weight, sum.y / weight, sum.z / weight, samples / weight);
    return found;
}

// load_shared_face (voxel_view_cache.wgsl:25-99): an admitted face (its
// incarnation recorded) always loads the resident authority; a face entering
// the view loads it when the authority holds at least as many samples, else
// keeps its own history (promotion). A face not yet shown takes the
// authority's presentation (or its mean). The owner then reserves this
// frame's path sequences on the authority.
__device__ __forceinline__ void load_shared_face(const ViewParams &p, const SharedCache &c, const ViewReceiver &r, ViewEstimate &e) {
    uint32_t slot = shared_slot(c, r, e.identity.y);
    if (slot == NONE) {
        e.identity.w = 0u;
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
    if (e.shown.w == 0.0f && count > 0u) {
        const float4 kept = make_float4(__uint_as_float(st[24]), __uint_as_float(st[25]), __uint_as_float(st[26]),
                                        __uint_as_float(st[27]));
        e.shown = kept.w > 0.0f ? kept : make_float4(e.mean.x, e.mean.y, e.mean.z, 1.0f);
    }
    e.identity.w = incarnation;
    const uint32_t first = max(next, e.identity.x);
    st[6] = first + p.samples;
    e.identity.x = first;
}

// load_shared_faces (voxel_view_cache.wgsl:188-220): the exchange, then a
// new, never-sampled face with the far tier on starts from the far cell at
// its centre with one unit of (borrowed) weight.
__global__ void view_load_kernel(ViewParams p, FarParams f, SharedCache c, ViewReceiver *__restrict__ receivers,
                                 ViewEstimate *__restrict__ estimates, const uint32_t *__restrict__ owners,
                                 const PixelHistory *__restrict__ previous_cells, uint32_t *__restrict__ counters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.pixels || owners[i] != i) return;
    const ViewReceiver r = receivers[i];
    ViewEstimate e = estimates[i];
    if (p.exchange) load_shared_face(p, c, r, e);
    if (e.mean.w == 0.0f && e.identity.x == 0u && f.flags != 0u) {
        uint32_t sequence;
        const float4 cell = previous_cell(p, f, previous_cells, mk(r.sx, r.sy, r.sz), r.face, sequence);
        if (cell.w > 0.0f) {
            e.mean = make_float4(cell.x, cell.y, cell.z, 1.0f);
            e.moment = make_float4(cell.x * cell.x, cell.y * cell.y, cell.z * cell.z, 0.0f);
            e.shown = make_float4(cell.x, cell.y, cell.z, 1.0f);
            atomicAdd(&counters[PIXEL_RECEIVERS_SEEDED], 1u);
        }
    }
    estimates[i] = e;
    receivers[i].sequence = e.identity.x;
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

// store_shared_faces (voxel_view_cache.wgsl:222-258): an admitted owner
// writes its estimate back as the resident authority (raw moments, count and
// pending flag, next sequence, pending change, presentation) and refreshes
// the slot's ranking record when its count or mean changed.
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
    // The presentation (which may hold borrowed support) goes to the slot's
    // display record and its copies in the result records, never into the
    // raw moments or the count the path tracer reads.
    st[24] = __float_as_uint(e.shown.x); st[25] = __float_as_uint(e.shown.y);
    st[26] = __float_as_uint(e.shown.z); st[27] = __float_as_uint(e.shown.w);
    st[7] = st[24];
    st[13] = st[25]; st[14] = st[26]; st[15] = st[27];
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

// ---------------------------------------------------------------- presentation
// reconstruct_view (voxel_view_runtime.wgsl:350-474): spatial pooling of
// coplanar owners, each estimate weighted by 1 / (variance of its mean +
// squared bias against this face). Owners are binned by 2x2-pixel screen
// tiles of their own pixel; a tile's members are visited in pixel order (the
// reference walks a linked list in arrival order, so its sums are not
// reproducible; here they are).
constexpr float kVariancePriorSamples = 64.0f;

__device__ __forceinline__ float luminance(float4 v) { return (v.x * 0.2126f + v.y * 0.7152f) + v.z * 0.0722f; }

// mean_variance (voxel_view_runtime.wgsl:380-394): the variance of a mean,
// per-channel deviations shrunk toward relative sigma one with 64 pseudo
// samples, plus an unseen-light floor (reference / n)^2.
__device__ __forceinline__ float mean_variance(float4 m, float4 moment, float reference) {
    const float n = fmx(m.w, 1.0f);
    const float mean = luminance(m);
    const float cx = fmx(moment.x - m.x * m.x, 0.0f), cy = fmx(moment.y - m.y * m.y, 0.0f), cz = fmx(moment.z - m.z * m.z, 0.0f);
    const float deviation = (fsqrt(cx) * 0.2126f + fsqrt(cy) * 0.7152f) + fsqrt(cz) * 0.0722f;
    const float prior = mean * mean;
    const float samples = fmx(n - 1.0f, 0.0f);
    const float shrunk = (deviation * deviation * samples + prior * kVariancePriorSamples) / (samples + kVariancePriorSamples);
    const float unseen = reference / n;
    return shrunk / n + unseen * unseen;
}

// coplanar_donor (voxel_view_runtime.wgsl:396-404), unit voxels (stride 1).
__device__ __forceinline__ bool coplanar_donor(const ViewReceiver &r, f3 normal, const ViewReceiver *__restrict__ receivers,
                                               const uint32_t *__restrict__ owners, uint32_t candidate) {
    const ViewReceiver d = receivers[candidate];
    if (d.face != r.face) return false;
    const f3 delta = mk(d.sx, d.sy, d.sz) - mk(r.sx, r.sy, r.sz);
    const uint32_t u = (r.face / 2u + 1u) % 3u;
    const float stride = 2.0f * ((u == 0u ? r.sx : (u == 1u ? r.sy : r.sz)) - (float)(u == 0u ? r.vx : (u == 1u ? r.vy : r.vz)));
    const float reach = 16.0f * stride;
    return fabsf(dot3(delta, normal)) <= 0.05f && dot3(delta, delta) <= reach * reach && owners[candidate] == candidate;
}

// Calls visit(candidate) for the owners-to-be of tile (tx, ty): its pixels
// in row-major order.
template <class Visit>
__device__ __forceinline__ void tile_members(const ViewParams &p, int tx, int ty, Visit visit) {
    for (int b = 0; b < 2; ++b) {
        const int y = ty * 2 + b;
        if (y >= (int)p.height) break;
        for (int a = 0; a < 2; ++a) {
            const int x = tx * 2 + a;
            if (x >= (int)p.width) break;
            visit((uint32_t)y * p.width + (uint32_t)x);
        }
    }
}

__global__ void view_reconstruct_kernel(ViewParams p, const ViewReceiver *__restrict__ receivers, ViewEstimate *__restrict__ estimates,
                                        const uint32_t *__restrict__ owners) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.pixels || owners[i] != i) return;
    const ViewReceiver r = receivers[i];
    const ViewEstimate center = estimates[i];
    const f3 normal = face_normal(r.face);
    const int tiles_x = (int)((p.width + 1u) / 2u), tiles_y = (int)((p.height + 1u) / 2u);
    const int tx = (int)((i % p.width) / 2u), ty = (int)((i / p.width) / 2u);
    const float center_luminance = luminance(center.mean);
    // The brightest sampled coplanar estimate nearby (3x3 tiles).
    float reference = center_luminance;
    for (int y = -1; y <= 1; ++y) {
        for (int x = -1; x <= 1; ++x) {
            const int qx = tx + x, qy = ty + y;
            if (qx < 0 || qy < 0 || qx >= tiles_x || qy >= tiles_y) continue;
            tile_members(p, qx, qy, [&](uint32_t candidate) {
                if (candidate == i || !coplanar_donor(r, normal, receivers, owners, candidate)) return;
                const float4 mean = estimates[candidate].mean;
                if (mean.w > 0.0f) reference = fmx(reference, luminance(mean));
            });
        }
    }
    const float center_variance = mean_variance(center.mean, center.moment, reference);
    const float tolerance = p.display[0] * center_luminance;
    // How many times more evidence this face needs, as a gather radius.
    const float deficit = center_variance / fmx(tolerance * tolerance, 1e-12f);
    const int radius = (int)fclamp(ceilf(fsqrt(deficit / 3.14159f) * 0.5f), 1.0f, 4.0f);
    const bool unsampled = center.mean.w <= 0.0f;
    f3 sum = mk(0.0f, 0.0f, 0.0f);
    float weight = 0.0f;
    if (!unsampled) {
        const float w = 1.0f / fmx(center_variance, 1e-12f);
        sum = mk(center.mean.x, center.mean.y, center.mean.z) * w;
        weight = w;
    }
    if ((unsampled || center_variance > tolerance * tolerance) && p.display[1] > 0.0f) {
        for (int y = -radius; y <= radius; ++y) {
            for (int x = -radius; x <= radius; ++x) {
                const int qx = tx + x, qy = ty + y;
                if (qx < 0 || qy < 0 || qx >= tiles_x || qy >= tiles_y) continue;
                tile_members(p, qx, qy, [&](uint32_t candidate) {
                    if (candidate == i || !coplanar_donor(r, normal, receivers, owners, candidate)) return;
                    // Only mean and moment: other threads write their own `shown`.
                    const float4 mean = estimates[candidate].mean;
                    if (mean.w <= 0.0f) return;
                    const float variance = mean_variance(mean, estimates[candidate].moment, reference);
                    float bias2 = 0.0f;
                    if (!unsampled) {
                        const float difference = luminance(mean) - center_luminance;
                        bias2 = fmx(difference * difference - variance - center_variance, 0.0f);
                    }
                    const float w = 1.0f / fmx(variance + bias2, 1e-12f);
                    sum = sum + mk(mean.x, mean.y, mean.z) * w;
                    weight = weight + w;
                });
            }
        }
    }
    if (weight <= 0.0f) return;
    const float t = unsampled ? 0.0f : 1.0f - p.display[1];
    const f3 pooled = mk(sum.x / weight, sum.y / weight, sum.z / weight);
    const f3 desired = pooled * (1.0f - t) + mk(center.mean.x, center.mean.y, center.mean.z) * t;
    const float blend = center.shown.w > 0.0f ? p.display[2] : 1.0f;
    const f3 shown = mk(center.shown.x, center.shown.y, center.shown.z) * (1.0f - blend) + desired * blend;
    estimates[i].shown = make_float4(shown.x, shown.y, shown.z, 1.0f);
}

// publish_view (voxel_view_runtime.wgsl:476-521): per pixel, an owner's
// shown irradiance (plus custom direct light, which no host light draws),
// the evidence behind it and its corner sun bytes; counters of history,
// uncovered and displayed faces.
__global__ void view_publish_kernel(ViewParams p, SceneView s, Optics o, const ViewReceiver *__restrict__ receivers,
                                    const ViewEstimate *__restrict__ estimates, const uint32_t *__restrict__ owners,
                                    ViewPixel *__restrict__ out, uint32_t *__restrict__ counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.pixels) return;
    ViewPixel px = {0.0f, 0.0f, 0.0f, NONE, 0.0f, 0.0f, 0.0f, 0.0f, make_uint4(0u, 0u, 0u, 0u)};
    const ViewReceiver r = receivers[i];
    if (r.face < 6u && owners[i] == i) {
        const ViewEstimate e = estimates[i];
        px.sx = r.sx; px.sy = r.sy; px.sz = r.sz;
        px.face = r.face;
        f3 irradiance = mk(e.shown.x, e.shown.y, e.shown.z);
        if (p.custom) {
            f3 dv, dc;
            NoTally tally;
            direct_at(s, o, mk(r.sx, r.sy, r.sz), face_normal(r.face), dv, dc, tally);
            irradiance = irradiance + dc;
        }
        px.ir = irradiance.x; px.ig = irradiance.y; px.ib = irradiance.z;
        const bool lit = e.shown.w > 0.0f || irradiance.x > 0.0f || irradiance.y > 0.0f || irradiance.z > 0.0f;
        px.evidence = lit ? fmx(e.mean.w, 0.001f) : 0.0f;
        uint32_t sky[4];
#pragma unroll
        for (uint32_t corner = 0u; corner < 4u; ++corner) {
            const uint32_t byte = (e.identity.z >> (corner * 8u)) & 255u;
            sky[corner] = (byte & 15u) | ((byte & 128u) << 24u);
        }
        px.sky = make_uint4(sky[0], sky[1], sky[2], sky[3]);
        const uint32_t history = (uint32_t)e.mean.w;
        atomicAdd(&counters[SHOW_HISTORY], history);
        atomicMax(&counters[SHOW_MAX_HISTORY], history);
        if (e.mean.w == 0.0f) atomicAdd(&counters[SHOW_UNCOVERED], 1u);
        if (px.evidence > 0.0f) atomicAdd(&counters[SHOW_DISPLAYED], 1u);
    }
    out[i] = px;
}

// ---------------------------------------------------------------- far tier
// write_output (voxel_view_pixels.wgsl:71-81): the material's packed cell,
// the mean plus custom direct light (a presentation term, never part of the
// history): half rg, half (b, count), face and stamp, the surface's
// coordinate along the normal axis.
__device__ __forceinline__ uint32_t pack_half2(float lo, float hi) {
    const _Float16 a = (_Float16)lo, b = (_Float16)hi;
    return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16u);
}

__device__ __forceinline__ void far_output(const FarParams &f, const SceneView &s, const Optics &o, uint32_t *__restrict__ output,
                                           uint32_t index, const PixelHistory &cell) {
    const uint32_t face = cell.face & 7u;
    f3 shown = mk(cell.mean.x, cell.mean.y, cell.mean.z);
    if (f.custom) {
        f3 dv, dc;
        BudgetTally tally;
        tally.limit = f.steps;
        direct_at(s, o, mk(cell.sx, cell.sy, cell.sz), face_normal(face), dv, dc, tally);
        shown = shown + dc;
    }
    const uint32_t axis = face / 2u;
    const float plane = axis == 0u ? cell.sx : (axis == 1u ? cell.sy : cell.sz);
    uint4 word = make_uint4(pack_half2(shown.x, shown.y), pack_half2(shown.z, cell.mean.w), cell.face, __float_as_uint(plane));
    *reinterpret_cast<uint4 *>(output + 16u + (size_t)index * 4u) = word;
}

// Wavefront-aggregated counting: one atomic per wavefront instead of one per
// lane (every lane of the wavefront must reach the call). wave_count adds the
// number of lanes whose flag is set; wave_sum adds the lanes' values.
__device__ __forceinline__ void wave_count(uint32_t *counter, bool flag) {
    const uint64_t mask = __ballot(flag);
    if (mask != 0ull && __lane_id() == (uint32_t)__ffsll((long long)mask) - 1u) atomicAdd(counter, (uint32_t)__popcll(mask));
}
__device__ __forceinline__ void wave_sum(uint32_t *counter, uint32_t value) {
    uint32_t sum = value;
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) sum += __shfl_xor(sum, offset, 64);
    if (__lane_id() == 0u && sum != 0u) atomicAdd(counter, sum);
}

// classify_pixels (voxel_view_pixels.wgsl:129-194): each cell samples one of
// its pixels this frame (cycling through all of them), keeps it if the
// pixel's face is a far face, takes its history from the previous cells
// (reprojected by world position) or, failing that, from the face's view
// receiver of the previous frame, and queues itself for a path when its
// history is short or one frame in `rate`; otherwise it writes its output.
// The queue is filled by wavefront-compacted reservations (ballot, prefix,
// one atomic per wavefront), so its order is the arrival order of
// wavefronts; each queued cell's path depends only on the cell.
__device__ __forceinline__ bool far_classify(const ViewParams &p, const FarParams &f, const SceneView &s, const float *__restrict__ depth,
                                             const float *__restrict__ normals, const ViewReceiver *__restrict__ prev_receivers,
                                             const ViewEstimate *__restrict__ prev_estimates, const uint32_t *__restrict__ prev_hash,
                                             const PixelHistory *__restrict__ previous, uint32_t id, PixelHistory &cell,
                                             bool &reprojected, bool &seeded) {
    const uint32_t size = f.size;
    const uint32_t cx = id % f.cells_x, cy = id / f.cells_x;
    // One pixel of the cell per frame, cycling through all of them (5 is
    // coprime with every cell area 1, 4, 9, 16).
    const uint32_t k = (f.frame * 5u + (cx ^ cy)) % (size * size);
    const uint32_t px = cx * size + k % size, py = cy * size + k / size;
    if (px >= p.width || py >= p.height) return false;
    const uint32_t pixel = py * p.width + px;
    const float z = depth[pixel];
    if (z <= 0.0f) return false;
    const f3 position = view_unproject(p, (float)px, (float)py, z);
    const f3 camera = mk(p.cam[0], p.cam[1], p.cam[2]);
    // The side test needs only the distance: most near cells stop here.
    if (((f.flags >> 1u) & 3u) == 0u && f.axis[3] > 1.05f * f.threshold * len3(position - camera)) return false;
    const uint32_t face = prepass_face(mk(normals[pixel * 3u], normals[pixel * 3u + 1u], normals[pixel * 3u + 2u]));
    if (face >= 6u) return false;
    const f3 n = face_normal(face);
    const uint32_t axis = face / 2u;
    float sv[3] = {position.x, position.y, position.z};
    sv[axis] = rintf(sv[axis]);
    const f3 surface = mk(sv[0], sv[1], sv[2]);
    const f3 inside = surface - n * 0.01f;
    const int vx = (int)floorf(inside.x), vy = (int)floorf(inside.y), vz = (int)floorf(inside.z);
    if (!in_interest(s, vx, vy, vz) || !pixel_owned(p, f, vx, vy, vz, face)) return false;
    uint32_t sequence;
    float4 history = previous_cell(p, f, previous, surface, face, sequence);
    if (history.w > 0.0f) {
        reprojected = true;
    } else {
        // Handoff: a face crossing into the far band starts from its view
        // receiver's history (the same estimator).
        ViewReceiver key;
        key.vx = vx; key.vy = vy; key.vz = vz;
        key.face = face;
        const f3 centre = face_centre(vx, vy, vz, n);
        key.sx = centre.x; key.sy = centre.y; key.sz = centre.z;
        key.sequence = 0u;
        const uint32_t owner = previous_owner(p, key, prev_receivers, prev_hash);
        if (owner != NONE && prev_estimates[owner].mean.w > 0.0f) {
            history = prev_estimates[owner].mean;
            sequence = prev_estimates[owner].identity.x;
            seeded = true;
        }
    }
    const float count = fmn(history.w, f.history);
    cell.mean = make_float4(history.x, history.y, history.z, count);
    cell.sx = surface.x; cell.sy = surface.y; cell.sz = surface.z;
    cell.face = face | (pixel_stamp(f.frame) << 3u);
    cell.sequence = sequence;
    cell.pad[0] = cell.pad[1] = cell.pad[2] = 0u;
    return true;
}

__global__ __launch_bounds__(256) void far_classify_kernel(ViewParams p, FarParams f, SceneView s, Optics o, const float *__restrict__ depth,
                                                           const float *__restrict__ normals,
                                                           const ViewReceiver *__restrict__ prev_receivers,
                                                           const ViewEstimate *__restrict__ prev_estimates,
                                                           const uint32_t *__restrict__ prev_hash, FarBuffers b,
                                                           uint32_t *__restrict__ counters) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id == 0u) {
        uint32_t *h = b.output;
        h[0] = __float_as_uint(p.cam[0]); h[1] = __float_as_uint(p.cam[1]); h[2] = __float_as_uint(p.cam[2]); h[3] = 0u;
        h[4] = __float_as_uint(f.axis[0]); h[5] = __float_as_uint(f.axis[1]); h[6] = __float_as_uint(f.axis[2]);
        h[7] = __float_as_uint(f.axis[3]);
        h[8] = f.cells_x; h[9] = f.cells_y; h[10] = f.size; h[11] = f.flags;
        h[12] = 0u; h[13] = 0u; h[14] = __float_as_uint(f.threshold); h[15] = pixel_stamp(f.frame);
    }
    PixelHistory cell;
    bool reprojected = false, seeded = false;
    const bool owned = id < f.cells &&
                       far_classify(p, f, s, depth, normals, prev_receivers, prev_estimates, prev_hash, b.previous, id, cell, reprojected,
                                    seeded);
    bool queued = false;
    if (owned) {
        b.cells[id] = cell;
        const uint32_t rate = max((uint32_t)f.rate, 1u);
        queued = cell.mean.w < 4.0f || (id + f.frame) % rate == 0u;
        if (!queued) far_output(f, s, o, b.output, id, cell);  // the trace writes a queued cell's output
    }
    // Every lane of the wavefront reaches the counting and the queue's
    // reservation: ballot, the lane's rank among the queued lanes, one
    // atomic per wavefront for the queue and for each counter.
    const uint64_t mask = __ballot(queued);
    if (mask != 0ull) {
        const uint32_t lane = __lane_id();
        const uint32_t leader = (uint32_t)__ffsll((long long)mask) - 1u;
        uint32_t base = 0u;
        if (lane == leader) base = atomicAdd(&b.queue[0], (uint32_t)__popcll(mask));
        base = __shfl(base, (int)leader, 64);
        if (queued) b.queue[1u + base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = id;
    }
    wave_count(&counters[PIXEL_OWNED], owned);
    wave_count(&counters[PIXEL_REPROJECTED], reprojected);
    wave_count(&counters[PIXEL_SEEDED], seeded);
    wave_sum(&counters[PIXEL_HISTORY], owned ? (uint32_t)cell.mean.w : 0u);
}

// trace_pixel_queue (voxel_view_pixels.wgsl:201-216): one path per queued
// cell under the far tier's ray budget, stream pcg(cell * 9781 + 0x9e3779b9),
// the cell's own sequence; a known sample enters the capped running mean.
__global__ void far_trace_kernel(FarParams f, SceneView s, Optics o, FarBuffers b, uint32_t *__restrict__ counters) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t queued = b.queue[0];
    wave_count(&counters[PIXEL_TRACED], id < queued);
    if (id >= queued) return;
    const uint32_t index = b.queue[1u + id];
    PixelHistory cell = b.cells[index];
    BudgetTally tally;
    tally.limit = f.steps;
    bool known;
    FirstHit first;
    const f3 value = trace_receiver<BudgetTally>(s, o, mk(cell.sx, cell.sy, cell.sz), face_normal(cell.face & 7u),
                                                 pcg(index * 9781u + 0x9e3779b9u), cell.sequence, known, first, tally);
    cell.sequence += 1u;
    if (known) {
        cell.mean.w = fmn(cell.mean.w + 1.0f, f.history);
        const float t = 1.0f / cell.mean.w, u = 1.0f - t;
        cell.mean.x = cell.mean.x * u + value.x * t;
        cell.mean.y = cell.mean.y * u + value.y * t;
        cell.mean.z = cell.mean.z * u + value.z * t;
    }
    b.cells[index] = cell;
    far_output(f, s, o, b.output, index, cell);
}

void launch_view_frame(const ViewParams &p, const FarParams &f, const SceneView &s, const Optics &o, const SharedCache &cache,
                       const float *depth, const float *normals, ViewReceiver *receivers, ViewEstimate *estimates, uint32_t *hash,
                       uint32_t *owners, const ViewReceiver *prev_receivers, const ViewEstimate *prev_estimates,
                       const uint32_t *prev_hash, uint32_t *queue, ViewPixel *pixels_out, uint32_t *counters, const FarBuffers &far,
                       hipStream_t stream, hipEvent_t before_trace, hipEvent_t after_trace, hipEvent_t before_far) {
    if (p.pixels == 0u) return;
    const dim3 block(256), grid((p.pixels + 255u) / 256u);
    hipLaunchKernelGGL(view_discover_kernel, grid, block, 0, stream, p, f, s, depth, normals, receivers, estimates);
    hipLaunchKernelGGL(view_index_kernel, grid, block, 0, stream, p, receivers, hash, queue);
    hipLaunchKernelGGL(view_owner_kernel, grid, block, 0, stream, p, receivers, hash, owners);
    hipLaunchKernelGGL(view_reproject_kernel, grid, block, 0, stream, p, receivers, estimates, owners, prev_receivers, prev_estimates,
                       prev_hash, queue);
    if (p.exchange || f.flags != 0u)
        hipLaunchKernelGGL(view_load_kernel, grid, block, 0, stream, p, f, cache, receivers, estimates, owners, far.previous, counters);
    if (before_trace) (void)hipEventRecord(before_trace, stream);
    uint32_t lanes = 1u;
    while (lanes < 8u && lanes < p.samples) lanes *= 2u;
    const dim3 trace_grid((uint32_t)(((uint64_t)p.pixels * lanes + 255u) / 256u));
    hipLaunchKernelGGL(view_trace_kernel, trace_grid, block, 0, stream, p, s, o, receivers, estimates, queue, lanes);
    if (after_trace) (void)hipEventRecord(after_trace, stream);
    hipLaunchKernelGGL(view_reconstruct_kernel, grid, block, 0, stream, p, receivers, estimates, owners);
    if (p.exchange) hipLaunchKernelGGL(view_store_kernel, grid, block, 0, stream, p, cache, receivers, estimates, owners);
    hipLaunchKernelGGL(view_publish_kernel, grid, block, 0, stream, p, s, o, receivers, estimates, owners, pixels_out, counters);
    if (before_far) (void)hipEventRecord(before_far, stream);
    if ((f.flags & 1u) != 0u && f.cells > 0u) {
        const dim3 cell_grid((f.cells + 255u) / 256u);
        hipLaunchKernelGGL(far_classify_kernel, cell_grid, block, 0, stream, p, f, s, o, depth, normals, prev_receivers, prev_estimates,
                           prev_hash, far, counters);
        hipLaunchKernelGGL(far_trace_kernel, cell_grid, block, 0, stream, f, s, o, far, counters);
    }
}

}  // namespace vgi
