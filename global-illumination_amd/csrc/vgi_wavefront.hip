// Wavefront path tracing of bound faces: the estimator of trace_faces_kernel
// (voxel_trace.wgsl:44-92 over voxel_path.wgsl:396-520) split into stages so
// that every lane of a wavefront does the same kind of work.
//
//   begin   one thread per (face, sample): the receiver's emitter sample and
//           first direction; queues its shadow ray and the first path ray.
//   trace   persistent: lanes take rays from the queue and walk them one
//           step at a time; a lane whose ray ended takes the next ray at
//           once (wave ballot, prefix count, one atomic per wave), so the
//           walk never idles behind the wavefront's longest ray.
//   shade   one thread per path: resolves the previous hit's deferred terms
//           (sun and emitter shadows, terminal connection) into the path's
//           energy, then shades the new hit and queues its shadow rays and
//           the next path ray (compacted: wave prefix sum + one atomic);
//           optionally the new hits' world-cache lookups are deduplicated
//           per wavefront first (ballot match, one probe per distinct face).
//   final   one thread per face: the samples' values summed in the
//           megakernel's lane order, the analytic direct term, the record.
//
// Every ray is walked by the same ray_start/ray_advance as the megakernel and
// every term is computed with the same operations in the same order, so the
// records are bit-identical to trace_faces_kernel's.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vgi_kernels.h"
#include "vgi_trace.h"
#include "vgi_wavefront.h"

namespace vgi {

namespace {

constexpr uint32_t F_PATH = 1u << 0;      // a path ray is queued (its hit is shaded next)
constexpr uint32_t F_RESOLVE = 1u << 1;   // the last hit's deferred terms are pending
constexpr uint32_t F_NEE0 = 1u << 2;      // the receiver's emitter shadow ray is queued
constexpr uint32_t F_UNKNOWN = 1u << 3;   // the path reached unloaded space
constexpr uint32_t F_STARTED = 1u << 4;   // the first direction left the face
constexpr uint32_t KIND_SHIFT = 8u;       // deferred emitter term: see K_*
constexpr uint32_t LIGHT_SHIFT = 12u;     // lights whose shadow ray was queued (5 bits)
constexpr uint32_t BOUNCE_SHIFT = 20u;    // bounce of the pending path ray (4 bits)

constexpr uint32_t WF_CHUNK = 512u;   // rays a wavefront takes at a time
constexpr uint32_t WF_REFILL = 16u;   // idle lanes that trigger a refill

constexpr uint32_t K_ZERO = 0u;       // no emitter term
constexpr uint32_t K_NEE = 1u;        // candidate x visibility (slot 1)
constexpr uint32_t K_TERM_NEE = 2u;   // (candidate x visibility) x 2 (slot 1)
constexpr uint32_t K_TERM_HIT = 3u;   // terminal connection through the slot-7 hit
constexpr uint32_t K_CACHED = 4u;     // the world cache's value, no emitter term

__device__ __forceinline__ f3 xyz(float4 v) { return mk(v.x, v.y, v.z); }
__device__ __forceinline__ float4 pack(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ float4 pack_bits(f3 v, uint32_t w) { return make_float4(v.x, v.y, v.z, __uint_as_float(w)); }

// Queue writes of a wavefront: each lane queues `n` rays; one atomic per
// wavefront reserves them, a prefix sum over the lanes places them.
__device__ __forceinline__ uint32_t queue_reserve(uint32_t *__restrict__ qcount, uint32_t n) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t inclusive = n;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t v = (uint32_t)__shfl_up((int)inclusive, off, 64);
        if (lane >= (uint32_t)off) inclusive += v;
    }
    uint32_t total = (uint32_t)__shfl((int)inclusive, 63, 64);
    uint32_t base = 0u;
    if (lane == 63u && total) base = atomicAdd(qcount, total);
    base = (uint32_t)__shfl((int)base, 63, 64);
    return base + inclusive - n;
}

struct Emit {
    WfRay r[7];
    uint32_t n = 0u;
    __device__ __forceinline__ void add(f3 o, f3 d, float max_distance, uint32_t ref) {
        r[n] = WfRay{o.x, o.y, o.z, max_distance, d.x, d.y, d.z, ref};
        n += 1u;
    }
};

__device__ __forceinline__ void emit_all(const Emit &e, WfRay *__restrict__ queue, uint32_t *__restrict__ qcount) {
    uint32_t at = queue_reserve(qcount, e.n);
    for (uint32_t k = 0u; k < e.n; ++k) queue[at + k] = e.r[k];
}

__device__ __forceinline__ void face_of_path(const uint4 *__restrict__ tasks, uint32_t face_index, uint32_t sample, uint32_t sequence_base,
                                             int &vx, int &vy, int &vz, f3 &normal, uint32_t &stream, uint32_t &sequence, bool &valid) {
    uint4 t0 = tasks[face_index * 2u], t1 = tasks[face_index * 2u + 1u];
    valid = t0.w != NONE;
    vx = (int)t0.x; vy = (int)t0.y; vz = (int)t0.z;
    normal = valid ? face_normal(t0.w) : mk(0.0f, 0.0f, 0.0f);
    stream = hash_voxel(vx, vy, vz) ^ pcg(t0.w + 1u);
    sequence = t1.x + sequence_base + sample;
}

// Hit-face deduplication of the world-cache lookups (voxel_path.wgsl:63-87)
// of one shading launch. Every lane of the wavefront calls it; `want` lanes
// hold a (voxel, face) hit. A match loop over the wavefront's distinct keys
// (ballot of the lanes equal to the lowest pending lane's key) gives every
// lane the lowest lane of its key; only those lanes probe the cache (in
// parallel), and the others take their answer across the wavefront. The
// lookup is a pure function of the key and the cache state, which no lane
// writes while shading, so every lane's answer is the one it would have
// probed itself. `stats` (optional) gains the lookups and the probes made.
__device__ __forceinline__ bool wave_cached_energy(const Optics &o, bool want, int vx, int vy, int vz, uint32_t face, f3 &value,
                                                   unsigned long long *__restrict__ stats) {
    const uint32_t lane = __lane_id();
    uint64_t pending = __ballot(want);
    uint32_t leader = lane;
    const uint64_t wanted = pending;
    while (pending != 0ull) {
        const int first = __ffsll((long long)pending) - 1;
        const int lx = __shfl(vx, first, 64), ly = __shfl(vy, first, 64), lz = __shfl(vz, first, 64);
        const uint32_t lf = (uint32_t)__shfl((int)face, first, 64);
        const bool same = want && vx == lx && vy == ly && vz == lz && face == lf;
        const uint64_t group = __ballot(same) & pending;
        if ((group >> lane) & 1ull) leader = (uint32_t)first;
        pending &= ~group;
    }
    const bool probe = want && leader == lane;
    bool found = false;
    f3 cached = mk(0.0f, 0.0f, 0.0f);
    if (probe) {
        uint32_t slot;
        found = cached_energy(o, vx, vy, vz, face, cached, slot);
    }
    const bool shared = __shfl((int)found, (int)leader, 64) != 0;
    value = mk(__shfl(cached.x, (int)leader, 64), __shfl(cached.y, (int)leader, 64), __shfl(cached.z, (int)leader, 64));
    const uint64_t probing = __ballot(probe);
    if (stats && lane == 0u && wanted != 0ull) {
        atomicAdd(&stats[5], (unsigned long long)__popcll(wanted));
        atomicAdd(&stats[6], (unsigned long long)__popcll(probing));
    }
    return want && shared;
}

}  // namespace

// ------------------------------------------------------------------ begin
// trace_receiver (voxel_path.wgsl:504-520) up to its first rays.
__global__ __launch_bounds__(256) void wf_begin_kernel(SceneView scene, Optics optics, const uint4 *__restrict__ tasks, uint32_t count,
                                                       uint32_t samples, uint32_t sequence_base, WfPaths paths, WfRay *__restrict__ queue,
                                                       uint32_t *__restrict__ qcount, const uint32_t *__restrict__ order) {
    (void)scene;
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    Emit e;
    if (p < count * samples) {
        uint32_t face_index = p / samples, sample = p % samples;
        if (order) face_index = order[face_index];
        int vx, vy, vz;
        f3 normal;
        uint32_t stream, sequence;
        bool valid;
        face_of_path(tasks, face_index, sample, sequence_base, vx, vy, vz, normal, stream, sequence, valid);
        uint32_t flags = 0u;
        f3 c0 = mk(0.0f, 0.0f, 0.0f);
        if (valid) {
            f3 surface = mk((float)vx, (float)vy, (float)vz) + mk(0.5f, 0.5f, 0.5f) + normal * 0.5f;
            f3 direction = hemisphere(normal, sample_1d(stream, sequence, 0u), sample_1d(stream, sequence, 1u));
            f3 so, sd;
            float sdist;
            if (emitter_candidate(optics, surface, normal, stream, sequence, 0u, c0, so, sd, sdist)) {
                flags |= F_NEE0;
                e.add(so, sd, sdist, p * 8u + 1u);
            }
            if (dot3(normal, direction) > 0.0f) {
                flags |= F_STARTED;
                f3 start = surface + normal * 0.001f;
                float previous_pdf = cosine_pdf(normal, direction);
                f3 previous_position = start - normal * 0.001f;
                paths.throughput[p] = pack(mk(1.0f, 1.0f, 1.0f), previous_pdf);
                paths.origin[p] = pack(start, 0.0f);
                paths.direction[p] = pack(direction, 0.0f);
                paths.previous[p] = pack_bits(previous_position, normal_code(normal));
                if (previous_pdf > 0.0f) {
                    flags |= F_PATH;
                    e.add(start, direction, connection_reach(optics), p * 8u);
                }
            }
        }
        paths.energy[p] = pack_bits(mk(0.0f, 0.0f, 0.0f), flags);
        paths.nee0[p] = pack(c0, 0.0f);
    }
    emit_all(e, queue, qcount);
}

// ------------------------------------------------------------------ trace
// Slot 0: the path ray (trace_diffuse, voxel_path.wgsl:271-282) -> hit0;
// slot 7: the terminal connection ray -> hit1; slots 1..6: shadow rays ->
// vis (1 when nothing was hit).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VGI_TRACE_WAVES))) void wf_trace_kernel(
    SceneView scene, Optics optics, const WfRay *__restrict__ queue, const uint32_t *__restrict__ qcount, uint32_t *__restrict__ head,
    WfHit *__restrict__ hit0, WfHit *__restrict__ hit1, uint8_t *__restrict__ vis, unsigned long long *__restrict__ rays_counter) {
    const uint32_t count = *qcount;
    const uint32_t lane = threadIdx.x & 63u;
    RayTally tally;
    RayWalk w;
    RayResult r;
    uint32_t ref = 0u;
    bool active = false, exhausted = count == 0u;
    // Each wavefront takes a chunk of consecutive rays (rays of neighbouring
    // faces, so its lanes stay coherent in the caches) with one atomic, and
    // hands a chunk's rays to its idle lanes until the chunk is empty.
    uint32_t chunk_next = 0u, chunk_end = 0u;
    for (;;) {
        bool ended = false;
        unsigned long long idle = __ballot(!active);
        uint32_t n_idle = (uint32_t)__popcll(idle);
        if (n_idle >= WF_REFILL || n_idle == 64u) {
            if (chunk_next >= chunk_end && !exhausted) {
                uint32_t base = 0u;
                if (lane == 0u) base = atomicAdd(head, WF_CHUNK);
                base = (uint32_t)__shfl((int)base, 0, 64);
                if (base >= count) exhausted = true;
                else { chunk_next = base; chunk_end = min(base + WF_CHUNK, count); }
            }
            uint32_t take = min(n_idle, chunk_end - chunk_next);
            if (!active) {
                uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (rank < take) {
                    WfRay q = queue[chunk_next + rank];
                    ref = q.ref;
                    w.origin = mk(q.ox, q.oy, q.oz);
                    w.direction = mk(q.dx, q.dy, q.dz);
                    w.max_distance = q.max_distance;
                    if (ray_start(scene, w.origin, w.direction, q.max_distance, w, r, tally)) ended = true;
                    else active = true;
                }
            }
            chunk_next += take;
        }
        if (active && ray_advance(scene, w, r, tally)) {
            active = false;
            ended = true;
        }
        if (ended) {
            uint32_t path = ref >> 3, slot = ref & 7u;
            if (slot == 0u || slot == 7u) {
                // trace_diffuse: with limited reach, an escape that still has
                // geometry ahead is cut (status 3).
                if (slot == 0u && optics.limited_reach && r.status == 0u &&
                    geometry_ahead(scene, w.origin + w.direction * w.max_distance, w.direction, optics.max_distance - w.max_distance))
                    r.status = 3u;
                WfHit h;
                h.vx = r.vx; h.vy = r.vy; h.vz = r.vz;
                h.material = r.material;
                h.px = r.position.x; h.py = r.position.y; h.pz = r.position.z;
                h.status_normal = r.status | (normal_code(r.normal) << 8);
                (slot == 0u ? hit0 : hit1)[path] = h;
            } else {
                vis[(size_t)path * 8u + slot] = r.status == 0u ? 1u : 0u;
            }
        }
        if (exhausted && chunk_next >= chunk_end && __ballot(active) == 0ull) break;
    }
    unsigned long long n = tally.rays;
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if (lane == 0u && rays_counter) atomicAdd(rays_counter, n);
}

// ------------------------------------------------------------------ shade
// kDedup: the new hits' world-cache lookups are deduplicated per wavefront
// (wave_cached_energy) before shading; the records are the same either way.
template <bool kDedup>
__global__ __launch_bounds__(256) void wf_shade_kernel(SceneView scene, Optics optics, const uint4 *__restrict__ tasks, uint32_t count,
                                                       uint32_t samples, uint32_t sequence_base, WfPaths paths,
                                                       const WfHit *__restrict__ hit0, const WfHit *__restrict__ hit1,
                                                       const uint8_t *__restrict__ vis, WfRay *__restrict__ queue,
                                                       uint32_t *__restrict__ qcount, const uint32_t *__restrict__ order,
                                                       unsigned long long *__restrict__ stats) {
    (void)scene;
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    Emit e;
    uint32_t flags = 0u;
    float4 en4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (p < count * samples) {
        en4 = paths.energy[p];
        flags = __float_as_uint(en4.w);
    }
    const uint32_t work = flags & (F_PATH | F_RESOLVE | F_NEE0);
    f3 deduped_value = mk(0.0f, 0.0f, 0.0f);
    bool deduped_found = false;
    if constexpr (kDedup) {
        bool want = false;
        int kx = 0, ky = 0, kz = 0;
        uint32_t kface = 6u;
        if ((flags & F_PATH) && optics.cache_on) {
            const WfHit h = hit0[p];
            if ((h.status_normal & 255u) == 1u) {
                want = true;
                kx = h.vx; ky = h.vy; kz = h.vz;
                kface = face_index_of(normal_of(h.status_normal >> 8));
            }
        }
        deduped_found = wave_cached_energy(optics, want, kx, ky, kz, kface, deduped_value, stats);
    }
    if (work) {
        uint32_t face_index = p / samples, sample = p % samples;
        if (order) face_index = order[face_index];
        int fx, fy, fz;
        f3 face_n;
        uint32_t stream, sequence;
        bool valid;
        face_of_path(tasks, face_index, sample, sequence_base, fx, fy, fz, face_n, stream, sequence, valid);
        const uint8_t *v = vis + (size_t)p * 8u;
        f3 energy = xyz(en4);
        // The receiver's emitter sample (trace_receiver's first term).
        if (flags & F_NEE0) {
            f3 c0 = xyz(paths.nee0[p]);
            paths.nee0[p] = pack(v[1] ? c0 : mk(0.0f, 0.0f, 0.0f), 0.0f);
            flags &= ~F_NEE0;
        }
        // The last hit's deferred terms: direct light and the emitter (or
        // cached) term, added as trace_path adds them.
        if (flags & F_RESOLVE) {
            float4 h4 = paths.hit[p];
            f3 hp = xyz(h4), hn = normal_of(__float_as_uint(h4.w));
            f3 dv = mk(0.0f, 0.0f, 0.0f);
            uint32_t issued = (flags >> LIGHT_SHIFT) & 31u;
            for (uint32_t i = 0u; i < optics.light_count; ++i) {
                if (!((issued >> i) & 1u) || !v[2u + i]) continue;
                f3 dir;
                float dist, weight;
                light_ray(optics, i, hp, hn, dir, dist, weight);
                float4 en = optics.lights[i * 3u + 1u];
                dv = dv + mk(en.x, en.y, en.z) * weight;
            }
            uint32_t kind = (flags >> KIND_SHIFT) & 7u;
            f3 c = xyz(paths.candidate[p]);
            f3 nee = mk(0.0f, 0.0f, 0.0f);
            if (kind == K_NEE) nee = v[1] ? c : mk(0.0f, 0.0f, 0.0f);
            else if (kind == K_TERM_NEE) nee = (v[1] ? c : mk(0.0f, 0.0f, 0.0f)) * 2.0f;
            else if (kind == K_TERM_HIT) {
                // sample_terminal_emission (voxel_path.wgsl:368-394) past its ray.
                WfHit th = hit1[p];
                uint32_t bounce = (flags >> BOUNCE_SHIFT) & 15u;
                f3 direction = hemisphere(hn, continuation_sample(stream, sequence, bounce, 0u), continuation_sample(stream, sequence, bounce, 1u));
                float reach = connection_reach(optics);
                if ((th.status_normal & 255u) == 1u) {
                    float4 md = optics.materials[th.material * 2u];
                    float4 me = optics.materials[th.material * 2u + 1u];
                    float power = md.w;
                    f3 tpos = mk(th.px, th.py, th.pz), tn = normal_of(th.status_normal >> 8);
                    f3 delta = tpos - hp;
                    float d2 = dot3(delta, delta);
                    if (!(power == 0.0f || d2 < 0.0001f || d2 > reach * reach)) {
                        float selection_pdf = power / fmx(optics.total_weight, 1e-20f);
                        if (optics.has_tree) selection_pdf = tree_probability(optics, hp, hn, th.vx, th.vy, th.vz, tn);
                        float light_pdf = selection_pdf * d2 / fmx(dot3(tn, -direction), 1e-8f);
                        float directional_pdf = cosine_pdf(hn, direction);
                        float factor = 2.0f * directional_pdf / (directional_pdf + light_pdf);
                        nee = mk(me.x, me.y, me.z) * factor;
                    }
                }
            }
            f3 T = xyz(paths.pending_t[p]), emm = xyz(paths.pending_e[p]), diffuse = xyz(paths.pending_d[p]);
            if (kind == K_CACHED) energy = energy + T * (emm + diffuse * (dv + c));
            else energy = energy + T * (emm + diffuse * (dv + nee));
            flags &= ~(F_RESOLVE | (7u << KIND_SHIFT) | (31u << LIGHT_SHIFT));
        }
        // The new hit of the path (one iteration of trace_path's loop).
        if (flags & F_PATH) {
            flags &= ~F_PATH;
            WfHit h = hit0[p];
            uint32_t status = h.status_normal & 255u;
            uint32_t bounce = (flags >> BOUNCE_SHIFT) & 15u;
            float4 tp4 = paths.throughput[p];
            f3 throughput = xyz(tp4);
            float previous_pdf = tp4.w;
            f3 direction = xyz(paths.direction[p]);
            if (status == 0u) {
                float horizon = fclamp(direction.y * 0.35f + 0.65f, 0.15f, 1.0f);
                energy = energy + throughput * mk(optics.env_r, optics.env_g, optics.env_b) * horizon;
            } else if (status == 2u) {
                flags |= F_UNKNOWN;
            } else if (status == 1u) {
                f3 hp = mk(h.px, h.py, h.pz), hn = normal_of(h.status_normal >> 8);
                f3 origin = xyz(paths.origin[p]);
                float4 prev4 = paths.previous[p];
                f3 previous_position = xyz(prev4), previous_normal = normal_of(__float_as_uint(prev4.w));
                float4 md = optics.materials[h.material * 2u];
                float4 me = optics.materials[h.material * 2u + 1u];
                f3 diffuse = mk(md.x, md.y, md.z);
                f3 emission = mk(me.x, me.y, me.z);
                // direct_at's shadow rays.
                uint32_t issued = 0u;
                f3 lo = hp + hn * 0.002f;
                for (uint32_t i = 0u; i < optics.light_count; ++i) {
                    f3 dir;
                    float dist, weight;
                    if (!light_ray(optics, i, hp, hn, dir, dist, weight)) continue;
                    issued |= 1u << i;
                    e.add(lo, dir, dist, p * 8u + 2u + i);
                }
                float emission_weight = md.w;
                f3 delta = hp - origin;
                float selection_pdf = emission_weight / fmx(optics.total_weight, 1e-20f);
                if (emission_weight > 0.0f && optics.emitter_faces != 0u && optics.has_tree)
                    selection_pdf = tree_probability(optics, previous_position, previous_normal, h.vx, h.vy, h.vz, hn);
                float light_pdf = selection_pdf * dot3(delta, delta) / fmx(dot3(hn, -direction), 1e-8f);
                float mis = 1.0f;
                if (light_pdf > 0.0f) mis = previous_pdf / (previous_pdf + light_pdf);
                uint32_t kind = K_ZERO;
                f3 c = mk(0.0f, 0.0f, 0.0f);
                f3 cached = deduped_value;
                bool found = deduped_found;
                if constexpr (!kDedup) {
                    uint32_t cslot;
                    found = cached_energy(optics, h.vx, h.vy, h.vz, face_index_of(hn), cached, cslot);
                }
                bool last = bounce + 1u == optics.bounces;
                bool stop = found || last;
                if (found) {
                    kind = K_CACHED;
                    c = cached;
                } else if (last) {
                    if (optics.emitter_faces != 0u) {
                        if (continuation_sample(stream, sequence, bounce, 2u) < 0.5f) {
                            f3 so, sd;
                            float sdist;
                            if (emitter_candidate(optics, hp, hn, stream, sequence, bounce + 1u, c, so, sd, sdist)) {
                                kind = K_TERM_NEE;
                                e.add(so, sd, sdist, p * 8u + 1u);
                            }
                        } else {
                            f3 tdir = hemisphere(hn, continuation_sample(stream, sequence, bounce, 0u),
                                                 continuation_sample(stream, sequence, bounce, 1u));
                            if (dot3(hn, tdir) > 0.0f) {
                                kind = K_TERM_HIT;
                                e.add(hp + hn * 0.002f, tdir, connection_reach(optics), p * 8u + 7u);
                            }
                        }
                    }
                } else {
                    f3 so, sd;
                    float sdist;
                    if (emitter_candidate(optics, hp, hn, stream, sequence, bounce + 1u, c, so, sd, sdist)) {
                        kind = K_NEE;
                        e.add(so, sd, sdist, p * 8u + 1u);
                    }
                }
                paths.pending_t[p] = pack(throughput, 0.0f);
                paths.pending_e[p] = pack(emission * mis, 0.0f);
                paths.pending_d[p] = pack(diffuse, 0.0f);
                paths.hit[p] = pack_bits(hp, normal_code(hn));
                paths.candidate[p] = pack(c, 0.0f);
                flags |= F_RESOLVE | (kind << KIND_SHIFT) | (issued << LIGHT_SHIFT);
                if (!stop) {
                    // The path goes on: throughput, roulette, next direction.
                    throughput = throughput * diffuse;
                    float maximum = fmx(throughput.x, fmx(throughput.y, throughput.z));
                    bool alive = maximum != 0.0f;
                    if (alive && bounce >= 1u) {
                        float survival = fclamp(maximum, 0.1f, 0.95f);
                        if (continuation_sample(stream, sequence, bounce, 2u) > survival) alive = false;
                        else throughput = throughput / survival;
                    }
                    if (alive) {
                        f3 next_origin = hp + hn * 0.002f;
                        f3 next_direction = hemisphere(hn, continuation_sample(stream, sequence, bounce, 0u),
                                                       continuation_sample(stream, sequence, bounce, 1u));
                        float next_pdf = cosine_pdf(hn, next_direction);
                        if (next_pdf > 0.0f) {
                            paths.throughput[p] = pack(throughput, next_pdf);
                            paths.origin[p] = pack(next_origin, 0.0f);
                            paths.direction[p] = pack(next_direction, 0.0f);
                            paths.previous[p] = pack_bits(hp, normal_code(hn));
                            flags = (flags & ~(15u << BOUNCE_SHIFT)) | ((bounce + 1u) << BOUNCE_SHIFT) | F_PATH;
                            e.add(next_origin, next_direction, connection_reach(optics), p * 8u);
                        }
                    }
                }
            }
        }
        paths.energy[p] = pack_bits(energy, flags);
    }
    emit_all(e, queue, qcount);
}

// ------------------------------------------------------------------ final
// trace_face's tail (voxel_trace.wgsl:72-92 + write_result :25-37): the
// samples' values in the megakernel's lane order.
__global__ __launch_bounds__(256) void wf_final_kernel(SceneView scene, Optics optics, const uint4 *__restrict__ tasks, uint32_t count,
                                                       uint32_t samples, uint32_t lanes, WfPaths paths, float4 *__restrict__ results,
                                                       unsigned long long *__restrict__ rays_counter, const uint32_t *__restrict__ order) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;  // processing slot; its paths are k * samples + i
    RayTally tally;
    if (k < count) {
        uint32_t f = order ? order[k] : k;
        uint4 t0 = tasks[f * 2u];
        if (t0.w != NONE) {
            f3 tmean = mk(0.0f, 0.0f, 0.0f), tsecond = mk(0.0f, 0.0f, 0.0f);
            uint32_t tknown = 0u;
            for (uint32_t lane = 0u; lane < lanes; ++lane) {
                f3 mean = mk(0.0f, 0.0f, 0.0f), second = mk(0.0f, 0.0f, 0.0f);
                uint32_t known = 0u;
                for (uint32_t i = lane; i < samples; i += lanes) {
                    size_t p = (size_t)k * samples + i;
                    float4 en4 = paths.energy[p];
                    uint32_t flags = __float_as_uint(en4.w);
                    if (flags & F_UNKNOWN) continue;
                    f3 value = xyz(paths.nee0[p]);
                    if (flags & F_STARTED) value = value + xyz(en4);
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
            r[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (!optics.direct_reuse) {
                int vx = (int)t0.x, vy = (int)t0.y, vz = (int)t0.z;
                f3 normal = face_normal(t0.w);
                f3 dv, dc;
                direct_at(scene, optics, mk((float)vx, (float)vy, (float)vz) + mk(0.5f, 0.5f, 0.5f) + normal * 0.5f, normal, dv, dc,
                          tally);
                r[4] = make_float4(dv.x, dv.y, dv.z, 0.0f);
                r[5] = make_float4(0.0f, dc.x, dc.y, dc.z);
            }
        }
    }
    if (!rays_counter) return;
    unsigned long long n = tally.rays;
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if ((threadIdx.x & 63u) == 0u) atomicAdd(rays_counter, n);
}

__global__ void wf_reset_kernel(uint32_t *__restrict__ words, uint32_t n) {
    uint32_t i = threadIdx.x;
    if (i < n) words[i] = 0u;
}

// ------------------------------------------------------------------ launcher
static inline uint32_t wf_blocks(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

void launch_wavefront(const SceneView &scene, const Optics &optics, const uint4 *tasks, uint32_t count, uint32_t samples, uint32_t lanes,
                      uint32_t sequence_base, const WfBuffers &b, uint32_t trace_grid, float4 *results, unsigned long long *rays,
                      hipStream_t stream, bool dedup) {
    uint64_t n = (uint64_t)count * samples;
    if (n == 0) return;
    // control words: [0] queue count, [1] queue head
    hipLaunchKernelGGL(wf_reset_kernel, dim3(1), dim3(64), 0, stream, b.control, 2u);
    hipLaunchKernelGGL(wf_begin_kernel, dim3(wf_blocks(n, 256)), dim3(256), 0, stream, scene, optics, tasks, count, samples, sequence_base,
                       b.paths, b.queue, b.control, b.order);
    for (uint32_t it = 0; it <= optics.bounces; ++it) {
        hipLaunchKernelGGL(wf_trace_kernel, dim3(trace_grid), dim3(256), 0, stream, scene, optics, b.queue, b.control, b.control + 1,
                           b.hit0, b.hit1, b.vis, rays);
        hipLaunchKernelGGL(wf_reset_kernel, dim3(1), dim3(64), 0, stream, b.control, 2u);
        if (dedup)
            hipLaunchKernelGGL(wf_shade_kernel<true>, dim3(wf_blocks(n, 256)), dim3(256), 0, stream, scene, optics, tasks, count, samples,
                               sequence_base, b.paths, b.hit0, b.hit1, b.vis, b.queue, b.control, b.order, rays);
        else
            hipLaunchKernelGGL(wf_shade_kernel<false>, dim3(wf_blocks(n, 256)), dim3(256), 0, stream, scene, optics, tasks, count, samples,
                               sequence_base, b.paths, b.hit0, b.hit1, b.vis, b.queue, b.control, b.order, rays);
    }
    hipLaunchKernelGGL(wf_final_kernel, dim3(wf_blocks(count, 256)), dim3(256), 0, stream, scene, optics, tasks, count, samples, lanes,
                       b.paths, results, rays, b.order);
}

}  // namespace vgi
