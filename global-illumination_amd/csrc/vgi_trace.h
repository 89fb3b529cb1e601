// Voxel traversal and the path estimator on the MI355X scene layout.
//
// Scene layout in HBM (one allocation each, all read-only while tracing):
//   headers  uint4 per region cell of the loaded regions' bounding box, dense,
//            x fastest: (kind, material-or-table-offset, top-level negative
//            bytes, top-level positive bytes). kind 0 = not loaded, 1 empty,
//            2 uniform solid, 3 brick table. One load replaces the
//            reference's hashed directory probe + header load
//            (voxel_traversal.wgsl:32-41 + :170).
//   cells    u32: 64-word brick tables (word: extents<<2|0 air brick,
//            mat<<2|1 uniform brick, ptr<<2|2 mixed brick), then 32-word
//            mixed bricks of two 16-bit cells (mat<<1|1 solid, extents<<1
//            air) -- the reference's cell encoding (gpu_scene.rs:199-236).
// The traversal reads exactly the values the reference's directory/header
// path yields, so boxes, steps and hits are identical.
#pragma once
#include "vgi_math.h"

namespace vgi {

struct SceneView {
    const uint4 *__restrict__ headers;
    const uint32_t *__restrict__ cells;
    int minx, miny, minz;  // region bounds, inclusive
    int maxx, maxy, maxz;
    int sx, sy;            // region box size x, y
};

struct Optics {
    const float4 *__restrict__ materials;   // 2 per material: (diffuse rgb, emission power), (emission rgb, 0)
    const float4 *__restrict__ lights;      // 3 per light (gpu.rs:136-141)
    const uint4 *__restrict__ emitters;     // 2 uint4 per emissive face, then 4 per tree node
    const uint4 *__restrict__ tree;         // 2 uint4 per tree node: the walk's terms (centre, radius, power, share, children)
    uint32_t light_count;
    uint32_t directional;                   // params.maximum.w: the first `directional` lights are suns
    uint32_t emitter_faces;                 // params.extra.w
    float total_weight;                     // params.extra.z
    uint32_t has_tree;
    float env_r, env_g, env_b;
    float max_distance;                     // params.extra.y
    float reach;                            // params.transport.x
    uint32_t limited_reach;
    uint32_t bounces;                       // DIFFUSE_HITS
    // world cache (voxel_path.wgsl:51-87)
    const uint4 *__restrict__ cache_index;  // open-addressed (voxel, face<<29|slot), no header record
    const uint32_t *__restrict__ cache_state;  // 36 words per slot
    uint32_t cache_mask;                    // capacity - 1, 0 with no cache
    uint32_t cache_on;
    // 1: every valid face's results row already holds its direct terms (words
    // 16-23) for this scene, these lights and this reach, so the trace leaves
    // them in place instead of tracing the face's light rays again.
    uint32_t direct_reuse;
};

// One scene of a batched launch: its scene and optics views and its bound
// faces; `tiles` queue tiles of this scene end the scene when all are done.
struct BatchScene {
    SceneView scene;
    Optics optics;
    const uint4 *__restrict__ tasks;
    float4 *__restrict__ results;
    const uint32_t *__restrict__ order;  // optional: faces in processing order
    uint32_t count;
    uint32_t tiles;
};

struct RayResult {
    int vx, vy, vz;
    uint32_t material;
    f3 position;
    uint32_t status;  // 0 escaped, 1 hit, 2 unknown, 3 cut at reach
    f3 normal;
};

// Per-thread work counts. The tracing functions are templates on the tally
// type, so the production launch (NoTally) carries no counting at all, the
// timed frames count rays only (RayTally), and the parity/roofline launch
// counts rays, traversal steps and the dependent gathers of the traversal --
// region headers (16 B), brick-table words (4 B), mixed-brick cell words (4 B).
// The tally also carries the ray budget of the path tracer's `trace`
// (voxel_path.wgsl:6-14,239-243): 2048 traversal steps with a ray out of
// steps unknown (kBudgeted false), or the far tier's budget (BudgetTally):
// `limit` steps, unloaded regions crossed as empty space, and a ray out of
// steps a miss.
struct NoTally {
    static constexpr bool kBudgeted = false;
    __device__ uint32_t step_limit() const { return 2048u; }
    __device__ void ray() {}
    __device__ void step() {}
    __device__ void header() {}
    __device__ void brick() {}
    __device__ void cell() {}
    __device__ uint32_t ray_count() const { return 0u; }
};
struct RayTally : NoTally {
    uint32_t rays = 0u;
    __device__ void ray() { rays += 1u; }
    __device__ uint32_t ray_count() const { return rays; }
};
struct FullTally {
    static constexpr bool kBudgeted = false;
    __device__ uint32_t step_limit() const { return 2048u; }
    uint32_t rays = 0u, steps = 0u, headers = 0u, bricks = 0u, cells = 0u;
    __device__ void ray() { rays += 1u; }
    __device__ void step() { steps += 1u; }
    __device__ void header() { headers += 1u; }
    __device__ void brick() { bricks += 1u; }
    __device__ void cell() { cells += 1u; }
    __device__ uint32_t ray_count() const { return rays; }
};
struct BudgetTally : NoTally {
    static constexpr bool kBudgeted = true;
    uint32_t limit = 128u;
    __device__ uint32_t step_limit() const { return limit; }
};

__device__ __forceinline__ bool in_bounds(const SceneView &s, int rx, int ry, int rz) {
    return rx >= s.minx && ry >= s.miny && rz >= s.minz && rx <= s.maxx && ry <= s.maxy && rz <= s.maxz;
}
// The region's header, kind 0 when not loaded (callers check bounds first).
__device__ __forceinline__ uint4 region_header(const SceneView &s, int rx, int ry, int rz) {
    size_t i = (size_t)(rx - s.minx) + (size_t)s.sx * ((size_t)(ry - s.miny) + (size_t)s.sy * (size_t)(rz - s.minz));
    return s.headers[i];
}
// (int)floorf(x) as one instruction, the floor and the conversion together
// (v_cvt_flr_i32_f32; the same value for every finite x in int range, which
// every traversal position is).
__device__ __forceinline__ int floor_to_int(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// The loaded box's index of region (rx, ry, rz), NONE outside the box: one
// unsigned compare per axis and 32-bit arithmetic (the box holds fewer than
// 2^32 regions: its headers alone would fill 64 GiB).
__device__ __forceinline__ uint32_t region_index(const SceneView &s, int rx, int ry, int rz) {
    const uint32_t dx = (uint32_t)(rx - s.minx), dy = (uint32_t)(ry - s.miny), dz = (uint32_t)(rz - s.minz);
    if (dx >= (uint32_t)s.sx || dy >= (uint32_t)s.sy || dz > (uint32_t)(s.maxz - s.minz)) return NONE;
    return (dz * (uint32_t)s.sy + dy) * (uint32_t)s.sx + dx;
}

// voxel_traversal.wgsl:48-70 (voxel_material / voxel_cell): a voxel's
// material, AIR for air and for voxels outside the loaded regions.
__device__ __forceinline__ uint32_t voxel_material(const SceneView &s, int vx, int vy, int vz) {
    int rx = vx >> 4, ry = vy >> 4, rz = vz >> 4;
    if (!in_bounds(s, rx, ry, rz)) return AIR;
    uint4 h = region_header(s, rx, ry, rz);
    if (h.x == 0u || h.x == 1u) return AIR;
    if (h.x != 3u) return h.y;
    uint32_t lx = (uint32_t)(vx & 15), ly = (uint32_t)(vy & 15), lz = (uint32_t)(vz & 15);
    uint32_t word = s.cells[h.y + ((lz >> 2) * 4u + (ly >> 2)) * 4u + (lx >> 2)];
    if ((word & 3u) == 0u) return AIR;
    if ((word & 3u) == 1u) return word >> 2u;
    uint32_t idx = ((lz & 3u) * 4u + (ly & 3u)) * 4u + (lx & 3u);
    uint32_t cell = (s.cells[(word >> 2u) + (idx >> 1)] >> ((idx & 1u) * 16u)) & 65535u;
    return (cell & 1u) != 0u ? cell >> 1u : AIR;
}

// n / d rounded to binary32 as IEEE division rounds it, from r = RN64(1/d)
// (binary64, correctly rounded, taken once per ray): RN32(RN64(n * r)).
// Why it is exact: RN64(n * r) = (n/d)(1 + e) with |e| <= 2^-52 (1 + 2^-53).
// A binary32 quotient n/d is never a midpoint between two binary32 numbers
// (a 25-bit odd significand times d's significand has more than 24
// significant bits unless d is a power of two, and then n/d is exact), and
// when it is not exact it lies at least 2^-49 (relative) from every
// midpoint: with n = N 2^a, d = D 2^b, midpoint M 2^c (N, D < 2^24,
// M < 2^25), n - d m is a nonzero multiple of 2^(b+c) and d m < 2^(49+b+c).
// So both roundings of the binary64 product land on the binary32 value of
// n/d (the final rounding is to nearest even, as IEEE division's). The
// quotients here stay far from binary32 overflow and subnormals (|n| <
// 2^20, 1e-8 < |d| <= 1). It costs two conversions and one binary64
// multiply instead of the ~10-instruction correctly rounded binary32 divide.
// The reciprocal need not be the correctly rounded one: the argument holds
// for any r within a few binary64 ulps of 1/d (the error budget is 2^-49
// against the product's ~2^-52), so r is the hardware's reciprocal estimate
// refined by two Newton steps (each squares the relative error; the last
// rounding leaves r within half an ulp plus 2^-100 of 1/d) instead of the
// full binary64 division sequence with its scaling and fix-up.
__device__ __forceinline__ double recip64(float d) {
    const double x = (double)d;
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ float div_by(float n, double r) { return (float)((double)n * r); }
// a / s for the three components through one reciprocal (s a normal binary32
// number, the quotients normal or zero).
__device__ __forceinline__ f3 div3(f3 a, float s) {
    const double r = recip64(s);
    return mk(div_by(a.x, r), div_by(a.y, r), div_by(a.z, r));
}

// The reciprocals a ray's box steps divide by: 1/d for an axis the ray moves
// along (|d| > 1e-8), 1 otherwise (voxel_traversal.wgsl:94-99 divides by 1).
struct RayInverse {
    double x, y, z;
};
__device__ __forceinline__ RayInverse ray_inverse(f3 direction) {
    return RayInverse{fabsf(direction.x) > 1e-8f ? recip64(direction.x) : 1.0,
                      fabsf(direction.y) > 1e-8f ? recip64(direction.y) : 1.0,
                      fabsf(direction.z) > 1e-8f ? recip64(direction.z) : 1.0};
}

// voxel_traversal.wgsl:90-107
__device__ __forceinline__ void advance_box(f3 position, f3 direction, const RayInverse &inv, int cx, int cy, int cz, int nx, int ny,
                                            int nz, int px, int py, int pz, float size, f3 &normal_out, float &dist_out) {
    float mnx = (float)(cx - nx) * size, mny = (float)(cy - ny) * size, mnz = (float)(cz - nz) * size;
    float mxx = (float)(cx + px + 1) * size, mxy = (float)(cy + py + 1) * size, mxz = (float)(cz + pz + 1) * size;
    float tr[3];
    const float dv[3] = {direction.x, direction.y, direction.z};
    const float pv[3] = {position.x, position.y, position.z};
    const double iv[3] = {inv.x, inv.y, inv.z};
    const float lo[3] = {mnx, mny, mnz}, hi[3] = {mxx, mxy, mxz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        bool moving = fabsf(dv[a]) > 1e-8f;
        float plane = dv[a] > 0.0f ? hi[a] : lo[a];
        float raw = div_by(plane - pv[a], iv[a]);  // (plane - p) / (moving ? d : 1)
        tr[a] = (moving && raw >= -0.001f && raw < 1e30f) ? raw : 1e30f;
    }
    bool choose_y = tr[1] < fmx(tr[0], 0.0f);
    float distance_xy = fmx(choose_y ? tr[1] : tr[0], 0.0f);
    bool choose_z = tr[2] < distance_xy;
    float distance = fmx(choose_z ? tr[2] : distance_xy, 0.0f);
    f3 axis = choose_z ? mk(0.0f, 0.0f, 1.0f) : (choose_y ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f));
    f3 sgn = mk(direction.x > 0.0f ? -1.0f : 1.0f, direction.y > 0.0f ? -1.0f : 1.0f, direction.z > 0.0f ? -1.0f : 1.0f);
    normal_out = distance < 1e30f ? axis * sgn : mk(0.0f, 0.0f, 0.0f);
    dist_out = fmx(distance + 0.001f, 0.001f);
}

// advance_box for a box given by its cell and packed extents (`field` bits
// per axis, `shift` apart; the negative extents in `neg`, the positive ones
// in `pos`). Per axis only the face the ray leaves through matters -- the
// positive side when it moves up the axis, else the negative side -- so the
// extent is taken from that word before it is unpacked. The planes, the
// divisions and the choice are advance_box's, operation for operation.
__device__ __forceinline__ void advance_packed(f3 position, f3 direction, const RayInverse &inv, int cx, int cy, int cz, uint32_t neg,
                                               uint32_t pos, uint32_t shift, uint32_t field, float size, f3 &normal_out,
                                               float &dist_out) {
    const float dv[3] = {direction.x, direction.y, direction.z};
    const float pv[3] = {position.x, position.y, position.z};
    const double iv[3] = {inv.x, inv.y, inv.z};
    const int cv[3] = {cx, cy, cz};
    float tr[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool up = dv[a] > 0.0f;
        const int e = (int)(((up ? pos : neg) >> (shift * (uint32_t)a)) & field);
        const float plane = (float)(up ? cv[a] + e + 1 : cv[a] - e) * size;
        const bool moving = fabsf(dv[a]) > 1e-8f;
        const float raw = div_by(plane - pv[a], iv[a]);  // (plane - p) / d
        tr[a] = (moving && raw >= -0.001f && raw < 1e30f) ? raw : 1e30f;
    }
    bool choose_y = tr[1] < fmx(tr[0], 0.0f);
    float distance_xy = fmx(choose_y ? tr[1] : tr[0], 0.0f);
    bool choose_z = tr[2] < distance_xy;
    float distance = fmx(choose_z ? tr[2] : distance_xy, 0.0f);
    f3 axis = choose_z ? mk(0.0f, 0.0f, 1.0f) : (choose_y ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f));
    f3 sgn = mk(direction.x > 0.0f ? -1.0f : 1.0f, direction.y > 0.0f ? -1.0f : 1.0f, direction.z > 0.0f ? -1.0f : 1.0f);
    normal_out = distance < 1e30f ? axis * sgn : mk(0.0f, 0.0f, 0.0f);
    dist_out = fmx(distance + 0.001f, 0.001f);
}

// voxel_traversal.wgsl:108-121
__device__ __forceinline__ bool geometry_ahead(const SceneView &s, f3 start, f3 direction, float max_distance) {
    f3 position = start;
    float remaining = max_distance;
    const RayInverse inv = ray_inverse(direction);
    for (uint32_t step = 0; step < 2048u; ++step) {
        int rx = (int)floorf(position.x / 16.0f), ry = (int)floorf(position.y / 16.0f), rz = (int)floorf(position.z / 16.0f);
        if (remaining < 0.0f || !in_bounds(s, rx, ry, rz)) return false;
        uint4 h = region_header(s, rx, ry, rz);
        if (h.x != 0u && h.x != 1u) return true;
        f3 n;
        float travel;
        advance_box(position, direction, inv, rx, ry, rz, 0, 0, 0, 0, 0, 0, 16.0f, n, travel);
        position = position + direction * travel;
        remaining = remaining - travel;
    }
    return true;
}

// voxel_traversal.wgsl:126-219 with visible_geometry = false, max_steps = 2048
// (the path tracer's `trace`, voxel_path.wgsl:239-243; a budgeted tally
// changes both, as the far tier's pipeline constants do), split into a start
// and single steps so that a lane can take up a new ray whenever its current
// one ends (the wavefront tracer); trace_ray runs one ray to its end.
struct RayWalk {
    f3 origin, direction;
    RayInverse inv;
    float max_distance;
    f3 position;
    float distance;
    f3 normal;
    uint32_t last_region;  // loaded-box index of the region `meta` belongs to (NONE: none)
    uint4 meta;
    uint32_t step;
};

// Clip to the scene box. Returns true when the ray already has its result.
template <class Tally>
__device__ __forceinline__ bool ray_start(const SceneView &s, f3 origin, f3 direction, float max_distance, RayWalk &w,
                                          RayResult &result, Tally &tally) {
    result.vx = result.vy = result.vz = 0;
    result.material = 0u;
    result.position = mk(0.0f, 0.0f, 0.0f);
    result.status = 0u;
    result.normal = mk(0.0f, 0.0f, 0.0f);
    tally.ray();
    float near_ = -1e30f, far_ = 1e30f;
    f3 normal = mk(0.0f, 0.0f, 0.0f);
    const RayInverse inv = ray_inverse(direction);
    const double iv[3] = {inv.x, inv.y, inv.z};
    const float dv[3] = {direction.x, direction.y, direction.z};
    const float ov[3] = {origin.x, origin.y, origin.z};
    const float lo[3] = {(float)s.minx * 16.0f, (float)s.miny * 16.0f, (float)s.minz * 16.0f};
    const float hi[3] = {(float)(s.maxx + 1) * 16.0f, (float)(s.maxy + 1) * 16.0f, (float)(s.maxz + 1) * 16.0f};
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        if (fabsf(dv[axis]) <= 1e-8f) {
            if (ov[axis] < lo[axis] || ov[axis] > hi[axis]) return true;
            continue;
        }
        float first = div_by(lo[axis] - ov[axis], iv[axis]);   // (lo - o) / d
        float second = div_by(hi[axis] - ov[axis], iv[axis]);  // (hi - o) / d
        if (fmn(first, second) > near_) {
            near_ = fmn(first, second);
            float sg = first < second ? -1.0f : 1.0f;
            normal = mk(axis == 0 ? sg : 0.0f, axis == 1 ? sg : 0.0f, axis == 2 ? sg : 0.0f);
        }
        far_ = fmn(far_, fmx(first, second));
    }
    if (far_ < fmx(near_, 0.0f)) return true;
    w.origin = origin;
    w.direction = direction;
    w.inv = inv;
    w.max_distance = max_distance;
    w.distance = fmx(near_, 0.0f) + 0.001f;
    w.position = origin + direction * w.distance;
    w.normal = normal;
    w.last_region = NONE;
    w.meta = make_uint4(0u, 0u, 0u, 0u);
    w.step = 0u;
    return false;
}

// One step of the walk. Returns true when the ray has ended, with its
// `status` (0 escaped or beyond its distance, 1 hit, 2 unknown) and a hit's
// material (NONE when the step limit ended the ray, else 0 for a miss). A hit
// leaves the walk where it hit (w.position, w.normal).
// kVisible (voxel_traversal.wgsl:160-166, visible_geometry = true): unloaded
// regions are walked through instead of ending the ray (change footprints
// look for the real occluders beyond gaps in the loaded set).
template <class Tally, bool kVisible = false>
__device__ __forceinline__ bool ray_step(const SceneView &s, RayWalk &w, uint32_t &status, uint32_t &hit_material, Tally &tally) {
    if (w.step >= tally.step_limit()) {
        status = 2u;
        hit_material = NONE;
        return true;
    }
    status = 0u;
    hit_material = 0u;
    w.step += 1u;
    tally.step();
    const f3 position = w.position, direction = w.direction;
    if (w.distance > w.max_distance) return true;
    int vx = floor_to_int(position.x), vy = floor_to_int(position.y), vz = floor_to_int(position.z);
    int rx = vx >> 4, ry = vy >> 4, rz = vz >> 4;
    const uint32_t region = region_index(s, rx, ry, rz);
    if (region == NONE) return true;
    if (region != w.last_region) {
        uint4 h = s.headers[region];
        // One 16-byte load: the kind decides the branch, the rest is used on
        // most paths; without this the compiler loads the kind first and
        // the other words after the branch, a second round trip.
        asm volatile("" : "+v"(h.y), "+v"(h.z), "+v"(h.w));
        tally.header();
        if (h.x == 0u) {
            if constexpr (kVisible || Tally::kBudgeted) {
                f3 n;
                float skip;
                advance_box(position, direction, w.inv, rx, ry, rz, 0, 0, 0, 0, 0, 0, 16.0f, n, skip);
                w.distance = w.distance + skip;
                w.position = position + direction * skip;
                w.normal = n;
                return false;
            }
            status = geometry_ahead(s, position, direction, w.max_distance - w.distance) ? 2u : 0u;
            return true;
        }
        w.meta = h;
        w.last_region = region;
    }
    const uint4 meta = w.meta;
    // The empty box to skip -- the region's top-level box (kind 1), an air
    // brick's box, or an air voxel's box -- found per lane, then one
    // advance_box for every lane (lanes at different levels share it).
    int cx = rx, cy = ry, cz = rz;
    uint32_t neg = meta.z, pos = meta.w;  // region level: 8 bits per axis
    uint32_t shift = 8u, field = 255u;
    float size = 16.0f;
    bool skip = meta.x == 1u;
    if (!skip) {
        uint32_t material = meta.y;
        if (meta.x == 3u) {
            uint32_t lx = (uint32_t)(vx & 15), ly = (uint32_t)(vy & 15), lz = (uint32_t)(vz & 15);
            uint32_t bx = lx >> 2, by = ly >> 2, bz = lz >> 2;
            const uint32_t brick = s.cells[meta.y + (bz * 4u + by) * 4u + bx];
            tally.brick();
            uint32_t st = brick & 3u;
            if (st == 0u) {
                uint32_t e = brick >> 2u;
                cx = rx * 4 + (int)bx; cy = ry * 4 + (int)by; cz = rz * 4 + (int)bz;
                neg = e; pos = e >> 2;  // 2 bits per axis, 4 apart
                shift = 4u; field = 3u;
                size = 4.0f;
                skip = true;
                material = AIR;
            } else if (st == 1u) {
                material = brick >> 2u;
            } else {
                uint32_t idx = ((lz & 3u) * 4u + (ly & 3u)) * 4u + (lx & 3u);
                uint32_t word = s.cells[(brick >> 2u) + (idx >> 1)];
                tally.cell();
                uint32_t cell = (word >> ((idx & 1u) * 16u)) & 65535u;
                material = (cell & 1u) != 0u ? cell >> 1u : AIR;
                if (material == AIR) {
                    uint32_t e = cell >> 1u;
                    cx = vx; cy = vy; cz = vz;
                    neg = e; pos = e >> 2;
                    shift = 4u; field = 3u;
                    size = 1.0f;
                    skip = true;
                }
            }
        }
        if (material != AIR) {
            status = 1u;
            hit_material = material;
            return true;
        }
    }
    f3 adv_n = mk(0.0f, 0.0f, 0.0f);
    float adv_w = 0.0f;
    if (skip) advance_packed(position, direction, w.inv, cx, cy, cz, neg, pos, shift, field, size, adv_n, adv_w);
    w.distance = w.distance + adv_w;
    w.position = position + direction * adv_w;
    w.normal = adv_n;
    return false;
}

// The ended ray's record as ray_start initialised it and ray_step ended it.
__device__ __forceinline__ void ray_record(const RayWalk &w, uint32_t status, uint32_t hit_material, RayResult &result) {
    result.status = status;
    if (status == 1u) {
        result.vx = (int)floorf(w.position.x);
        result.vy = (int)floorf(w.position.y);
        result.vz = (int)floorf(w.position.z);
        result.material = hit_material;
        result.position = w.position;
        result.normal = w.normal;
    } else if (hit_material == NONE) {
        result.material = NONE;
    }
}

// One step of the walk. Returns true when the ray has its result.
template <class Tally, bool kVisible = false>
__device__ __forceinline__ bool ray_advance(const SceneView &s, RayWalk &w, RayResult &result, Tally &tally) {
    uint32_t status, hit_material;
    if (!ray_step<Tally, kVisible>(s, w, status, hit_material, tally)) return false;
    ray_record(w, status, hit_material, result);
    return true;
}

template <class Tally, bool kVisible = false>
__device__ __forceinline__ RayResult trace_ray(const SceneView &s, f3 origin, f3 direction, float max_distance, Tally &tally) {
    RayResult result;
    RayWalk w;
    if (ray_start(s, origin, direction, max_distance, w, result, tally)) return result;
    while (!ray_advance<Tally, kVisible>(s, w, result, tally)) {
    }
    // A budgeted `trace` counts a ray out of steps as a miss (voxel_path.wgsl:241).
    if constexpr (Tally::kBudgeted && !kVisible)
        if (result.status == 2u && result.material == NONE) result.status = 0u;
    return result;
}

// ---------------------------------------------------------------- estimator (voxel_path.wgsl)
struct Emitter {
    int vx, vy, vz;
    uint32_t face;
    float cumulative, probability, material, tree;
};
__device__ __forceinline__ Emitter load_emitter(const Optics &o, uint32_t index) {
    uint4 a = o.emitters[index * 2u];
    uint4 b = o.emitters[index * 2u + 1u];
    return Emitter{(int)a.x, (int)a.y, (int)a.z, a.w, __uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z),
                   __uint_as_float(b.w)};
}
// A tree node as the walk reads it: the per-node terms of
// voxel_path.wgsl:108-131 that depend on the node alone, rounded once when
// the table is built ((min + max) * 0.5, len(max - min) * 0.5 and the left
// child's share of the power), so each visit computes only what depends on
// the shading point.
struct WalkNode {
    f3 centre;
    float radius, power, share;
    uint32_t left, right, normals;
};
__device__ __forceinline__ WalkNode load_walk(const Optics &o, uint32_t index) {
    uint4 a = o.tree[index * 2u], b = o.tree[index * 2u + 1u];
    WalkNode n;
    n.centre = mk(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
    n.radius = __uint_as_float(a.w);
    n.power = __uint_as_float(b.x);
    n.share = __uint_as_float(b.y);
    n.left = b.z;
    n.right = b.w & 0x3ffffffu;
    n.normals = b.w >> 26u;
    return n;
}
// voxel_path.wgsl:108-122. The four quotients by the distance go through
// one binary64 reciprocal, RN32(RN64(n * RN64(1/d))), which equals binary32
// division RN32(n/d) for every binary32 n and d whose quotient is normal
// (the division check in tests/host covers it).
__device__ __forceinline__ float node_importance(const WalkNode &node, f3 position, f3 normal) {
    f3 delta = node.centre - position;
    float distance = fmx(len3(delta), 1e-5f);
    double inv = recip64(distance);
    f3 direction = mk(div_by(delta.x, inv), div_by(delta.y, inv), div_by(delta.z, inv));
    float spread = fmn(div_by(node.radius, inv), 1.0f);
    float receiver = fclamp(dot3(normal, direction) + spread, 0.0f, 1.0f);
    float ppx = (node.normals & 1u) ? -direction.x : -1.0f, ppy = (node.normals & 4u) ? -direction.y : -1.0f,
          ppz = (node.normals & 16u) ? -direction.z : -1.0f;
    float pnx = (node.normals & 2u) ? direction.x : -1.0f, pny = (node.normals & 8u) ? direction.y : -1.0f,
          pnz = (node.normals & 32u) ? direction.z : -1.0f;
    float qx = fmx(ppx, pnx), qy = fmx(ppy, pny), qz = fmx(ppz, pnz);
    float emitter = fclamp(fmx(fmx(qx, qy), qz) + spread, 0.0f, 1.0f);
    return node.power * receiver * emitter / (distance * distance + node.radius * node.radius + 0.25f);
}
// voxel_path.wgsl:123-131
__device__ __forceinline__ float node_left_probability(float share, const WalkNode &left, const WalkNode &right, f3 position,
                                                       f3 normal) {
    float a = node_importance(left, position, normal);
    float b = node_importance(right, position, normal);
    if (a + b <= 1e-20f) return fclamp(share, 1e-6f, 1.0f - 1e-6f);
    return fclamp(0.05f * share + 0.95f * a / (a + b), 1e-6f, 1.0f - 1e-6f);
}
// voxel_path.wgsl:132-149. Between levels the walk keeps only what the
// next level needs of the child it took: its children and its power share.
__device__ __forceinline__ Emitter tree_emitter(const Optics &o, f3 position, f3 normal, float choice) {
    float uniform = choice;
    float probability = 1.0f;
    WalkNode root = load_walk(o, 0u);
    uint32_t left_index = root.left, right_index = root.right;
    float node_share = root.share;
    for (;;) {
        if (left_index == NONE) {
            Emitter e = load_emitter(o, right_index);
            e.probability = probability;
            return e;
        }
        WalkNode left = load_walk(o, left_index), right = load_walk(o, right_index);
        float p = node_left_probability(node_share, left, right, position, normal);
        // The branch taken rescales the choice: uniform / p on the left,
        // (uniform - p) / (1 - p) on the right, as one division of the
        // selected operands (the same binary32 operations, no divergence).
        bool go_left = uniform < p;
        float share = go_left ? p : 1.0f - p;
        float offset = go_left ? uniform : uniform - p;
        probability = probability * share;
        uniform = fmn(offset / share, __uint_as_float(0x3f7fffffu));
        left_index = go_left ? left.left : right.left;
        right_index = go_left ? left.right : right.right;
        node_share = go_left ? left.share : right.share;
    }
}
// voxel_path.wgsl:158-174 (key_less :150-157); the split key (pivot voxel,
// face, axis) comes from the node's emitter-table words.
__device__ __forceinline__ float tree_probability(const Optics &o, f3 position, f3 normal, int vx, int vy, int vz, f3 hit_normal) {
    uint32_t face = 0u;
    for (uint32_t i = 0u; i < 6u; ++i) {
        f3 fn = face_normal(i);
        if (fn.x == hit_normal.x && fn.y == hit_normal.y && fn.z == hit_normal.z) face = i;
    }
    uint32_t index = 0u;
    float probability = 1.0f;
    WalkNode root = load_walk(o, 0u);
    uint32_t left_index = root.left, right_index = root.right;
    float node_share = root.share;
    for (;;) {
        if (left_index == NONE) {
            Emitter f = load_emitter(o, right_index);
            return (f.vx == vx && f.vy == vy && f.vz == vz && f.face == face) ? probability : 0.0f;
        }
        WalkNode left = load_walk(o, left_index), right = load_walk(o, right_index);
        float p = node_left_probability(node_share, left, right, position, normal);
        uint32_t off = o.emitter_faces * 2u + index * 4u;
        uint4 key = o.emitters[off + 2u];
        uint32_t axis = o.emitters[off + 3u].z;
        const int v[3] = {vx, vy, vz};
        const int pv[3] = {(int)key.x, (int)key.y, (int)key.z};
        bool less = face < key.w;
        for (uint32_t j = 0u; j < 3u; ++j) {
            uint32_t i = (axis + j) % 3u;
            if (v[i] != pv[i]) { less = v[i] < pv[i]; break; }
        }
        probability = probability * (less ? p : 1.0f - p);
        index = less ? left_index : right_index;
        node_share = less ? left.share : right.share;
        uint32_t next_left = less ? left.left : right.left;
        right_index = less ? left.right : right.right;
        left_index = next_left;
    }
}
// voxel_path.wgsl:309-324
__device__ __forceinline__ Emitter choose_emitter(const Optics &o, f3 position, f3 normal, float choice) {
    if (o.has_tree) return tree_emitter(o, position, normal, choice);
    uint32_t low = 0u, high = o.emitter_faces;
    while (low < high) {
        uint32_t middle = (low + high) / 2u;
        if (__uint_as_float(o.emitters[middle * 2u + 1u].x) <= choice) low = middle + 1u;
        else high = middle;
    }
    return load_emitter(o, min(low, o.emitter_faces - 1u));
}
__device__ __forceinline__ float connection_reach(const Optics &o) { return o.limited_reach ? o.reach : o.max_distance; }

// voxel_path.wgsl:332-364
template <class Tally>
__device__ __forceinline__ f3 sample_emitter(const SceneView &s, const Optics &o, f3 position, f3 normal, uint32_t source_stream,
                                          uint32_t sequence, uint32_t bounce, Tally &tally) {
    f3 zero = mk(0.0f, 0.0f, 0.0f);
    if (o.emitter_faces == 0u) return zero;
    uint32_t stream = source_stream ^ pcg(bounce + 0xa341316cu);
    Emitter e = choose_emitter(o, position, normal, sample_1d(stream, sequence, 12u));
    f3 en = face_normal(e.face);
    uint32_t axis = e.face / 2u;
    f3 tangent = axis == 0u ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    f3 bitangent = axis == 2u ? mk(0.0f, 1.0f, 0.0f) : mk(0.0f, 0.0f, 1.0f);
    float u13 = sample_1d(stream, sequence, 13u);
    float u14 = sample_1d(stream, sequence, 14u);
    f3 light_position = mk((float)e.vx, (float)e.vy, (float)e.vz) + mk(0.5f, 0.5f, 0.5f) + en * 0.5f +
                        tangent * (u13 - 0.5f) + bitangent * (u14 - 0.5f);
    f3 delta = light_position - position;
    float d2 = dot3(delta, delta);
    float reach = connection_reach(o);
    if (d2 < 0.0001f || d2 > reach * reach) return zero;
    float distance = fsqrt(d2);
    f3 direction = div3(delta, distance);
    float receiver_cosine = fmx(dot3(normal, direction), 0.0f);
    float emitter_cosine = fmx(dot3(en, -direction), 0.0f);
    if (receiver_cosine <= 0.0f || emitter_cosine <= 0.0f || e.probability <= 0.0f) return zero;
    float light_pdf = e.probability * d2 / emitter_cosine;
    float directional_pdf = receiver_cosine / 3.14159265359f;
    f3 shadow_origin = position + normal * 0.002f;
    f3 shadow_delta = light_position + en * 0.002f - shadow_origin;
    float shadow_distance = len3(shadow_delta);
    if (trace_ray(s, shadow_origin, div3(shadow_delta, shadow_distance), shadow_distance, tally).status != 0u) return zero;
    float4 em = o.materials[(uint32_t)e.material * 2u + 1u];
    float factor = directional_pdf / (light_pdf + directional_pdf);
    return mk(em.x, em.y, em.z) * factor;
}

// voxel_path.wgsl:283-307
template <class Tally>
__device__ __forceinline__ void direct_at(const SceneView &s, const Optics &o, f3 position, f3 normal, f3 &value, f3 &custom,
                                       Tally &tally) {
    value = mk(0.0f, 0.0f, 0.0f);
    custom = mk(0.0f, 0.0f, 0.0f);
    f3 origin = position + normal * 0.002f;
    for (uint32_t i = 0u; i < o.light_count; ++i) {
        float4 pr = o.lights[i * 3u], en = o.lights[i * 3u + 1u], ex = o.lights[i * 3u + 2u];
        f3 lp = mk(pr.x, pr.y, pr.z);
        f3 direction = normalize3(lp);
        float distance = o.max_distance;
        float attenuation = 1.0f;
        if (pr.w >= 0.0f) {
            f3 delta = lp - position;
            float d2 = dot3(delta, delta);
            if (d2 <= 0.0000001f || d2 > pr.w * pr.w) continue;
            distance = fsqrt(d2);
            direction = div3(delta, distance);
            distance = distance - 0.004f;
            attenuation = 1.0f / fmx(d2, 1.0f);
        }
        float cosine = fmx(dot3(normal, direction), 0.0f);
        if (cosine > 0.0f && trace_ray(s, origin, direction, distance, tally).status == 0u) {
            float sc = cosine * attenuation;
            value = value + mk(en.x, en.y, en.z) * sc;
            custom = custom + mk(ex.y, ex.z, ex.w) * sc;
        }
    }
}

// voxel_path.wgsl:66-87: a cached face ends the path once it holds 16 samples
// and no pending-change bit.
__device__ __forceinline__ bool cached_energy(const Optics &o, int vx, int vy, int vz, uint32_t face, f3 &value, uint32_t &slot_out) {
    slot_out = NONE;
    if (!o.cache_on || face > 5u) return false;
    uint32_t mask = o.cache_mask;
    uint32_t slot = (hash_voxel(vx, vy, vz) ^ pcg(face + 1u)) & mask;
    for (uint32_t attempt = 0u; attempt < 64u; ++attempt) {
        uint4 e = o.cache_index[slot];
        if (e.w == NONE) return false;
        if ((int)e.x == vx && (int)e.y == vy && (int)e.z == vz && (e.w >> 29u) == face) {
            uint32_t s = e.w & 0x1fffffffu;
            slot_out = s;
            const uint32_t *st = o.cache_state + (size_t)s * 36u;
            uint32_t count = st[5];
            if (count < 16u || (count & 0x80000000u) != 0u) return false;
            value = mk(__uint_as_float(st[0]), __uint_as_float(st[1]), __uint_as_float(st[2]));
            return true;
        }
        slot = (slot + 1u) & mask;
    }
    return false;
}

// voxel_path.wgsl:271-282
template <class Tally>
__device__ __forceinline__ RayResult trace_diffuse(const SceneView &s, const Optics &o, f3 origin, f3 direction, Tally &tally) {
    float reach = connection_reach(o);
    RayResult r = trace_ray(s, origin, direction, reach, tally);
    if (o.limited_reach && r.status == 0u &&
        geometry_ahead(s, origin + direction * reach, direction, o.max_distance - reach))
        r.status = 3u;
    return r;
}

// voxel_path.wgsl:368-394
template <class Tally>
__device__ __forceinline__ f3 sample_terminal_emission(const SceneView &s, const Optics &o, f3 position, f3 normal, uint32_t stream,
                                                    uint32_t sequence, uint32_t bounce, Tally &tally) {
    f3 zero = mk(0.0f, 0.0f, 0.0f);
    if (o.emitter_faces == 0u) return zero;
    if (continuation_sample(stream, sequence, bounce, 2u) < 0.5f) {
        f3 r = sample_emitter(s, o, position, normal, stream, sequence, bounce + 1u, tally);
        return r * 2.0f;
    }
    f3 direction = hemisphere(normal, continuation_sample(stream, sequence, bounce, 0u),
                              continuation_sample(stream, sequence, bounce, 1u));
    if (dot3(normal, direction) <= 0.0f) return zero;
    float reach = connection_reach(o);
    RayResult hit = trace_ray(s, position + normal * 0.002f, direction, reach, tally);
    if (hit.status != 1u) return zero;
    float4 md = o.materials[hit.material * 2u];
    float4 me = o.materials[hit.material * 2u + 1u];
    float power = md.w;
    f3 delta = hit.position - position;
    float d2 = dot3(delta, delta);
    if (power == 0.0f || d2 < 0.0001f || d2 > reach * reach) return zero;
    float selection_pdf = power / fmx(o.total_weight, 1e-20f);
    if (o.has_tree) selection_pdf = tree_probability(o, position, normal, hit.vx, hit.vy, hit.vz, hit.normal);
    float light_pdf = selection_pdf * d2 / fmx(dot3(hit.normal, -direction), 1e-8f);
    float directional_pdf = cosine_pdf(normal, direction);
    float factor = 2.0f * directional_pdf / (directional_pdf + light_pdf);
    return mk(me.x, me.y, me.z) * factor;
}

// ---------------------------------------------------------------- estimator pieces for split tracers
// A hit normal as 6 bits (per axis: nonzero, sign bit) and back; the round
// trip keeps every value advance_box and ray_start produce, signed zeros
// included.
__device__ __forceinline__ uint32_t normal_code(f3 n) {
    auto c = [](float v) { return (v != 0.0f ? 1u : 0u) | ((__float_as_uint(v) >> 31) << 1); };
    return c(n.x) | (c(n.y) << 2) | (c(n.z) << 4);
}
__device__ __forceinline__ f3 normal_of(uint32_t code) {
    auto v = [](uint32_t c) {
        float x = (c & 1u) ? 1.0f : 0.0f;
        return (c & 2u) ? -x : x;
    };
    return mk(v(code & 3u), v((code >> 2) & 3u), v((code >> 4) & 3u));
}
// The emitter sample of sample_emitter (voxel_path.wgsl:332-364) up to its
// shadow ray: false when the sample contributes nothing; else the candidate
// value (counted if the shadow ray is unoccluded) and the shadow ray.
__device__ __forceinline__ bool emitter_candidate(const Optics &o, f3 position, f3 normal, uint32_t source_stream, uint32_t sequence,
                                                  uint32_t bounce, f3 &candidate, f3 &origin, f3 &direction, float &distance) {
    if (o.emitter_faces == 0u) return false;
    uint32_t stream = source_stream ^ pcg(bounce + 0xa341316cu);
    Emitter e = choose_emitter(o, position, normal, sample_1d(stream, sequence, 12u));
    f3 en = face_normal(e.face);
    uint32_t axis = e.face / 2u;
    f3 tangent = axis == 0u ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    f3 bitangent = axis == 2u ? mk(0.0f, 1.0f, 0.0f) : mk(0.0f, 0.0f, 1.0f);
    float u13 = sample_1d(stream, sequence, 13u);
    float u14 = sample_1d(stream, sequence, 14u);
    f3 light_position = mk((float)e.vx, (float)e.vy, (float)e.vz) + mk(0.5f, 0.5f, 0.5f) + en * 0.5f + tangent * (u13 - 0.5f) +
                        bitangent * (u14 - 0.5f);
    f3 delta = light_position - position;
    float d2 = dot3(delta, delta);
    float reach = connection_reach(o);
    if (d2 < 0.0001f || d2 > reach * reach) return false;
    float dist = fsqrt(d2);
    f3 dir = div3(delta, dist);
    float receiver_cosine = fmx(dot3(normal, dir), 0.0f);
    float emitter_cosine = fmx(dot3(en, -dir), 0.0f);
    if (receiver_cosine <= 0.0f || emitter_cosine <= 0.0f || e.probability <= 0.0f) return false;
    float light_pdf = e.probability * d2 / emitter_cosine;
    float directional_pdf = receiver_cosine / 3.14159265359f;
    f3 shadow_origin = position + normal * 0.002f;
    f3 shadow_delta = light_position + en * 0.002f - shadow_origin;
    float shadow_distance = len3(shadow_delta);
    float4 em = o.materials[(uint32_t)e.material * 2u + 1u];
    float factor = directional_pdf / (light_pdf + directional_pdf);
    candidate = mk(em.x, em.y, em.z) * factor;
    origin = shadow_origin;
    direction = div3(shadow_delta, shadow_distance);
    distance = shadow_distance;
    return true;
}

// direct_at's lights (voxel_path.wgsl:283-307): light i's shadow ray and its
// weight; false when the light cannot reach the point.
__device__ __forceinline__ bool light_ray(const Optics &o, uint32_t i, f3 position, f3 normal, f3 &direction, float &distance,
                                          float &weight) {
    float4 pr = o.lights[i * 3u];
    f3 lp = mk(pr.x, pr.y, pr.z);
    direction = normalize3(lp);
    distance = o.max_distance;
    float attenuation = 1.0f;
    if (pr.w >= 0.0f) {
        f3 delta = lp - position;
        float d2 = dot3(delta, delta);
        if (d2 <= 0.0000001f || d2 > pr.w * pr.w) return false;
        distance = fsqrt(d2);
        direction = div3(delta, distance);
        distance = distance - 0.004f;
        attenuation = 1.0f / fmx(d2, 1.0f);
    }
    float cosine = fmx(dot3(normal, direction), 0.0f);
    if (!(cosine > 0.0f)) return false;
    weight = cosine * attenuation;
    return true;
}

struct FirstHit {
    int vx, vy, vz;
    uint32_t face, status, material, slot;
};

// Learned mode: where the path stopped into the cache, and the factor that
// multiplies the cache's predicted incident radiance there.
struct CacheQuery {
    int vx, vy, vz;
    uint32_t face_material;  // NONE: the path ended without a query
    f3 factor;
};

// A path of trace_path between bounces (voxel_path.wgsl:418-496): the energy
// gathered so far, the throughput, the next ray, and the previous vertex the
// next hit's emitter MIS weight is taken against. `bounce` counts the hits
// shaded so far (trace_path's step).
struct PathState {
    f3 energy, throughput, origin, direction, previous_normal, previous_position;
    float previous_pdf;
    uint32_t bounce;
};

// begin_path: false when the path gathers nothing (no bounce budget or a
// direction below the surface).
__device__ __forceinline__ bool path_begin(const Optics &o, f3 start, f3 first_direction, f3 initial_normal, PathState &p) {
    p.energy = mk(0.0f, 0.0f, 0.0f);
    p.throughput = mk(1.0f, 1.0f, 1.0f);
    p.origin = start;
    p.direction = first_direction;
    p.previous_normal = initial_normal;
    p.previous_pdf = cosine_pdf(initial_normal, first_direction);
    p.previous_position = start - initial_normal * 0.001f;
    p.bounce = 0u;
    return p.previous_pdf > 0.0f && o.bounces != 0u;
}

// One iteration of advance_path's loop (terminal = true): traces the path ray,
// shades its hit (direct light, emission with its MIS weight, the emitter or
// terminal sample, or the world cache's value) and draws the next direction.
// Returns true when the path goes on.
// kLearned: the path ends at its first explicit hit into the learned cache
// (the world-cache termination rule of voxel_path.wgsl:459-463 with the
// cached value replaced by the network's prediction of the hit face's
// incident radiance beyond its direct sun): the energy gathered is everything
// but the prediction term, whose factor goes to `query`.
template <class Tally, bool kLearned = false>
__device__ __forceinline__ bool path_bounce(const SceneView &s, const Optics &o, PathState &p, uint32_t stream, uint32_t sequence,
                                            bool &unknown, FirstHit &first, Tally &tally, CacheQuery *query = nullptr) {
    const uint32_t bounce = p.bounce;
    RayResult hit = trace_diffuse(s, o, p.origin, p.direction, tally);
    if (bounce == 0u) {
        first.vx = hit.vx; first.vy = hit.vy; first.vz = hit.vz;
        first.status = hit.status;
        first.material = hit.material;
        first.face = hit.status == 1u ? face_index_of(hit.normal) : 6u;
    }
    if (hit.status == 0u) {
        float horizon = fclamp(p.direction.y * 0.35f + 0.65f, 0.15f, 1.0f);
        p.energy = p.energy + p.throughput * mk(o.env_r, o.env_g, o.env_b) * horizon;
        return false;
    }
    if (hit.status == 2u) {
        unknown = true;
        return false;
    }
    if (hit.status == 3u) return false;
    float4 md = o.materials[hit.material * 2u];
    float4 me = o.materials[hit.material * 2u + 1u];
    f3 diffuse = mk(md.x, md.y, md.z);
    f3 emission = mk(me.x, me.y, me.z);
    f3 dv, dc;
    direct_at(s, o, hit.position, hit.normal, dv, dc, tally);
    float emission_weight = md.w;
    f3 delta = hit.position - p.origin;
    float selection_pdf = emission_weight / fmx(o.total_weight, 1e-20f);
    if (emission_weight > 0.0f && o.emitter_faces != 0u && o.has_tree)
        selection_pdf = tree_probability(o, p.previous_position, p.previous_normal, hit.vx, hit.vy, hit.vz, hit.normal);
    float light_pdf = selection_pdf * dot3(delta, delta) / fmx(dot3(hit.normal, -p.direction), 1e-8f);
    float mis = 1.0f;
    if (light_pdf > 0.0f) mis = p.previous_pdf / (p.previous_pdf + light_pdf);
    if constexpr (kLearned) {
        p.energy = p.energy + p.throughput * (emission * mis + diffuse * dv);
        query->vx = hit.vx; query->vy = hit.vy; query->vz = hit.vz;
        query->face_material = face_index_of(hit.normal) | (hit.material << 8);
        query->factor = p.throughput * diffuse;
        return false;
    }
    f3 cached;
    uint32_t cslot;
    bool found = cached_energy(o, hit.vx, hit.vy, hit.vz, face_index_of(hit.normal), cached, cslot);
    if (bounce == 0u) first.slot = cslot;
    if (found) {
        p.energy = p.energy + p.throughput * (emission * mis + diffuse * (dv + cached));
        return false;
    }
    bool last = bounce + 1u == o.bounces;
    f3 nee = last ? sample_terminal_emission(s, o, hit.position, hit.normal, stream, sequence, bounce, tally)
                  : sample_emitter(s, o, hit.position, hit.normal, stream, sequence, bounce + 1u, tally);
    p.energy = p.energy + p.throughput * (emission * mis + diffuse * (dv + nee));
    if (last) return false;
    p.throughput = p.throughput * diffuse;
    float maximum = fmx(p.throughput.x, fmx(p.throughput.y, p.throughput.z));
    if (maximum == 0.0f) return false;
    if (bounce >= 1u) {
        float survival = fclamp(maximum, 0.1f, 0.95f);
        if (continuation_sample(stream, sequence, bounce, 2u) > survival) return false;
        p.throughput = p.throughput / survival;
    }
    p.origin = hit.position + hit.normal * 0.002f;
    p.previous_position = hit.position;
    p.previous_normal = hit.normal;
    p.direction = hemisphere(hit.normal, continuation_sample(stream, sequence, bounce, 0u),
                             continuation_sample(stream, sequence, bounce, 1u));
    p.previous_pdf = cosine_pdf(hit.normal, p.direction);
    if (p.previous_pdf <= 0.0f) return false;
    p.bounce = bounce + 1u;
    return true;
}

// voxel_path.wgsl:418-496 (begin_path + advance_path with terminal = true)
template <class Tally, bool kLearned = false>
__device__ __forceinline__ f3 trace_path(const SceneView &s, const Optics &o, f3 start, f3 first_direction, f3 initial_normal,
                                         uint32_t stream, uint32_t sequence, bool &unknown, FirstHit &first, Tally &tally,
                                         CacheQuery *query = nullptr) {
    unknown = false;
    PathState p;
    if (!path_begin(o, start, first_direction, initial_normal, p)) return p.energy;
    while (path_bounce<Tally, kLearned>(s, o, p, stream, sequence, unknown, first, tally, query)) {
    }
    return p.energy;
}

// voxel_path.wgsl:504-520
template <class Tally, bool kLearned = false>
__device__ __forceinline__ f3 trace_receiver(const SceneView &s, const Optics &o, f3 surface, f3 normal, uint32_t stream,
                                             uint32_t sequence, bool &known, FirstHit &first, Tally &tally,
                                             CacheQuery *query = nullptr) {
    f3 direction = hemisphere(normal, sample_1d(stream, sequence, 0u), sample_1d(stream, sequence, 1u));
    f3 value = sample_emitter(s, o, surface, normal, stream, sequence, 0u, tally);
    bool unknown = false;
    first.vx = first.vy = first.vz = 0;
    first.face = 6u;
    first.status = 4u;
    first.material = 0u;
    first.slot = NONE;
    if (dot3(normal, direction) > 0.0f) {
        f3 c = trace_path<Tally, kLearned>(s, o, surface + normal * 0.001f, direction, normal, stream, sequence, unknown, first,
                                           tally, query);
        value = value + c;
    }
    known = !unknown;
    return value;
}

}  // namespace vgi
