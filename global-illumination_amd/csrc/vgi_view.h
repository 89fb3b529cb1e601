// The view tier: camera-visible faces discovered from a depth/normal prepass
// (voxel_view_discovery.wgsl:4-130), deduplicated into one owner per face by
// a per-camera CAS hash (index_receivers, voxel_view_runtime.wgsl:128-142),
// reprojected by face identity (reproject_paths :148-217), exchanged both
// ways with the resident world cache (load_shared_faces / store_shared_faces,
// voxel_view_cache.wgsl:1-99,188-258) and traced with the world tier's
// estimator (trace_visible_paths, voxel_view_runtime.wgsl:273-324), then
// presented (reconstruct_view and publish_view, voxel_view_runtime.wgsl:
// 350-521). Faces small on screen belong to the far tier instead
// (voxel_view_pixels.wgsl): per-pixel cells with a reprojected history.
//
// Layout per prepass pixel p (row-major, x fastest): the pixel's receiver
// (ViewReceiver, 32 B), its estimate (ViewEstimate, 80 B; meaningful for an
// owner), its owner (u32: the pixel index of the face's owner) and its
// published record (ViewPixel, 48 B). The face hash has a power-of-two
// capacity of at least twice the pixels; each entry is the owner's pixel
// index + 1 (0 empty). Two generations of receivers, estimates and hash (the
// frame's and the previous frame's) alternate, as do two generations of far
// cells (PixelHistory, 48 B per cell of size x size pixels).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vgi_trace.h"

namespace vgi {

struct ViewReceiver {
    int32_t vx, vy, vz;
    uint32_t face;       // 0xffffffff: no receiver
    float sx, sy, sz;    // face centre
    uint32_t sequence;   // next path sequence
};

// voxel_view_runtime.wgsl:30-35 ViewEstimate plus the face's analytic direct
// term of this frame.
struct ViewEstimate {
    float4 mean;      // raw RGB, effective history weight (count)
    float4 moment;    // raw second moment, pending predicted change
    uint4 identity;   // next path sequence, material, corner sun word (0), resident incarnation
    float4 direct;    // direct rgb at the face (direct_at), 0
    float4 shown;     // reconstructed presentation rgb, valid (1) or not yet shown (0)
};

// publish_view's record per pixel (FarPixel, voxel_view_runtime.wgsl:40-46).
struct ViewPixel {
    float sx, sy, sz;
    uint32_t face;       // 0xffffffff: not an owner
    float ir, ig, ib;    // shown rgb (+ custom direct rgb)
    float evidence;      // sample weight behind it, 0: nothing to show
    uint4 sky;           // corner sun bytes (bits 0-3 lit lights, bit 31 known)
};

// The far tier (voxel_view_pixels.wgsl): one cell of `size` x `size`
// prepass pixels samples one of its pixels per frame and keeps a history
// reprojected by world position. PixelHistory is the reference's 48-byte
// cell: mean (rgb, count), surface point, face | stamp << 3, next path
// sequence.
struct PixelHistory {
    float4 mean;
    float sx, sy, sz;
    uint32_t face;
    uint32_t sequence;
    uint32_t pad[3];
};

struct FarParams {
    float clip[16];       // previous frame's clip_from_world, column-major
    float axis[4];        // view forward xyz, prepass pixels per unit at view depth 1
    uint32_t cells_x, cells_y, size;
    uint32_t flags;       // bit 0: on; bits 1-2: measure (0 voxel side, 1 voxel area, 2 block area)
    float threshold;      // voxel_pixels
    float history;        // history cap in samples
    float rate;           // trace one frame in `rate` once a cell holds 4 samples
    uint32_t previous;    // stamp of the frame that wrote the previous cells, NONE: none
    uint32_t frame;       // this frame's number (its cells' stamp)
    uint32_t custom;      // some light has custom visible rgb (params.transport.y)
    uint32_t steps;       // traversal steps per far ray (PATH_STEPS)
    uint32_t cells;       // cells_x * cells_y
};

// Words of the far tier's counters (voxel_view_pixels.wgsl:34-39) and of the
// presentation's (voxel_view_runtime.wgsl:51-64), one buffer of 16 words.
enum : uint32_t {
    PIXEL_OWNED = 0u, PIXEL_TRACED = 1u, PIXEL_HISTORY = 2u, PIXEL_REPROJECTED = 3u, PIXEL_SEEDED = 4u,
    PIXEL_RECEIVERS_SEEDED = 5u,
    SHOW_HISTORY = 8u, SHOW_MAX_HISTORY = 9u, SHOW_UNCOVERED = 10u, SHOW_DISPLAYED = 11u,
};

struct ViewParams {
    float m[16];          // world_from_clip, column-major
    float cam[4];         // camera position, inner overlap distance
    uint32_t width, height, pixels;
    uint32_t mask;        // this frame's hash capacity - 1
    uint32_t prev_pixels; // previous frame's pixels (0: no history)
    uint32_t prev_mask;
    uint32_t samples;     // paths per owner this frame
    float retention;      // measured retention (the keep rule's c)
    uint32_t clock_ms;
    uint32_t exchange;    // 1: resident world cache bound for the exchange
    float display[3];     // noise threshold, spatial strength, temporal blend (view.display.xyz)
    uint32_t custom;      // some light has custom visible rgb (params.transport.y)
};

// The resident world cache as the exchange sees it.
struct SharedCache {
    const uint4 *__restrict__ index;  // (voxel, face << 29 | slot), NONE in w: empty
    uint32_t *__restrict__ state;     // 36 words per slot
    uint32_t *__restrict__ meta;      // 12 words per slot
    uint32_t mask;                    // index capacity - 1
};

// The far tier's buffers of one camera: two generations of cells (the
// frame's and the previous frame's), the material's output (16 header words,
// then 4 words per cell), the trace queue (a count, then cells).
struct FarBuffers {
    const PixelHistory *previous;
    PixelHistory *cells;
    uint32_t *output;
    uint32_t *queue;
};

// One view frame: discovery, index, owners, reprojection, the exchange with
// the world cache (and a new face's seed from the far cells), tracing,
// presentation (reconstruct, store, publish), then the far tier's classify
// and trace. `counters` are 16 words, cleared by the caller.
void launch_view_frame(const ViewParams &p, const FarParams &f, const SceneView &s, const Optics &o, const SharedCache &cache,
                       const float *depth, const float *normals, ViewReceiver *receivers, ViewEstimate *estimates, uint32_t *hash,
                       uint32_t *owners, const ViewReceiver *prev_receivers, const ViewEstimate *prev_estimates,
                       const uint32_t *prev_hash, uint32_t *queue, ViewPixel *pixels_out, uint32_t *counters, const FarBuffers &far,
                       hipStream_t stream, hipEvent_t before_trace = nullptr, hipEvent_t after_trace = nullptr,
                       hipEvent_t before_far = nullptr);

}  // namespace vgi
