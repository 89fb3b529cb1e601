// The view tier: camera-visible faces discovered from a depth/normal prepass
// (voxel_view_discovery.wgsl:4-130), deduplicated into one owner per face by
// a per-camera CAS hash (index_receivers, voxel_view_runtime.wgsl:128-142),
// reprojected by face identity (reproject_paths :148-217), exchanged both
// ways with the resident world cache (load_shared_faces / store_shared_faces,
// voxel_view_cache.wgsl:1-99,188-258) and traced with the world tier's
// estimator (trace_visible_paths, voxel_view_runtime.wgsl:273-324).
//
// Layout per prepass pixel p (row-major, x fastest): the pixel's receiver
// (ViewReceiver, 32 B), its estimate (ViewEstimate, 64 B; meaningful for an
// owner), its owner (u32: the pixel index of the face's owner). The face
// hash has a power-of-two capacity of at least twice the pixels; each entry
// is the owner's pixel index + 1 (0 empty). Two generations of receivers,
// estimates and hash (the frame's and the previous frame's) alternate.
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

// voxel_view_runtime.wgsl:30-35 ViewEstimate without the presentation word
// (`shown`), plus the face's analytic direct term of this frame.
struct ViewEstimate {
    float4 mean;      // raw RGB, effective history weight (count)
    float4 moment;    // raw second moment, pending predicted change
    uint4 identity;   // next path sequence, material, corner sun word (0), resident incarnation
    float4 direct;    // direct rgb at the face (direct_at), 0
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
};

// The resident world cache as the exchange sees it.
struct SharedCache {
    const uint4 *__restrict__ index;  // (voxel, face << 29 | slot), NONE in w: empty
    uint32_t *__restrict__ state;     // 36 words per slot
    uint32_t *__restrict__ meta;      // 12 words per slot
    uint32_t mask;                    // index capacity - 1
};

void launch_view_frame(const ViewParams &p, const SceneView &s, const Optics &o, const SharedCache &cache, const float *depth,
                       const float *normals, ViewReceiver *receivers, ViewEstimate *estimates, uint32_t *hash, uint32_t *owners,
                       const ViewReceiver *prev_receivers, const ViewEstimate *prev_estimates, const uint32_t *prev_hash,
                       uint32_t *queue, hipStream_t stream, hipEvent_t before_trace = nullptr, hipEvent_t after_trace = nullptr);

}  // namespace vgi
