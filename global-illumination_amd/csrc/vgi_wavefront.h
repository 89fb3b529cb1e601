// Wavefront path tracing of bound faces: buffers and launcher
// (vgi_wavefront.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vgi {

struct SceneView;
struct Optics;

// A queued ray: origin, length, direction and its path << 3 | slot (slot 0
// path ray, 1 emitter shadow ray, 2..6 light shadow rays, 7 terminal
// connection ray).
struct WfRay {
    float ox, oy, oz, max_distance;
    float dx, dy, dz;
    uint32_t ref;
};

// A closest hit: voxel, material, position, status | normal code << 8 (each
// normal component as nonzero | sign << 1, so signed zeros survive).
struct WfHit {
    int vx, vy, vz;
    uint32_t material;
    float px, py, pz;
    uint32_t status_normal;
};

// Path state, one float4 per path per field (structure of arrays).
struct WfPaths {
    float4 *throughput;  // rgb, previous pdf
    float4 *energy;      // rgb, flags (bits)
    float4 *origin;      // origin of the pending path ray
    float4 *direction;   // direction of the pending path ray
    float4 *previous;    // previous position, previous normal code (bits)
    float4 *nee0;        // the receiver's emitter sample (candidate, then value)
    float4 *pending_t;   // throughput at the last hit
    float4 *pending_e;   // emission x MIS weight at the last hit
    float4 *pending_d;   // albedo at the last hit
    float4 *hit;         // last hit position, normal code (bits)
    float4 *candidate;   // deferred emitter candidate or cached value
};
constexpr uint32_t WF_PATH_FIELDS = 11;
constexpr uint32_t WF_MAX_LIGHTS = 5;

struct WfBuffers {
    WfPaths paths;
    WfRay *queue;       // capacity: paths x (2 + lights)
    WfHit *hit0, *hit1; // per path
    uint8_t *vis;       // 8 per path
    uint32_t *control;  // [queue count, queue head]
    const uint32_t *order;  // optional: bound faces in processing order
};

// One frame of the bound faces (records as trace_faces_kernel writes them);
// `rays` (optional) counts traversals.
void launch_wavefront(const SceneView &scene, const Optics &optics, const uint4 *tasks, uint32_t count, uint32_t samples, uint32_t lanes,
                      uint32_t sequence_base, const WfBuffers &buffers, uint32_t trace_grid, float4 *results, unsigned long long *rays,
                      hipStream_t stream, bool dedup = false);

}  // namespace vgi
