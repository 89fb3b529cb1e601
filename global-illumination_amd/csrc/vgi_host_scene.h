// Host-side scene bookkeeping of the tracer: the top-level region field grown
// in the reference's FIFO order, the emitter distribution and light tree, and
// the (voxel, face) -> slot index. These run on the CPU in the reference too
// (world.rs scene worker, emission.rs, gpu/face_index.rs); the brick and voxel
// NAADF fields and all tracing run on the GPU.
#pragma once
#include <stdint.h>

#include <array>
#include <cstddef>
#include <deque>
#include <map>
#include <tuple>
#include <vector>

namespace vgi_host {

constexpr uint16_t AIR = 256;
constexpr uint32_t NONE = 0xffffffffu;

struct Int3 {
    int32_t x, y, z;
};
inline bool operator==(const Int3 &a, const Int3 &b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

inline uint32_t pcg(uint32_t value) {
    uint32_t state = value * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
inline uint32_t hash_cell(int32_t x, int32_t y, int32_t z) {
    return pcg(pcg(pcg((uint32_t)x) ^ (uint32_t)y) ^ (uint32_t)z);
}

// Six u8 empty-box extents: -x, +x, -y, +y, -z, +z (naadf.rs:3-15).
using Extents = std::array<uint8_t, 6>;

// The top-level field over the loaded regions' bounding box
// (TopLevelAadf, world.rs:428-713): barriers are unloaded or solid regions.
struct TopLevelField {
    Int3 minimum{0, 0, 0}, maximum{0, 0, 0};
    int sx = 0, sy = 0, sz = 0;
    bool enabled = false;
    std::vector<Extents> fields;
    std::vector<uint8_t> blocked;

    // Rebuild or update after the `changed` regions (sorted z, y, x) flipped
    // residency or emptiness. `barrier` is the new state over [mn, mx]
    // (x fastest). Follows TopLevelAadf::updated (world.rs:477-583).
    void update(Int3 mn, Int3 mx, const std::vector<Int3> &changed, const std::vector<uint8_t> &barrier);
    Extents at(Int3 c) const;
};

// Emissive faces with their power CDF and light tree, packed as the words the
// kernels read (gpu.rs:686-713, tree.rs:170-192).
struct EmitterTable {
    std::vector<uint32_t> words;
    // The light tree as the device walks it, 8 words per node: centre xyz and
    // radius of its box; power, its left child's share of its power, left
    // child (NONE for a leaf), right child (a leaf's emitter face) with the
    // node's normal bits above bit 26. Derived from `words` by the same
    // binary32 operations the walk would otherwise repeat at every visit.
    std::vector<uint32_t> walk;
    uint32_t faces = 0;
    uint32_t nodes = 0;
    float total_weight = 0.0f;
};

// Material lookup over the loaded scene: AIR for air, 0xffff outside it.
struct VoxelSource {
    const uint16_t *grid;  // dense box, x fastest
    Int3 origin;           // voxel coordinate of grid[0]
    int gx, gy, gz;
    const uint8_t *loaded; // per region of the box
    Int3 region_min;
    int rsx, rsy, rsz;
    uint32_t material_at(int32_t x, int32_t y, int32_t z) const;
};

// emission.rs:91-165 + tree.rs:28-74 over the regions listed in z, y, x
// order (those holding an emitting material).
void build_emitters(const VoxelSource &scene, const std::vector<Int3> &regions, const float *emission_rgb, EmitterTable &out);

// gpu/face_index.rs: open addressing, linear probing, home =
// stream_id(voxel, face) & mask, capacity doubling at load 1/2.
struct FaceSlots {
    std::vector<std::array<uint32_t, 4>> entries;
    size_t live = 0;
    bool dirty = true;                // the whole table needs an upload (new or grown)
    std::vector<uint32_t> touched;    // entries written since the last upload
    FaceSlots();
    long find(int32_t x, int32_t y, int32_t z, uint32_t face) const;
    void insert(int32_t x, int32_t y, int32_t z, uint32_t face, uint32_t slot);
};

// ---------------------------------------------------------------- edit retention
// One changed box of an edit with its measured-retention weight
// (geometry.rs:110-229 ChangeWeight; retention.rs:65-74 RetentionEvent).
struct ChangeBox {
    Int3 min, max;
    float flux;       // the changed voxels' own emission change (measured edits
                      // add the brightest outgoing luminance on the GPU)
    uint32_t voxels;  // changed voxels
    uint32_t kind;    // 0 unknown flux, 1 unload, 2 load, 3 edit, 6 edit measured on the GPU
    bool urgent;      // holds an edited voxel
    int32_t top = 0;  // load: the region's top solid voxel row (INT32_MIN: air only)
};

struct RetentionEvent {
    Int3 union_min, union_max;
    std::vector<ChangeBox> boxes;  // sorted by (min, max)
};

// A region's 16^3 materials (x fastest) before and after an edit or a page
// upload; a region not loaded on one side holds air there.
struct RegionDelta {
    Int3 coord;
    std::vector<uint16_t> before, after;
    bool solid_before, solid_after;
    bool loaded_before = true, loaded_after = true;
};

// The scene before the edit, as the event needs it: the bounding box of its
// solid regions and which 32^3 pages hold a solid region.
struct PreviousScene {
    bool any_solid;
    Int3 solid_min, solid_max;  // region coordinates
    Int3 page_min;              // page coordinate of page_solid[0]
    int psx, psy, psz;
    const uint8_t *page_solid;  // x fastest
    bool page_occupied(Int3 page) const;
};

// The retention event of one edit: SceneChanges::with_materials
// (geometry.rs:110-229) over the changed regions (sorted z, y, x),
// GeometryChange::from_bounds (geometry.rs:299-407) at the influence radius,
// finalize_change_flux (publication.rs:423-454) and SpatialHistory::event
// (retention.rs:110-167). Returns false when nothing within reach changed.
class WorkPool;
// `pool` (optional) runs the per-region scans on host threads; the result does
// not depend on it.
bool edit_event(const std::vector<RegionDelta> &deltas, const PreviousScene &previous, const float *emission_rgb,
                const std::vector<Int3> &edits, RetentionEvent &out, WorkPool *pool = nullptr);
// The same for any scene change, page loads and unloads included
// (SceneChanges::with_materials geometry.rs:190-215): a region that was not
// loaded before is a Load of its whole box, its top solid row kept, with flux
// `sky` x 16 (finalize_change_flux publication.rs:438: sky = the brighter of
// the environment's luminance and half of each directional light's); one not
// loaded after is an Unload of its whole box, flux 0; the rest are edits.
bool scene_event(const std::vector<RegionDelta> &deltas, const PreviousScene &previous, const float *emission_rgb, float sky,
                 const std::vector<Int3> &edits, RetentionEvent &out, WorkPool *pool = nullptr);

// encode_events (gpu/resident.rs:1319-1417): event headers, then change
// records; measures are three uint4 per edit whose flux the GPU measures.
struct EncodedEvents {
    std::vector<uint32_t> words, measures;
    uint32_t events = 0, entries = 0, positions = 0;
};
void encode_events(const std::vector<RetentionEvent> &events, EncodedEvents &out);

// Change footprints (gpu/resident.rs:1215-1310): plan_footprints reserves one
// Light event (kind 4, empty bound) per edited box of `events` (edits first,
// then boxes of unknown flux; at most 64 a batch) and returns each box with
// its event's index; encode_footprints lays out the GPU tasks (a count, then
// per task (min, header record) (max, bound record) (first thread, grid, 0,
// 0)) for the events as encode_events lays them out, and returns the threads.
constexpr uint32_t MAX_FOOTPRINTS = 64;
struct Footprint {
    Int3 min, max;
    size_t event;
};
std::vector<Footprint> plan_footprints(std::vector<RetentionEvent> &events);
uint32_t encode_footprints(const std::vector<Footprint> &footprints, const std::vector<RetentionEvent> &events,
                           std::vector<int32_t> &words);

// Sun flips (voxel_view_runtime.wgsl:222-268 report_sun_flips): boxes whose
// corner sunlight changed, with the light reflected there.
struct SunFlip {
    Int3 min, max;
    float flux;
};
// decode_sun_flips (gpu/view/timing.rs:22-49): a flip region (a count, up to
// `records` (voxel, flux bits) records from word 4, then the overflow box as
// biased maxima and its flux x 64 at word 4 + 4 records) as boxes.
void decode_sun_flips(const uint32_t *region, uint32_t records, std::vector<SunFlip> &out);
// SunChanges (sun_changes.rs:18-69): flips binned per 16^3 region of their
// minimum corner (bounds widened, flux summed in arrival order), published
// together at most every 250 ms of the tracer clock.
struct SunChanges {
    std::map<std::tuple<int, int, int>, SunFlip> buckets;  // (z, y, x) region
    bool published = false;
    uint32_t published_ms = 0;
    void add(const std::vector<SunFlip> &flips);
    bool take(uint32_t now_ms, std::vector<SunFlip> &out);
};
// The published flips as a Light retention event (GeometryChange::from_bounds
// over `scene`'s solid pages, then SpatialHistory::event): boxes within reach,
// kind 4, sorted; false when none is in reach.
bool light_event(const std::vector<SunFlip> &flips, const PreviousScene &scene, RetentionEvent &out);

constexpr int INFLUENCE_RADIUS = 192;     // retention.rs:32-39 at HISTORY_LIMIT 65536
constexpr uint32_t FLUX_SCAN_BUDGET = 16384;  // publication.rs:413

}  // namespace vgi_host
