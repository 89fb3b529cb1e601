// Kernel launch interface shared by the host library and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Waves per SIMD the trace kernels are compiled for (their register budget:
// 7 waves leave 72 VGPRs a lane, 8 leave 64). Measured on the 512^3 cave's
// plain frame with one-wavefront workgroups: 6 waves 51.1 ms, 7 waves 47.4
// ms, 8 waves 51.7 ms (256-thread workgroups: 58.4, 53.0 and 56.9 ms).
#ifndef VGI_TRACE_WAVES
#define VGI_TRACE_WAVES 7
#endif

namespace vgi {

struct SceneView;
struct Optics;
struct MlpQuery;
struct BatchScene;

// A queue tile of a batched frame: `faces` bound faces of one scene.
struct BatchTile {
    uint32_t scene, first, faces, pad;
};
// Device-side frame state of a batched launch (written only by kernels, so a
// captured graph replays it): the tile queue head, this frame's sequence base
// and the next one, the start timestamp.
struct BatchControl {
    uint32_t head, frame_seq, next_seq, samples;
    unsigned long long start, rays;
};

// Per listed region: the classification of world.rs:243-319 plus the packing
// offsets the scan assigns (words into the cells buffer).
struct RegionInfo {
    uint32_t storage;  // 0 empty, 1 uniform, 2 dense, 3 all bricks uniform
    uint32_t uniform;
    uint32_t any_solid;
    uint32_t table;    // 1 when the region packs a 64-word brick table
    uint32_t mixed_lo, mixed_hi;        // bricks that need a 32-word payload
    uint32_t occupied_lo, occupied_hi;  // bricks holding a solid voxel
    uint32_t mixed_count, table_rank, brick_rank, table_offset;
    uint32_t brick_offset, pad0, pad1, pad2;
    uint16_t brick_material[64];  // first voxel's material per brick
};

struct AccumulateParams {
    uint32_t history_limit;
    uint32_t samples;
    uint32_t clock_ms;
    float recency;
    float physical;  // measured retention c
    float response;  // display response seconds
    // record_stability: the transport revision, the relative error and the
    // samples a face needs to count as stable; picked: GPU-picked faces (their
    // host mirror goes stale)
    uint32_t revision;
    float stable_error;
    uint32_t stable_samples;
    uint32_t picked;
};

// GPU work selection (vgi_select.hip). Scratch words: four 256-bucket key
// histograms (coverage, refinement, sweep, urgent), then per class the
// threshold bucket, the remainder taken at it and the ties seen at it, then
// the pick total and the mirror-record count (resident.rs SELECTION_WORDS).
constexpr uint32_t SEL_THRESHOLD = 1024u, SEL_REMAINDER = 1028u, SEL_TIES = 1032u, SEL_TOTAL = 1036u, SEL_MIRRORS = 1037u;
constexpr uint32_t SEL_WORDS = 1038u;
struct SelectParams {
    int fx, fy, fz;           // focus voxel
    uint32_t budget;          // GPU picks this batch
    uint32_t batch;           // batch id (28 bits, never 0)
    uint32_t now_ms;          // the merge clock
    uint32_t revision;        // transport revision
    uint32_t stable_samples;  // samples a stable face needs
    uint32_t audit_ms;        // age at which a stable face is audited
    uint32_t first, end;      // this batch's classified slots (a rotating quarter)
    uint32_t slots;           // high-water mark
    uint32_t mirror_budget;   // statistics-mirror records per batch
};
// Clear the scratch and the pick list, then classify, histogram, threshold and
// emit: pick_tasks (2 uint4 per pick, face task layout) and pick_slots
// (NONE: no pick) hold `budget` entries, mirror `mirror_budget` records of 7 uint4.
// `ranges`: SELECT_RANGES x 16 words of emit scratch (per-range counts and offsets).
constexpr uint32_t SELECT_RANGES = 2048u;
void launch_select(const SelectParams &p, uint32_t *meta, const uint32_t *state, uint32_t *keys, uint32_t *selection, uint4 *mirror,
                   uint4 *pick_tasks, uint32_t *pick_slots, uint32_t *ranges, uint32_t histogram_blocks, hipStream_t stream);

// Slot ranking flags (voxel_accumulate.wgsl:460-461, resident.rs META_BYTES):
// an edit's change binds the face's history; its host mirror is stale.
constexpr uint32_t RANK_URGENT = 0x40000000u;
constexpr uint32_t MIRROR_STALE = 0x20000000u;

// `order` (optional): the bound faces in the order threads take them (a
// spatially coherent order); records still go to each face's own index.
void launch_trace(const SceneView &scene, const Optics &optics, const uint4 *tasks, uint32_t count, uint32_t samples,
                  uint32_t lanes, uint32_t sequence_base, float4 *results, uint32_t *hits, unsigned long long *counters,
                  bool full_counters, hipStream_t stream, const uint32_t *order = nullptr);
void launch_trace_learned(const SceneView &scene, const Optics &optics, const uint4 *tasks, uint32_t count, uint32_t samples,
                          uint32_t lanes, uint32_t sequence_base, float4 *results, unsigned long long *counters, float4 *learned,
                          MlpQuery *queries, hipStream_t stream, const uint32_t *order = nullptr);
void launch_learned_combine(const float4 *learned, const float *pred, uint32_t count, uint32_t samples, uint32_t lanes,
                            const uint4 *tasks, float4 *results, hipStream_t stream);
void launch_train_gather(const uint4 *bound, const uint32_t *perm, const uint32_t *face_material, uint32_t n_bound, uint32_t start,
                         uint32_t count, uint4 *tasks, MlpQuery *queries, hipStream_t stream);
void launch_train_targets(const float4 *results, uint32_t count, float4 *targets, hipStream_t stream);
void launch_fp64_accumulate(const float4 *results, uint32_t count, double *sums, hipStream_t stream);
void launch_classify(const uint16_t *grid, int gx, int gy, int gz, int vx, int vy, int vz, const int4 *list, uint32_t n,
                     RegionInfo *info, hipStream_t stream);
void launch_scan(RegionInfo *info, uint32_t n, uint32_t cells_base, uint32_t *totals, hipStream_t stream);
void launch_pack(const uint16_t *grid, int gx, int gy, int gz, int vx, int vy, int vz, const int4 *list, uint32_t n,
                 const RegionInfo *info, uint32_t *cells, hipStream_t stream);
void launch_headers(const int4 *list, uint32_t n, const RegionInfo *info, const uint32_t *top, uint4 *headers, int rx, int ry,
                    int rz, int sx, int sy, hipStream_t stream);
void launch_top(const uint32_t *top, uint32_t cells, uint4 *headers, hipStream_t stream);
void launch_scatter_voxels(uint16_t *grid, const uint32_t *index, const uint16_t *values, uint32_t n, hipStream_t stream);
// batch != 0: host-requested faces of batch `batch` (claimed from GPU
// selection; their selection keys, when `keys` is given, are cleared).
void launch_prepare(uint4 *tasks, const uint32_t *slots, uint32_t n, uint32_t *state, uint32_t *meta, float physical,
                    uint32_t batch, uint32_t *keys, hipStream_t stream);
void launch_accumulate(const uint32_t *slots, uint32_t n, float4 *results, uint32_t *state, uint32_t *checkpoint, uint32_t *meta,
                       const AccumulateParams &p, hipStream_t stream);
void launch_apply_events(const uint4 *events, uint32_t event_count, uint32_t high_water, uint32_t *state, uint32_t *checkpoint,
                         uint32_t *meta, float physical, hipStream_t stream);

// Edit flux from the resident statistics (measure_edits + finish_edits);
// measure_max must be cleared first (one word per entry).
void launch_measure_edits(const uint4 *measures, uint32_t entries, uint32_t positions, const uint4 *index, uint32_t mask,
                          const uint32_t *state, const uint32_t *meta, const float2 *material_light, uint32_t *measure_max,
                          uint4 *events, hipStream_t stream);

void launch_scatter_bound(const uint32_t *index, const uint4 *tasks, const uint32_t *slots, uint32_t n, uint4 *bound_tasks,
                          uint32_t *bound_slots, hipStream_t stream);
void launch_scatter_words(const uint2 *writes, uint32_t n, uint32_t *words, hipStream_t stream);

// One batched frame: reset the queue (and advance the sequence), then a
// persistent launch of `grid` workgroups of `block` threads that take tiles
// until the queue is empty; per scene, the wall clock when its last tile ended.
void launch_batch_frame(const BatchScene *scenes, uint32_t n_scenes, const BatchTile *tiles, uint32_t n_tiles, uint32_t lanes,
                        uint32_t block, uint32_t grid, BatchControl *control, uint32_t *done, unsigned long long *finish,
                        hipStream_t stream);

// Change footprints (voxel_trace.wgsl:94-202): trace the tasks' tubes into
// `sums` (8 words per task, initialised to min 0xffffffff x3, max 0 x3, flux
// 0, spare 0) and write each landing into its reserved Light event.
void launch_footprints(const SceneView &scene, const Optics &optics, const int4 *tasks, uint32_t task_count, uint32_t threads,
                       uint32_t *sums, uint4 *events, hipStream_t stream);

// Sun corner words (voxel_path.wgsl:245-266; voxel_view_runtime.wgsl:168-268)
// of `count` bound faces (tasks: 2 uint4 each; slots: their resident slots):
// `corners` holds one word per slot (0 = none yet); `flips` (one float per
// bound face), `queue` (1 + count words: the faces that trace) and `ranges`
// (10 x SUN_RANGES words: per-range counts, offsets, overflow parts) are scratch; the tracing pass runs at most
// `trace_blocks` workgroups of 256 over the queue. The flip region
// gets the count (word 0), up to SUN_FLIP_RECORDS records (voxel, flux bits)
// from word 4 in bound-face order, then the overflow box (biased maxima of
// 2^30 - min and max + 2^30) and its flux x 64 at SUN_FLIP_OVERFLOW.
constexpr uint32_t SUN_FLIP_RECORDS = 2048u;
constexpr uint32_t SUN_FLIP_OVERFLOW = 4u + 4u * SUN_FLIP_RECORDS;
constexpr uint32_t SUN_FLIP_WORDS = SUN_FLIP_OVERFLOW + 8u;
constexpr int SUN_FLIP_BIAS = 1073741824;
constexpr uint32_t SUN_RANGES = 4096u;
void launch_sun_corners(const SceneView &scene, const Optics &optics, const uint4 *tasks, const uint32_t *slots, uint32_t count,
                        const uint32_t *meta, uint32_t *corners, float *flips, uint32_t *queue, uint32_t trace_blocks, uint32_t frame,
                        uint32_t refresh, uint32_t *ranges, uint32_t *region, hipStream_t stream);

void launch_copy(const void *src, void *dst, size_t bytes, hipStream_t stream);
void launch_fill_words(uint32_t *words, uint64_t n, uint32_t value, hipStream_t stream);

}  // namespace vgi
