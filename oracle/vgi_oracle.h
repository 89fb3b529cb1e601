/*
 * CPU oracle for the voxel-GI lighting hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker, never the product: only tests/, the smoke
 * entry and the bench's cpu_baseline leg may load it. It restates, in scalar
 * C++ with IEEE single precision (-ffp-contract=off, no fast-math), the
 * reference passes of bevy_voxel_gi (crates/bevy_voxel_gi/...):
 *   naadf.rs, world.rs (NaadfRegion, TopLevelAadf, Growth, cut_boxes),
 *   gpu_scene.rs (directory, headers, cells), emission.rs + emission/tree.rs,
 *   sampling.rs, shaders/voxel_traversal.wgsl, voxel_path.wgsl,
 *   voxel_trace.wgsl, voxel_accumulate.wgsl, voxel_retention.wgsl,
 *   gpu/face_index.rs, gpu/resident.rs (slot allocation).
 * Each function cites the file:line it follows.
 *
 * Transcendentals: the reference leaves cos/sin/exp to the shader compiler
 * and driver (not correctly rounded anywhere). The oracle and the HIP kernels
 * both use this project's own sin/cos/exp (Cody-Waite reduction + Taylor
 * polynomials, see vgi_oracle.cpp), so both sides are bit-identical to each
 * other; exactness against a driver's own cos/sin is not claimed.
 */
#ifndef VGI_ORACLE_H
#define VGI_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- hashes and samplers (sampling.rs:5-13, voxel_path.wgsl:178-238) ---- */
uint32_t orc_pcg(uint32_t value);
uint32_t orc_hash_voxel(int32_t x, int32_t y, int32_t z);
uint32_t orc_stream_id(int32_t x, int32_t y, int32_t z, uint32_t face);
float orc_sample_1d(uint32_t stream, uint32_t sequence, uint32_t dimension);
float orc_sinf(float x);
float orc_cosf(float x);
float orc_expf(float x);
void orc_hemisphere(const float n[3], float u, float v, float out[3]);
void orc_advance_box(const float position[3], const float direction[3], const int32_t cell[3],
                     const int32_t negative[3], const int32_t positive[3], float size,
                     float out[4]);

/* ---- NAADF (naadf.rs:37-148, naadf/construction_tests.rs:6-95) ---- */
/* blocked: size.x*size.y*size.z bytes (x fastest); out: 6 bytes per cell. */
int orc_build_aadf(uint32_t sx, uint32_t sy, uint32_t sz, const uint8_t *blocked, uint8_t *out);
int orc_full_sweep_aadf(uint32_t sx, uint32_t sy, uint32_t sz, const uint8_t *blocked,
                        uint8_t *out);

/* ---- scene (world.rs, gpu_scene.rs) ---- */
void *orc_scene_new(void);
void orc_scene_free(void *scene);
/* pages: 32^3 u16 voxels each (x fastest), AIR = 256. */
int orc_scene_load_pages(void *scene, const int32_t *page_xyz, uint32_t pages,
                         const uint16_t *voxels, const int32_t *removed_xyz, uint32_t removed);
/* Voxel edits: each edit sets one voxel (value AIR clears). Edited pages are
 * rebuilt as the reference's page overlay does (pages.rs:498). */
int orc_scene_edit(void *scene, const int32_t *xyz, const uint16_t *values, uint32_t count);
/* Region bounds (min xyz, max xyz) of loaded regions; returns 0 if empty. */
int orc_scene_bounds(void *scene, int32_t out[6]);
/* The top-level field: size xyz in out_size, 6 bytes per cell in out (may be
 * NULL to query the size). Returns cell count, 0 when the field is disabled. */
int64_t orc_scene_top_level(void *scene, int32_t out_size[3], uint8_t *out);
/* 0xffff outside the loaded scene, AIR (256) for air. */
uint32_t orc_scene_material_at(void *scene, int32_t x, int32_t y, int32_t z);
/* Canonical (layout-free) packed form of one region, as the traversal sees it:
 * out[0]=kind (0 unloaded,1 empty,2 uniform,3 mixed), out[1]=material,
 * out[2]=top-level negative bytes, out[3]=top-level positive bytes,
 * out[4..68)=brick words with the pointer replaced by brick ordinal tag 2,
 * out[68..68+64*32)=32 cell words per brick (zero for non-mixed bricks). */
int orc_scene_region_canonical(void *scene, int32_t rx, int32_t ry, int32_t rz, uint32_t *out);
/* Region any_solid flag (world.rs:183), -1 when not loaded. */
int orc_scene_region_solid(void *scene, int32_t rx, int32_t ry, int32_t rz);
/* Number of loaded regions / packed statistics for the bench. */
int64_t orc_scene_stats(void *scene, int64_t out[8]);

/* ---- transport (voxel_path.wgsl, voxel_trace.wgsl) ---- */
typedef struct orc_light {
    float position_range[4]; /* xyz direction toward the light (w < 0) or position, w range */
    float energy[4];
    float extra[4]; /* (0, custom visible rgb) */
} orc_light;

typedef struct orc_transport {
    uint32_t bounces;      /* DIFFUSE_HITS (1..8) */
    uint32_t samples;      /* samples per face per dispatch */
    uint32_t lanes;        /* PATH_LANES: 0 = the reference's choice (4 if spp<16 else 8) */
    uint32_t limited_reach;/* LIMITED_REACH */
    float reach;           /* transport.x when limited */
    float max_distance;    /* extra.y, MAX_PATH_DISTANCE 768 */
    float environment[3];
} orc_transport;

void *orc_tracer_new(void *scene);
void orc_tracer_free(void *tracer);
/* 256 materials: diffuse rgb, emission rgb (already bounded by the caller). */
int orc_tracer_set_materials(void *tracer, const float *diffuse, const float *emission);
int orc_tracer_set_lights(void *tracer, const orc_light *lights, uint32_t count,
                          uint32_t directional);
int orc_tracer_set_transport(void *tracer, const orc_transport *transport);
/* Rebuild packed scene + emitter distribution after scene changes. */
int orc_tracer_sync(void *tracer);
/* Emitter table: faces, tree nodes. */
int64_t orc_tracer_emitters(void *tracer, uint32_t *out_words, int64_t capacity);
/* tasks: 8 u32 per face (voxel xyz, face, sequence, 0,0,0).
 * results: 24 floats per face (voxel_trace.wgsl:25-37 layout).
 * hits (optional): per face-sample 8 u32 (first-hit voxel xyz, face|status<<8,
 *   material, cache slot, rays traced by this sample, 0).
 * counters (optional): [rays, steps, region loads, brick loads, cell loads, paths, 0, 0]. */
int orc_trace_faces(void *tracer, const uint32_t *tasks, uint32_t count, float *results,
                    uint32_t *hits, uint64_t *counters, int threads);
/* A learned frame's paths (ending at the first hit into the network): per
 * face-sample the partial value + known flag (4 floats), the prediction
 * factor (3 floats) and the query (4 u32). */
int orc_trace_learned(void *tracer, const uint32_t *tasks, uint32_t count, float *partial, float *factor, uint32_t *queries,
                      int threads);

/* ---- resident world cache (voxel_accumulate.wgsl, face_index.rs, resident.rs) ---- */
/* plain != 0: slots and accumulation as usual, but paths never end at cached faces. */
int orc_cache_enable(void *tracer, uint32_t history_limit, float recency, float measured_retention, uint32_t plain);
/* Registers faces (allocating slots as resident.rs:770-812) and writes their slots. */
int orc_cache_register(void *tracer, const uint32_t *tasks, uint32_t count, uint32_t *slots);
/* One batch: sequence from each slot's state, trace, accumulate. results as above
 * (the accumulated state's first six records after the merge). */
int orc_cache_batch(void *tracer, uint32_t *tasks, const uint32_t *slots, uint32_t count,
                    float *results, uint32_t *hits, uint64_t *counters, int threads,
                    uint32_t clock_ms);
/* Raw state words (9 uvec4 per slot) of `count` slots from `first`. */
int orc_cache_state(void *tracer, uint32_t first, uint32_t count, uint32_t *out);
/* Retention events: see voxel_retention.wgsl:29-96 for the layout. */
int orc_cache_meta(void *tracer, uint32_t first, uint32_t count, uint32_t *out);
int orc_cache_apply_events(void *tracer, const uint32_t *events, uint32_t event_count,
                           uint32_t words);
/* ---- Sun corner words (voxel_path.wgsl:245-266; voxel_view_runtime.wgsl:168-268;
 * sun_changes.rs; gpu/view/timing.rs:22-49) ---- */
int64_t orc_sun_corners(void *tracer, const uint32_t *tasks, const uint32_t *slots, uint32_t count, uint32_t frame, uint32_t refresh,
                        uint32_t clock_ms, uint32_t *region, uint32_t *events, uint32_t capacity);
int orc_sun_decode(const uint32_t *region, uint32_t records, uint32_t *out, uint32_t capacity);
int orc_sun_corner_words(void *tracer, uint32_t first, uint32_t count, uint32_t *out);
int orc_sun_add(void *tracer, const uint32_t *boxes, uint32_t n);
int orc_sun_take(void *tracer, uint32_t now_ms, uint32_t *out, uint32_t capacity);
/* Change footprints on (default) or off for edit_event (gpu.rs:419-425,
 * without_footprints: tests of the CPU-planned events alone). */
int orc_set_footprints(void *tracer, uint32_t on);
/* ---- GPU work selection (voxel_accumulate.wgsl:99-130, 448-695) ---- */
/* Scheduling settings record_stability and classification read. */
int orc_cache_schedule(void *tracer, const int32_t focus[3], uint32_t revision, float stable_error, uint32_t stable_samples,
                       uint32_t audit_ms, uint32_t mirror_budget);
/* One pick batch's selection passes: `selection` gets the 1038 scratch words,
 * `pick_list` (optional, `picks` words) the pick list (accepted keys in slot
 * order, ties up to each class's remainder; 0xffffffff where unused); returns
 * the number of mirror records, their slots (in slot order) copied to
 * `mirrored` (at most `capacity`). The picks are traced by orc_cache_picks. */
int orc_select(void *tracer, uint32_t picks, uint32_t clock_ms, uint32_t *selection, uint32_t *mirrored, uint32_t capacity,
               uint32_t *pick_list);
/* The selection's log2 (the bucket formulas' logarithm). */
float orc_select_log2(float x);
/* A slot's selection key (class << 8 | bucket, or 0xffffffff) for batch
 * `batch` at `now_ms` under the tracer's schedule (voxel_accumulate.wgsl:474-516). */
uint32_t orc_select_key(void *tracer, uint32_t slot, uint32_t batch, uint32_t now_ms);
/* select_threshold on given histograms (4 x 256): per class the threshold bucket (256: none) and remainder. */
int orc_select_threshold(const uint32_t *histograms, uint32_t picks, uint32_t *thresholds, uint32_t *remainders);
/* Overwrite slot metadata (12 words per slot) of allocated slots. */
int orc_cache_set_meta(void *tracer, uint32_t first, uint32_t count, const uint32_t *words);
/* Selection keys of slots [first, first+count). */
int orc_cache_keys(void *tracer, uint32_t first, uint32_t count, uint32_t *out);
/* The pick list (resident slots, 0xffffffff skipped) traced and merged as GPU
 * picks; results (optional): 24 floats per traced pick. Returns picks traced. */
int orc_cache_picks(void *tracer, const uint32_t *pick_slots, uint32_t count, float *results, int threads, uint32_t clock_ms);
/* voxel_retention.wgsl:9-12 */
/* Edit the scene and build its retention event (one event's encoded words,
 * measured against this tracer's resident statistics); returns the word
 * count, 0 when nothing within reach changed. orc_last_event copies the
 * words (at most `capacity`). */
int64_t orc_edit_event(void *tracer, const int32_t *xyz, const uint16_t *values, uint32_t count);
/* Load and unload pages (as orc_scene_load_pages) and build the upload's
 * retention event (Load / Unload / edit changes), read with orc_last_event. */
int64_t orc_load_event(void *tracer, const int32_t *page_xyz, uint32_t pages, const uint16_t *voxels, const int32_t *removed_xyz,
                       uint32_t removed);
int64_t orc_last_event(void *tracer, uint32_t *out, int64_t capacity);
uint32_t orc_curve_limit(uint32_t d2);
/* One traversal (voxel_traversal.wgsl:126-219): out_i = voxel xyz, material,
 * status, face; out_f = position xyz, normal xyz. */
int orc_trace_ray(void *tracer, const float origin[3], const float direction[3], float max_distance, int32_t *out_i,
                  float *out_f);

/* ---- learned cache (FP64 checker of the BF16 MFMA network): weights are the
 * 14,528 FP32 master values (W1 64x32, W2..W4 64x64, W5 3x64 row-major);
 * queries 4 u32 (voxel xyz, face | material << 8); scene = (lo xyz, 1/extent);
 * targets 4 floats (rgb, weight). ---- */
int orc_mlp_forward(const float *w, const float *table, const uint32_t *queries, uint32_t n, const float *diffuse, const float scene[4],
                    double *out);
int orc_mlp_gradient(const float *w, const float *table, const uint32_t *queries, const float *targets, uint32_t n, const float *diffuse,
                     const float scene[4], double *grad, double *table_grad, double *loss);
/* The hash-grid encoding alone: 16 features per query, 8 corner entries and
 * weights per level; the cells per level for a scene. */
int orc_mlp_hash(const float *table, const uint32_t *queries, uint32_t n, const float scene[4], float *features, uint32_t *index,
                 float *weight);
int orc_mlp_levels(const float scene[4], uint32_t *res);
/* Exact mode on/off: no BF16 rounding and double hash features (the gradient
 * is then the exact derivative of the double-precision forward). */
int orc_mlp_exact(int on);
/* 0: the hash-grid encoding; 1: the frequency encoding (sin/cos of 2^k pi u,
 * k = 0..3, no table). */
int orc_mlp_encoding(int mode);

/* ---- view tier (voxel_view_discovery.wgsl, voxel_view_runtime.wgsl,
 * voxel_view_cache.wgsl, voxel_view_pixels.wgsl): one camera frame over a
 * depth/normal prepass; the camera record has vgi_view_camera's layout. Per
 * pixel: the owner (pixel index, 0xffffffff none) and a 28-word face record
 * (vgi_view_face). ---- */
typedef struct orc_view_camera {
    float world_from_clip[16];
    float position[3];
    float inner_distance;
    uint32_t width, height;
    uint32_t samples;
    float measured_retention;
    float clip_from_world[16];
    float forward[3];
    float pixels_per_unit;
    uint32_t frame;
} orc_view_camera;
/* vgi_far_settings' layout. */
typedef struct orc_far_settings {
    float voxel_pixels;
    uint32_t measure, cell;
    float history;
    uint32_t rate, steps;
} orc_far_settings;
int orc_view_frame(void *tracer, const orc_view_camera *camera, const float *depth, const float *normals, uint32_t clock_ms,
                   uint32_t *owners, void *faces, uint32_t *owner_count, int threads);
int orc_view_reset(void *tracer);
int orc_view_far(void *tracer, const orc_far_settings *settings);
int orc_view_display(void *tracer, float noise_threshold, float spatial, float blend);
/* The last view frame's published records (12 words per pixel, vgi_view_pixel),
 * counters (16 words) and far tier (output: 16 + 4 words per cell; cells: 12
 * words each); any pointer may be NULL. Returns the number of far cells (0:
 * the tier did not run). */
int64_t orc_view_read(void *tracer, void *pixels, uint32_t *counters, uint32_t *far_output, uint32_t *far_cells);
/* pack2x16float with round-to-nearest-even: lo in bits 0-15, hi in 16-31. */
uint32_t orc_pack_half2(float lo, float hi);
/* The 32 encoded input features of each query (n x 32 doubles). */
int orc_mlp_encode(const float *table, const uint32_t *queries, uint32_t n, const float *diffuse, const float scene[4], double *out);

/* Packed scene of a tracer: [cell words, regions, directory capacity,
 * interned mixed bricks, brick references, 0, 0, 0]; a voxel's material
 * decoded from the packed tables (0xffff outside them). */
int orc_packed_info(void *tracer, int64_t out[8]);
uint32_t orc_packed_material(void *tracer, int32_t x, int32_t y, int32_t z);
int64_t orc_packed_words(void *tracer, int which, uint32_t *out, int64_t capacity);
/* A stand-alone face index (insert, backward-shift remove, get -> slot or -1). */
void *orc_face_index_new(uint32_t capacity);
void orc_face_index_free(void *index);
void orc_face_index_insert(void *index, int32_t x, int32_t y, int32_t z, uint32_t face, uint32_t slot);
void orc_face_index_remove(void *index, int32_t x, int32_t y, int32_t z, uint32_t face);
int64_t orc_face_index_get(void *index, int32_t x, int32_t y, int32_t z, uint32_t face);
int64_t orc_face_index_len(void *index);
int64_t orc_face_index_max_probe(void *index);
int64_t orc_face_index_capacity(void *index);
/* Raw resident state writes; one accumulate merge of a 24-word record into a slot. */
int orc_cache_set_state(void *tracer, uint32_t first, uint32_t count, const uint32_t *words);
int orc_cache_accumulate(void *tracer, uint32_t slot, float *record, uint32_t samples, uint32_t clock_ms);

#ifdef __cplusplus
}
#endif
#endif
