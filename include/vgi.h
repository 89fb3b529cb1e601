/*
 * vgi.h -- C ABI of the MI355X-native voxel GI lighting hot path.
 *
 * The library replaces the reference's GPU transport seam, `GpuTracer`
 * (crates/bevy_voxel_gi/src/gpu.rs:309-1236), together with the scene
 * construction it consumes (`TransportScene::with_page_contents`,
 * world.rs:982-1046; `PackedScene::update`, gpu_scene.rs:72-309) and the
 * resident world cache (`ResidentAccumulator`, gpu/resident.rs:319-1213).
 * Each entry point below names the reference interface it stands in for.
 *
 * Conventions: integer status codes (0 ok, negative error) and no exceptions
 * across the ABI; `vgi_last_error()` describes the last failure on the calling
 * thread. Host memory is caller-owned, device memory library-owned. One
 * thread at a time per tracer (the reference's `&mut self` contract,
 * dispatch.rs:167-271). No torch types, no HIP types in signatures.
 */
#ifndef VGI_H
#define VGI_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define VGI_AIR 256u            /* world.rs:19 AIR: u16 material sentinel */
#define VGI_PAGE_EDGE 32u       /* world.rs:22 */
#define VGI_NO_FACE 0xffffffffu /* voxel_trace.wgsl:64 unused pick capacity */

typedef struct vgi_tracer vgi_tracer;
typedef struct vgi_batch vgi_batch;

/* gpu.rs:136-141 GpuLight: position_range.w < 0 marks a directional source
 * whose xyz is the unit direction toward the light (sources.rs:130). */
typedef struct vgi_light {
    float position_range[4];
    float energy[4];
    float extra[4]; /* (0, custom visible rgb) */
} vgi_light;

/* gpu.rs:144-177 / voxel_trace.wgsl:5-12 FaceTask, 32 bytes. */
typedef struct vgi_face_task {
    int32_t voxel[3];
    uint32_t face;     /* 0..5 = +x,-x,+y,-y,+z,-z (world.rs:50-57); VGI_NO_FACE skips */
    uint32_t sequence; /* first QMC sequence index of this batch */
    uint32_t padding[3];
} vgi_face_task;

/* voxel_trace.wgsl:13-20 FaceResult, 96 bytes (gpu.rs:1072-1106 decode). */
typedef struct vgi_face_result {
    float mean[4];
    float mean_extra[4]; /* y: known sample count (f32); w: presentation (resident) */
    float second[4];
    float second_extra[4];
    float direct[4];
    float direct_extra[4]; /* yzw: custom visible direct */
} vgi_face_result;

/* TransportSettings (settings.rs:391-426) + TraceWork history fields
 * (work.rs:11-55) + the WGSL override constants (voxel_path.wgsl:4-14). */
typedef struct vgi_transport {
    uint32_t bounces;        /* DIFFUSE_HITS; the reference clamps its public setting to 1..3, 1..8 here */
    uint32_t samples;        /* samples per face per batch */
    uint32_t lanes;          /* PATH_LANES (0: 4 if samples < 16 else 8, gpu.rs:1179-1181) */
    uint32_t limited_reach;  /* LIMITED_REACH */
    float reach;             /* transport.x */
    float max_distance;      /* extra.y = MAX_PATH_DISTANCE (tracer.rs:3), 0 -> 768 */
    float environment[3];
    uint32_t history_limit;  /* 0 -> HISTORY_LIMIT 65536 (resident.rs:715-719) */
    float history_recency;   /* >= 1 (resident.rs:721-725) */
    float measured_retention;/* settings.rs:262 */
    uint32_t plain;          /* 1: paths never end at cached faces (plain multi-bounce tracing);
                                0: the reference's world-cache termination (voxel_path.wgsl:459-463) */
} vgi_transport;

/* ---- lifecycle: GpuTracer::new / enable_resident (gpu.rs:309-366) ---- */
int vgi_create(int device_ordinal, vgi_tracer **out);
void vgi_destroy(vgi_tracer *tracer);
int vgi_enable_resident(vgi_tracer *tracer);
/* The resident clock (ms) that the display EMA reads; negative restores the
 * wall clock. Mirrors the reference's test clock (gpu/resident.rs:708-710). */
int vgi_set_clock(vgi_tracer *tracer, int64_t milliseconds);

/* ---- scene: VoxelGiInput::{page_loaded, page_unloaded} (input.rs:54-112) feeding
 * TransportScene::with_page_contents (world.rs:982-1046). Pages are 32^3 u16
 * voxels, x fastest, VGI_AIR for air. The NAADF brick/voxel fields are built on
 * the GPU; the top level is grown in exact FIFO order on the host. With the
 * resident cache, an upload that changes a loaded scene queues its retention
 * event for the next batch or frame: regions that load are Load changes of
 * their whole box with their top solid row and the sky's flux, regions that
 * unload are Unload changes, regions of a page sent again with new content
 * are edits measured on the GPU (SceneChanges::with_materials
 * geometry.rs:110-229, finalize_change_flux publication.rs:423-454). ---- */
int vgi_upload_pages(vgi_tracer *tracer, const int32_t *page_xyz, uint32_t pages, const uint16_t *voxels,
                     const int32_t *removed_xyz, uint32_t removed);
/* VoxelGiInput::edit_voxel (input.rs:90-93): set voxels (in order; a later
 * edit of a voxel wins); edited regions are re-packed on the GPU and the top
 * level is updated incrementally (TopLevelAadf::updated, world.rs:477-583).
 * With the resident cache, the edit's retention event is built
 * (SceneChanges::with_materials geometry.rs:110-229, GeometryChange::from_bounds
 * geometry.rs:299-407, finalize_change_flux publication.rs:423-454,
 * SpatialHistory::event retention.rs:110-167) and queued; the next batch or
 * frame measures its flux on the GPU (measure_edits / finish_edits,
 * voxel_accumulate.wgsl:375-430) and applies it to every slot
 * (apply_events, voxel_accumulate.wgsl:295-332) before its prepare. With
 * vgi_bind_exposed, bound faces follow the edit (pages.rs:451-458 rescan). */
int vgi_edit_voxels(vgi_tracer *tracer, const int32_t *xyz, const uint16_t *values, uint32_t count);
/* (An edit whose bound-face update fails, e.g. at the resident slot limit,
 * is still applied to the scene; it returns the error, and frames refuse
 * the bound set until faces are bound again.) */

/* ---- optics and sources: VoxelGiMaterials (world.rs:90-174) as packed by
 * gpu.rs:714-727; TransportSources (sources.rs:105-160) as packed by
 * gpu.rs:728-760 ---- */
int vgi_set_materials(vgi_tracer *tracer, const float *diffuse_rgb, const float *emission_rgb);
int vgi_set_lights(vgi_tracer *tracer, const vgi_light *lights, uint32_t count, uint32_t directional);
int vgi_set_transport(vgi_tracer *tracer, const vgi_transport *transport);

/* ---- batches: GpuTracer::submit_compact (gpu.rs:647-1022). Non-resident:
 * trace `count` tasks as given. With the resident cache on, each face gets
 * its slot (allocating as resident.rs:770-812), takes the slot's sequence
 * (prepare, voxel_accumulate.wgsl:83-133) and is merged (accumulate,
 * voxel_accumulate.wgsl:135-227). Flags: VGI_RECORD_HITS keeps the per
 * face-sample first-hit records; VGI_COUNT_RAYS counts rays and steps. ---- */
#define VGI_RECORD_HITS 1u
#define VGI_COUNT_RAYS 2u
#define VGI_COUNT_GATHERS 4u /* frames: also steps and traversal gathers (tickets always count all) */
int vgi_submit_faces(vgi_tracer *tracer, const vgi_face_task *tasks, uint32_t count, uint32_t flags,
                     uint64_t *ticket);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture elsewhere in the process, tight
 * serving loops, harnesses that forbid hipMalloc): the trace of
 * vgi_submit_faces + vgi_read_faces on a tracer without the resident cache,
 * synchronous, with the tasks and records staged in a fixed device scratch
 * region (VGI_NOALLOC_MAX_FACES faces, 128 bytes each: 32 MiB per GPU, held
 * until the process ends) instead of buffers the call allocates. The same
 * records bit for bit. It allocates nothing: more faces than the maximum, a
 * resident tracer, or an emitter table that would have to grow are errors
 * (set-up calls allocate; run one vgi_submit_faces after a scene or material
 * change). Every call on one GPU shares that GPU's scratch: calls that
 * overlap on one GPU (several threads, several tracers) race. */
#define VGI_NOALLOC_MAX_FACES 262144u
int vgi_trace_faces_noalloc(vgi_tracer *tracer, const vgi_face_task *tasks, uint32_t count, vgi_face_result *out);
/* GpuTracer::snapshot_ready (gpu.rs:1024-1049) */
int vgi_poll(vgi_tracer *tracer, uint64_t ticket, int *ready);
/* GpuTracer::complete_readback (gpu.rs:1050-1170): 96-byte records. */
int vgi_read_faces(vgi_tracer *tracer, uint64_t ticket, vgi_face_result *out, uint32_t count);
/* First-hit record per face-sample (cf. voxel_view_trace.wgsl:21-38): 8 u32:
 * voxel xyz, face | status << 8, material, cache slot, rays of the sample,
 * known. Requires VGI_RECORD_HITS. */
int vgi_read_hits(vgi_tracer *tracer, uint64_t ticket, uint32_t *out, uint64_t words);
/* [rays, steps, header loads, brick-word loads, cell-word loads, 0...] of the ticket (VGI_COUNT_RAYS). */
int vgi_read_counters(vgi_tracer *tracer, uint64_t ticket, uint64_t *out, uint32_t count);
/* GPU timestamps (gpu.rs:486-497,1050-1070): ms of [prepare, trace, accumulate,
 * build]; count up to 8 adds the last edit's host ms: [regions + GPU re-pack,
 * top level, retention event, exposure update]. */
int vgi_timings(vgi_tracer *tracer, float *out, uint32_t count);
/* Resident slots of the last batch's faces (face_index.rs:71-97 lookups). */
int vgi_read_slots(vgi_tracer *tracer, uint32_t *out, uint32_t count);

/* ---- device-resident frames (no host transfer): the face set is bound once
 * and each frame runs prepare -> trace -> accumulate on it, as the reference's
 * worker does batch after batch. ---- */
int vgi_bind_faces(vgi_tracer *tracer, const vgi_face_task *tasks, uint32_t count);
int vgi_run_frame(vgi_tracer *tracer, uint32_t flags);
int vgi_sync(vgi_tracer *tracer);
/* Time of the last frame's trace launch on its stream (ms), valid after vgi_sync. */
float vgi_last_trace_ms(vgi_tracer *tracer);
/* GPU time of the whole last frame (prepare, trace or learned trace +
 * inference, accumulate, and the learned cache's training), valid after
 * vgi_sync. */
float vgi_last_frame_ms(vgi_tracer *tracer);
/* Bound-face frames by wavefront path tracing (on) or by the one-kernel
 * estimator (off): the same records bit for bit. Frames that count gathers
 * (VGI_COUNT_GATHERS), learned frames and more than 5 lights use the
 * one-kernel estimator either way. */
int vgi_set_wavefront(vgi_tracer *tracer, int on);
/* The wavefront tracer's hit-face deduplication (on) or not (off, the
 * default): before a shading pass, the world-cache lookups of its new hits
 * (cached_energy, voxel_path.wgsl:63-87) are matched per wavefront (ballot
 * over the distinct (voxel, face) keys) and each distinct face is probed
 * once, its answer shared across the wavefront. The same records bit for
 * bit; with VGI_COUNT_RAYS the frame's counters 5 and 6 receive the lookups
 * and the probes made. */
int vgi_set_dedup(vgi_tracer *tracer, int on);
/* The bound faces' 96-byte records after the last frame (out may be NULL with
 * count 0 to read only the counters); counters (optional)
 * receive [rays, steps, header loads, brick-word loads, cell-word loads,
 * deduplicated world-cache lookups, their probes] when the frame ran with
 * VGI_COUNT_RAYS. Records follow the faces bound now: after
 * an edit that moved bound faces (vgi_bind_exposed), including one issued
 * while the frame still ran, each face reads its record of the last frame,
 * and a face bound since reads zeros -- the same records in either call
 * order. */
int vgi_read_bound(vgi_tracer *tracer, vgi_face_result *out, uint32_t count, uint64_t *counters);
/* First-hit records of the last bound-face frame run with VGI_RECORD_HITS (as
 * vgi_read_hits: 8 u32 per face-sample, bound face i's samples at
 * (i * samples + s) * 8; cf. voxel_view_trace.wgsl:21-38, the reference's
 * debug trace of the same frame). Learned frames keep none. */
int vgi_read_bound_hits(vgi_tracer *tracer, uint32_t *out, uint64_t words);
/* Bind every exposed face of the loaded scene (a solid voxel with loaded air
 * across the face; z, y, x, face order) and keep the bound set equal to the
 * exposed set through later edits: covered faces leave (the last bound face
 * takes the place), newly exposed faces join at the end. Needs the resident
 * cache. The reference discovers these faces by scanning changed pages for
 * surface probes (world.rs scan_surface_probe_pages, pages.rs:451-458). */
int vgi_bind_exposed(vgi_tracer *tracer);
uint32_t vgi_bound_count(vgi_tracer *tracer);
/* The bound faces' 32-byte tasks in bound order (sequence words as the last
 * prepare left them). */
int vgi_read_bound_tasks(vgi_tracer *tracer, vgi_face_task *out, uint32_t count);
/* The encoded retention events the last resident batch or frame applied
 * (edits, page loads and unloads, change footprints, sun flips; none: 0
 * words), after the GPU's finish_edits filled in measured flux
 * (resident.rs:1326-1417 layout); returns the word count, copies at most
 * `capacity` words when out is not NULL. */
int64_t vgi_read_events(vgi_tracer *tracer, uint32_t *out, uint32_t capacity);
/* Change footprints (resident.rs:1215-1310, voxel_trace.wgsl:94-202): each
 * edit's retention events carry one Light event per edited box (at most 64),
 * whose bound and flux the GPU traces along the lights before the events
 * apply. On by default; off plans the CPU's events alone (gpu.rs:419-425,
 * without_footprints). */
int vgi_set_footprints(vgi_tracer *tracer, uint32_t on);
/* Sun corner words of the bound faces (sky_corner_visibility,
 * voxel_path.wgsl:245-266; the view tier's per-face words and
 * report_sun_flips, voxel_view_runtime.wgsl:168-268, here for the world's
 * resident faces): per face four corner bytes (bit 7 known, bit i sees
 * coloured directional light i, at most four lights), kept per resident slot.
 * A face whose word is known is re-traced every `refresh_frames` calls
 * (staggered by bound index: (frame + index) % refresh_frames == 0; 0 or 1:
 * every call); a face facing no directional light gets 0x80808080. Faces whose
 * known corners flipped report the light gained or lost (flux); the flips are
 * binned per 16^3 region and published as one Light retention event at most
 * every 250 ms of the tracer clock (sun_changes.rs:18-69), applied with the
 * next resident batch. Synchronous. Requires vgi_enable_resident and bound
 * faces. */
int vgi_sun_corners(vgi_tracer *tracer, uint32_t frame, uint32_t refresh_frames);
/* The last vgi_sun_corners call's flip region (gpu/view/runtime.rs:556-561
 * layout: count, 2,048 records (voxel xyz, flux bits) from word 4 in
 * bound-face order, then the overflow box (2^30 - min, max + 2^30 as maxima)
 * and its flux x 64): returns its word count (8,204), copies at most
 * `capacity` words when out is not NULL. */
int64_t vgi_read_sun_flips(vgi_tracer *tracer, uint32_t *out, uint32_t capacity);
/* The last vgi_sun_corners call (as vgi_timings for batches, gpu.rs:486-497):
 * [0] GPU ms of its passes (corner words and the flip compaction), [1] host
 * ms of decoding, binning and the event. */
int vgi_sun_timings(vgi_tracer *tracer, float *out, uint32_t n);
/* Corner words of resident slots first .. first + count - 1 (0: not traced;
 * the view tier keeps the same word per receiver, estimate.identity.z,
 * voxel_view_runtime.wgsl:175-198). */
int vgi_read_sun_corners(vgi_tracer *tracer, uint32_t first, uint32_t count, uint32_t *out);
/* ---- view tier: camera-visible faces (voxel_view_discovery.wgsl:4-130,
 * voxel_view_runtime.wgsl:94-324, voxel_view_cache.wgsl:1-258; render_view
 * gpu/view/runtime.rs:749,1254-1266). A frame takes one camera's depth and
 * normal prepass (host memory, row-major, x fastest), discovers each pixel's
 * exposed voxel face (the decoded point with its normal axis snapped to the
 * face plane; within one pixel's footprint of an in-plane voxel boundary the
 * exposed neighbour), gives every face one owner -- the smallest pixel index
 * among its pixels -- through a per-camera CAS face hash, carries each
 * owner's estimate from the previous frame's owner of the same face (same
 * material), exchanges it both ways with the resident world cache
 * (load_shared_faces / store_shared_faces: an admitted face always loads the
 * authority, a face entering the view loads it when the authority holds at
 * least as many samples, and an admitted face stores its estimate back), and
 * traces `samples` paths per owner with the world tier's estimator (running
 * mean with weight 1 / min(count + 1, 65536), the measured-retention keep
 * rule on a pending change) plus its direct term. Non-blocking, ordered after
 * the tracer's earlier work. Not covered: coarse LOD blocks, the traced
 * discovery fallback for foreign meshes (such pixels own no receiver),
 * pending geometry, corner sun words and sun flips of view faces, topology
 * priors. After tracing, the presentation (reconstruct_view,
 * voxel_view_runtime.wgsl:350-474) pools coplanar owners into each face's
 * shown irradiance and the frame publishes one record per pixel
 * (publish_view :476-521). With the far tier on (vgi_view_far), faces small on
 * screen own no receiver and are lit per pixel instead
 * (voxel_view_pixels.wgsl): cells of size x size pixels sample one pixel a
 * frame, reproject their history by world position and trace one path when
 * due. ---- */
typedef struct vgi_view_camera {
    float world_from_clip[16]; /* column-major: element (row r, column c) at [4 c + r]; clip y up, depth z > 0 valid */
    float position[3];         /* camera position (world) */
    float inner_distance;      /* points closer than this own no receiver (0: none) */
    uint32_t width;            /* prepass pixels per row */
    uint32_t height;           /* prepass rows */
    uint32_t samples;          /* paths per visible face this frame (<= 1024) */
    float measured_retention;  /* the keep rule's c (settings.rs:262) */
    float clip_from_world[16]; /* this frame's world -> clip (the next frame reprojects far cells with it), column-major */
    float forward[3];          /* view forward (world, unit): the far tier's distance test */
    float pixels_per_unit;     /* prepass pixels per world unit at view depth 1 (clip_from_view[1][1] * height / 2) */
    uint32_t frame;            /* frame number: far cells are stamped with it (gpu/view/runtime.rs:994-1002) */
} vgi_view_camera;
/* One prepass pixel's receiver and, for an owner, its face's estimate. */
typedef struct vgi_view_face {
    int32_t voxel[3];
    uint32_t face;             /* 0xffffffff: the pixel has no receiver */
    float surface[3];          /* face centre */
    uint32_t sequence;         /* next path sequence */
    float mean[4];             /* raw incident rgb, effective sample weight */
    float moment[4];           /* raw second moment rgb, pending predicted change */
    uint32_t identity[4];      /* next sequence, material, corner word (0), resident incarnation (0: not admitted) */
    float direct[4];           /* direct rgb at the face centre, 0 */
    float shown[4];            /* presented rgb (reconstructed), valid: 1 (0: never shown) */
} vgi_view_face;
/* publish_view's record per pixel (FarPixel, voxel_view_runtime.wgsl:40-46). */
typedef struct vgi_view_pixel {
    float surface[3];          /* owner's face centre */
    uint32_t face;             /* 0xffffffff: the pixel is no face's owner */
    float irradiance[3];       /* shown rgb plus custom direct rgb */
    float evidence;            /* the sample weight behind it (>= 0.001), 0: nothing to show */
    uint32_t sky[4];           /* corner sun words: lit lights (bits 0-3), known (bit 31) */
} vgi_view_pixel;
/* The far tier (FarTierSettings, settings.rs:95-160; gpu/view/runtime.rs:
 * 912-1000). All zero (the default at vgi_create) is off. */
typedef struct vgi_far_settings {
    float voxel_pixels;        /* a face is far under this many pixels across (measure 0) or covered (1, 2); 0: off */
    uint32_t measure;          /* 0 voxel side, 1 voxel area, 2 block area (unit voxels: as 1) */
    uint32_t cell;             /* pixels per cell side 1-4; 0: round(height / 480) clamped to 1-4 */
    float history;             /* history cap in samples (1-1024) */
    uint32_t rate;             /* once a cell holds 4 samples, it traces one frame in `rate` (1-8) */
    uint32_t steps;            /* traversal steps per far ray (1-2048); a ray out of steps is a miss */
} vgi_far_settings;
#define VGI_VIEW_MAX_PIXELS 16777216u
int vgi_view_frame(vgi_tracer *tracer, const vgi_view_camera *camera, const float *depth, const float *normals);
/* GPU ms of the last view frame: [all its passes, the owners' tracing, the far tier]. */
int vgi_view_timings(vgi_tracer *tracer, float out[3]);
/* The far tier's settings for the next view frames (the renderer's
 * settings.far_tier, normalized; a NULL pointer or voxel_pixels 0 turns it off). */
int vgi_view_far(vgi_tracer *tracer, const vgi_far_settings *settings);
/* The presentation's settings (view.display.xyz, gpu/view/runtime.rs:962-967):
 * the relative standard error under which a face is shown alone
 * (spatial_noise_threshold, default 0.031), the spatial reconstruction
 * strength (default 1), the temporal blend 1 - exp(-frame seconds / display
 * response) (default 0.1993: 60 frames a second, 75 ms). */
int vgi_view_display(vgi_tracer *tracer, float noise_threshold, float spatial, float blend);
/* Published records of pixels [first, first + count) of the last view frame. */
int vgi_view_read_pixels(vgi_tracer *tracer, uint32_t first, uint32_t count, vgi_view_pixel *out);
/* The last view frame's counters: far tier (owned, traced, history, reprojected,
 * seeded from receivers, receivers seeded from cells, 0, 0; voxel_view_pixels.wgsl:34-39),
 * presentation (history, max history, uncovered, displayed, 0, 0, 0, 0;
 * voxel_view_runtime.wgsl:55-58). */
int vgi_view_read_counters(vgi_tracer *tracer, uint32_t out[16]);
/* The far tier's output of the last view frame (the material's buffer: 16
 * header words -- camera, axis, grid (cells x, y, size, flags), origin
 * (0, 0, threshold bits, stamp) -- then per cell half rg, half (b, count),
 * face | stamp << 3, surface coordinate along the normal axis) and the
 * frame's cells (12 words each: mean rgb count, surface, face | stamp << 3,
 * sequence, 0, 0, 0). Cells not written this frame keep older contents (the
 * stamp tells). Either pointer may be NULL. Returns the number of cells. */
int64_t vgi_view_read_far(vgi_tracer *tracer, uint32_t *output, uint64_t output_words, uint32_t *cells, uint64_t cell_words);
/* Forget the view history (the next frame reprojects nothing). */
int vgi_view_reset(vgi_tracer *tracer);
/* The last view frame's owner per pixel (pixel index of the face's owner,
 * 0xffffffff: none) for the first `pixels` pixels; returns the number of
 * owners (faces); *overflow (optional) counts faces the hash could not place. */
int64_t vgi_view_read_owners(vgi_tracer *tracer, uint32_t *out, uint32_t pixels, uint32_t *overflow);
/* Receivers and estimates of pixels [first, first + count) of the last view frame. */
int vgi_view_read_faces(vgi_tracer *tracer, uint32_t first, uint32_t count, vgi_view_face *out);

/* ---- GPU work selection (voxel_accumulate.wgsl:448-695; resident.rs:1093-1164) ---- */
/* Scheduling settings (TraceWork.{focus, transport_revision, stable_error,
 * stable_samples, audit_ms, mirror_budget}, work.rs:38-50; the plugin fills
 * them from settings.rs:337-357). Every resident batch's accumulate ranks its
 * faces with them (record_stability, voxel_accumulate.wgsl:99-130) and GPU
 * pick batches classify by them. All zero until set (TraceWork::default()). */
typedef struct vgi_schedule {
    int32_t focus[3];
    uint32_t transport_revision; /* low 32 bits */
    float stable_error;
    uint32_t stable_samples;
    uint32_t audit_ms;
    uint32_t mirror_budget;      /* statistics-mirror records per pick batch (at least 1) */
} vgi_schedule;
int vgi_set_schedule(vgi_tracer *tracer, const vgi_schedule *schedule);
/* A resident batch whose faces the GPU picks itself (TraceWork.picks with no
 * host faces): after queued retention, the selection passes rank every
 * resident slot (classification re-ranks one quarter per batch) and pick up
 * to `picks` faces (urgent first, then coverage 1/4, refinement 1/2, sweep
 * 1/4); the picks are traced and merged like host faces. Non-blocking;
 * flags as vgi_run_frame (counters read with vgi_read_bound's counters). */
int vgi_select_frame(vgi_tracer *tracer, uint32_t picks, uint32_t flags);
/* GPU ms of the last pick batch: [selection passes, trace, accumulate]. */
int vgi_select_timings(vgi_tracer *tracer, float *out);
/* The last pick batch's pick list in pick order: resident slots, 0xffffffff
 * for an unused place; *count = the budget. */
int vgi_read_picks(vgi_tracer *tracer, uint32_t *slots, uint32_t capacity, uint32_t *count);
/* The selection scratch of the last pick batch (up to 1038 words): four
 * 256-bucket histograms (coverage, refinement, sweep, urgent), per class the
 * threshold bucket (256: none), the remainder taken at it and the ties seen
 * at it, then the pick total and the mirror-record count. */
int vgi_read_selection(vgi_tracer *tracer, uint32_t *words, uint32_t count);
/* Selection keys (class << 8 | bucket, 0xffffffff: none) of slots [first, first+count). */
int vgi_read_keys(vgi_tracer *tracer, uint32_t first, uint32_t count, uint32_t *keys);
/* Statistics-mirror records of the last pick batch, 28 words each: (slot,
 * incarnation, revision tag, 0) and the face's six 16-byte result words, the
 * settled flag in bit 31 of word 5 (voxel_accumulate.wgsl:537-550). */
int vgi_read_mirror(vgi_tracer *tracer, uint32_t *words, uint32_t capacity, uint32_t *count);
/* Slot metadata (3 uint4 per slot: identity, schedule, ranking; voxel_accumulate.wgsl:36-41) of slots [first, first+count). */
int vgi_cache_meta(vgi_tracer *tracer, uint32_t first, uint32_t count, uint32_t *out);
/* Raw resident state (9 uint4 per slot, voxel_accumulate.wgsl:1-4) of slots [first, first+count). */
int vgi_cache_state(vgi_tracer *tracer, uint32_t first, uint32_t count, uint32_t *out);

/* ---- error harness: the long-horizon reference. Traces the bound faces
 * `launches` times (sequences first_sequence + l * samples) with plain
 * multi-bounce paths (no cache termination) and sums each face's per-launch
 * moments in FP64 on the GPU: out = 8 doubles per bound face (sum of mean*known
 * rgb, sum of second*known rgb, sum of known, 0). The analytic direct term is
 * deterministic and read from any record. ---- */
int vgi_converged(vgi_tracer *tracer, uint32_t launches, uint32_t first_sequence, double *out);

/* ---- learned radiance cache (no counterpart in the reference): a
 * 32-64-64-64-64-3 ReLU MLP (BF16 MFMA, FP32 accumulation, Adam) that predicts
 * a face's incident radiance from a 32-dim encoding of its centre (a trainable
 * multiresolution hash grid, 8 levels x 2 features, and sin/cos of pi u per
 * axis), its normal and albedo. With it on, bound-face frames end every path at its
 * first hit into the network (the world-cache termination rule of
 * voxel_path.wgsl:459-463 with the prediction as the cached value) and train
 * it on `train_per_frame` plain multi-bounce samples per frame in `steps`
 * Adam steps. Queries are 4 u32 (voxel xyz, face | material << 8). ---- */
int vgi_mlp_enable(vgi_tracer *tracer, uint32_t seed, uint32_t train_per_frame, uint32_t steps, float lr);
int vgi_mlp_set_learned(vgi_tracer *tracer, int on);
/* Read (out) and/or write (in) the 14,528 FP32 master weights: W1 64x32,
 * W2..W4 64x64, W5 3x64, row-major (out x in), in that order. */
int vgi_mlp_weights(vgi_tracer *tracer, float *out, const float *in);
int vgi_mlp_infer(vgi_tracer *tracer, const uint32_t *queries, uint32_t count, float *out_rgb);
/* One forward/backward on (query, target rgb + weight) pairs, count a multiple
 * of 16; update 0 returns the summed gradient (relative L2 loss), 1 takes an
 * Adam step. */
int vgi_mlp_train(vgi_tracer *tracer, const uint32_t *queries, const float *targets, uint32_t count, int update,
                  float *grad_out, float *loss_out);
/* The hash grid: read (out) and/or write (in) its 67,108,864 FP32 entries
 * (level-major, 2^22 entries of 2 features per level), and copy the table
 * gradient of the last vgi_mlp_train without update (grad_out). */
int vgi_mlp_table(vgi_tracer *tracer, float *out, const float *in, float *grad_out);
/* The hash table's Adam learning rate (ten times the network's by default). */
int vgi_mlp_table_lr(vgi_tracer *tracer, float lr);
/* Paths averaged into each training target (1..64; 1 by default). */
int vgi_mlp_train_samples(vgi_tracer *tracer, uint32_t samples);
/* Change the training schedule of an enabled cache without touching the
 * network or its Adam moments: training targets and Adam steps per frame and
 * the network's rate (the table's becomes ten times it). 0 targets freezes
 * the network: learned frames then only infer. */
int vgi_mlp_training(vgi_tracer *tracer, uint32_t train_per_frame, uint32_t steps, float lr);
/* The position encoding of an enabled cache: 0 (the default) the hash grid
 * described above; 1 frequencies only -- sin and cos of 2^k pi u per axis,
 * k = 0..3, beside the normal and albedo: the 32-dim frequency encoding of
 * the network stated for the 512^3 cave configuration; the hash table is
 * then neither read nor trained. Kept across vgi_mlp_enable. */
#define VGI_MLP_ENCODING_HASH 0u
#define VGI_MLP_ENCODING_FREQUENCY 1u
int vgi_mlp_encoding(vgi_tracer *tracer, uint32_t mode);
/* Training-only frames on the bound faces (targets traced, Adam steps taken,
 * bound faces not traced): the network's warm-up. */
int vgi_mlp_pretrain(vgi_tracer *tracer, uint32_t frames);
float vgi_mlp_last_loss(vgi_tracer *tracer);
/* GPU time of the last learned frame's network work (valid after it ran):
 * [inference ms, training MFMA ms (forward/backward and weight-gradient
 * passes of every step), training ms (the steps with their Adam updates),
 * queries inferred, training samples (all steps), steps]. */
int vgi_mlp_timings(vgi_tracer *tracer, float out[6]);

/* ---- retention events (resident.rs:1326-1417 encode_events layout;
 * voxel_retention.wgsl:29-96, apply_events voxel_accumulate.wgsl:295-332) ---- */
int vgi_apply_events(vgi_tracer *tracer, const uint32_t *events, uint32_t event_count, uint32_t words);

/* ---- batched frames (no counterpart in the reference, which runs one world;
 * docs/architecture.md:294): many tracers -- separate scenes, or chunks of
 * uneven face counts bound to several tracers -- traced in ONE persistent
 * launch per frame that takes (scene, face range) tiles from a device queue,
 * captured once into a HIP graph and replayed. Each scene's bound faces are
 * traced exactly as vgi_run_frame would trace them (non-resident: records go
 * to the tracer's bound records, read with vgi_read_bound), with sequence
 * base set by vgi_batch_set_sequence and advancing by the samples per face
 * every frame. All scenes trace the same samples per face. ---- */
int vgi_batch_create(int device_ordinal, vgi_batch **out);
void vgi_batch_destroy(vgi_batch *batch);
/* Add a tracer whose faces are bound (vgi_bind_faces / vgi_bind_exposed). */
int vgi_batch_add(vgi_batch *batch, vgi_tracer *tracer);
/* Workgroup size (64/128/256), lanes per face (1, 2, 4 or 8: one lane, at
 * least one lane per sample, or the reference's count, so sums keep the
 * reference's order), queue tile in faces, workgroups (0: enough to fill
 * every CU at the kernel's occupancy). */
int vgi_batch_configure(vgi_batch *batch, uint32_t block, uint32_t lanes, uint32_t tile, uint32_t grid);
int vgi_batch_set_sequence(vgi_batch *batch, uint32_t base);
int vgi_batch_run_frame(vgi_batch *batch);
int vgi_batch_sync(vgi_batch *batch);
/* GPU time of the last frame (ms), valid after vgi_batch_sync. */
float vgi_batch_last_ms(vgi_batch *batch);
/* Per scene in the order added: ms from the frame's start until its last
 * tile ended; rays traced in the frame (optional). */
int vgi_batch_latency(vgi_batch *batch, float *out_ms, uint32_t count, uint64_t *rays);
/* Time every allowed (workgroup, lanes, tile) on `frames` frames after a
 * warm-up and keep the fastest: out = block, lanes, tile, median ms (float
 * bits), candidates tried. */
int vgi_batch_autotune(vgi_batch *batch, uint32_t frames, uint32_t out[5]);

/* ---- measurement plumbing: a plain copy kernel of `bytes` run `repeats`
 * times (average ms per launch in *ms), the known-size reference that scales
 * hardware byte counters (not part of the reference seam). ---- */
int vgi_probe_copy(vgi_tracer *tracer, uint64_t bytes, uint32_t repeats, float *ms);

/* ---- scene inspection for parity checks ---- */
int vgi_scene_bounds(vgi_tracer *tracer, int32_t out[6]);
/* The canonical per-region form (see oracle): kind, material, top-level
 * bytes, 64 brick words (pointer replaced by tag 2), 64x32 brick cell words. */
int vgi_scene_region_canonical(vgi_tracer *tracer, int32_t rx, int32_t ry, int32_t rz, uint32_t *out);
/* [loaded regions, solid regions, cell words, top-level cells, resident
 * slots, emissive faces, light-tree nodes, pinned staging (allocations << 32 |
 * times the region started over)] */
int64_t vgi_scene_stats(vgi_tracer *tracer, int64_t out[8]);
int64_t vgi_emitter_words(vgi_tracer *tracer, uint32_t *out, int64_t capacity);

const char *vgi_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
