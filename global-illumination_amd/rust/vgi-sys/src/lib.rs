// Generated by global-illumination_amd/rust/gen_sys.py from include/vgi.h; do not edit.
//! Raw bindings of `libvgi.so`, the MI355X backend of bevy_voxel_gi's GPU
//! transport seam (`GpuTracer`, crates/bevy_voxel_gi/src/gpu.rs:309-1236).
//! Every item mirrors include/vgi.h one to one; the safe adapter lives in
//! the `bevy_voxel_gi_amd` crate.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int};

/// world.rs:19 AIR: u16 material sentinel
pub const VGI_AIR: u32 = 256;

/// world.rs:22
pub const VGI_PAGE_EDGE: u32 = 32;

/// voxel_trace.wgsl:64 unused pick capacity
pub const VGI_NO_FACE: u32 = 0xffffffff;

#[repr(C)]
pub struct vgi_tracer {
    _private: [u8; 0],
}

#[repr(C)]
pub struct vgi_batch {
    _private: [u8; 0],
}

/// gpu.rs:136-141 GpuLight: position_range.w < 0 marks a directional source
/// whose xyz is the unit direction toward the light (sources.rs:130).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct vgi_light {
    pub position_range: [f32; 4],
    pub energy: [f32; 4],
    pub extra: [f32; 4],
}

/// gpu.rs:144-177 / voxel_trace.wgsl:5-12 FaceTask, 32 bytes.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct vgi_face_task {
    pub voxel: [i32; 3],
    pub face: u32,
    pub sequence: u32,
    pub padding: [u32; 3],
}

/// voxel_trace.wgsl:13-20 FaceResult, 96 bytes (gpu.rs:1072-1106 decode).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct vgi_face_result {
    pub mean: [f32; 4],
    pub mean_extra: [f32; 4],
    pub second: [f32; 4],
    pub second_extra: [f32; 4],
    pub direct: [f32; 4],
    pub direct_extra: [f32; 4],
}

/// TransportSettings (settings.rs:391-426) + TraceWork history fields
/// (work.rs:11-55) + the WGSL override constants (voxel_path.wgsl:4-14).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct vgi_transport {
    pub bounces: u32,
    pub samples: u32,
    pub lanes: u32,
    pub limited_reach: u32,
    pub reach: f32,
    pub max_distance: f32,
    pub environment: [f32; 3],
    pub history_limit: u32,
    pub history_recency: f32,
    pub measured_retention: f32,
    pub plain: u32,
}

pub const VGI_RECORD_HITS: u32 = 1;

pub const VGI_COUNT_RAYS: u32 = 2;

/// frames: also steps and traversal gathers (tickets always count all)
pub const VGI_COUNT_GATHERS: u32 = 4;

pub const VGI_NOALLOC_MAX_FACES: u32 = 0x40000;

/// view tier: camera-visible faces (voxel_view_discovery.wgsl:4-130,
/// voxel_view_runtime.wgsl:94-324, voxel_view_cache.wgsl:1-258; render_view
/// gpu/view/runtime.rs:749,1254-1266). A frame takes one camera's depth and
/// normal prepass (host memory, row-major, x fastest), discovers each pixel's
/// exposed voxel face (the decoded point with its normal axis snapped to the
/// face plane; within one pixel's footprint of an in-plane voxel boundary the
/// exposed neighbour), gives every face one owner -- the smallest pixel index
/// among its pixels -- through a per-camera CAS face hash, carries each
/// owner's estimate from the previous frame's owner of the same face (same
/// material), exchanges it both ways with the resident world cache
/// (load_shared_faces / store_shared_faces: an admitted face always loads the
/// authority, a face entering the view loads it when the authority holds at
/// least as many samples, and an admitted face stores its estimate back), and
/// traces `samples` paths per owner with the world tier's estimator (running
/// mean with weight 1 / min(count + 1, 65536), the measured-retention keep
/// rule on a pending change) plus its direct term. Non-blocking, ordered after
/// the tracer's earlier work. Not covered: coarse LOD blocks, the traced
/// discovery fallback for foreign meshes (such pixels own no receiver),
/// pending geometry, corner sun words and sun flips of view faces, topology
/// priors. After tracing, the presentation (reconstruct_view,
/// voxel_view_runtime.wgsl:350-474) pools coplanar owners into each face's
/// shown irradiance and the frame publishes one record per pixel
/// (publish_view :476-521). With the far tier on (vgi_view_far), faces small on
/// screen own no receiver and are lit per pixel instead
/// (voxel_view_pixels.wgsl): cells of size x size pixels sample one pixel a
/// frame, reproject their history by world position and trace one path when
/// due.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct vgi_view_camera {
    pub world_from_clip: [f32; 16],
    pub position: [f32; 3],
    pub inner_distance: f32,
    pub width: u32,
    pub height: u32,
    pub samples: u32,
    pub measured_retention: f32,
    pub clip_from_world: [f32; 16],
    pub forward: [f32; 3],
    pub pixels_per_unit: f32,
    pub frame: u32,
}

/// One prepass pixel's receiver and, for an owner, its face's estimate.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct vgi_view_face {
    pub voxel: [i32; 3],
    pub face: u32,
    pub surface: [f32; 3],
    pub sequence: u32,
    pub mean: [f32; 4],
    pub moment: [f32; 4],
    pub identity: [u32; 4],
    pub direct: [f32; 4],
    pub shown: [f32; 4],
}

/// publish_view's record per pixel (FarPixel, voxel_view_runtime.wgsl:40-46).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct vgi_view_pixel {
    pub surface: [f32; 3],
    pub face: u32,
    pub irradiance: [f32; 3],
    pub evidence: f32,
    pub sky: [u32; 4],
}

/// The far tier (FarTierSettings, settings.rs:95-160; gpu/view/runtime.rs:
/// 912-1000). All zero (the default at vgi_create) is off.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct vgi_far_settings {
    pub voxel_pixels: f32,
    pub measure: u32,
    pub cell: u32,
    pub history: f32,
    pub rate: u32,
    pub steps: u32,
}

pub const VGI_VIEW_MAX_PIXELS: u32 = 0x1000000;

/// Scheduling settings (TraceWork.{focus, transport_revision, stable_error,
/// stable_samples, audit_ms, mirror_budget}, work.rs:38-50; the plugin fills
/// them from settings.rs:337-357). Every resident batch's accumulate ranks its
/// faces with them (record_stability, voxel_accumulate.wgsl:99-130) and GPU
/// pick batches classify by them. All zero until set (TraceWork::default()).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct vgi_schedule {
    pub focus: [i32; 3],
    pub transport_revision: u32,
    pub stable_error: f32,
    pub stable_samples: u32,
    pub audit_ms: u32,
    pub mirror_budget: u32,
}

pub const VGI_MLP_ENCODING_HASH: u32 = 0;

pub const VGI_MLP_ENCODING_FREQUENCY: u32 = 1;

#[link(name = "vgi")]
extern "C" {
    /// lifecycle: GpuTracer::new / enable_resident (gpu.rs:309-366)
    pub fn vgi_create(device_ordinal: c_int, out: *mut *mut vgi_tracer) -> c_int;

    pub fn vgi_destroy(tracer: *mut vgi_tracer);

    pub fn vgi_enable_resident(tracer: *mut vgi_tracer) -> c_int;

    /// The resident clock (ms) that the display EMA reads; negative restores the
    /// wall clock. Mirrors the reference's test clock (gpu/resident.rs:708-710).
    pub fn vgi_set_clock(tracer: *mut vgi_tracer, milliseconds: i64) -> c_int;

    /// scene: VoxelGiInput::{page_loaded, page_unloaded} (input.rs:54-112) feeding
    /// TransportScene::with_page_contents (world.rs:982-1046). Pages are 32^3 u16
    /// voxels, x fastest, VGI_AIR for air. The NAADF brick/voxel fields are built on
    /// the GPU; the top level is grown in exact FIFO order on the host. With the
    /// resident cache, an upload that changes a loaded scene queues its retention
    /// event for the next batch or frame: regions that load are Load changes of
    /// their whole box with their top solid row and the sky's flux, regions that
    /// unload are Unload changes, regions of a page sent again with new content
    /// are edits measured on the GPU (SceneChanges::with_materials
    /// geometry.rs:110-229, finalize_change_flux publication.rs:423-454).
    pub fn vgi_upload_pages(tracer: *mut vgi_tracer, page_xyz: *const i32, pages: u32, voxels: *const u16, removed_xyz: *const i32, removed: u32) -> c_int;

    /// VoxelGiInput::edit_voxel (input.rs:90-93): set voxels (in order; a later
    /// edit of a voxel wins); edited regions are re-packed on the GPU and the top
    /// level is updated incrementally (TopLevelAadf::updated, world.rs:477-583).
    /// With the resident cache, the edit's retention event is built
    /// (SceneChanges::with_materials geometry.rs:110-229, GeometryChange::from_bounds
    /// geometry.rs:299-407, finalize_change_flux publication.rs:423-454,
    /// SpatialHistory::event retention.rs:110-167) and queued; the next batch or
    /// frame measures its flux on the GPU (measure_edits / finish_edits,
    /// voxel_accumulate.wgsl:375-430) and applies it to every slot
    /// (apply_events, voxel_accumulate.wgsl:295-332) before its prepare. With
    /// vgi_bind_exposed, bound faces follow the edit (pages.rs:451-458 rescan).
    pub fn vgi_edit_voxels(tracer: *mut vgi_tracer, xyz: *const i32, values: *const u16, count: u32) -> c_int;

    /// optics and sources: VoxelGiMaterials (world.rs:90-174) as packed by
    /// gpu.rs:714-727; TransportSources (sources.rs:105-160) as packed by
    /// gpu.rs:728-760
    pub fn vgi_set_materials(tracer: *mut vgi_tracer, diffuse_rgb: *const f32, emission_rgb: *const f32) -> c_int;

    pub fn vgi_set_lights(tracer: *mut vgi_tracer, lights: *const vgi_light, count: u32, directional: u32) -> c_int;

    pub fn vgi_set_transport(tracer: *mut vgi_tracer, transport: *const vgi_transport) -> c_int;

    /// batches: GpuTracer::submit_compact (gpu.rs:647-1022). Non-resident:
    /// trace `count` tasks as given. With the resident cache on, each face gets
    /// its slot (allocating as resident.rs:770-812), takes the slot's sequence
    /// (prepare, voxel_accumulate.wgsl:83-133) and is merged (accumulate,
    /// voxel_accumulate.wgsl:135-227). Flags: VGI_RECORD_HITS keeps the per
    /// face-sample first-hit records; VGI_COUNT_RAYS counts rays and steps.
    pub fn vgi_submit_faces(tracer: *mut vgi_tracer, tasks: *const vgi_face_task, count: u32, flags: u32, ticket: *mut u64) -> c_int;

    /// No counterpart in the reference. For callers that cannot allocate device
    /// memory (code under HIP graph capture elsewhere in the process, tight
    /// serving loops, harnesses that forbid hipMalloc): the trace of
    /// vgi_submit_faces + vgi_read_faces on a tracer without the resident cache,
    /// synchronous, with the tasks and records staged in a fixed device scratch
    /// region (VGI_NOALLOC_MAX_FACES faces, 128 bytes each: 32 MiB per GPU, held
    /// until the process ends) instead of buffers the call allocates. The same
    /// records bit for bit. It allocates nothing: more faces than the maximum, a
    /// resident tracer, or an emitter table that would have to grow are errors
    /// (set-up calls allocate; run one vgi_submit_faces after a scene or material
    /// change). Every call on one GPU shares that GPU's scratch: calls that
    /// overlap on one GPU (several threads, several tracers) race.
    pub fn vgi_trace_faces_noalloc(tracer: *mut vgi_tracer, tasks: *const vgi_face_task, count: u32, out: *mut vgi_face_result) -> c_int;

    /// GpuTracer::snapshot_ready (gpu.rs:1024-1049)
    pub fn vgi_poll(tracer: *mut vgi_tracer, ticket: u64, ready: *mut c_int) -> c_int;

    /// GpuTracer::complete_readback (gpu.rs:1050-1170): 96-byte records.
    pub fn vgi_read_faces(tracer: *mut vgi_tracer, ticket: u64, out: *mut vgi_face_result, count: u32) -> c_int;

    /// First-hit record per face-sample (cf. voxel_view_trace.wgsl:21-38): 8 u32:
    /// voxel xyz, face | status << 8, material, cache slot, rays of the sample,
    /// known. Requires VGI_RECORD_HITS.
    pub fn vgi_read_hits(tracer: *mut vgi_tracer, ticket: u64, out: *mut u32, words: u64) -> c_int;

    /// [rays, steps, header loads, brick-word loads, cell-word loads, 0...] of the ticket (VGI_COUNT_RAYS).
    pub fn vgi_read_counters(tracer: *mut vgi_tracer, ticket: u64, out: *mut u64, count: u32) -> c_int;

    /// GPU timestamps (gpu.rs:486-497,1050-1070): ms of [prepare, trace, accumulate,
    /// build]; count up to 8 adds the last edit's host ms: [regions + GPU re-pack,
    /// top level, retention event, exposure update].
    pub fn vgi_timings(tracer: *mut vgi_tracer, out: *mut f32, count: u32) -> c_int;

    /// Resident slots of the last batch's faces (face_index.rs:71-97 lookups).
    pub fn vgi_read_slots(tracer: *mut vgi_tracer, out: *mut u32, count: u32) -> c_int;

    /// device-resident frames (no host transfer): the face set is bound once
    /// and each frame runs prepare -> trace -> accumulate on it, as the reference's
    /// worker does batch after batch.
    pub fn vgi_bind_faces(tracer: *mut vgi_tracer, tasks: *const vgi_face_task, count: u32) -> c_int;

    pub fn vgi_run_frame(tracer: *mut vgi_tracer, flags: u32) -> c_int;

    pub fn vgi_sync(tracer: *mut vgi_tracer) -> c_int;

    /// Time of the last frame's trace launch on its stream (ms), valid after vgi_sync.
    pub fn vgi_last_trace_ms(tracer: *mut vgi_tracer) -> f32;

    /// GPU time of the whole last frame (prepare, trace or learned trace +
    /// inference, accumulate, and the learned cache's training), valid after
    /// vgi_sync.
    pub fn vgi_last_frame_ms(tracer: *mut vgi_tracer) -> f32;

    /// Bound-face frames by wavefront path tracing (on) or by the one-kernel
    /// estimator (off): the same records bit for bit. Frames that count gathers
    /// (VGI_COUNT_GATHERS), learned frames and more than 5 lights use the
    /// one-kernel estimator either way.
    pub fn vgi_set_wavefront(tracer: *mut vgi_tracer, on: c_int) -> c_int;

    /// The wavefront tracer's hit-face deduplication (on) or not (off, the
    /// default): before a shading pass, the world-cache lookups of its new hits
    /// (cached_energy, voxel_path.wgsl:63-87) are matched per wavefront (ballot
    /// over the distinct (voxel, face) keys) and each distinct face is probed
    /// once, its answer shared across the wavefront. The same records bit for
    /// bit; with VGI_COUNT_RAYS the frame's counters 5 and 6 receive the lookups
    /// and the probes made.
    pub fn vgi_set_dedup(tracer: *mut vgi_tracer, on: c_int) -> c_int;

    /// The bound faces' 96-byte records after the last frame (out may be NULL with
    /// count 0 to read only the counters); counters (optional)
    /// receive [rays, steps, header loads, brick-word loads, cell-word loads,
    /// deduplicated world-cache lookups, their probes] when the frame ran with
    /// VGI_COUNT_RAYS. Records follow the faces bound now: after
    /// an edit that moved bound faces (vgi_bind_exposed), including one issued
    /// while the frame still ran, each face reads its record of the last frame,
    /// and a face bound since reads zeros -- the same records in either call
    /// order.
    pub fn vgi_read_bound(tracer: *mut vgi_tracer, out: *mut vgi_face_result, count: u32, counters: *mut u64) -> c_int;

    /// First-hit records of the last bound-face frame run with VGI_RECORD_HITS (as
    /// vgi_read_hits: 8 u32 per face-sample, bound face i's samples at
    /// (i * samples + s) * 8; cf. voxel_view_trace.wgsl:21-38, the reference's
    /// debug trace of the same frame). Learned frames keep none.
    pub fn vgi_read_bound_hits(tracer: *mut vgi_tracer, out: *mut u32, words: u64) -> c_int;

    /// Bind every exposed face of the loaded scene (a solid voxel with loaded air
    /// across the face; z, y, x, face order) and keep the bound set equal to the
    /// exposed set through later edits: covered faces leave (the last bound face
    /// takes the place), newly exposed faces join at the end. Needs the resident
    /// cache. The reference discovers these faces by scanning changed pages for
    /// surface probes (world.rs scan_surface_probe_pages, pages.rs:451-458).
    pub fn vgi_bind_exposed(tracer: *mut vgi_tracer) -> c_int;

    pub fn vgi_bound_count(tracer: *mut vgi_tracer) -> u32;

    /// The bound faces' 32-byte tasks in bound order (sequence words as the last
    /// prepare left them).
    pub fn vgi_read_bound_tasks(tracer: *mut vgi_tracer, out: *mut vgi_face_task, count: u32) -> c_int;

    /// The encoded retention events the last resident batch or frame applied
    /// (edits, page loads and unloads, change footprints, sun flips; none: 0
    /// words), after the GPU's finish_edits filled in measured flux
    /// (resident.rs:1326-1417 layout); returns the word count, copies at most
    /// `capacity` words when out is not NULL.
    pub fn vgi_read_events(tracer: *mut vgi_tracer, out: *mut u32, capacity: u32) -> i64;

    /// Change footprints (resident.rs:1215-1310, voxel_trace.wgsl:94-202): each
    /// edit's retention events carry one Light event per edited box (at most 64),
    /// whose bound and flux the GPU traces along the lights before the events
    /// apply. On by default; off plans the CPU's events alone (gpu.rs:419-425,
    /// without_footprints).
    pub fn vgi_set_footprints(tracer: *mut vgi_tracer, on: u32) -> c_int;

    /// Sun corner words of the bound faces (sky_corner_visibility,
    /// voxel_path.wgsl:245-266; the view tier's per-face words and
    /// report_sun_flips, voxel_view_runtime.wgsl:168-268, here for the world's
    /// resident faces): per face four corner bytes (bit 7 known, bit i sees
    /// coloured directional light i, at most four lights), kept per resident slot.
    /// A face whose word is known is re-traced every `refresh_frames` calls
    /// (staggered by bound index: (frame + index) % refresh_frames == 0; 0 or 1:
    /// every call); a face facing no directional light gets 0x80808080. Faces whose
    /// known corners flipped report the light gained or lost (flux); the flips are
    /// binned per 16^3 region and published as one Light retention event at most
    /// every 250 ms of the tracer clock (sun_changes.rs:18-69), applied with the
    /// next resident batch. Synchronous. Requires vgi_enable_resident and bound
    /// faces.
    pub fn vgi_sun_corners(tracer: *mut vgi_tracer, frame: u32, refresh_frames: u32) -> c_int;

    /// The last vgi_sun_corners call's flip region (gpu/view/runtime.rs:556-561
    /// layout: count, 2,048 records (voxel xyz, flux bits) from word 4 in
    /// bound-face order, then the overflow box (2^30 - min, max + 2^30 as maxima)
    /// and its flux x 64): returns its word count (8,204), copies at most
    /// `capacity` words when out is not NULL.
    pub fn vgi_read_sun_flips(tracer: *mut vgi_tracer, out: *mut u32, capacity: u32) -> i64;

    /// The last vgi_sun_corners call (as vgi_timings for batches, gpu.rs:486-497):
    /// [0] GPU ms of its passes (corner words and the flip compaction), [1] host
    /// ms of decoding, binning and the event.
    pub fn vgi_sun_timings(tracer: *mut vgi_tracer, out: *mut f32, n: u32) -> c_int;

    /// Corner words of resident slots first .. first + count - 1 (0: not traced;
    /// the view tier keeps the same word per receiver, estimate.identity.z,
    /// voxel_view_runtime.wgsl:175-198).
    pub fn vgi_read_sun_corners(tracer: *mut vgi_tracer, first: u32, count: u32, out: *mut u32) -> c_int;

    pub fn vgi_view_frame(tracer: *mut vgi_tracer, camera: *const vgi_view_camera, depth: *const f32, normals: *const f32) -> c_int;

    /// GPU ms of the last view frame: [all its passes, the owners' tracing, the far tier].
    pub fn vgi_view_timings(tracer: *mut vgi_tracer, out: *mut f32) -> c_int;

    /// The far tier's settings for the next view frames (the renderer's
    /// settings.far_tier, normalized; a NULL pointer or voxel_pixels 0 turns it off).
    pub fn vgi_view_far(tracer: *mut vgi_tracer, settings: *const vgi_far_settings) -> c_int;

    /// The presentation's settings (view.display.xyz, gpu/view/runtime.rs:962-967):
    /// the relative standard error under which a face is shown alone
    /// (spatial_noise_threshold, default 0.031), the spatial reconstruction
    /// strength (default 1), the temporal blend 1 - exp(-frame seconds / display
    /// response) (default 0.1993: 60 frames a second, 75 ms).
    pub fn vgi_view_display(tracer: *mut vgi_tracer, noise_threshold: f32, spatial: f32, blend: f32) -> c_int;

    /// Published records of pixels [first, first + count) of the last view frame.
    pub fn vgi_view_read_pixels(tracer: *mut vgi_tracer, first: u32, count: u32, out: *mut vgi_view_pixel) -> c_int;

    /// The last view frame's counters: far tier (owned, traced, history, reprojected,
    /// seeded from receivers, receivers seeded from cells, 0, 0; voxel_view_pixels.wgsl:34-39),
    /// presentation (history, max history, uncovered, displayed, 0, 0, 0, 0;
    /// voxel_view_runtime.wgsl:55-58).
    pub fn vgi_view_read_counters(tracer: *mut vgi_tracer, out: *mut u32) -> c_int;

    /// The far tier's output of the last view frame (the material's buffer: 16
    /// header words -- camera, axis, grid (cells x, y, size, flags), origin
    /// (0, 0, threshold bits, stamp) -- then per cell half rg, half (b, count),
    /// face | stamp << 3, surface coordinate along the normal axis) and the
    /// frame's cells (12 words each: mean rgb count, surface, face | stamp << 3,
    /// sequence, 0, 0, 0). Cells not written this frame keep older contents (the
    /// stamp tells). Either pointer may be NULL. Returns the number of cells.
    pub fn vgi_view_read_far(tracer: *mut vgi_tracer, output: *mut u32, output_words: u64, cells: *mut u32, cell_words: u64) -> i64;

    /// Forget the view history (the next frame reprojects nothing).
    pub fn vgi_view_reset(tracer: *mut vgi_tracer) -> c_int;

    /// The last view frame's owner per pixel (pixel index of the face's owner,
    /// 0xffffffff: none) for the first `pixels` pixels; returns the number of
    /// owners (faces); *overflow (optional) counts faces the hash could not place.
    pub fn vgi_view_read_owners(tracer: *mut vgi_tracer, out: *mut u32, pixels: u32, overflow: *mut u32) -> i64;

    /// Receivers and estimates of pixels [first, first + count) of the last view frame.
    pub fn vgi_view_read_faces(tracer: *mut vgi_tracer, first: u32, count: u32, out: *mut vgi_view_face) -> c_int;

    pub fn vgi_set_schedule(tracer: *mut vgi_tracer, schedule: *const vgi_schedule) -> c_int;

    /// A resident batch whose faces the GPU picks itself (TraceWork.picks with no
    /// host faces): after queued retention, the selection passes rank every
    /// resident slot (classification re-ranks one quarter per batch) and pick up
    /// to `picks` faces (urgent first, then coverage 1/4, refinement 1/2, sweep
    /// 1/4); the picks are traced and merged like host faces. Non-blocking;
    /// flags as vgi_run_frame (counters read with vgi_read_bound's counters).
    pub fn vgi_select_frame(tracer: *mut vgi_tracer, picks: u32, flags: u32) -> c_int;

    /// GPU ms of the last pick batch: [selection passes, trace, accumulate].
    pub fn vgi_select_timings(tracer: *mut vgi_tracer, out: *mut f32) -> c_int;

    /// The last pick batch's pick list in pick order: resident slots, 0xffffffff
    /// for an unused place; *count = the budget.
    pub fn vgi_read_picks(tracer: *mut vgi_tracer, slots: *mut u32, capacity: u32, count: *mut u32) -> c_int;

    /// The selection scratch of the last pick batch (up to 1038 words): four
    /// 256-bucket histograms (coverage, refinement, sweep, urgent), per class the
    /// threshold bucket (256: none), the remainder taken at it and the ties seen
    /// at it, then the pick total and the mirror-record count.
    pub fn vgi_read_selection(tracer: *mut vgi_tracer, words: *mut u32, count: u32) -> c_int;

    /// Selection keys (class << 8 | bucket, 0xffffffff: none) of slots [first, first+count).
    pub fn vgi_read_keys(tracer: *mut vgi_tracer, first: u32, count: u32, keys: *mut u32) -> c_int;

    /// Statistics-mirror records of the last pick batch, 28 words each: (slot,
    /// incarnation, revision tag, 0) and the face's six 16-byte result words, the
    /// settled flag in bit 31 of word 5 (voxel_accumulate.wgsl:537-550).
    pub fn vgi_read_mirror(tracer: *mut vgi_tracer, words: *mut u32, capacity: u32, count: *mut u32) -> c_int;

    /// Slot metadata (3 uint4 per slot: identity, schedule, ranking; voxel_accumulate.wgsl:36-41) of slots [first, first+count).
    pub fn vgi_cache_meta(tracer: *mut vgi_tracer, first: u32, count: u32, out: *mut u32) -> c_int;

    /// Raw resident state (9 uint4 per slot, voxel_accumulate.wgsl:1-4) of slots [first, first+count).
    pub fn vgi_cache_state(tracer: *mut vgi_tracer, first: u32, count: u32, out: *mut u32) -> c_int;

    /// error harness: the long-horizon reference. Traces the bound faces
    /// `launches` times (sequences first_sequence + l * samples) with plain
    /// multi-bounce paths (no cache termination) and sums each face's per-launch
    /// moments in FP64 on the GPU: out = 8 doubles per bound face (sum of mean*known
    /// rgb, sum of second*known rgb, sum of known, 0). The analytic direct term is
    /// deterministic and read from any record.
    pub fn vgi_converged(tracer: *mut vgi_tracer, launches: u32, first_sequence: u32, out: *mut f64) -> c_int;

    /// learned radiance cache (no counterpart in the reference): a
    /// 32-64-64-64-64-3 ReLU MLP (BF16 MFMA, FP32 accumulation, Adam) that predicts
    /// a face's incident radiance from a 32-dim encoding of its centre (a trainable
    /// multiresolution hash grid, 8 levels x 2 features, and sin/cos of pi u per
    /// axis), its normal and albedo. With it on, bound-face frames end every path at its
    /// first hit into the network (the world-cache termination rule of
    /// voxel_path.wgsl:459-463 with the prediction as the cached value) and train
    /// it on `train_per_frame` plain multi-bounce samples per frame in `steps`
    /// Adam steps. Queries are 4 u32 (voxel xyz, face | material << 8).
    pub fn vgi_mlp_enable(tracer: *mut vgi_tracer, seed: u32, train_per_frame: u32, steps: u32, lr: f32) -> c_int;

    pub fn vgi_mlp_set_learned(tracer: *mut vgi_tracer, on: c_int) -> c_int;

    /// Read (out) and/or write (in) the 14,528 FP32 master weights: W1 64x32,
    /// W2..W4 64x64, W5 3x64, row-major (out x in), in that order.
    pub fn vgi_mlp_weights(tracer: *mut vgi_tracer, out: *mut f32, in_: *const f32) -> c_int;

    pub fn vgi_mlp_infer(tracer: *mut vgi_tracer, queries: *const u32, count: u32, out_rgb: *mut f32) -> c_int;

    /// One forward/backward on (query, target rgb + weight) pairs, count a multiple
    /// of 16; update 0 returns the summed gradient (relative L2 loss), 1 takes an
    /// Adam step.
    pub fn vgi_mlp_train(tracer: *mut vgi_tracer, queries: *const u32, targets: *const f32, count: u32, update: c_int, grad_out: *mut f32, loss_out: *mut f32) -> c_int;

    /// The hash grid: read (out) and/or write (in) its 67,108,864 FP32 entries
    /// (level-major, 2^22 entries of 2 features per level), and copy the table
    /// gradient of the last vgi_mlp_train without update (grad_out).
    pub fn vgi_mlp_table(tracer: *mut vgi_tracer, out: *mut f32, in_: *const f32, grad_out: *mut f32) -> c_int;

    /// The hash table's Adam learning rate (ten times the network's by default).
    pub fn vgi_mlp_table_lr(tracer: *mut vgi_tracer, lr: f32) -> c_int;

    /// Paths averaged into each training target (1..64; 1 by default).
    pub fn vgi_mlp_train_samples(tracer: *mut vgi_tracer, samples: u32) -> c_int;

    /// Change the training schedule of an enabled cache without touching the
    /// network or its Adam moments: training targets and Adam steps per frame and
    /// the network's rate (the table's becomes ten times it). 0 targets freezes
    /// the network: learned frames then only infer.
    pub fn vgi_mlp_training(tracer: *mut vgi_tracer, train_per_frame: u32, steps: u32, lr: f32) -> c_int;

    /// The position encoding of an enabled cache: 0 (the default) the hash grid
    /// described above; 1 frequencies only -- sin and cos of 2^k pi u per axis,
    /// k = 0..3, beside the normal and albedo: the 32-dim frequency encoding of
    /// the network stated for the 512^3 cave configuration; the hash table is
    /// then neither read nor trained. Kept across vgi_mlp_enable.
    pub fn vgi_mlp_encoding(tracer: *mut vgi_tracer, mode: u32) -> c_int;

    /// Training-only frames on the bound faces (targets traced, Adam steps taken,
    /// bound faces not traced): the network's warm-up.
    pub fn vgi_mlp_pretrain(tracer: *mut vgi_tracer, frames: u32) -> c_int;

    pub fn vgi_mlp_last_loss(tracer: *mut vgi_tracer) -> f32;

    /// GPU time of the last learned frame's network work (valid after it ran):
    /// [inference ms, training MFMA ms (forward/backward and weight-gradient
    /// passes of every step), training ms (the steps with their Adam updates),
    /// queries inferred, training samples (all steps), steps].
    pub fn vgi_mlp_timings(tracer: *mut vgi_tracer, out: *mut f32) -> c_int;

    /// retention events (resident.rs:1326-1417 encode_events layout;
    /// voxel_retention.wgsl:29-96, apply_events voxel_accumulate.wgsl:295-332)
    pub fn vgi_apply_events(tracer: *mut vgi_tracer, events: *const u32, event_count: u32, words: u32) -> c_int;

    /// batched frames (no counterpart in the reference, which runs one world;
    /// docs/architecture.md:294): many tracers -- separate scenes, or chunks of
    /// uneven face counts bound to several tracers -- traced in ONE persistent
    /// launch per frame that takes (scene, face range) tiles from a device queue,
    /// captured once into a HIP graph and replayed. Each scene's bound faces are
    /// traced exactly as vgi_run_frame would trace them (non-resident: records go
    /// to the tracer's bound records, read with vgi_read_bound), with sequence
    /// base set by vgi_batch_set_sequence and advancing by the samples per face
    /// every frame. All scenes trace the same samples per face.
    pub fn vgi_batch_create(device_ordinal: c_int, out: *mut *mut vgi_batch) -> c_int;

    pub fn vgi_batch_destroy(batch: *mut vgi_batch);

    /// Add a tracer whose faces are bound (vgi_bind_faces / vgi_bind_exposed).
    pub fn vgi_batch_add(batch: *mut vgi_batch, tracer: *mut vgi_tracer) -> c_int;

    /// Workgroup size (64/128/256), lanes per face (1, 2, 4 or 8: one lane, at
    /// least one lane per sample, or the reference's count, so sums keep the
    /// reference's order), queue tile in faces, workgroups (0: enough to fill
    /// every CU at the kernel's occupancy).
    pub fn vgi_batch_configure(batch: *mut vgi_batch, block: u32, lanes: u32, tile: u32, grid: u32) -> c_int;

    pub fn vgi_batch_set_sequence(batch: *mut vgi_batch, base: u32) -> c_int;

    pub fn vgi_batch_run_frame(batch: *mut vgi_batch) -> c_int;

    pub fn vgi_batch_sync(batch: *mut vgi_batch) -> c_int;

    /// GPU time of the last frame (ms), valid after vgi_batch_sync.
    pub fn vgi_batch_last_ms(batch: *mut vgi_batch) -> f32;

    /// Per scene in the order added: ms from the frame's start until its last
    /// tile ended; rays traced in the frame (optional).
    pub fn vgi_batch_latency(batch: *mut vgi_batch, out_ms: *mut f32, count: u32, rays: *mut u64) -> c_int;

    /// Time every allowed (workgroup, lanes, tile) on `frames` frames after a
    /// warm-up and keep the fastest: out = block, lanes, tile, median ms (float
    /// bits), candidates tried.
    pub fn vgi_batch_autotune(batch: *mut vgi_batch, frames: u32, out: *mut u32) -> c_int;

    /// measurement plumbing: a plain copy kernel of `bytes` run `repeats`
    /// times (average ms per launch in *ms), the known-size reference that scales
    /// hardware byte counters (not part of the reference seam).
    pub fn vgi_probe_copy(tracer: *mut vgi_tracer, bytes: u64, repeats: u32, ms: *mut f32) -> c_int;

    /// scene inspection for parity checks
    pub fn vgi_scene_bounds(tracer: *mut vgi_tracer, out: *mut i32) -> c_int;

    /// The canonical per-region form (see oracle): kind, material, top-level
    /// bytes, 64 brick words (pointer replaced by tag 2), 64x32 brick cell words.
    pub fn vgi_scene_region_canonical(tracer: *mut vgi_tracer, rx: i32, ry: i32, rz: i32, out: *mut u32) -> c_int;

    /// [loaded regions, solid regions, cell words, top-level cells, resident
    /// slots, emissive faces, light-tree nodes, pinned staging (allocations << 32 |
    /// times the region started over)]
    pub fn vgi_scene_stats(tracer: *mut vgi_tracer, out: *mut i64) -> i64;

    pub fn vgi_emitter_words(tracer: *mut vgi_tracer, out: *mut u32, capacity: i64) -> i64;

    pub fn vgi_last_error() -> *const c_char;
}
