//! `GpuTracer` over `libvgi.so`: the MI355X transport backend for
//! bevy_voxel_gi's trace worker.
//!
//! The reference's worker closure (crates/bevy_voxel_gi/src/dispatch.rs:167-271)
//! owns one `GpuTracer` (gpu.rs:309-1236) and drives it one call at a time:
//! `new` + `enable_resident`, `queue_retention`, `snapshot_ready`,
//! `complete_readback`, `submit_compact`. This crate keeps those names and
//! their `Result<_, String>` errors over the C ABI of include/vgi.h, and adds
//! the entry points that have no counterpart there (device-resident frames,
//! the learned cache, batched scenes, measurement). Every function of the
//! `vgi-sys` binding is reached from exactly one method here.
//!
//! The plugin's own types (`TransportScene`, `TraceWork`, `VoxelGiMaterials`)
//! stay in the plugin: its `gpu.rs` converts them into the plain records below
//! (`FaceTask` is `encode_tasks`' 32-byte record, gpu.rs:153-177; `FaceResult`
//! is the 96-byte record `complete_readback` decodes, gpu.rs:1072-1106).
use std::ffi::CStr;
use std::marker::PhantomData;
use std::os::raw::c_int;
use std::ptr::{self, NonNull};

pub use vgi_sys as sys;
pub use vgi_sys::{
    vgi_face_result as FaceResult, vgi_face_task as FaceTask, vgi_light as Light, vgi_schedule as Schedule,
    vgi_far_settings as FarTierSettings, vgi_transport as Transport, vgi_view_camera as ViewCamera, vgi_view_face as ViewFace,
    vgi_view_pixel as FarPixel, VGI_AIR as AIR, VGI_COUNT_GATHERS as COUNT_GATHERS, VGI_COUNT_RAYS as COUNT_RAYS,
    VGI_NOALLOC_MAX_FACES as NOALLOC_MAX_FACES, VGI_NO_FACE as NO_FACE, VGI_PAGE_EDGE as PAGE_EDGE,
    VGI_RECORD_HITS as RECORD_HITS,
};

/// Voxels in one 32^3 page (world.rs:22).
pub const PAGE_VOXELS: usize = (PAGE_EDGE * PAGE_EDGE * PAGE_EDGE) as usize;
/// u32 words of one first-hit record (vgi_read_hits).
pub const HIT_WORDS: usize = 8;
/// u32 words of one statistics-mirror record (vgi_read_mirror).
pub const MIRROR_WORDS: usize = 28;
/// u32 words of one slot's resident state / metadata (vgi_cache_state, vgi_cache_meta).
pub const STATE_WORDS: usize = 36;
pub const META_WORDS: usize = 12;
/// FP32 master weights of the learned cache's network (vgi_mlp_weights).
pub const MLP_WEIGHTS: usize = 14_528;
/// FP32 entries of the learned cache's hash grid (vgi_mlp_table).
pub const MLP_TABLE: usize = 67_108_864;
/// Words of the sun-flip region (vgi_read_sun_flips).
pub const SUN_FLIP_WORDS: usize = 8_204;
/// Words of one canonical region (vgi_scene_region_canonical): the 4 header
/// words (kind, material, top-level bytes), 64 brick words, 64 x 32 cell words.
pub const REGION_CANONICAL_WORDS: usize = 4 + 64 + 64 * 32;

pub type Result<T> = std::result::Result<T, String>;

fn last_error() -> String {
    // SAFETY: vgi_last_error returns a NUL-terminated thread-local string (or NULL).
    unsafe {
        let text = sys::vgi_last_error();
        if text.is_null() {
            return "libvgi: unknown error".to_string();
        }
        CStr::from_ptr(text).to_string_lossy().into_owned()
    }
}

fn check(rc: c_int) -> Result<()> {
    if rc < 0 {
        Err(last_error())
    } else {
        Ok(())
    }
}

fn check_count(rc: i64) -> Result<usize> {
    if rc < 0 {
        Err(last_error())
    } else {
        Ok(rc as usize)
    }
}

fn len_u32(n: usize, what: &str) -> Result<u32> {
    u32::try_from(n).map_err(|_| format!("{what}: {n} exceeds u32"))
}

fn xyz_ptr(v: &[[i32; 3]]) -> *const i32 {
    if v.is_empty() {
        ptr::null()
    } else {
        v.as_ptr() as *const i32
    }
}

/// An in-flight batch: the reference's `Readback` snapshot (gpu/readback.rs).
#[derive(Debug)]
pub struct Readback {
    ticket: u64,
    faces: u32,
}

impl Readback {
    pub fn faces(&self) -> usize {
        self.faces as usize
    }
}

/// GPU ms of one batch: [prepare, trace, accumulate, build] and, after an
/// edit, its host ms [regions + re-pack, top level, retention event, exposure].
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct Timings {
    pub gpu: [f32; 4],
    pub edit_host: [f32; 4],
}

/// The learned cache's network timings of the last learned frame (vgi_mlp_timings).
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct MlpTimings {
    pub inference_ms: f32,
    pub training_mfma_ms: f32,
    pub training_ms: f32,
    pub queries: f32,
    pub training_samples: f32,
    pub steps: f32,
}

/// The reference's `GpuTracer` (gpu.rs:309-366): one world's transport on one GPU.
/// Not `Sync`: one thread at a time, as the reference's `&mut self` contract.
pub struct GpuTracer {
    raw: NonNull<sys::vgi_tracer>,
    resident: bool,
    /// GPU compute time of the last completed batch (gpu.rs:356-358).
    pub last_gpu_ms: Option<f32>,
    /// The scheduling estimate of the last completed batch (gpu.rs:355).
    pub last_dispatch_ms: f32,
}

// SAFETY: the library keeps no thread affinity for a tracer; callers move it
// between worker threads as the reference moves its `GpuTracer`.
unsafe impl Send for GpuTracer {}

impl Drop for GpuTracer {
    fn drop(&mut self) {
        // SAFETY: raw came from vgi_create and is destroyed once.
        unsafe { sys::vgi_destroy(self.raw.as_ptr()) }
    }
}

impl GpuTracer {
    fn raw(&self) -> *mut sys::vgi_tracer {
        self.raw.as_ptr()
    }

    // ---- lifecycle (gpu.rs:368-450) ----

    /// `GpuTracer::new` (gpu.rs:448): a tracer on HIP device `device`.
    pub fn new(device: i32) -> Result<Self> {
        let mut out: *mut sys::vgi_tracer = ptr::null_mut();
        check(unsafe { sys::vgi_create(device, &mut out) })?;
        let raw = NonNull::new(out).ok_or_else(|| "vgi_create returned no tracer".to_string())?;
        Ok(GpuTracer { raw, resident: false, last_gpu_ms: None, last_dispatch_ms: 0.0 })
    }

    /// gpu.rs:369-375.
    pub fn backend_name(&self) -> &'static str {
        if self.resident {
            "gpu-resident"
        } else {
            "gpu-sync"
        }
    }

    /// `GpuTracer::enable_resident` (gpu.rs:427-432).
    pub fn enable_resident(&mut self) -> Result<()> {
        check(unsafe { sys::vgi_enable_resident(self.raw()) })?;
        self.resident = true;
        Ok(())
    }

    /// The resident clock in ms (None: wall clock), gpu/resident.rs:708-710.
    pub fn set_clock(&mut self, milliseconds: Option<i64>) -> Result<()> {
        check(unsafe { sys::vgi_set_clock(self.raw(), milliseconds.unwrap_or(-1)) })
    }

    // ---- scene (TransportScene::with_page_contents world.rs:982-1046; PackedScene::update gpu_scene.rs:72-309) ----

    /// Pages loaded (32^3 voxels each, x fastest) and unloaded, in page coordinates.
    pub fn upload_pages(&mut self, loaded: &[[i32; 3]], voxels: &[u16], removed: &[[i32; 3]]) -> Result<()> {
        if voxels.len() != loaded.len() * PAGE_VOXELS {
            return Err(format!("upload_pages: {} voxels for {} pages", voxels.len(), loaded.len()));
        }
        let pages = len_u32(loaded.len(), "upload_pages")?;
        let gone = len_u32(removed.len(), "upload_pages")?;
        let v = if voxels.is_empty() { ptr::null() } else { voxels.as_ptr() };
        check(unsafe { sys::vgi_upload_pages(self.raw(), xyz_ptr(loaded), pages, v, xyz_ptr(removed), gone) })
    }

    /// `VoxelGiInput::edit_voxel` (input.rs:90-93) for a list of voxels; a later edit of a voxel wins.
    pub fn edit_voxels(&mut self, xyz: &[[i32; 3]], values: &[u16]) -> Result<()> {
        if xyz.len() != values.len() {
            return Err(format!("edit_voxels: {} positions, {} values", xyz.len(), values.len()));
        }
        let n = len_u32(xyz.len(), "edit_voxels")?;
        check(unsafe { sys::vgi_edit_voxels(self.raw(), xyz_ptr(xyz), values.as_ptr(), n) })
    }

    /// The material table (gpu.rs:714-727): diffuse and emission rgb of materials 0..255.
    pub fn set_materials(&mut self, diffuse: &[[f32; 3]; 256], emission: &[[f32; 3]; 256]) -> Result<()> {
        check(unsafe {
            sys::vgi_set_materials(self.raw(), diffuse.as_ptr() as *const f32, emission.as_ptr() as *const f32)
        })
    }

    /// The sources (gpu.rs:728-760): `directional` of `lights` come first.
    pub fn set_lights(&mut self, lights: &[Light], directional: u32) -> Result<()> {
        let n = len_u32(lights.len(), "set_lights")?;
        check(unsafe { sys::vgi_set_lights(self.raw(), lights.as_ptr(), n, directional) })
    }

    /// `TraceWork.transport` and history settings (work.rs:11-55, settings.rs:391-426).
    pub fn set_transport(&mut self, transport: &Transport) -> Result<()> {
        check(unsafe { sys::vgi_set_transport(self.raw(), transport) })
    }

    // ---- batches: submit_compact / snapshot_ready / complete_readback (gpu.rs:647-1170) ----

    /// `GpuTracer::queue_retention` (gpu.rs:411-417): encoded retention events
    /// (encode_events layout, gpu/resident.rs:1326-1417), applied before the
    /// next resident batch. A tracer without the resident cache drops them.
    pub fn queue_retention(&mut self, words: &mut Vec<u32>, event_count: u32) -> Result<()> {
        if self.resident && event_count > 0 {
            let n = len_u32(words.len(), "queue_retention")?;
            check(unsafe { sys::vgi_apply_events(self.raw(), words.as_ptr(), event_count, n) })?;
        }
        words.clear();
        Ok(())
    }

    /// `GpuTracer::submit_compact` (gpu.rs:647-1022): the batch's faces as
    /// 32-byte tasks; None when there is nothing to trace (maintenance).
    pub fn submit_compact(&mut self, tasks: &[FaceTask], flags: u32) -> Result<Option<Readback>> {
        if tasks.is_empty() {
            return Ok(None);
        }
        let n = len_u32(tasks.len(), "submit_compact")?;
        let mut ticket = 0u64;
        check(unsafe { sys::vgi_submit_faces(self.raw(), tasks.as_ptr(), n, flags, &mut ticket) })?;
        Ok(Some(Readback { ticket, faces: n }))
    }

    /// `GpuTracer::snapshot_ready` (gpu.rs:382-384).
    pub fn snapshot_ready(&self, snapshot: &Readback) -> Result<bool> {
        let mut ready: c_int = 0;
        check(unsafe { sys::vgi_poll(self.raw(), snapshot.ticket, &mut ready) })?;
        Ok(ready != 0)
    }

    /// `GpuTracer::complete_readback` (gpu.rs:1024-1170): the batch's 96-byte
    /// records, one per submitted task; sets `last_gpu_ms`.
    pub fn complete_readback(&mut self, snapshot: Readback, out: &mut [FaceResult]) -> Result<()> {
        if out.len() < snapshot.faces() {
            return Err(format!("complete_readback: {} records for {} faces", out.len(), snapshot.faces));
        }
        check(unsafe { sys::vgi_read_faces(self.raw(), snapshot.ticket, out.as_mut_ptr(), snapshot.faces) })?;
        let t = self.timings()?;
        self.last_gpu_ms = Some(t.gpu[0] + t.gpu[1] + t.gpu[2]);
        self.last_dispatch_ms = t.gpu.iter().sum();
        Ok(())
    }

    /// First-hit records (8 words per face-sample) of a batch submitted with RECORD_HITS.
    pub fn read_hits(&self, snapshot: &Readback, out: &mut [u32]) -> Result<()> {
        if out.is_empty() {
            return Err("read_hits: empty output".to_string());
        }
        check(unsafe { sys::vgi_read_hits(self.raw(), snapshot.ticket, out.as_mut_ptr(), out.len() as u64) })
    }

    /// [rays, steps, header loads, brick-word loads, cell-word loads, ...] of a batch.
    pub fn read_counters(&self, snapshot: &Readback, out: &mut [u64]) -> Result<()> {
        let n = len_u32(out.len(), "read_counters")?;
        check(unsafe { sys::vgi_read_counters(self.raw(), snapshot.ticket, out.as_mut_ptr(), n) })
    }

    /// GPU timestamps (gpu.rs:486-497) and the last edit's host times.
    pub fn timings(&self) -> Result<Timings> {
        let mut words = [0f32; 8];
        check(unsafe { sys::vgi_timings(self.raw(), words.as_mut_ptr(), 8) })?;
        let mut t = Timings::default();
        t.gpu.copy_from_slice(&words[0..4]);
        t.edit_host.copy_from_slice(&words[4..8]);
        Ok(t)
    }

    /// Resident slots of the last batch's faces (face_index.rs:71-97).
    pub fn read_slots(&self, out: &mut [u32]) -> Result<()> {
        let n = len_u32(out.len(), "read_slots")?;
        check(unsafe { sys::vgi_read_slots(self.raw(), out.as_mut_ptr(), n) })
    }

    /// Synchronous trace of at most NOALLOC_MAX_FACES faces that allocates no
    /// device memory (no counterpart in the reference; for callers that cannot
    /// allocate). Calls that overlap on one GPU race.
    pub fn trace_faces_noalloc(&mut self, tasks: &[FaceTask], out: &mut [FaceResult]) -> Result<()> {
        if out.len() < tasks.len() {
            return Err(format!("trace_faces_noalloc: {} records for {} faces", out.len(), tasks.len()));
        }
        let n = len_u32(tasks.len(), "trace_faces_noalloc")?;
        check(unsafe { sys::vgi_trace_faces_noalloc(self.raw(), tasks.as_ptr(), n, out.as_mut_ptr()) })
    }

    // ---- device-resident frames (no counterpart: the bound face set stays on the GPU) ----

    pub fn bind_faces(&mut self, tasks: &[FaceTask]) -> Result<()> {
        let n = len_u32(tasks.len(), "bind_faces")?;
        check(unsafe { sys::vgi_bind_faces(self.raw(), tasks.as_ptr(), n) })
    }

    /// Bind every exposed face and follow the exposed surface through edits
    /// (the reference rescans changed pages for probes, pages.rs:451-458).
    pub fn bind_exposed(&mut self) -> Result<()> {
        check(unsafe { sys::vgi_bind_exposed(self.raw()) })
    }

    pub fn bound_count(&self) -> usize {
        unsafe { sys::vgi_bound_count(self.raw()) as usize }
    }

    pub fn bound_tasks(&self) -> Result<Vec<FaceTask>> {
        let mut out = vec![FaceTask::default(); self.bound_count()];
        let n = len_u32(out.len(), "bound_tasks")?;
        check(unsafe { sys::vgi_read_bound_tasks(self.raw(), out.as_mut_ptr(), n) })?;
        Ok(out)
    }

    /// One frame (prepare, trace, accumulate) over the bound faces; non-blocking.
    pub fn run_frame(&mut self, flags: u32) -> Result<()> {
        check(unsafe { sys::vgi_run_frame(self.raw(), flags) })
    }

    pub fn sync(&mut self) -> Result<()> {
        check(unsafe { sys::vgi_sync(self.raw()) })
    }

    /// ms of the last frame's trace launch (after `sync`).
    pub fn last_trace_ms(&self) -> f32 {
        unsafe { sys::vgi_last_trace_ms(self.raw()) }
    }

    /// GPU ms of the whole last frame (after `sync`).
    pub fn last_frame_ms(&self) -> f32 {
        unsafe { sys::vgi_last_frame_ms(self.raw()) }
    }

    /// Frames by wavefront path tracing (true) or the one-kernel estimator; same records.
    pub fn set_wavefront(&mut self, on: bool) -> Result<()> {
        check(unsafe { sys::vgi_set_wavefront(self.raw(), on as c_int) })
    }

    /// The wavefront tracer's per-wavefront hit-face deduplication of
    /// world-cache lookups (the same records either way).
    pub fn set_dedup(&mut self, on: bool) -> Result<()> {
        check(unsafe { sys::vgi_set_dedup(self.raw(), on as c_int) })
    }

    /// The bound faces' records after the last frame; `counters` receives
    /// [rays, steps, header, brick-word, cell-word loads, deduplicated
    /// world-cache lookups, their probes] of a counted frame.
    pub fn read_bound(&self, out: &mut [FaceResult], counters: Option<&mut [u64; 7]>) -> Result<()> {
        let n = len_u32(out.len(), "read_bound")?;
        let records = if out.is_empty() { ptr::null_mut() } else { out.as_mut_ptr() };
        let c = counters.map_or(ptr::null_mut(), |c| c.as_mut_ptr());
        check(unsafe { sys::vgi_read_bound(self.raw(), records, n, c) })
    }

    /// First-hit records of the last frame run with RECORD_HITS (8 words per face-sample).
    pub fn read_bound_hits(&self, out: &mut [u32]) -> Result<()> {
        if out.is_empty() {
            return Err("read_bound_hits: empty output".to_string());
        }
        check(unsafe { sys::vgi_read_bound_hits(self.raw(), out.as_mut_ptr(), out.len() as u64) })
    }

    /// The encoded retention events the last resident batch or frame applied,
    /// with measured flux filled in (gpu/resident.rs:1326-1417 layout).
    pub fn applied_events(&self) -> Result<Vec<u32>> {
        let n = check_count(unsafe { sys::vgi_read_events(self.raw(), ptr::null_mut(), 0) })?;
        let mut out = vec![0u32; n];
        let cap = len_u32(n, "applied_events")?;
        check_count(unsafe { sys::vgi_read_events(self.raw(), out.as_mut_ptr(), cap) })?;
        Ok(out)
    }

    /// Change footprints on (default) or off: `GpuTracer::without_footprints` (gpu.rs:419-425).
    pub fn set_footprints(&mut self, on: bool) -> Result<()> {
        check(unsafe { sys::vgi_set_footprints(self.raw(), on as u32) })
    }

    pub fn without_footprints(&mut self) -> Result<()> {
        self.set_footprints(false)
    }

    // ---- sun corners (voxel_view_runtime.wgsl:168-268, sun_changes.rs) ----

    pub fn sun_corners(&mut self, frame: u32, refresh_frames: u32) -> Result<()> {
        check(unsafe { sys::vgi_sun_corners(self.raw(), frame, refresh_frames) })
    }

    /// The last `sun_corners` call's flip region (SUN_FLIP_WORDS words).
    pub fn sun_flips(&self) -> Result<Vec<u32>> {
        let n = check_count(unsafe { sys::vgi_read_sun_flips(self.raw(), ptr::null_mut(), 0) })?;
        let mut out = vec![0u32; n];
        let cap = len_u32(n, "sun_flips")?;
        check_count(unsafe { sys::vgi_read_sun_flips(self.raw(), out.as_mut_ptr(), cap) })?;
        Ok(out)
    }

    /// [GPU ms of the corner passes, host ms of decoding and the event].
    pub fn sun_timings(&self) -> Result<[f32; 2]> {
        let mut out = [0f32; 2];
        check(unsafe { sys::vgi_sun_timings(self.raw(), out.as_mut_ptr(), 2) })?;
        Ok(out)
    }

    pub fn read_sun_corners(&self, first: u32, out: &mut [u32]) -> Result<()> {
        let n = len_u32(out.len(), "read_sun_corners")?;
        check(unsafe { sys::vgi_read_sun_corners(self.raw(), first, n, out.as_mut_ptr()) })
    }

    // ---- view tier (render_view, gpu/view/runtime.rs:749) ----

    /// One camera's view frame over its depth and normal prepass (row-major,
    /// width x height depths and 3 x width x height normal components).
    pub fn view_frame(&mut self, camera: &ViewCamera, depth: &[f32], normals: &[f32]) -> Result<()> {
        let pixels = camera.width as usize * camera.height as usize;
        if depth.len() != pixels || normals.len() != 3 * pixels {
            return Err(format!("view_frame: {} depths and {} normal components for {pixels} pixels", depth.len(), normals.len()));
        }
        check(unsafe { sys::vgi_view_frame(self.raw(), camera, depth.as_ptr(), normals.as_ptr()) })
    }

    /// GPU ms of the last view frame: [all its passes, the owners' tracing, the far tier].
    pub fn view_timings(&self) -> Result<[f32; 3]> {
        let mut out = [0f32; 3];
        check(unsafe { sys::vgi_view_timings(self.raw(), out.as_mut_ptr()) })?;
        Ok(out)
    }

    /// `settings.far_tier` (FarTierSettings, settings.rs:95-160); `None` turns the far tier off.
    pub fn set_far_tier(&mut self, settings: Option<&FarTierSettings>) -> Result<()> {
        let p = settings.map_or(ptr::null(), |s| s as *const FarTierSettings);
        check(unsafe { sys::vgi_view_far(self.raw(), p) })
    }

    /// `view.display.xyz` (gpu/view/runtime.rs:962-967): the noise threshold,
    /// the spatial strength and the temporal blend of the presentation.
    pub fn set_display(&mut self, noise_threshold: f32, spatial: f32, blend: f32) -> Result<()> {
        check(unsafe { sys::vgi_view_display(self.raw(), noise_threshold, spatial, blend) })
    }

    /// publish_view's records of pixels first .. first + out.len() of the last view frame.
    pub fn view_pixels(&self, first: u32, out: &mut [FarPixel]) -> Result<()> {
        let n = len_u32(out.len(), "view_pixels")?;
        check(unsafe { sys::vgi_view_read_pixels(self.raw(), first, n, out.as_mut_ptr()) })
    }

    /// The last view frame's far tier and presentation counters.
    pub fn view_counters(&self) -> Result<[u32; 16]> {
        let mut out = [0u32; 16];
        check(unsafe { sys::vgi_view_read_counters(self.raw(), out.as_mut_ptr()) })?;
        Ok(out)
    }

    /// The far tier's output (the material's buffer) and cells of the last
    /// view frame; returns the number of cells.
    pub fn far_output(&self, output: &mut [u32], cells: &mut [u32]) -> Result<usize> {
        check_count(unsafe {
            sys::vgi_view_read_far(self.raw(), output.as_mut_ptr(), output.len() as u64, cells.as_mut_ptr(), cells.len() as u64)
        })
    }

    /// Forget the view history.
    pub fn view_reset(&mut self) -> Result<()> {
        check(unsafe { sys::vgi_view_reset(self.raw()) })
    }

    /// The last view frame's owner per pixel; returns (faces, hash overflows).
    pub fn view_owners(&self, out: &mut [u32]) -> Result<(usize, u32)> {
        let n = len_u32(out.len(), "view_owners")?;
        let mut overflow = 0u32;
        let owners = check_count(unsafe { sys::vgi_view_read_owners(self.raw(), out.as_mut_ptr(), n, &mut overflow) })?;
        Ok((owners, overflow))
    }

    /// Receivers and estimates of pixels first .. first + out.len() of the last view frame.
    pub fn view_faces(&self, first: u32, out: &mut [ViewFace]) -> Result<()> {
        let n = len_u32(out.len(), "view_faces")?;
        check(unsafe { sys::vgi_view_read_faces(self.raw(), first, n, out.as_mut_ptr()) })
    }

    // ---- GPU work selection (voxel_accumulate.wgsl:448-695; gpu/resident.rs:1093-1164) ----

    /// `TraceWork.{focus, transport_revision, stable_error, ...}` (work.rs:38-50).
    pub fn set_schedule(&mut self, schedule: &Schedule) -> Result<()> {
        check(unsafe { sys::vgi_set_schedule(self.raw(), schedule) })
    }

    /// A batch of `picks` GPU-picked faces (TraceWork.picks); non-blocking.
    pub fn select_frame(&mut self, picks: u32, flags: u32) -> Result<()> {
        check(unsafe { sys::vgi_select_frame(self.raw(), picks, flags) })
    }

    /// [selection passes, trace, accumulate] GPU ms of the last pick batch.
    pub fn select_timings(&self) -> Result<[f32; 3]> {
        let mut out = [0f32; 3];
        check(unsafe { sys::vgi_select_timings(self.raw(), out.as_mut_ptr()) })?;
        Ok(out)
    }

    /// The last pick list (resident slots, NO_FACE unused); returns the budget.
    pub fn read_picks(&self, slots: &mut [u32]) -> Result<u32> {
        let mut count = 0u32;
        let n = len_u32(slots.len(), "read_picks")?;
        check(unsafe { sys::vgi_read_picks(self.raw(), slots.as_mut_ptr(), n, &mut count) })?;
        Ok(count)
    }

    /// The selection scratch of the last pick batch (up to 1038 words).
    pub fn read_selection(&self, words: &mut [u32]) -> Result<()> {
        let n = len_u32(words.len(), "read_selection")?;
        check(unsafe { sys::vgi_read_selection(self.raw(), words.as_mut_ptr(), n) })
    }

    pub fn read_keys(&self, first: u32, keys: &mut [u32]) -> Result<()> {
        let n = len_u32(keys.len(), "read_keys")?;
        check(unsafe { sys::vgi_read_keys(self.raw(), first, n, keys.as_mut_ptr()) })
    }

    /// Statistics-mirror records (MIRROR_WORDS each) that `work.picked` decodes; returns their count.
    pub fn read_mirror(&self, words: &mut [u32]) -> Result<u32> {
        let mut count = 0u32;
        let n = len_u32(words.len(), "read_mirror")?;
        check(unsafe { sys::vgi_read_mirror(self.raw(), words.as_mut_ptr(), n, &mut count) })?;
        Ok(count)
    }

    /// Slot metadata, META_WORDS per slot.
    pub fn cache_meta(&self, first: u32, out: &mut [u32]) -> Result<()> {
        let n = len_u32(out.len() / META_WORDS, "cache_meta")?;
        check(unsafe { sys::vgi_cache_meta(self.raw(), first, n, out.as_mut_ptr()) })
    }

    /// Raw resident state, STATE_WORDS per slot.
    pub fn cache_state(&self, first: u32, out: &mut [u32]) -> Result<()> {
        let n = len_u32(out.len() / STATE_WORDS, "cache_state")?;
        check(unsafe { sys::vgi_cache_state(self.raw(), first, n, out.as_mut_ptr()) })
    }

    /// FP64 long-horizon sums of the bound faces (8 doubles per face).
    pub fn converged(&mut self, launches: u32, first_sequence: u32) -> Result<Vec<f64>> {
        let mut out = vec![0f64; self.bound_count() * 8];
        check(unsafe { sys::vgi_converged(self.raw(), launches, first_sequence, out.as_mut_ptr()) })?;
        Ok(out)
    }

    // ---- learned radiance cache (no counterpart in the reference) ----

    pub fn mlp_enable(&mut self, seed: u32, train_per_frame: u32, steps: u32, lr: f32) -> Result<()> {
        check(unsafe { sys::vgi_mlp_enable(self.raw(), seed, train_per_frame, steps, lr) })
    }

    pub fn mlp_set_learned(&mut self, on: bool) -> Result<()> {
        check(unsafe { sys::vgi_mlp_set_learned(self.raw(), on as c_int) })
    }

    /// Read and/or write the MLP_WEIGHTS master weights.
    pub fn mlp_weights(&mut self, out: Option<&mut [f32]>, input: Option<&[f32]>) -> Result<()> {
        for n in out.as_ref().map(|o| o.len()).into_iter().chain(input.map(|i| i.len())) {
            if n != MLP_WEIGHTS {
                return Err(format!("mlp_weights: {n} values, want {MLP_WEIGHTS}"));
            }
        }
        let o = out.map_or(ptr::null_mut(), |o| o.as_mut_ptr());
        let i = input.map_or(ptr::null(), |i| i.as_ptr());
        check(unsafe { sys::vgi_mlp_weights(self.raw(), o, i) })
    }

    /// Predicted radiance (rgb) of queries (4 u32 each: voxel xyz, face | material << 8).
    pub fn mlp_infer(&mut self, queries: &[[u32; 4]], out_rgb: &mut [[f32; 3]]) -> Result<()> {
        if out_rgb.len() < queries.len() {
            return Err("mlp_infer: output shorter than the queries".to_string());
        }
        let n = len_u32(queries.len(), "mlp_infer")?;
        check(unsafe {
            sys::vgi_mlp_infer(self.raw(), queries.as_ptr() as *const u32, n, out_rgb.as_mut_ptr() as *mut f32)
        })
    }

    /// One forward/backward on (query, target rgb + weight) pairs (a multiple
    /// of 16); `update` takes an Adam step, else `grad_out` receives the
    /// summed gradient (MLP_WEIGHTS values). Returns the loss.
    pub fn mlp_train(&mut self, queries: &[[u32; 4]], targets: &[[f32; 4]], update: bool, grad_out: Option<&mut [f32]>) -> Result<f32> {
        if queries.len() != targets.len() {
            return Err("mlp_train: queries and targets differ in length".to_string());
        }
        if grad_out.as_ref().map_or(false, |g| g.len() != MLP_WEIGHTS) {
            return Err(format!("mlp_train: the gradient holds {MLP_WEIGHTS} values"));
        }
        let n = len_u32(queries.len(), "mlp_train")?;
        let g = grad_out.map_or(ptr::null_mut(), |g| g.as_mut_ptr());
        let mut loss = 0f32;
        check(unsafe {
            sys::vgi_mlp_train(
                self.raw(),
                queries.as_ptr() as *const u32,
                targets.as_ptr() as *const f32,
                n,
                update as c_int,
                g,
                &mut loss,
            )
        })?;
        Ok(loss)
    }

    /// Read and/or write the MLP_TABLE hash-grid entries; copy the last
    /// gradient-only train's table gradient.
    pub fn mlp_table(&mut self, out: Option<&mut [f32]>, input: Option<&[f32]>, grad_out: Option<&mut [f32]>) -> Result<()> {
        let sizes = [out.as_ref().map(|o| o.len()), input.map(|i| i.len()), grad_out.as_ref().map(|g| g.len())];
        if sizes.iter().flatten().any(|&n| n != MLP_TABLE) {
            return Err(format!("mlp_table: every buffer holds {MLP_TABLE} values"));
        }
        let o = out.map_or(ptr::null_mut(), |o| o.as_mut_ptr());
        let i = input.map_or(ptr::null(), |i| i.as_ptr());
        let g = grad_out.map_or(ptr::null_mut(), |g| g.as_mut_ptr());
        check(unsafe { sys::vgi_mlp_table(self.raw(), o, i, g) })
    }

    pub fn mlp_table_lr(&mut self, lr: f32) -> Result<()> {
        check(unsafe { sys::vgi_mlp_table_lr(self.raw(), lr) })
    }

    pub fn mlp_train_samples(&mut self, samples: u32) -> Result<()> {
        check(unsafe { sys::vgi_mlp_train_samples(self.raw(), samples) })
    }

    /// A new training schedule for the enabled cache, network kept; 0 targets freezes it.
    pub fn mlp_training(&mut self, train_per_frame: u32, steps: u32, lr: f32) -> Result<()> {
        check(unsafe { sys::vgi_mlp_training(self.raw(), train_per_frame, steps, lr) })
    }

    /// The position encoding: sys::VGI_MLP_ENCODING_HASH (default) or
    /// sys::VGI_MLP_ENCODING_FREQUENCY (no hash table).
    pub fn mlp_encoding(&mut self, mode: u32) -> Result<()> {
        check(unsafe { sys::vgi_mlp_encoding(self.raw(), mode) })
    }

    pub fn mlp_pretrain(&mut self, frames: u32) -> Result<()> {
        check(unsafe { sys::vgi_mlp_pretrain(self.raw(), frames) })
    }

    pub fn mlp_last_loss(&self) -> f32 {
        unsafe { sys::vgi_mlp_last_loss(self.raw()) }
    }

    pub fn mlp_timings(&self) -> Result<MlpTimings> {
        let mut w = [0f32; 6];
        check(unsafe { sys::vgi_mlp_timings(self.raw(), w.as_mut_ptr()) })?;
        Ok(MlpTimings {
            inference_ms: w[0],
            training_mfma_ms: w[1],
            training_ms: w[2],
            queries: w[3],
            training_samples: w[4],
            steps: w[5],
        })
    }

    // ---- measurement and inspection ----

    /// Average ms per launch of a plain copy of `bytes` (the known-size
    /// reference that scales hardware byte counters).
    pub fn probe_copy(&mut self, bytes: u64, repeats: u32) -> Result<f32> {
        let mut ms = 0f32;
        check(unsafe { sys::vgi_probe_copy(self.raw(), bytes, repeats, &mut ms) })?;
        Ok(ms)
    }

    /// The loaded scene's region bounds (min xyz, max xyz); None before a scene loads.
    pub fn scene_bounds(&self) -> Result<Option<[i32; 6]>> {
        let mut out = [0i32; 6];
        let rc = unsafe { sys::vgi_scene_bounds(self.raw(), out.as_mut_ptr()) };
        check(rc)?;
        Ok(if rc > 0 { Some(out) } else { None })
    }

    /// The canonical form of region (rx, ry, rz), REGION_CANONICAL_WORDS
    /// words; None for a region that is not loaded.
    pub fn scene_region_canonical(&self, region: [i32; 3]) -> Result<Option<Vec<u32>>> {
        let mut out = vec![0u32; REGION_CANONICAL_WORDS];
        let rc = unsafe { sys::vgi_scene_region_canonical(self.raw(), region[0], region[1], region[2], out.as_mut_ptr()) };
        check(rc)?;
        Ok(if rc > 0 { Some(out) } else { None })
    }

    /// [loaded regions, solid regions, cell words, top-level cells, resident
    /// slots, emissive faces, light-tree nodes, pinned staging].
    pub fn scene_stats(&self) -> Result<[i64; 8]> {
        let mut out = [0i64; 8];
        check_count(unsafe { sys::vgi_scene_stats(self.raw(), out.as_mut_ptr()) })?;
        Ok(out)
    }

    /// `GpuTracer::geometry_storage_bytes` (gpu.rs:1231-1236): the NAADF
    /// cell words resident on the GPU.
    pub fn geometry_storage_bytes(&self) -> Result<usize> {
        Ok(self.scene_stats()?[2] as usize * 4)
    }

    /// The emitter table's words (emission distribution of the loaded scene).
    pub fn emitter_words(&self) -> Result<Vec<u32>> {
        let n = check_count(unsafe { sys::vgi_emitter_words(self.raw(), ptr::null_mut(), 0) })?;
        let mut out = vec![0u32; n];
        check_count(unsafe { sys::vgi_emitter_words(self.raw(), out.as_mut_ptr(), n as i64) })?;
        Ok(out)
    }
}

/// Many tracers' bound faces traced in one persistent launch per frame,
/// captured into a HIP graph (no counterpart in the reference, which runs one
/// world). The batch borrows its tracers for its lifetime.
pub struct GpuBatch<'a> {
    raw: NonNull<sys::vgi_batch>,
    scenes: usize,
    _tracers: PhantomData<&'a mut GpuTracer>,
}

unsafe impl Send for GpuBatch<'_> {}

impl Drop for GpuBatch<'_> {
    fn drop(&mut self) {
        // SAFETY: raw came from vgi_batch_create and is destroyed once.
        unsafe { sys::vgi_batch_destroy(self.raw.as_ptr()) }
    }
}

/// The winner of `GpuBatch::autotune`.
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct BatchShape {
    pub block: u32,
    pub lanes: u32,
    pub tile: u32,
    pub median_ms: f32,
    pub candidates: u32,
}

impl<'a> GpuBatch<'a> {
    fn raw(&self) -> *mut sys::vgi_batch {
        self.raw.as_ptr()
    }

    pub fn new(device: i32) -> Result<Self> {
        let mut out: *mut sys::vgi_batch = ptr::null_mut();
        check(unsafe { sys::vgi_batch_create(device, &mut out) })?;
        let raw = NonNull::new(out).ok_or_else(|| "vgi_batch_create returned no batch".to_string())?;
        Ok(GpuBatch { raw, scenes: 0, _tracers: PhantomData })
    }

    /// Add a tracer whose faces are bound.
    pub fn add(&mut self, tracer: &'a mut GpuTracer) -> Result<()> {
        check(unsafe { sys::vgi_batch_add(self.raw(), tracer.raw()) })?;
        self.scenes += 1;
        Ok(())
    }

    /// Workgroup size, lanes per face, queue tile in faces, workgroups (0: fill the chip).
    pub fn configure(&mut self, block: u32, lanes: u32, tile: u32, grid: u32) -> Result<()> {
        check(unsafe { sys::vgi_batch_configure(self.raw(), block, lanes, tile, grid) })
    }

    pub fn autotune(&mut self, frames: u32) -> Result<BatchShape> {
        let mut out = [0u32; 5];
        check(unsafe { sys::vgi_batch_autotune(self.raw(), frames, out.as_mut_ptr()) })?;
        Ok(BatchShape { block: out[0], lanes: out[1], tile: out[2], median_ms: f32::from_bits(out[3]), candidates: out[4] })
    }

    pub fn set_sequence(&mut self, base: u32) -> Result<()> {
        check(unsafe { sys::vgi_batch_set_sequence(self.raw(), base) })
    }

    pub fn run_frame(&mut self) -> Result<()> {
        check(unsafe { sys::vgi_batch_run_frame(self.raw()) })
    }

    pub fn sync(&mut self) -> Result<()> {
        check(unsafe { sys::vgi_batch_sync(self.raw()) })
    }

    pub fn last_ms(&self) -> f32 {
        unsafe { sys::vgi_batch_last_ms(self.raw()) }
    }

    /// Per scene in the order added: ms until its last tile ended, and its rays.
    pub fn latency(&self) -> Result<(Vec<f32>, Vec<u64>)> {
        let mut ms = vec![0f32; self.scenes];
        let mut rays = vec![0u64; self.scenes];
        let n = len_u32(self.scenes, "latency")?;
        check(unsafe { sys::vgi_batch_latency(self.raw(), ms.as_mut_ptr(), n, rays.as_mut_ptr()) })?;
        Ok((ms, rays))
    }
}
