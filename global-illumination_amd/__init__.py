"""Python host layer over libvgi.so (include/vgi.h): the MI355X voxel GI
lighting hot path.

`GpuTracer` mirrors the reference's transport seam `GpuTracer`
(crates/bevy_voxel_gi/src/gpu.rs:309-1236): `new`, `enable_resident`,
`submit_compact`, `snapshot_ready`, `complete_readback`, plus the scene input
of `VoxelGiInput` (input.rs:54-112) and the optics of `VoxelGiMaterials`
(world.rs:90-174). Errors raise `VoxelGiError` carrying the library's
message, as the reference returns `Err(String)`.

Everything runs in the HIP library; there is no CPU fallback. Loading fails
loudly when libvgi.so is missing.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvgi.so")

AIR = 256
NO_FACE = 0xFFFFFFFF
RECORD_HITS = 1
COUNT_RAYS = 2
COUNT_GATHERS = 4
MLP_TABLE = 8 * (1 << 22) * 2  # hash-grid entries (level-major, 2 features each)
MLP_PARAMS = 14528
NOALLOC_MAX_FACES = 262144

FACE_TASK = np.dtype([("voxel", "<i4", (3,)), ("face", "<u4"), ("sequence", "<u4"), ("padding", "<u4", (3,))])
assert FACE_TASK.itemsize == 32


class VoxelGiError(RuntimeError):
    pass


class Light(ctypes.Structure):
    _fields_ = [("position_range", ctypes.c_float * 4), ("energy", ctypes.c_float * 4), ("extra", ctypes.c_float * 4)]


class Transport(ctypes.Structure):
    _fields_ = [
        ("bounces", ctypes.c_uint32),
        ("samples", ctypes.c_uint32),
        ("lanes", ctypes.c_uint32),
        ("limited_reach", ctypes.c_uint32),
        ("reach", ctypes.c_float),
        ("max_distance", ctypes.c_float),
        ("environment", ctypes.c_float * 3),
        ("history_limit", ctypes.c_uint32),
        ("history_recency", ctypes.c_float),
        ("measured_retention", ctypes.c_float),
        ("plain", ctypes.c_uint32),
    ]


class Schedule(ctypes.Structure):
    _fields_ = [
        ("focus", ctypes.c_int32 * 3),
        ("transport_revision", ctypes.c_uint32),
        ("stable_error", ctypes.c_float),
        ("stable_samples", ctypes.c_uint32),
        ("audit_ms", ctypes.c_uint32),
        ("mirror_budget", ctypes.c_uint32),
    ]


class ViewCamera(ctypes.Structure):
    """vgi_view_camera: matrices column-major (element (r, c) at 4c + r)."""
    _fields_ = [
        ("world_from_clip", ctypes.c_float * 16),
        ("position", ctypes.c_float * 3),
        ("inner_distance", ctypes.c_float),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("samples", ctypes.c_uint32),
        ("measured_retention", ctypes.c_float),
        ("clip_from_world", ctypes.c_float * 16),
        ("forward", ctypes.c_float * 3),
        ("pixels_per_unit", ctypes.c_float),
        ("frame", ctypes.c_uint32),
    ]


class FarSettings(ctypes.Structure):
    """vgi_far_settings (FarTierSettings, settings.rs:95-138)."""
    _fields_ = [
        ("voxel_pixels", ctypes.c_float),
        ("measure", ctypes.c_uint32),
        ("cell", ctypes.c_uint32),
        ("history", ctypes.c_float),
        ("rate", ctypes.c_uint32),
        ("steps", ctypes.c_uint32),
    ]


def view_camera(world_from_clip, position, width, height, samples=1, measured_retention=1.0, inner_distance=0.0,
                clip_from_world=None, forward=(0.0, 0.0, -1.0), pixels_per_unit=0.0, frame=0):
    """A ViewCamera from 4x4 row-major numpy matrices (clip -> world, and
    world -> clip, by default the float64 inverse of the first)."""
    c = ViewCamera()
    m = np.asarray(world_from_clip, np.float32)
    if clip_from_world is None:
        try:
            clip_from_world = np.linalg.inv(np.asarray(world_from_clip, np.float64))
        except np.linalg.LinAlgError:  # a designed prepass with no camera behind it
            clip_from_world = np.zeros((4, 4))
    mi = np.asarray(clip_from_world, np.float32)
    for col in range(4):
        for row in range(4):
            c.world_from_clip[col * 4 + row] = float(m[row, col])
            c.clip_from_world[col * 4 + row] = float(mi[row, col])
    for i in range(3):
        c.position[i] = float(position[i])
        c.forward[i] = float(forward[i])
    c.inner_distance = inner_distance
    c.width, c.height, c.samples, c.measured_retention = width, height, samples, measured_retention
    c.pixels_per_unit, c.frame = pixels_per_unit, frame
    return c


def far_settings(voxel_pixels=8.0, measure=0, cell=0, history=32.0, rate=4, steps=128):
    """FarSettings with the reference's defaults (settings.rs:127-138)."""
    f = FarSettings()
    f.voxel_pixels, f.measure, f.cell, f.history, f.rate, f.steps = voxel_pixels, measure, cell, history, rate, steps
    return f


# vgi_read_bound's counters: rays, steps, header, brick-word and cell-word
# loads, deduplicated world-cache lookups, their probes (the library writes
# all seven).
BOUND_COUNTERS = 7

VIEW_FACE = np.dtype([("voxel", "<i4", (3,)), ("face", "<u4"), ("surface", "<f4", (3,)), ("sequence", "<u4"), ("mean", "<f4", (4,)),
                      ("moment", "<f4", (4,)), ("identity", "<u4", (4,)), ("direct", "<f4", (4,)), ("shown", "<f4", (4,))])
VIEW_PIXEL = np.dtype([("surface", "<f4", (3,)), ("face", "<u4"), ("irradiance", "<f4", (3,)), ("evidence", "<f4"),
                       ("sky", "<u4", (4,))])

SEL_THRESHOLD, SEL_REMAINDER, SEL_TIES, SEL_TOTAL, SEL_MIRRORS, SEL_WORDS = 1024, 1028, 1032, 1036, 1037, 1038

_lib = None


def library():
    """The loaded libvgi.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VoxelGiError("libvgi.so is not built: run __graft_entry__.build() (or global-illumination_amd/build.py)")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    u32, i32, u64, i64, f32 = ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int64, ctypes.c_float
    sig = {
        "vgi_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(P)]),
        "vgi_destroy": (None, [P]),
        "vgi_enable_resident": (ctypes.c_int, [P]),
        "vgi_set_clock": (ctypes.c_int, [P, i64]),
        "vgi_upload_pages": (ctypes.c_int, [P, P, u32, P, P, u32]),
        "vgi_edit_voxels": (ctypes.c_int, [P, P, P, u32]),
        "vgi_set_materials": (ctypes.c_int, [P, P, P]),
        "vgi_set_lights": (ctypes.c_int, [P, P, u32, u32]),
        "vgi_set_transport": (ctypes.c_int, [P, ctypes.POINTER(Transport)]),
        "vgi_submit_faces": (ctypes.c_int, [P, P, u32, u32, ctypes.POINTER(u64)]),
        "vgi_poll": (ctypes.c_int, [P, u64, ctypes.POINTER(ctypes.c_int)]),
        "vgi_read_faces": (ctypes.c_int, [P, u64, P, u32]),
        "vgi_trace_faces_noalloc": (ctypes.c_int, [P, P, u32, P]),
        "vgi_read_hits": (ctypes.c_int, [P, u64, P, u64]),
        "vgi_read_counters": (ctypes.c_int, [P, u64, P, u32]),
        "vgi_timings": (ctypes.c_int, [P, P, u32]),
        "vgi_read_slots": (ctypes.c_int, [P, P, u32]),
        "vgi_bind_faces": (ctypes.c_int, [P, P, u32]),
        "vgi_run_frame": (ctypes.c_int, [P, u32]),
        "vgi_sync": (ctypes.c_int, [P]),
        "vgi_last_trace_ms": (f32, [P]),
        "vgi_last_frame_ms": (f32, [P]),
        "vgi_set_wavefront": (i32, [P, i32]),
        "vgi_bind_exposed": (i32, [P]),
        "vgi_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(P)]),
        "vgi_batch_destroy": (None, [P]),
        "vgi_batch_add": (i32, [P, P]),
        "vgi_batch_configure": (i32, [P, u32, u32, u32, u32]),
        "vgi_batch_set_sequence": (i32, [P, u32]),
        "vgi_batch_run_frame": (i32, [P]),
        "vgi_batch_sync": (i32, [P]),
        "vgi_batch_last_ms": (f32, [P]),
        "vgi_batch_latency": (i32, [P, P, u32, P]),
        "vgi_batch_autotune": (i32, [P, u32, P]),
        "vgi_bound_count": (u32, [P]),
        "vgi_read_bound_tasks": (i32, [P, P, u32]),
        "vgi_read_events": (i64, [P, P, u32]),
        "vgi_read_bound": (ctypes.c_int, [P, P, u32, P]),
        "vgi_read_bound_hits": (ctypes.c_int, [P, P, u64]),
        "vgi_cache_state": (ctypes.c_int, [P, u32, u32, P]),
        "vgi_cache_meta": (ctypes.c_int, [P, u32, u32, P]),
        "vgi_set_schedule": (ctypes.c_int, [P, ctypes.POINTER(Schedule)]),
        "vgi_set_footprints": (ctypes.c_int, [P, ctypes.c_uint32]),
        "vgi_sun_corners": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32]),
        "vgi_view_frame": (ctypes.c_int, [P, P, P, P]),
        "vgi_view_reset": (ctypes.c_int, [P]),
        "vgi_view_timings": (ctypes.c_int, [P, P]),
        "vgi_view_read_owners": (i64, [P, P, ctypes.c_uint32, P]),
        "vgi_view_read_faces": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, P]),
        "vgi_view_far": (ctypes.c_int, [P, P]),
        "vgi_set_dedup": (ctypes.c_int, [P, ctypes.c_int]),
        "vgi_view_display": (ctypes.c_int, [P, f32, f32, f32]),
        "vgi_view_read_pixels": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, P]),
        "vgi_view_read_counters": (ctypes.c_int, [P, P]),
        "vgi_view_read_far": (i64, [P, P, u64, P, u64]),
        "vgi_read_sun_flips": (ctypes.c_int64, [P, P, ctypes.c_uint32]),
        "vgi_read_sun_corners": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, P]),
        "vgi_sun_timings": (ctypes.c_int, [P, P, ctypes.c_uint32]),
        "vgi_select_frame": (ctypes.c_int, [P, u32, u32]),
        "vgi_select_timings": (ctypes.c_int, [P, P]),
        "vgi_read_picks": (ctypes.c_int, [P, P, u32, ctypes.POINTER(u32)]),
        "vgi_read_selection": (ctypes.c_int, [P, P, u32]),
        "vgi_read_keys": (ctypes.c_int, [P, u32, u32, P]),
        "vgi_read_mirror": (ctypes.c_int, [P, P, u32, ctypes.POINTER(u32)]),
        "vgi_apply_events": (ctypes.c_int, [P, P, u32, u32]),
        "vgi_probe_copy": (ctypes.c_int, [P, u64, u32, P]),
        "vgi_converged": (ctypes.c_int, [P, u32, u32, P]),
        "vgi_mlp_enable": (ctypes.c_int, [P, u32, u32, u32, f32]),
        "vgi_mlp_set_learned": (ctypes.c_int, [P, ctypes.c_int]),
        "vgi_mlp_weights": (ctypes.c_int, [P, P, P]),
        "vgi_mlp_infer": (ctypes.c_int, [P, P, u32, P]),
        "vgi_mlp_train": (ctypes.c_int, [P, P, P, u32, ctypes.c_int, P, P]),
        "vgi_mlp_table": (ctypes.c_int, [P, P, P, P]),
        "vgi_mlp_table_lr": (ctypes.c_int, [P, ctypes.c_float]),
        "vgi_mlp_train_samples": (ctypes.c_int, [P, ctypes.c_uint32]),
        "vgi_mlp_training": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float]),
        "vgi_mlp_encoding": (ctypes.c_int, [P, ctypes.c_uint32]),
        "vgi_mlp_last_loss": (f32, [P]),
        "vgi_mlp_timings": (ctypes.c_int, [P, P]),
        "vgi_mlp_pretrain": (ctypes.c_int, [P, u32]),
        "vgi_scene_bounds": (ctypes.c_int, [P, P]),
        "vgi_scene_region_canonical": (ctypes.c_int, [P, i32, i32, i32, P]),
        "vgi_scene_stats": (i64, [P, P]),
        "vgi_emitter_words": (i64, [P, P, i64]),
        "vgi_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    """Names the C ABI declares (include/vgi.h)."""
    return [
        "vgi_create", "vgi_destroy", "vgi_enable_resident", "vgi_set_clock", "vgi_upload_pages", "vgi_edit_voxels",
        "vgi_set_materials", "vgi_set_lights", "vgi_set_transport", "vgi_submit_faces", "vgi_trace_faces_noalloc", "vgi_poll",
        "vgi_read_faces", "vgi_read_hits", "vgi_read_counters", "vgi_timings", "vgi_read_slots",
        "vgi_bind_faces", "vgi_run_frame", "vgi_sync", "vgi_last_trace_ms", "vgi_last_frame_ms", "vgi_set_wavefront", "vgi_read_bound", "vgi_read_bound_hits",
        "vgi_bind_exposed", "vgi_bound_count", "vgi_read_bound_tasks", "vgi_read_events",
        "vgi_cache_state", "vgi_apply_events", "vgi_probe_copy", "vgi_converged", "vgi_mlp_enable",
        "vgi_mlp_set_learned", "vgi_mlp_weights", "vgi_mlp_infer", "vgi_mlp_train", "vgi_mlp_table", "vgi_mlp_table_lr", "vgi_mlp_train_samples", "vgi_mlp_training", "vgi_mlp_encoding", "vgi_mlp_last_loss", "vgi_mlp_timings", "vgi_mlp_pretrain", "vgi_scene_bounds", "vgi_scene_region_canonical",
        "vgi_scene_stats", "vgi_emitter_words", "vgi_last_error", "vgi_batch_create", "vgi_batch_destroy",
        "vgi_batch_add", "vgi_batch_configure", "vgi_batch_set_sequence", "vgi_batch_run_frame", "vgi_batch_sync",
        "vgi_batch_last_ms", "vgi_batch_latency", "vgi_batch_autotune", "vgi_cache_meta", "vgi_set_schedule", "vgi_set_footprints",
        "vgi_sun_corners", "vgi_read_sun_flips", "vgi_read_sun_corners", "vgi_sun_timings",
        "vgi_view_frame", "vgi_view_reset", "vgi_view_timings", "vgi_view_read_owners", "vgi_view_read_faces",
        "vgi_set_dedup", "vgi_view_far", "vgi_view_display", "vgi_view_read_pixels", "vgi_view_read_counters", "vgi_view_read_far",
        "vgi_select_frame", "vgi_select_timings", "vgi_read_picks", "vgi_read_selection", "vgi_read_keys", "vgi_read_mirror",
    ]


MLP_ENCODING_HASH = 0
MLP_ENCODING_FREQUENCY = 1

WORKLOADS_PATH = os.path.join(HERE, "libvgi_workloads.so")
_workloads = None


def workloads():
    """libvgi_workloads.so (include/vgi_workloads.h): the seeded scenes, optics
    and edit streams of the tests and the benchmark -- not the product."""
    global _workloads
    if _workloads is not None:
        return _workloads
    if not os.path.exists(WORKLOADS_PATH):
        raise VoxelGiError("libvgi_workloads.so is not built: run __graft_entry__.build() (or global-illumination_amd/build.py)")
    lib = ctypes.CDLL(WORKLOADS_PATH)
    P = ctypes.c_void_p
    u32, i64 = ctypes.c_uint32, ctypes.c_int64
    sig = {
        "vgi_scene_generate": (ctypes.c_int, [ctypes.c_int, u32, u32, P, P]),
        "vgi_scene_optics": (ctypes.c_int, [ctypes.c_int, u32, P, P, P, P, P]),
        "vgi_scene_faces": (i64, [P, P, ctypes.c_int, P, i64]),
        "vgi_scene_edits": (i64, [u32, u32, u32, u32, P, P, i64]),
        "vgi_scene_prepass": (i64, [P, P, P, P, ctypes.c_float, u32, u32, P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _workloads = lib
    return lib


def workload_symbols():
    """Names include/vgi_workloads.h declares (libvgi_workloads.so)."""
    return ["vgi_scene_generate", "vgi_scene_optics", "vgi_scene_faces", "vgi_scene_edits", "vgi_scene_prepass"]


def _check_workload(rc, what):
    if rc < 0:
        raise VoxelGiError("%s failed (%d)" % (what, rc))
    return rc


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(rc):
    if rc < 0:
        raise VoxelGiError(library().vgi_last_error().decode())
    return rc


# ---------------------------------------------------------------- scenes
def generate(workload, seed, edge=0):
    """Dense u16 voxel grid (z, y, x) of a seeded workload; AIR = 256."""
    lib = workloads()
    dims = np.zeros(3, np.int32)
    _check_workload(lib.vgi_scene_generate(workload, seed, edge, None, _ptr(dims)), "vgi_scene_generate")
    grid = np.empty((dims[2], dims[1], dims[0]), np.uint16)
    _check_workload(lib.vgi_scene_generate(workload, seed, edge, _ptr(grid), _ptr(dims)), "vgi_scene_generate")
    return grid


def optics(workload, seed):
    """(diffuse[256,3], emission[256,3], lights (n,12) float32, environment[3])."""
    lib = workloads()
    diffuse = np.zeros((256, 3), np.float32)
    emission = np.zeros((256, 3), np.float32)
    lights = np.zeros((16, 12), np.float32)
    count = np.zeros(1, np.uint32)
    env = np.zeros(3, np.float32)
    _check_workload(lib.vgi_scene_optics(workload, seed, _ptr(diffuse), _ptr(emission), _ptr(lights), _ptr(count), _ptr(env)), "vgi_scene_optics")
    return diffuse, emission, lights[: int(count[0])].copy(), env


def exposed_faces(grid, interior_only=False):
    """Exposed faces (solid voxel, air neighbour inside the grid) as FACE_TASK records."""
    lib = workloads()
    grid = np.ascontiguousarray(grid, np.uint16)
    dims = np.array([grid.shape[2], grid.shape[1], grid.shape[0]], np.int32)
    n = _check_workload(lib.vgi_scene_faces(_ptr(grid), _ptr(dims), int(interior_only), None, 0), "vgi_scene_faces")
    out = np.zeros(n, FACE_TASK)
    lib.vgi_scene_faces(_ptr(grid), _ptr(dims), int(interior_only), _ptr(out), n)
    return out


def scene_edits(seed, frame, count, edge):
    """One frame of the seeded edit stream: (xyz int32 [n,3], values uint16 [n])."""
    lib = workloads()
    n = _check_workload(lib.vgi_scene_edits(seed, frame, count, edge, None, None, 0), "vgi_scene_edits")
    xyz = np.zeros((n, 3), np.int32)
    values = np.zeros(n, np.uint16)
    lib.vgi_scene_edits(seed, frame, count, edge, _ptr(xyz), _ptr(values), n)
    return xyz, values


def scene_prepass(grid, eye, target, width, height, fov_y=1.0471975511965976):
    """(world_from_clip [4, 4] row-major, depth [h, w], normals [h, w, 3], pixels hit)
    of a dense grid seen from `eye` towards `target` (workload library)."""
    lib = workloads()
    grid = np.ascontiguousarray(grid, np.uint16)
    dims = np.array([grid.shape[2], grid.shape[1], grid.shape[0]], np.int32)
    wfc = np.zeros(16, np.float32)
    depth = np.zeros((height, width), np.float32)
    normals = np.zeros((height, width, 3), np.float32)
    e = np.asarray(eye, np.float32)
    t = np.asarray(target, np.float32)
    n = _check_workload(lib.vgi_scene_prepass(_ptr(grid), _ptr(dims), _ptr(e), _ptr(t), fov_y, width, height, _ptr(wfc), _ptr(depth),
                                              _ptr(normals)), "vgi_scene_prepass")
    return wfc.reshape(4, 4).T.copy(), depth, normals, n


def pages_of(grid, origin=(0, 0, 0)):
    """Split a dense (z, y, x) grid with edges multiple of 32 into pages:
    (page xyz int32 [n,3], voxels uint16 [n, 32768] with x fastest)."""
    z, y, x = grid.shape
    assert z % 32 == 0 and y % 32 == 0 and x % 32 == 0
    g = grid.reshape(z // 32, 32, y // 32, 32, x // 32, 32).transpose(0, 2, 4, 1, 3, 5)
    pz, py, px = np.meshgrid(np.arange(z // 32), np.arange(y // 32), np.arange(x // 32), indexing="ij")
    coords = np.stack([px.ravel(), py.ravel(), pz.ravel()], axis=1).astype(np.int32) + np.asarray(origin, np.int32) // 32
    return coords, np.ascontiguousarray(g.reshape(-1, 32768))


def _pcg(v):
    v = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    w = (((v >> ((v >> 28) + 4)) ^ v) * 277803737) & 0xFFFFFFFF
    return (w >> 22) ^ w


def batch_edges(count=32, seed=0x5EED0003):
    """Edges of the batched-scenes workload: 8 * round(exp(U[ln 32, ln 512]) / 8)."""
    out = []
    for i in range(count):
        u = (_pcg(seed ^ _pcg(i + 0x9E3779B9 & 0xFFFFFFFF)) >> 8) / float(1 << 24)
        edge = math.exp(math.log(32.0) + u * (math.log(512.0) - math.log(32.0)))
        out.append(int(8 * round(edge / 8.0)))
    return out


def padded(grid, multiple=32):
    """The grid padded with air to edges that are multiples of `multiple`."""
    z, y, x = grid.shape
    shape = tuple(-(-d // multiple) * multiple for d in (z, y, x))
    if shape == grid.shape:
        return grid
    out = np.full(shape, AIR, np.uint16)
    out[:z, :y, :x] = grid
    return out


# ---------------------------------------------------------------- tracer
class GpuTracer:
    """The reference's GpuTracer seam over the HIP library (gpu.rs:309-1236)."""

    def __init__(self, device=0):
        lib = library()
        h = ctypes.c_void_p()
        _check(lib.vgi_create(device, ctypes.byref(h)))
        self._h = h
        self._lib = lib

    @classmethod
    def new(cls, device=0):
        return cls(device)

    def close(self):
        if self._h:
            self._lib.vgi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def enable_resident(self):
        _check(self._lib.vgi_enable_resident(self._h))

    def set_clock(self, milliseconds):
        _check(self._lib.vgi_set_clock(self._h, milliseconds))

    # ---- scene input (VoxelGiInput page_loaded / page_unloaded / edit_voxel)
    def upload_pages(self, page_xyz, voxels, removed=None):
        page_xyz = np.ascontiguousarray(page_xyz, np.int32).reshape(-1, 3)
        voxels = np.ascontiguousarray(voxels, np.uint16).reshape(-1, 32768)
        removed = np.zeros((0, 3), np.int32) if removed is None else np.ascontiguousarray(removed, np.int32).reshape(-1, 3)
        _check(self._lib.vgi_upload_pages(self._h, _ptr(page_xyz), len(page_xyz), _ptr(voxels), _ptr(removed), len(removed)))

    def upload_grid(self, grid, origin=(0, 0, 0)):
        coords, pages = pages_of(grid, origin)
        self.upload_pages(coords, pages)

    def edit_voxels(self, xyz, values):
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        values = np.ascontiguousarray(values, np.uint16).reshape(-1)
        _check(self._lib.vgi_edit_voxels(self._h, _ptr(xyz), _ptr(values), len(values)))

    # ---- optics and sources
    def set_materials(self, diffuse, emission):
        d = np.ascontiguousarray(diffuse, np.float32).reshape(256, 3)
        e = np.ascontiguousarray(emission, np.float32).reshape(256, 3)
        _check(self._lib.vgi_set_materials(self._h, _ptr(d), _ptr(e)))

    def set_lights(self, lights, directional=None):
        lights = np.ascontiguousarray(lights, np.float32).reshape(-1, 12)
        if directional is None:
            directional = int((lights[:, 3] < 0).sum()) if len(lights) else 0
        _check(self._lib.vgi_set_lights(self._h, _ptr(lights) if len(lights) else None, len(lights), directional))

    def set_transport(self, bounces=3, samples=16, lanes=0, environment=(0.0, 0.0, 0.0), history_limit=0,
                      history_recency=1.0, measured_retention=1.0, limited_reach=False, reach=768.0, max_distance=768.0,
                      plain=False):
        t = Transport()
        t.bounces, t.samples, t.lanes, t.limited_reach = bounces, samples, lanes, int(limited_reach)
        t.reach, t.max_distance = reach, max_distance
        for i in range(3):
            t.environment[i] = environment[i]
        t.history_limit, t.history_recency, t.measured_retention = history_limit, history_recency, measured_retention
        t.plain = int(plain)
        self.transport = t
        _check(self._lib.vgi_set_transport(self._h, ctypes.byref(t)))

    # ---- batches (submit_compact / snapshot_ready / complete_readback)
    def submit_compact(self, tasks, flags=0):
        tasks = np.ascontiguousarray(tasks, FACE_TASK)
        ticket = ctypes.c_uint64()
        _check(self._lib.vgi_submit_faces(self._h, _ptr(tasks), len(tasks), flags, ctypes.byref(ticket)))
        self._last_count = len(tasks)
        return ticket.value

    def trace_faces_noalloc(self, tasks):
        """The records of `tasks` traced without any device allocation
        (at most NOALLOC_MAX_FACES faces; non-resident tracers only)."""
        tasks = np.ascontiguousarray(tasks, FACE_TASK)
        out = np.zeros((len(tasks), 24), np.float32)
        _check(self._lib.vgi_trace_faces_noalloc(self._h, _ptr(tasks), len(tasks), _ptr(out)))
        return out

    def snapshot_ready(self, ticket):
        ready = ctypes.c_int()
        _check(self._lib.vgi_poll(self._h, ticket, ctypes.byref(ready)))
        return bool(ready.value)

    def complete_readback(self, ticket, count):
        out = np.zeros((count, 24), np.float32)
        _check(self._lib.vgi_read_faces(self._h, ticket, _ptr(out), count))
        return out

    def read_hits(self, ticket, count, samples):
        out = np.zeros((count * samples, 8), np.uint32)
        _check(self._lib.vgi_read_hits(self._h, ticket, _ptr(out), out.size))
        return out

    def read_counters(self, ticket):
        out = np.zeros(8, np.uint64)
        _check(self._lib.vgi_read_counters(self._h, ticket, _ptr(out), 8))
        return out

    def timings(self):
        """[prepare, trace, accumulate, build] GPU ms, then the last edit's host ms
        [regions + re-pack, top level, retention event, exposure update]."""
        out = np.zeros(8, np.float32)
        _check(self._lib.vgi_timings(self._h, _ptr(out), 8))
        return out

    def slots(self, count):
        out = np.zeros(count, np.uint32)
        _check(self._lib.vgi_read_slots(self._h, _ptr(out), count))
        return out

    # ---- device-resident frames
    def bind_faces(self, tasks):
        tasks = np.ascontiguousarray(tasks, FACE_TASK)
        _check(self._lib.vgi_bind_faces(self._h, _ptr(tasks), len(tasks)))
        self._bound = len(tasks)

    def run_frame(self, flags=0):
        _check(self._lib.vgi_run_frame(self._h, flags))

    def sync(self):
        _check(self._lib.vgi_sync(self._h))

    def last_trace_ms(self):
        return float(self._lib.vgi_last_trace_ms(self._h))

    def set_wavefront(self, on):
        _check(self._lib.vgi_set_wavefront(self._h, int(on)))

    def set_dedup(self, on):
        """The wavefront tracer's per-wavefront hit-face deduplication."""
        _check(self._lib.vgi_set_dedup(self._h, int(on)))

    def last_frame_ms(self):
        return float(self._lib.vgi_last_frame_ms(self._h))

    def bind_exposed(self):
        """Bind every exposed face and follow the exposed surface through edits."""
        _check(self._lib.vgi_bind_exposed(self._h))
        self._bound = int(self._lib.vgi_bound_count(self._h))

    def bound_count(self):
        self._bound = int(self._lib.vgi_bound_count(self._h))
        return self._bound

    def bound_tasks(self):
        self._bound = int(self._lib.vgi_bound_count(self._h))
        out = np.zeros(self._bound, FACE_TASK)
        _check(self._lib.vgi_read_bound_tasks(self._h, _ptr(out), self._bound))
        return out

    def events(self):
        """Encoded retention events of the last applied edits (uint32 words)."""
        n = int(self._lib.vgi_read_events(self._h, None, 0))
        out = np.zeros(n, np.uint32)
        if n:
            _check(self._lib.vgi_read_events(self._h, _ptr(out), n))
        return out

    def read_bound(self, count=None):
        """Records of the first `count` bound faces (all by default)."""
        count = self._bound if count is None else int(count)
        out = np.zeros((count, 24), np.float32)
        counters = np.zeros(BOUND_COUNTERS, np.uint64)
        _check(self._lib.vgi_read_bound(self._h, _ptr(out), count, _ptr(counters)))
        return out, counters

    def read_bound_hits(self):
        """First-hit records (count * samples, 8) of the last frame run with RECORD_HITS."""
        out = np.zeros((self._bound * self.transport.samples, 8), np.uint32)
        _check(self._lib.vgi_read_bound_hits(self._h, _ptr(out), out.size))
        return out

    def read_counters_bound(self):
        counters = np.zeros(BOUND_COUNTERS, np.uint64)
        _check(self._lib.vgi_read_bound(self._h, None, 0, _ptr(counters)))
        return counters

    def cache_state(self, first, count):
        out = np.zeros((count, 36), np.uint32)
        _check(self._lib.vgi_cache_state(self._h, first, count, _ptr(out)))
        return out

    def cache_meta(self, first, count):
        out = np.zeros((count, 12), np.uint32)
        _check(self._lib.vgi_cache_meta(self._h, first, count, _ptr(out)))
        return out

    # ---- GPU work selection
    def set_schedule(self, focus=(0, 0, 0), transport_revision=0, stable_error=0.0, stable_samples=0, audit_ms=0,
                     mirror_budget=0):
        sc = Schedule()
        for i in range(3):
            sc.focus[i] = int(focus[i])
        sc.transport_revision, sc.stable_error, sc.stable_samples = transport_revision, stable_error, stable_samples
        sc.audit_ms, sc.mirror_budget = audit_ms, mirror_budget
        _check(self._lib.vgi_set_schedule(self._h, ctypes.byref(sc)))

    def set_footprints(self, on):
        _check(self._lib.vgi_set_footprints(self._h, 1 if on else 0))

    # ---- view tier
    def view_frame(self, camera, depth, normals):
        """One view frame: camera (ViewCamera), depth [h, w] float32, normals [h, w, 3] float32."""
        d = np.ascontiguousarray(depth, np.float32)
        n = np.ascontiguousarray(normals, np.float32)
        if d.size != camera.width * camera.height or n.size != 3 * d.size:
            raise VoxelGiError("prepass buffers do not match the camera's size")
        _check(self._lib.vgi_view_frame(self._h, ctypes.byref(camera), _ptr(d), _ptr(n)))

    def view_reset(self):
        _check(self._lib.vgi_view_reset(self._h))

    def view_timings(self):
        """GPU ms of the last view frame: [all passes, owners' tracing, far tier]."""
        out = np.zeros(3, np.float32)
        _check(self._lib.vgi_view_timings(self._h, _ptr(out)))
        return out

    def view_owners(self, pixels):
        """(owner per pixel, owner count, hash overflow count) of the last view frame."""
        out = np.zeros(pixels, np.uint32)
        overflow = np.zeros(1, np.uint32)
        n = _check(self._lib.vgi_view_read_owners(self._h, _ptr(out), pixels, _ptr(overflow)))
        return out, int(n), int(overflow[0])

    def view_faces(self, first, count):
        out = np.zeros(count, VIEW_FACE)
        _check(self._lib.vgi_view_read_faces(self._h, first, count, _ptr(out)))
        return out

    def view_far(self, settings):
        """The far tier's settings (FarSettings, or None: off)."""
        _check(self._lib.vgi_view_far(self._h, ctypes.byref(settings) if settings is not None else None))

    def view_display(self, noise_threshold=0.031, spatial=1.0, blend=0.19926):
        _check(self._lib.vgi_view_display(self._h, noise_threshold, spatial, blend))

    def view_pixels(self, first, count):
        """Published records (VIEW_PIXEL) of pixels [first, first + count)."""
        out = np.zeros(count, VIEW_PIXEL)
        _check(self._lib.vgi_view_read_pixels(self._h, first, count, _ptr(out)))
        return out

    def view_counters(self):
        out = np.zeros(16, np.uint32)
        _check(self._lib.vgi_view_read_counters(self._h, _ptr(out)))
        return out

    def view_far_output(self, cells):
        """(output words [16 + 4 cells], cell words [cells, 12]) of the last view frame."""
        output = np.zeros(16 + 4 * cells, np.uint32)
        history = np.zeros((cells, 12), np.uint32)
        n = _check(self._lib.vgi_view_read_far(self._h, _ptr(output), output.size, _ptr(history), history.size))
        if n != cells:
            raise VoxelGiError("the last view frame had %d far cells, not %d" % (n, cells))
        return output, history

    # ---- sun corner words
    def sun_corners(self, frame=0, refresh_frames=1):
        _check(self._lib.vgi_sun_corners(self._h, frame, refresh_frames))

    def sun_flips(self):
        n = _check(self._lib.vgi_read_sun_flips(self._h, None, 0))
        out = np.zeros(n, np.uint32)
        self._lib.vgi_read_sun_flips(self._h, _ptr(out), n)
        return out

    def sun_timings(self):
        out = np.zeros(2, np.float32)
        _check(self._lib.vgi_sun_timings(self._h, _ptr(out), 2))
        return out

    def sun_corner_words(self, first, count):
        out = np.zeros(count, np.uint32)
        _check(self._lib.vgi_read_sun_corners(self._h, first, count, _ptr(out)))
        return out

    def select_frame(self, picks, flags=0):
        _check(self._lib.vgi_select_frame(self._h, picks, flags))

    def select_timings(self):
        out = np.zeros(3, np.float32)
        _check(self._lib.vgi_select_timings(self._h, _ptr(out)))
        return out

    def picks(self):
        count = ctypes.c_uint32()
        _check(self._lib.vgi_read_picks(self._h, None, 0, ctypes.byref(count)))
        out = np.zeros(count.value, np.uint32)
        _check(self._lib.vgi_read_picks(self._h, _ptr(out), len(out), ctypes.byref(count)))
        return out

    def selection(self):
        out = np.zeros(SEL_WORDS, np.uint32)
        _check(self._lib.vgi_read_selection(self._h, _ptr(out), SEL_WORDS))
        return out

    def keys(self, first, count):
        out = np.zeros(count, np.uint32)
        _check(self._lib.vgi_read_keys(self._h, first, count, _ptr(out)))
        return out

    def mirror(self):
        count = ctypes.c_uint32()
        _check(self._lib.vgi_read_mirror(self._h, None, 0, ctypes.byref(count)))
        out = np.zeros((count.value, 28), np.uint32)
        _check(self._lib.vgi_read_mirror(self._h, _ptr(out), count.value, ctypes.byref(count)))
        return out

    def apply_events(self, events, event_count):
        events = np.ascontiguousarray(events, np.uint32).reshape(-1)
        _check(self._lib.vgi_apply_events(self._h, _ptr(events), event_count, events.size))

    def converged(self, launches, first_sequence=1 << 20):
        """FP64 sums over `launches` plain launches of the bound faces: (n, 8)."""
        out = np.zeros((self._bound, 8), np.float64)
        _check(self._lib.vgi_converged(self._h, launches, first_sequence, _ptr(out)))
        return out

    # ---- learned radiance cache
    def mlp_enable(self, seed=1, train_per_frame=65536, steps=4, lr=1e-3):
        _check(self._lib.vgi_mlp_enable(self._h, seed, train_per_frame, steps, lr))

    def mlp_set_learned(self, on):
        _check(self._lib.vgi_mlp_set_learned(self._h, int(on)))

    def mlp_weights(self, new=None):
        out = np.zeros(MLP_PARAMS, np.float32)
        w = None if new is None else np.ascontiguousarray(new, np.float32)
        _check(self._lib.vgi_mlp_weights(self._h, _ptr(out), _ptr(w)))
        return out

    def mlp_infer(self, queries):
        q = np.ascontiguousarray(queries, np.uint32).reshape(-1, 4)
        out = np.zeros((len(q), 3), np.float32)
        _check(self._lib.vgi_mlp_infer(self._h, _ptr(q), len(q), _ptr(out)))
        return out

    def mlp_train(self, queries, targets, update=False):
        q = np.ascontiguousarray(queries, np.uint32).reshape(-1, 4)
        y = np.ascontiguousarray(targets, np.float32).reshape(-1, 4)
        grad = np.zeros(MLP_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        _check(self._lib.vgi_mlp_train(self._h, _ptr(q), _ptr(y), len(q), int(update), _ptr(grad), _ptr(loss)))
        return grad, float(loss[0])

    def mlp_table(self, new=None):
        out = np.zeros(MLP_TABLE, np.float32)
        w = None if new is None else np.ascontiguousarray(new, np.float32)
        _check(self._lib.vgi_mlp_table(self._h, _ptr(out), _ptr(w), None))
        return out

    def mlp_table_lr(self, lr):
        _check(self._lib.vgi_mlp_table_lr(self._h, lr))

    def mlp_train_samples(self, samples):
        _check(self._lib.vgi_mlp_train_samples(self._h, samples))

    def mlp_training(self, train_per_frame, steps, lr):
        """New training schedule, network kept; 0 targets freezes it."""
        _check(self._lib.vgi_mlp_training(self._h, train_per_frame, steps, lr))

    def mlp_encoding(self, mode):
        """MLP_ENCODING_HASH (default) or MLP_ENCODING_FREQUENCY."""
        _check(self._lib.vgi_mlp_encoding(self._h, mode))

    def mlp_table_grad(self):
        out = np.zeros(MLP_TABLE, np.float32)
        _check(self._lib.vgi_mlp_table(self._h, None, None, _ptr(out)))
        return out

    def mlp_pretrain(self, frames):
        _check(self._lib.vgi_mlp_pretrain(self._h, frames))

    def mlp_timings(self):
        """The last learned frame: {infer_ms, train_mfma_ms, train_ms, queries, train_samples, steps}."""
        out = np.zeros(6, np.float32)
        _check(self._lib.vgi_mlp_timings(self._h, _ptr(out)))
        keys = ("infer_ms", "train_mfma_ms", "train_ms", "queries", "train_samples", "steps")
        return {k: float(v) for k, v in zip(keys, out)}

    def mlp_last_loss(self):
        return float(self._lib.vgi_mlp_last_loss(self._h))

    def probe_copy(self, nbytes, repeats=5):
        ms = np.zeros(1, np.float32)
        _check(self._lib.vgi_probe_copy(self._h, nbytes, repeats, _ptr(ms)))
        return float(ms[0])

    # ---- inspection
    def bounds(self):
        out = np.zeros(6, np.int32)
        return out if _check(self._lib.vgi_scene_bounds(self._h, _ptr(out))) else None

    def region_canonical(self, rx, ry, rz):
        out = np.zeros(4 + 64 + 64 * 32, np.uint32)
        _check(self._lib.vgi_scene_region_canonical(self._h, rx, ry, rz, _ptr(out)))
        return out

    def stats(self):
        out = np.zeros(8, np.int64)
        _check(self._lib.vgi_scene_stats(self._h, _ptr(out)))
        return out

    def emitter_words(self):
        n = _check(self._lib.vgi_emitter_words(self._h, None, 0))
        out = np.zeros(n, np.uint32)
        _check(self._lib.vgi_emitter_words(self._h, _ptr(out), n))
        return out


# ---------------------------------------------------------------- error harness
def converged_irradiance(sums, direct):
    """Per-face converged irradiance (incident mean + analytic direct) from the
    FP64 sums, and the per-face variance of that mean (the reference's own
    noise, removed from the error estimate)."""
    known = np.maximum(sums[:, 6:7], 1.0)
    mean = sums[:, 0:3] / known
    second = sums[:, 3:6] / known
    var_of_mean = np.maximum(second - mean * mean, 0.0) / known
    return mean + direct.astype(np.float64), var_of_mean


def relative_rmse(estimate, reference, reference_var=None):
    """sqrt(mean over faces and RGB of (estimate - reference)^2) / mean(reference),
    with the reference's own sampling variance subtracted when given."""
    err2 = np.mean((estimate.astype(np.float64) - reference) ** 2)
    if reference_var is not None:
        err2 = max(err2 - float(np.mean(reference_var)), 0.0)
    return float(np.sqrt(err2) / np.mean(reference))


class GpuBatch:
    """Many bound tracers traced in one persistent launch per frame, captured
    into a HIP graph (vgi_batch_*)."""

    def __init__(self, device=0):
        self._lib = library()
        h = ctypes.c_void_p()
        _check(self._lib.vgi_batch_create(device, ctypes.byref(h)))
        self._h = h
        self._scenes = []

    def close(self):
        if self._h:
            self._lib.vgi_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, tracer):
        _check(self._lib.vgi_batch_add(self._h, tracer._h))
        self._scenes.append(tracer)

    def configure(self, block=256, lanes=1, tile=1024, grid=0):
        _check(self._lib.vgi_batch_configure(self._h, block, lanes, tile, grid))

    def set_sequence(self, base):
        _check(self._lib.vgi_batch_set_sequence(self._h, base))

    def run_frame(self):
        _check(self._lib.vgi_batch_run_frame(self._h))

    def sync(self):
        _check(self._lib.vgi_batch_sync(self._h))

    def last_ms(self):
        return float(self._lib.vgi_batch_last_ms(self._h))

    def latency(self):
        """(per-scene ms from frame start to the scene's last tile, rays)."""
        out = np.zeros(len(self._scenes), np.float32)
        rays = ctypes.c_uint64()
        _check(self._lib.vgi_batch_latency(self._h, _ptr(out), len(out), ctypes.byref(rays)))
        return out, int(rays.value)

    def autotune(self, frames=3):
        out = np.zeros(5, np.uint32)
        _check(self._lib.vgi_batch_autotune(self._h, frames, _ptr(out)))
        return {"block": int(out[0]), "lanes": int(out[1]), "tile": int(out[2]),
                "ms": float(out[3:4].view(np.float32)[0]), "candidates": int(out[4])}
