"""ctypes access to the CPU oracle (oracle/liboracle.so) for tests and the
bench's CPU baseline. The oracle is the checker, never the product."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


class OrcTransport(ctypes.Structure):
    _fields_ = [
        ("bounces", ctypes.c_uint32),
        ("samples", ctypes.c_uint32),
        ("lanes", ctypes.c_uint32),
        ("limited_reach", ctypes.c_uint32),
        ("reach", ctypes.c_float),
        ("max_distance", ctypes.c_float),
        ("environment", ctypes.c_float * 3),
    ]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        build()
    L = ctypes.CDLL(ORACLE_LIB)
    P = ctypes.c_void_p
    u32, i32, u64, i64, f32 = ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int64, ctypes.c_float
    sig = {
        "orc_pcg": (u32, [u32]),
        "orc_hash_voxel": (u32, [i32, i32, i32]),
        "orc_stream_id": (u32, [i32, i32, i32, u32]),
        "orc_sample_1d": (f32, [u32, u32, u32]),
        "orc_sinf": (f32, [f32]),
        "orc_cosf": (f32, [f32]),
        "orc_expf": (f32, [f32]),
        "orc_hemisphere": (None, [P, f32, f32, P]),
        "orc_advance_box": (None, [P, P, P, P, P, f32, P]),
        "orc_build_aadf": (ctypes.c_int, [u32, u32, u32, P, P]),
        "orc_full_sweep_aadf": (ctypes.c_int, [u32, u32, u32, P, P]),
        "orc_scene_new": (P, []),
        "orc_scene_free": (None, [P]),
        "orc_scene_load_pages": (ctypes.c_int, [P, P, u32, P, P, u32]),
        "orc_scene_edit": (ctypes.c_int, [P, P, P, u32]),
        "orc_scene_bounds": (ctypes.c_int, [P, P]),
        "orc_scene_top_level": (i64, [P, P, P]),
        "orc_scene_material_at": (u32, [P, i32, i32, i32]),
        "orc_scene_region_canonical": (ctypes.c_int, [P, i32, i32, i32, P]),
        "orc_scene_region_solid": (ctypes.c_int, [P, i32, i32, i32]),
        "orc_scene_stats": (i64, [P, P]),
        "orc_tracer_new": (P, [P]),
        "orc_tracer_free": (None, [P]),
        "orc_tracer_set_materials": (ctypes.c_int, [P, P, P]),
        "orc_tracer_set_lights": (ctypes.c_int, [P, P, u32, u32]),
        "orc_tracer_set_transport": (ctypes.c_int, [P, ctypes.POINTER(OrcTransport)]),
        "orc_tracer_sync": (ctypes.c_int, [P]),
        "orc_tracer_emitters": (i64, [P, P, i64]),
        "orc_trace_faces": (ctypes.c_int, [P, P, u32, P, P, P, ctypes.c_int]),
        "orc_trace_learned": (ctypes.c_int, [P, P, u32, P, P, P, ctypes.c_int]),
        "orc_view_frame": (ctypes.c_int, [P, P, P, P, u32, P, P, P, ctypes.c_int]),
        "orc_view_reset": (ctypes.c_int, [P]),
        "orc_view_far": (ctypes.c_int, [P, P]),
        "orc_view_display": (ctypes.c_int, [P, f32, f32, f32]),
        "orc_view_read": (ctypes.c_int64, [P, P, P, P, P]),
        "orc_pack_half2": (u32, [f32, f32]),
        "orc_cache_enable": (ctypes.c_int, [P, u32, f32, f32, u32]),
        "orc_cache_register": (ctypes.c_int, [P, P, u32, P]),
        "orc_cache_batch": (ctypes.c_int, [P, P, P, u32, P, P, P, ctypes.c_int, u32]),
        "orc_cache_state": (ctypes.c_int, [P, u32, u32, P]),
        "orc_cache_apply_events": (ctypes.c_int, [P, P, u32, u32]),
        "orc_curve_limit": (u32, [u32]),
        "orc_edit_event": (i64, [P, P, P, u32]),
        "orc_load_event": (i64, [P, P, u32, P, P, u32]),
        "orc_last_event": (i64, [P, P, i64]),
        "orc_cache_meta": (ctypes.c_int, [P, u32, u32, P]),
        "orc_cache_schedule": (ctypes.c_int, [P, P, u32, f32, u32, u32, u32]),
        "orc_set_footprints": (ctypes.c_int, [P, u32]),
        "orc_sun_corners": (ctypes.c_int64, [P, P, P, u32, u32, u32, u32, P, P, u32]),
        "orc_sun_decode": (ctypes.c_int, [P, u32, P, u32]),
        "orc_sun_corner_words": (ctypes.c_int, [P, u32, u32, P]),
        "orc_sun_add": (ctypes.c_int, [P, P, u32]),
        "orc_sun_take": (ctypes.c_int, [P, u32, P, u32]),
        "orc_select": (ctypes.c_int, [P, u32, u32, P, P, u32, P]),
        "orc_cache_keys": (ctypes.c_int, [P, u32, u32, P]),
        "orc_cache_picks": (ctypes.c_int, [P, P, u32, P, ctypes.c_int, u32]),
        "orc_select_log2": (f32, [f32]),
        "orc_select_key": (u32, [P, u32, u32, u32]),
        "orc_select_threshold": (ctypes.c_int, [P, u32, P, P]),
        "orc_cache_set_meta": (ctypes.c_int, [P, u32, u32, P]),
        "orc_trace_ray": (ctypes.c_int, [P, P, P, f32, P, P]),
        "orc_mlp_forward": (ctypes.c_int, [P, P, P, u32, P, P, P]),
        "orc_mlp_gradient": (ctypes.c_int, [P, P, P, P, u32, P, P, P, P, P]),
        "orc_mlp_hash": (ctypes.c_int, [P, P, u32, P, P, P, P]),
        "orc_mlp_levels": (ctypes.c_int, [P, P]),
        "orc_mlp_exact": (ctypes.c_int, [ctypes.c_int]),
        "orc_mlp_encoding": (ctypes.c_int, [ctypes.c_int]),
        "orc_mlp_encode": (ctypes.c_int, [P, P, u32, P, P, P]),
        "orc_packed_info": (ctypes.c_int, [P, P]),
        "orc_packed_material": (u32, [P, i32, i32, i32]),
        "orc_packed_words": (i64, [P, ctypes.c_int, P, i64]),
        "orc_face_index_new": (P, [u32]),
        "orc_face_index_free": (None, [P]),
        "orc_face_index_insert": (None, [P, i32, i32, i32, u32, u32]),
        "orc_face_index_remove": (None, [P, i32, i32, i32, u32]),
        "orc_face_index_get": (i64, [P, i32, i32, i32, u32]),
        "orc_face_index_len": (i64, [P]),
        "orc_face_index_max_probe": (i64, [P]),
        "orc_face_index_capacity": (i64, [P]),
        "orc_cache_set_state": (ctypes.c_int, [P, u32, u32, P]),
        "orc_cache_accumulate": (ctypes.c_int, [P, u32, P, u32, u32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def build_aadf(blocked, full_sweep=False):
    """blocked: bool array (z, y, x). Returns uint8 (z, y, x, 6)."""
    blocked = np.ascontiguousarray(blocked, np.uint8)
    sz, sy, sx = blocked.shape
    out = np.zeros((sz, sy, sx, 6), np.uint8)
    fn = lib().orc_full_sweep_aadf if full_sweep else lib().orc_build_aadf
    assert fn(sx, sy, sz, ptr(blocked), ptr(out)) == 0
    return out


class Scene:
    """An oracle scene built from 32^3 pages."""

    def __init__(self):
        self.h = lib().orc_scene_new()

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_free(self.h)
            self.h = None

    def load_pages(self, coords, pages, removed=None):
        coords = np.ascontiguousarray(coords, np.int32)
        pages = np.ascontiguousarray(pages, np.uint16)
        removed = np.zeros((0, 3), np.int32) if removed is None else np.ascontiguousarray(removed, np.int32)
        assert lib().orc_scene_load_pages(self.h, ptr(coords), len(coords), ptr(pages), ptr(removed), len(removed)) == 0

    def edit(self, xyz, values):
        xyz = np.ascontiguousarray(xyz, np.int32)
        values = np.ascontiguousarray(values, np.uint16)
        assert lib().orc_scene_edit(self.h, ptr(xyz), ptr(values), len(values)) == 0

    def bounds(self):
        out = np.zeros(6, np.int32)
        return out if lib().orc_scene_bounds(self.h, ptr(out)) else None

    def top_level(self):
        size = np.zeros(3, np.int32)
        n = lib().orc_scene_top_level(self.h, ptr(size), None)
        if n == 0:
            return None, size
        out = np.zeros((size[2], size[1], size[0], 6), np.uint8)
        lib().orc_scene_top_level(self.h, ptr(size), ptr(out))
        return out, size

    def region_canonical(self, rx, ry, rz):
        out = np.zeros(4 + 64 + 64 * 32, np.uint32)
        lib().orc_scene_region_canonical(self.h, rx, ry, rz, ptr(out))
        return out

    def material_at(self, x, y, z):
        return lib().orc_scene_material_at(self.h, x, y, z)


class Tracer:
    def __init__(self, scene):
        self.scene = scene
        self.h = lib().orc_tracer_new(scene.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tracer_free(self.h)
            self.h = None

    def set_materials(self, diffuse, emission):
        d = np.ascontiguousarray(diffuse, np.float32)
        e = np.ascontiguousarray(emission, np.float32)
        lib().orc_tracer_set_materials(self.h, ptr(d), ptr(e))

    def set_lights(self, lights):
        lights = np.ascontiguousarray(lights, np.float32).reshape(-1, 12)
        directional = int((lights[:, 3] < 0).sum()) if len(lights) else 0
        lib().orc_tracer_set_lights(self.h, ptr(lights) if len(lights) else None, len(lights), directional)

    def set_transport(self, bounces=3, samples=16, lanes=0, environment=(0.0, 0.0, 0.0), limited_reach=False, reach=768.0,
                      max_distance=768.0):
        t = OrcTransport()
        t.bounces, t.samples, t.lanes, t.limited_reach = bounces, samples, lanes, int(limited_reach)
        t.reach, t.max_distance = reach, max_distance
        for i in range(3):
            t.environment[i] = environment[i]
        self.samples = samples
        lib().orc_tracer_set_transport(self.h, ctypes.byref(t))

    def emitter_words(self):
        n = lib().orc_tracer_emitters(self.h, None, 0)
        out = np.zeros(n, np.uint32)
        lib().orc_tracer_emitters(self.h, ptr(out), n)
        return out

    def trace_faces(self, tasks, hits=False, counters=False, threads=8):
        tasks = np.ascontiguousarray(tasks).view(np.uint32).reshape(-1, 8).copy()
        n = len(tasks)
        results = np.zeros((n, 24), np.float32)
        h = np.zeros((n * self.samples, 8), np.uint32) if hits else None
        c = np.zeros(8, np.uint64) if counters else None
        lib().orc_trace_faces(self.h, ptr(tasks), n, ptr(results), ptr(h), ptr(c), threads)
        return results, h, c

    def trace_learned(self, tasks, threads=8):
        """A learned frame's paths: (partial [n*spp, 4], factor [n*spp, 3], queries [n*spp, 4])."""
        tasks = np.ascontiguousarray(tasks).view(np.uint32).reshape(-1, 8).copy()
        n = len(tasks) * self.samples
        partial = np.zeros((n, 4), np.float32)
        factor = np.zeros((n, 3), np.float32)
        queries = np.zeros((n, 4), np.uint32)
        lib().orc_trace_learned(self.h, ptr(tasks), len(tasks), ptr(partial), ptr(factor), ptr(queries), threads)
        return partial, factor, queries

    def view_frame(self, camera, depth, normals, clock_ms=0, threads=8):
        """The oracle's view frame: (owner per pixel, VIEW_FACE records, owner count).
        `camera` is the package's ViewCamera (the same layout)."""
        d = np.ascontiguousarray(depth, np.float32)
        n = np.ascontiguousarray(normals, np.float32)
        pixels = camera.width * camera.height
        owners = np.zeros(pixels, np.uint32)
        faces = np.zeros(pixels * 28, np.uint32)
        count = np.zeros(1, np.uint32)
        lib().orc_view_frame(self.h, ctypes.byref(camera), ptr(d), ptr(n), clock_ms, ptr(owners), ptr(faces), ptr(count), threads)
        return owners, faces, int(count[0])

    def view_reset(self):
        lib().orc_view_reset(self.h)

    def view_far(self, settings):
        """The far tier's settings (the package's FarSettings, the same layout; None: off)."""
        lib().orc_view_far(self.h, ctypes.byref(settings) if settings is not None else None)

    def view_display(self, noise_threshold=0.031, spatial=1.0, blend=0.19926):
        lib().orc_view_display(self.h, noise_threshold, spatial, blend)

    def view_read(self, pixels):
        """(published records [pixels, 12] words, counters [16], far output words, far cells [n, 12]) of the last frame."""
        records = np.zeros((pixels, 12), np.uint32)
        counters = np.zeros(16, np.uint32)
        n = lib().orc_view_read(self.h, None, None, None, None)
        output = np.zeros(16 + 4 * n, np.uint32)
        cells = np.zeros((max(n, 1), 12), np.uint32)
        lib().orc_view_read(self.h, ptr(records), ptr(counters), ptr(output), ptr(cells))
        return records, counters, output, cells[:n]

    def trace_ray(self, origin, direction, max_distance=768.0):
        o = np.asarray(origin, np.float32)
        d = np.asarray(direction, np.float32)
        oi = np.zeros(6, np.int32)
        of = np.zeros(6, np.float32)
        lib().orc_trace_ray(self.h, ptr(o), ptr(d), max_distance, ptr(oi), ptr(of))
        return oi, of

    # resident cache
    def cache_enable(self, history_limit=0, recency=1.0, measured_retention=1.0, plain=False):
        lib().orc_cache_enable(self.h, history_limit, recency, measured_retention, int(plain))

    def cache_register(self, tasks):
        tasks = np.ascontiguousarray(tasks).view(np.uint32).reshape(-1, 8).copy()
        slots = np.zeros(len(tasks), np.uint32)
        lib().orc_cache_register(self.h, ptr(tasks), len(tasks), ptr(slots))
        return slots

    def cache_batch(self, tasks, slots, hits=False, threads=8, clock_ms=0, counters=False):
        """prepare -> trace -> accumulate of `tasks` on their slots: (records,
        hits or None) and, with counters=True, the counters [rays, steps,
        header loads, brick loads, cell loads, paths, write_result direct
        rays, 0] as a third item."""
        tasks = np.ascontiguousarray(tasks).view(np.uint32).reshape(-1, 8).copy()
        slots = np.ascontiguousarray(slots, np.uint32)
        n = len(tasks)
        results = np.zeros((n, 24), np.float32)
        h = np.zeros((n * self.samples, 8), np.uint32) if hits else None
        c = np.zeros(8, np.uint64) if counters else None
        lib().orc_cache_batch(self.h, ptr(tasks), ptr(slots), n, ptr(results), ptr(h), ptr(c), threads, clock_ms)
        return (results, h, c) if counters else (results, h)

    def packed_info(self):
        """[cell words, regions, directory capacity, interned bricks, brick references]."""
        out = np.zeros(8, np.int64)
        lib().orc_packed_info(self.h, ptr(out))
        return out[:5]

    def packed_words(self, which):
        """0: directory, 1: headers, 2: cells (u32 words)."""
        n = lib().orc_packed_words(self.h, which, None, 0)
        out = np.zeros(n, np.uint32)
        lib().orc_packed_words(self.h, which, ptr(out), n)
        return out

    def packed_material(self, x, y, z):
        return int(lib().orc_packed_material(self.h, x, y, z))

    def set_state(self, first, words):
        words = np.ascontiguousarray(words, np.uint32).reshape(-1, 36)
        assert lib().orc_cache_set_state(self.h, first, len(words), ptr(words)) == 0

    def accumulate(self, slot, record, samples, clock_ms=0):
        """Merge a 24-float record into `slot`; returns the merged record."""
        rec = np.ascontiguousarray(record, np.float32).reshape(24).copy()
        assert lib().orc_cache_accumulate(self.h, slot, ptr(rec), samples, clock_ms) == 0
        return rec

    def cache_state(self, first, count):
        out = np.zeros((count, 36), np.uint32)
        assert lib().orc_cache_state(self.h, first, count, ptr(out)) == 0
        return out

    def edit_event(self, xyz, values):
        """Edit the oracle scene; the edit's encoded retention event (measured
        against this tracer's resident statistics), empty when none."""
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        values = np.ascontiguousarray(values, np.uint16).reshape(-1)
        n = lib().orc_edit_event(self.h, ptr(xyz), ptr(values), len(values))
        out = np.zeros(n, np.uint32)
        if n:
            lib().orc_last_event(self.h, ptr(out), n)
        return out

    def load_event(self, coords, pages, removed=None):
        """Load (and remove) pages in the oracle scene; the upload's encoded
        retention event, empty when none."""
        coords = np.ascontiguousarray(coords, np.int32).reshape(-1, 3)
        pages = np.ascontiguousarray(pages, np.uint16).reshape(-1, 32768)
        removed = np.zeros((0, 3), np.int32) if removed is None else np.ascontiguousarray(removed, np.int32).reshape(-1, 3)
        n = lib().orc_load_event(self.h, ptr(coords), len(coords), ptr(pages), ptr(removed), len(removed))
        assert n >= 0
        out = np.zeros(n, np.uint32)
        if n:
            lib().orc_last_event(self.h, ptr(out), n)
        return out

    def cache_meta(self, first, count):
        out = np.zeros((count, 12), np.uint32)
        assert lib().orc_cache_meta(self.h, first, count, ptr(out)) == 0
        return out

    def apply_events(self, events, event_count):
        events = np.ascontiguousarray(events, np.uint32).reshape(-1)
        lib().orc_cache_apply_events(self.h, ptr(events), event_count, events.size)

    def set_schedule(self, focus=(0, 0, 0), transport_revision=0, stable_error=0.0, stable_samples=0, audit_ms=0,
                     mirror_budget=0):
        f = np.asarray(focus, np.int32)
        lib().orc_cache_schedule(self.h, ptr(f), transport_revision, stable_error, stable_samples, audit_ms, mirror_budget)

    def set_footprints(self, on):
        lib().orc_set_footprints(self.h, 1 if on else 0)

    def sun_corners(self, tasks, slots, frame=0, refresh=1, clock_ms=0):
        """Corner words of the bound faces (tasks in bound order, their
        slots): (flip region words, Light event words or empty)."""
        tasks = np.ascontiguousarray(tasks).view(np.uint32).reshape(-1, 8)
        slots = np.ascontiguousarray(slots, np.uint32)
        region = np.zeros(SUN_WORDS, np.uint32)
        cap = 1 << 20
        events = np.zeros(cap, np.uint32)
        n = lib().orc_sun_corners(self.h, ptr(tasks), ptr(slots), len(slots), frame, refresh, clock_ms, ptr(region), ptr(events), cap)
        return region, events[:n].copy()

    def sun_corner_words(self, first, count):
        out = np.zeros(count, np.uint32)
        lib().orc_sun_corner_words(self.h, first, count, ptr(out))
        return out

    def sun_add(self, boxes):
        """Bin flip boxes ((min xyz, max xyz, flux) each)."""
        words = sun_box_words(boxes)
        lib().orc_sun_add(self.h, ptr(words), len(boxes))

    def sun_take(self, now_ms):
        """The bins due at now_ms as (min, max, flux) boxes, or None."""
        out = np.zeros(7 * 4096, np.uint32)
        n = lib().orc_sun_take(self.h, now_ms, ptr(out), 4096)
        return None if n < 0 else sun_boxes_of(out[:7 * n])

    def select(self, picks, clock_ms=0):
        """One pick batch's selection passes: (1038 scratch words, mirrored
        slots, pick list)."""
        sel = np.zeros(1038, np.uint32)
        cap = 1 << 22
        mirrored = np.zeros(cap, np.uint32)
        pick_list = np.zeros(picks, np.uint32)
        n = lib().orc_select(self.h, picks, clock_ms, ptr(sel), ptr(mirrored), cap, ptr(pick_list))
        return sel, mirrored[:min(n, cap)].copy(), pick_list

    def select_key(self, slot, batch, now_ms):
        return int(lib().orc_select_key(self.h, slot, batch, now_ms))

    def set_meta(self, first, words):
        words = np.ascontiguousarray(words, np.uint32).reshape(-1, 12)
        assert lib().orc_cache_set_meta(self.h, first, len(words), ptr(words)) == 0

    def cache_keys(self, first, count):
        out = np.zeros(count, np.uint32)
        assert lib().orc_cache_keys(self.h, first, count, ptr(out)) == 0
        return out

    def cache_picks(self, pick_slots, threads=8, clock_ms=0):
        pick_slots = np.ascontiguousarray(pick_slots, np.uint32)
        return lib().orc_cache_picks(self.h, ptr(pick_slots), len(pick_slots), None, threads, clock_ms)


SUN_RECORDS = 2048
SUN_WORDS = 4 + 4 * SUN_RECORDS + 8


def sun_box_words(boxes):
    w = np.zeros((len(boxes), 7), np.uint32)
    for k, (mn, mx, flux) in enumerate(boxes):
        w[k, 0:3] = np.asarray(mn, np.int32).view(np.uint32)
        w[k, 3:6] = np.asarray(mx, np.int32).view(np.uint32)
        w[k, 6] = np.float32(flux).view(np.uint32)
    return np.ascontiguousarray(w.ravel())


def sun_boxes_of(words):
    w = np.asarray(words, np.uint32).reshape(-1, 7)
    return [(tuple(r[0:3].view(np.int32).tolist()), tuple(r[3:6].view(np.int32).tolist()), float(r[6:7].view(np.float32)[0]))
            for r in w]


def sun_decode(region, records=SUN_RECORDS):
    """A flip region as (min, max, flux) boxes (gpu/view/timing.rs:22-49)."""
    region = np.ascontiguousarray(region, np.uint32)
    out = np.zeros(7 * (records + 1), np.uint32)
    n = lib().orc_sun_decode(ptr(region), records, ptr(out), records + 1)
    return sun_boxes_of(out[:7 * n])


def event_count(words):
    """Events in an encoding of one edit's event followed by its change
    footprints' (one box each): records = 3 E + 2 (boxes + E - 1)."""
    if len(words) == 0:
        return 0
    w = np.asarray(words, np.uint32).reshape(-1, 4)
    boxes = int(w[2, 3])
    return (len(w) - 2 * boxes + 2) // 5


def select_log2(x):
    return float(lib().orc_select_log2(float(x)))


def select_threshold(histograms, picks):
    h = np.ascontiguousarray(histograms, np.uint32).reshape(1024)
    thr = np.zeros(4, np.uint32)
    rem = np.zeros(4, np.uint32)
    lib().orc_select_threshold(ptr(h), picks, ptr(thr), ptr(rem))
    return thr, rem


MLP_TABLE = 8 * (1 << 22) * 2


def mlp_forward(weights, table, queries, diffuse, scene):
    w = np.ascontiguousarray(weights, np.float32)
    t = np.ascontiguousarray(table, np.float32)
    q = np.ascontiguousarray(queries, np.uint32).reshape(-1, 4)
    d = np.ascontiguousarray(diffuse, np.float32)
    sc = np.ascontiguousarray(scene, np.float32)
    out = np.zeros((len(q), 3), np.float64)
    lib().orc_mlp_forward(ptr(w), ptr(t), ptr(q), len(q), ptr(d), ptr(sc), ptr(out))
    return out


def mlp_gradient(weights, table, queries, targets, diffuse, scene):
    """(weight gradient, hash-table gradient, loss)."""
    w = np.ascontiguousarray(weights, np.float32)
    t = np.ascontiguousarray(table, np.float32)
    q = np.ascontiguousarray(queries, np.uint32).reshape(-1, 4)
    y = np.ascontiguousarray(targets, np.float32).reshape(-1, 4)
    d = np.ascontiguousarray(diffuse, np.float32)
    sc = np.ascontiguousarray(scene, np.float32)
    grad = np.zeros(14528, np.float64)
    table_grad = np.zeros(MLP_TABLE, np.float64)
    loss = np.zeros(1, np.float64)
    lib().orc_mlp_gradient(ptr(w), ptr(t), ptr(q), ptr(y), len(q), ptr(d), ptr(sc), ptr(grad), ptr(table_grad), ptr(loss))
    return grad, table_grad, float(loss[0])


def mlp_hash(table, queries, scene):
    """Hash-grid features (n x 16) and per level the corner entries and
    weights (n x 8 x 8 each)."""
    t = np.ascontiguousarray(table, np.float32)
    q = np.ascontiguousarray(queries, np.uint32).reshape(-1, 4)
    sc = np.ascontiguousarray(scene, np.float32)
    feats = np.zeros((len(q), 16), np.float32)
    index = np.zeros((len(q), 8, 8), np.uint32)
    weight = np.zeros((len(q), 8, 8), np.float32)
    lib().orc_mlp_hash(ptr(t), ptr(q), len(q), ptr(sc), ptr(feats), ptr(index), ptr(weight))
    return feats, index, weight


def mlp_exact(on):
    lib().orc_mlp_exact(1 if on else 0)


def mlp_encoding(mode):
    """0: hash grid (default); 1: frequency encoding."""
    lib().orc_mlp_encoding(int(mode))


def mlp_encode(table, queries, diffuse, scene):
    """The 32 input features of each query (float64, before BF16 rounding)."""
    t = np.ascontiguousarray(table, np.float32)
    q = np.ascontiguousarray(queries, np.uint32).reshape(-1, 4)
    d = np.ascontiguousarray(diffuse, np.float32)
    sc = np.ascontiguousarray(scene, np.float32)
    out = np.zeros((len(q), 32), np.float64)
    lib().orc_mlp_encode(ptr(t), ptr(q), len(q), ptr(d), ptr(sc), ptr(out))
    return out


def mlp_levels(scene):
    sc = np.ascontiguousarray(scene, np.float32)
    res = np.zeros(8, np.uint32)
    lib().orc_mlp_levels(ptr(sc), ptr(res))
    return res


class FaceIndex:
    """The oracle's face index (gpu/face_index.rs) on its own."""

    def __init__(self, capacity=16):
        self.h = lib().orc_face_index_new(capacity)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_face_index_free(self.h)
            self.h = None

    def insert(self, voxel, face, slot):
        lib().orc_face_index_insert(self.h, int(voxel[0]), int(voxel[1]), int(voxel[2]), face, slot)

    def remove(self, voxel, face):
        lib().orc_face_index_remove(self.h, int(voxel[0]), int(voxel[1]), int(voxel[2]), face)

    def get(self, voxel, face):
        v = lib().orc_face_index_get(self.h, int(voxel[0]), int(voxel[1]), int(voxel[2]), face)
        return None if v < 0 else int(v)

    def __len__(self):
        return int(lib().orc_face_index_len(self.h))

    def max_probe(self):
        return int(lib().orc_face_index_max_probe(self.h))

    def capacity(self):
        return int(lib().orc_face_index_capacity(self.h))
