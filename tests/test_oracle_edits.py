"""Pins of the oracle's edit retention events (no GPU needed): the changed
boxes, flux, scan budget, reach filter and GPU-style measurement of an edit,
against the reference's own tests of that pipeline.

Event words (gpu/resident.rs:1319-1417): three uint4 headers per event
(union min, union max, (0, 0, first box, boxes)), then two uint4 per box
((min, flux bits), (max, kind | urgent << 3)); kinds 0 unknown flux, 3 edit,
4 light (a change footprint), 6 edit to be measured. The edit's event comes
first, then one footprint event per edited box (gpu/resident.rs:1215-1270).
"""
import math

import numpy as np

import oracle_lib as orc

f32 = np.float32
AIR = 256
W = (f32(0.2126), f32(0.7152), f32(0.0722))


def lum(rgb):
    return f32(f32(W[0] * f32(rgb[0])) + f32(W[1] * f32(rgb[1]))) + f32(W[2] * f32(rgb[2]))


def pages_of(grid):
    z, y, x = grid.shape
    g = grid.reshape(z // 32, 32, y // 32, 32, x // 32, 32).transpose(0, 2, 4, 1, 3, 5)
    pz, py, px = np.meshgrid(np.arange(z // 32), np.arange(y // 32), np.arange(x // 32), indexing="ij")
    coords = np.stack([px.ravel(), py.ravel(), pz.ravel()], axis=1).astype(np.int32)
    return coords, np.ascontiguousarray(g.reshape(-1, 32768))


def floor_scene(edge=64, emitters=()):
    """An open floor (y 0..3, material 1) with emissive voxels (material 7)."""
    grid = np.full((edge, edge, edge), AIR, np.uint16)
    grid[:, 0:4, :] = 1
    for x, y, z in emitters:
        grid[z, y, x] = 7
    diffuse = np.full((256, 3), 0.5, np.float32)
    diffuse[2] = (0.3, 0.6, 0.9)
    emission = np.zeros((256, 3), np.float32)
    emission[7] = (1.0, 2.0, 3.0)
    scene = orc.Scene()
    scene.load_pages(*pages_of(grid))
    tracer = orc.Tracer(scene)
    tracer.set_materials(diffuse, emission)
    sun = np.zeros(12, np.float32)
    sun[0:3] = np.array([0.3, 0.8, -0.5], np.float32) / np.linalg.norm([0.3, 0.8, -0.5])
    sun[3] = -1.0
    sun[4:7] = 3.0
    tracer.set_lights(sun.reshape(1, 12))
    tracer.set_transport(bounces=2, samples=16)
    tracer.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    return grid, scene, tracer, diffuse, emission


def boxes(words):
    """(min, max, flux, kind, urgent) of each box of the edit's event (the
    first of the encoding; the rest are its change footprints)."""
    w = words.reshape(-1, 4)
    n = int(w[2, 3])
    first = 3 * orc.event_count(words)
    out = []
    for b in range(n):
        lo, hi = w[first + 2 * b], w[first + 1 + 2 * b]
        out.append((tuple(lo[:3].view(np.int32)), tuple(hi[:3].view(np.int32)), lo[3:4].view(np.float32)[0], int(hi[3] & 7),
                    bool(hi[3] & 8)))
    return w, out


def test_single_voxel_edit_is_one_urgent_box():
    """geometry.rs:140-190 (tight box of the changed voxels in one region),
    retention.rs:151-158 (a box holding the edit is urgent),
    voxel_accumulate.wgsl:415-421 (a measured edit no resident face observed
    falls back to unknown flux)."""
    _, _, tracer, _, _ = floor_scene()
    words = tracer.edit_event([[20, 4, 20]], [2])
    w, b = boxes(words)
    assert len(words) == 4 * (2 * 3 + 2 * 2)  # the edit's event and its footprint's
    assert tuple(w[0, :3].view(np.int32)) == (20, 4, 20) and tuple(w[1, :3].view(np.int32)) == (20, 4, 20)
    assert b == [((20, 4, 20), (20, 4, 20), f32(0.0), 0, True)]


def test_emission_flux_is_six_faces_per_changed_voxel():
    """geometry.rs:160-167: |change of emission luminance| x 6 faces x volume,
    summed over the region's changed voxels."""
    _, _, tracer, _, emission = floor_scene(emitters=[(10, 4, 10), (11, 4, 10)])
    words = tracer.edit_event([[10, 4, 10], [11, 4, 10], [12, 4, 10]], [AIR, AIR, AIR])  # (12,4,10) was air already
    _, b = boxes(words)
    assert len(b) == 1
    (mn, mx, flux, kind, urgent), = b
    assert mn == (10, 4, 10) and mx == (11, 4, 10)
    e = lum(emission[7])
    expected = f32(f32(abs(e) * f32(6.0)) * f32(1.0))
    assert flux == f32(expected + expected)
    assert kind == 0 and urgent  # measured, unobserved: unknown flux in the kind bits


def test_scan_budget_leaves_later_boxes_unknown():
    """publication.rs:413-454 and its test edits_are_measured_on_the_gpu_within
    _the_scan_budget: an edit is measured while its one-voxel-margin scan fits
    the 16384-position budget; past it the flux is unknown (0, kind 0).
    Three whole emissive regions cleared: 18^3 = 5832 positions each."""
    grid = np.full((64, 64, 64), AIR, np.uint16)
    grid[:, 0:4, :] = 1
    grid[0:16, 16:32, 0:48] = 7  # regions (0,1,0), (1,1,0), (2,1,0) emissive
    diffuse = np.full((256, 3), 0.5, np.float32)
    emission = np.zeros((256, 3), np.float32)
    emission[7] = (1.0, 2.0, 3.0)
    scene = orc.Scene()
    scene.load_pages(*pages_of(grid))
    tracer = orc.Tracer(scene)
    tracer.set_materials(diffuse, emission)
    tracer.set_transport(bounces=1, samples=1)
    tracer.cache_enable()
    zz, yy, xx = np.nonzero(grid == 7)
    xyz = np.stack([xx, yy, zz], axis=1).astype(np.int32)
    words = tracer.edit_event(xyz, np.full(len(xyz), AIR, np.uint16))
    _, b = boxes(words)
    assert [x[0] for x in b] == [(0, 16, 0), (16, 16, 0), (32, 16, 0)]
    # Uniform before and after: one record of the whole region
    # (geometry.rs:170-172), volume 4096 in a single product.
    per_region = f32(f32(lum(emission[7]) * f32(6.0)) * f32(4096.0))
    assert b[0][2] == per_region and b[1][2] == per_region  # measured (unobserved): base flux kept
    assert b[2][2] == f32(0.0) and b[2][3] == 0  # over budget: unknown


def test_reach_filter_keeps_changes_near_solid_pages():
    """geometry.rs:299-403: a change is kept only within the influence radius
    (192, retention.rs:32-39) of a page that held a solid region."""
    grid = np.full((256, 256, 256), AIR, np.uint16)
    grid[0:8, 0:8, 0:8] = 1
    diffuse = np.full((256, 3), 0.5, np.float32)
    emission = np.zeros((256, 3), np.float32)
    scene = orc.Scene()
    scene.load_pages(*pages_of(grid))
    tracer = orc.Tracer(scene)
    tracer.set_materials(diffuse, emission)
    tracer.set_transport(bounces=1, samples=1)
    tracer.cache_enable()
    # Page (0,0,0) spans 0..31. Kept edits make their pages solid, so the
    # far cases go first. (224,31,31) is 193 away, (250,250,250) 379.
    assert len(tracer.edit_event([[224, 31, 31]], [1])) == 0
    assert len(tracer.edit_event([[250, 250, 250]], [1])) == 0
    # (223,31,31): exactly 192 from the page, still in reach; (100,10,10): 69.
    # Kept: the edit's event (three headers, one box) and its footprint's.
    assert len(tracer.edit_event([[223, 31, 31]], [1])) == 4 * 10
    assert len(tracer.edit_event([[100, 10, 10]], [1])) == 4 * 10


def test_measured_flux_and_urgency_follow_the_reference_statement():
    """gpu/resident.rs:2406-2500 (gpu_measures_edit_flux_and_marks_faces_whose
    _history_it_binds): after four batches on an open floor, an edit at
    (8,4,8) is measured as base + 2 x the brightest outgoing luminance
    (albedo luminance x cached incident + emission luminance) of any resident
    face within one voxel; retention then marks urgent exactly the faces
    within 64 voxels whose history the change binds."""
    grid, scene, tracer, diffuse, emission = floor_scene()
    tracer.set_footprints(False)  # as the reference's test: the CPU-planned event alone
    tracer.set_transport(bounces=2, samples=16, environment=(0.4, 0.45, 0.5))  # a sky: faces are lit
    faces = []
    for z in range(0, 32):
        for x in range(0, 32):
            faces.append((x, 3, z, 2, 0, 0, 0, 0))  # +y face of the floor top
    tasks = np.array(faces, np.uint32)
    slots = tracer.cache_register(tasks)
    for frame in range(4):
        tracer.cache_batch(tasks, slots, clock_ms=16 * frame)
    state = tracer.cache_state(0, len(slots)).view(np.float32)
    count = tracer.cache_state(0, len(slots))[:, 5] & 0x7FFFFFFF
    edit = np.array([8, 4, 8])
    light_d = lum(diffuse[1])
    brightest = f32(0.0)
    for i, f in enumerate(faces):
        v = np.array(f[:3])
        if np.abs(v - edit).max() <= 1 and count[i] > 0:
            incident = f32(lum(state[i, 0:3]) + lum(state[i, 16:19]))
            brightest = max(brightest, f32(f32(light_d * incident) + f32(0.0)))
    assert brightest > 0
    flux = f32(0.0) + f32(f32(brightest * f32(2.0)) * f32(1.0))
    words = tracer.edit_event([edit], [2])
    w, b = boxes(words)
    assert len(b) == 1
    measured, kind = b[0][2], b[0][3]
    assert abs(measured - flux) <= 1e-6 * flux, (measured, flux)
    assert kind == 3
    tracer.apply_events(words, 1)
    marks = (tracer.cache_meta(0, len(slots))[:, 11] & 0x40000000) != 0
    urgent = checked = 0
    for i, f in enumerate(faces):
        d2 = float(((np.array(f[:3]) - edit) ** 2).sum())
        change = f32(flux / f32(math.pi * max(d2, 1.0)))
        mean = state[i, 0:3]
        variance = lum(np.maximum(state[i, 8:11] - mean * mean, 0))
        binds = lum(mean) < f32(0.05) * change or variance / (change * change) < f32(count[i])
        expected = d2 <= 64.0 * 64.0 and binds
        assert marks[i] == expected, (f, d2)
        urgent += int(marks[i])
        checked += 1
    assert 0 < urgent < checked
