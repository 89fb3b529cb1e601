"""Pins of the oracle's sun corner words and flips (no GPU needed):
sky_corner_visibility (voxel_path.wgsl:245-266), the per-face words and flip
flux (voxel_view_runtime.wgsl:168-268), the flip region's decoding
(gpu/view/timing.rs:22-49) and SunChanges' binning and publishing
(sun_changes.rs), against the reference's own tests of them
(gpu/view/timing.rs:420-449, sun_changes.rs:75-108, sky.rs:100-115) and
against the contact-shadow fixture (sky.rs:62-94) worked out by hand.

A corner word holds one byte per face corner (corner & 1 steps along the
face's first axis, corner >> 1 along its second): bit 7 known, bit i "sees
coloured directional light i".
"""
import math

import numpy as np

import oracle_lib as orc
from test_oracle_edits import AIR, lum, pages_of

f32 = np.float32
BIAS = 1 << 30


def test_flip_region_overflow_decodes_to_one_bounding_box():
    # gpu/view/timing.rs:420-449, word for word.
    region = np.zeros(orc.SUN_WORDS, np.uint32)
    region[0] = orc.SUN_RECORDS + 5
    region[4:8] = [np.uint32(np.int32(-3).view(np.uint32)), 4, 5, np.float32(0.5).view(np.uint32)]
    overflow = 4 + 4 * orc.SUN_RECORDS
    region[overflow:overflow + 7] = [BIAS + 10, BIAS - 2, BIAS, BIAS + 7, BIAS + 9, BIAS + 1, 128]
    flips = orc.sun_decode(region)
    assert len(flips) == orc.SUN_RECORDS + 1
    assert flips[0] == ((-3, 4, 5), (-3, 4, 5), 0.5)
    assert flips[-1] == ((-10, 2, 0), (7, 9, 1), 2.0)
    region[0] = 3
    assert len(orc.sun_decode(region)) == 3


def empty_tracer():
    scene = orc.Scene()
    scene.load_pages(np.zeros((1, 3), np.int32), np.full((1, 32 ** 3), AIR, np.uint16))
    return orc.Tracer(scene)


def test_flips_merge_per_region_and_publish_a_few_times_per_second():
    # sun_changes.rs:75-108 on the tracer clock (ms).
    t = empty_tracer()
    now = 1000
    assert t.sun_take(now) is None, "nothing flipped"
    t.sun_add([((1, 2, 3), (1, 2, 3), 0.5), ((4, 2, 3), (4, 2, 3), 0.25), ((40, 2, 3), (40, 2, 3), 1.0)])
    due = t.sun_take(now)
    assert len(due) == 2, "one bound per region"
    assert ((1, 2, 3), (4, 2, 3), 0.75) in due and ((40, 2, 3), (40, 2, 3), 1.0) in due
    t.sun_add([((0, 0, 0), (80, 80, 80), 1.0)])
    assert t.sun_take(now + 100) is None, "merging"
    assert t.sun_take(now + 250) == [((0, 0, 0), (80, 80, 80), 1.0)]


def contact_scene(direction=(-0.35, 1.0, 0.0)):
    """sky.rs:62-94: a 5x5 floor (x, z 7..11, y 4) with a wall along x = 7
    (y 5), one sun of radiance 1."""
    grid = np.full((32, 32, 32), AIR, np.uint16)
    grid[7:12, 4, 7:12] = 2
    grid[7:12, 5, 7] = 2
    diffuse = np.full((256, 3), 0.5, np.float32)
    diffuse[2] = (0.3, 0.6, 0.9)
    scene = orc.Scene()
    scene.load_pages(*pages_of(grid))
    t = orc.Tracer(scene)
    t.set_materials(diffuse, np.zeros((256, 3), np.float32))
    set_sun(t, direction)
    t.set_transport(bounces=1, samples=1)
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    return grid, t, diffuse


def set_sun(t, direction):
    sun = np.zeros((1, 12), np.float32)
    d = np.array(direction, np.float64)
    sun[0, 0:3] = d / np.linalg.norm(d)
    sun[0, 3] = -1.0
    sun[0, 4:7] = 1.0
    t.set_lights(sun)


FACE_STEP = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


def exposed(grid):
    faces = []
    for z, y, x in zip(*np.nonzero(grid < AIR)):
        for f, (dx, dy, dz) in enumerate(FACE_STEP):
            nx, ny, nz = x + dx, y + dy, z + dz
            inside = 0 <= nx < 32 and 0 <= ny < 32 and 0 <= nz < 32
            if not inside or grid[nz, ny, nx] >= AIR:
                faces.append((int(x), int(y), int(z), f, 0, 0, 0, 0))
    return np.array(faces, np.uint32)


def corner_position(voxel, face, corner):
    # sky.rs:43-58 (face_axes, corner_position).
    u, v = [((0, 1, 0), (0, 0, 1)), ((1, 0, 0), (0, 0, 1)), ((1, 0, 0), (0, 1, 0))][face // 2]
    off = [max(c, 0) for c in FACE_STEP[face]]
    return tuple(voxel[a] + off[a] + u[a] * (corner & 1) + v[a] * ((corner >> 1) & 1) for a in range(3))


def test_contact_shadow_corner_words():
    grid, t, _ = contact_scene()
    tasks = exposed(grid)
    slots = t.cache_register(tasks)
    region, events = t.sun_corners(tasks, slots)
    assert region[0] == 0 and len(events) == 0, "first words: nothing to flip against"
    words = dict(zip(map(tuple, tasks[:, :4].tolist()), t.sun_corner_words(0, len(tasks)).tolist()))
    # The sun comes from -x and above: a floor corner at x = 8 starts inside
    # the wall (x 7..8, y 5..6) and sees nothing; one at x = 9 clears the
    # wall's top (its ray is at y 7.9 when it passes x = 8).
    assert words[(8, 4, 8, 2)] == 0x81808180
    assert words[(8, 4, 9, 2)] == 0x81808180
    assert words[(9, 4, 8, 2)] == 0x81818181
    assert words[(7, 5, 8, 2)] == 0x81818181  # the wall's top
    assert words[(7, 5, 8, 1)] == 0x81818181  # its -x side faces the sun
    assert words[(7, 5, 8, 0)] == 0x80808080  # its +x side does not: no trace, known unlit
    assert words[(8, 4, 8, 3)] == 0x80808080  # underside
    # sky.rs:100-115: corners are lattice points, so every face corner at the
    # same point (same side of the surface) holds the same byte.
    seen = {}
    for (x, y, z, f), w in words.items():
        if w == 0x80808080:
            continue
        for c in range(4):
            key = (corner_position((x, y, z), f, c), f)
            b = (w >> (8 * c)) & 255
            assert seen.setdefault(key, b) == b, key


def test_moving_sun_flips_and_light_event():
    grid, t, diffuse = contact_scene()
    tasks = exposed(grid)
    slots = t.cache_register(tasks)
    t.sun_corners(tasks, slots, clock_ms=0)
    before = t.sun_corner_words(0, len(tasks))
    set_sun(t, (0.35, 1.0, 0.0))
    region, events = t.sun_corners(tasks, slots, clock_ms=16)
    after = t.sun_corner_words(0, len(tasks))
    flips = orc.sun_decode(region)
    n = int(region[0])
    assert 0 < n < orc.SUN_RECORDS and len(flips) == n
    # Under the new sun (from +x) the floor corners at x = 8 are lit: corners
    # 0 and 2 of the top faces of (8, 4, 8) and (8, 4, 9) flip, each a quarter
    # face of white light at cos = 1/|(0.35, 1, 0)| times the floor's albedo.
    d = np.array([0.35, 1.0, 0.0], np.float32) / np.float32(np.linalg.norm([0.35, 1.0, 0.0]))
    d = d.astype(np.float32)
    norm = f32(math.sqrt(float(f32(f32(d[0] * d[0]) + f32(d[1] * d[1])) + f32(d[2] * d[2]))))
    cosine = f32(d[1] / norm)
    quarter = f32(f32(f32(0.25) * lum([1.0, 1.0, 1.0])) * cosine)
    expected = f32(f32(quarter + quarter) * lum(diffuse[2]))
    rec = {f[0]: f[2] for f in flips}
    assert rec[(8, 4, 8)] == expected and rec[(8, 4, 9)] == expected
    # Records: exactly the faces whose known corners changed, in bound order.
    def flipped(i):
        a, b = int(before[i]), int(after[i])
        facing = np.dot(FACE_STEP[int(tasks[i, 3])], d) > 0  # traced under the new sun
        return facing and any((a >> (8 * c)) & (b >> (8 * c)) & 128 and ((a ^ b) >> (8 * c)) & 1 for c in range(4))

    changed = [i for i in range(len(tasks)) if flipped(i)]
    assert [f[0] for f in flips] == [tuple(tasks[i, :3].astype(np.int32).tolist()) for i in changed]
    # One Light event: the flips binned per 16^3 region, kind 4, flux summed.
    w = events.reshape(-1, 4)
    assert orc.event_count(events) == 1 and int(w[2, 3]) == 1, "all flips lie in region (0, 0, 0)"
    lo, hi = w[3], w[4]
    mins = np.array([f[0] for f in flips]).min(axis=0)
    maxs = np.array([f[0] for f in flips]).max(axis=0)
    assert tuple(lo[:3].view(np.int32)) == tuple(mins) and tuple(hi[:3].view(np.int32)) == tuple(maxs)
    total = f32(0.0)
    for f in flips:
        total = f32(total + f32(f[2]))
    assert lo[3:4].view(np.float32)[0] == total and hi[3] == 4
    # The sun back where it was: its flips are binned but not published
    # within 250 ms of the last publication; a later call with no new flips
    # publishes them.
    set_sun(t, (-0.35, 1.0, 0.0))
    region, events = t.sun_corners(tasks, slots, clock_ms=100)
    assert int(region[0]) > 0 and len(events) == 0
    held = orc.sun_decode(region)
    assert np.array_equal(t.sun_corner_words(0, len(tasks)), before)
    quiet, events = t.sun_corners(tasks, slots, clock_ms=300)
    assert int(quiet[0]) == 0, "the sun did not move"
    total = f32(0.0)
    for f in held:
        total = f32(total + f32(f[2]))
    w = events.reshape(-1, 4)
    assert orc.event_count(events) == 1 and w[3, 3:4].view(np.float32)[0] == total


def test_refresh_is_staggered_by_bound_index():
    grid, t, _ = contact_scene()
    tasks = exposed(grid)
    slots = t.cache_register(tasks)
    t.sun_corners(tasks, slots, frame=0, refresh=4)  # unknown words trace at once
    before = t.sun_corner_words(0, len(tasks))
    set_sun(t, (0.35, 1.0, 0.0))
    t.sun_corners(tasks, slots, frame=1, refresh=4)
    after = t.sun_corner_words(0, len(tasks))
    _, t2, _ = contact_scene((0.35, 1.0, 0.0))
    s2 = t2.cache_register(tasks)
    t2.sun_corners(tasks, s2)
    target = t2.sun_corner_words(0, len(tasks))
    for i in range(len(tasks)):
        due = (1 + i) % 4 == 0
        assert after[i] == (target[i] if due else before[i]), i
    assert any(before[i] != target[i] for i in range(len(tasks)) if (1 + i) % 4 == 0)
