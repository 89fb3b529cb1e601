"""Sun corner words on the GPU against the oracle: every bound face's corner
word (sky_corner_visibility, voxel_path.wgsl:245-266, at the face's four
lattice corners), the flips of a moving sun compacted in bound-face order
(records, then the overflow box past 2,048), the Light retention event the
host publishes from them (sun_changes.rs) and the resident batch it applies
to (voxel_view_runtime.wgsl:168-268 for the world's faces).

Bar: corner words, flip regions and event words bit-identical; the next
batch's results and resident state bit-identical.
"""
import math

import numpy as np
import pytest

from conftest import gpu_tracer, make_scene
from test_oracle_sun import contact_scene, exposed

pytestmark = pytest.mark.gpu


def sun_light(elevation_deg, azimuth_deg, radiance=3.0):
    e, a = math.radians(elevation_deg), math.radians(azimuth_deg)
    light = np.zeros((1, 12), np.float32)
    light[0, 0:3] = (math.cos(e) * math.cos(a), math.sin(e), math.cos(e) * math.sin(a))
    light[0, 3] = -1.0
    light[0, 4:7] = radiance
    return light


def run_steps(vgi, orc, setup, tasks, steps, name):
    """Per step (light, frame, refresh, clock): corner words on both sides,
    then a resident batch that applies any published Light event."""
    t = setup["tracer"]
    t.set_transport(bounces=1, samples=1, environment=(0.1, 0.12, 0.15))
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=1, samples=1, environment=(0.1, 0.12, 0.15), measured_retention=1.0)
    g.bind_faces(tasks)
    slots = t.cache_register(tasks)
    assert np.array_equal(g.slots(len(tasks)), slots)
    n = int(slots.max()) + 1
    most = 0
    published = 0
    for k, (light, frame, refresh, clock) in enumerate(steps):
        print("%s step %d: frame %d refresh %d clock %d" % (name, k, frame, refresh, clock), flush=True)
        g.set_lights(light)
        t.set_lights(light)
        g.set_clock(clock)
        g.sun_corners(frame, refresh)
        region, events = t.sun_corners(tasks, slots, frame, refresh, clock_ms=clock)
        assert np.array_equal(g.sun_corner_words(0, n), t.sun_corner_words(0, n)), "%s step %d corner words" % (name, k)
        got = g.sun_flips()
        assert np.array_equal(got, region), "%s step %d flip region differs at %s" % (name, k, np.nonzero(got != region)[0][:8].tolist())
        most = max(most, int(region[0]))
        if len(events):
            t.apply_events(events, orc.event_count(events))
            published += 1
        ticket = g.submit_compact(tasks)
        gpu = g.complete_readback(ticket, len(tasks))
        cpu, _ = t.cache_batch(tasks, slots, clock_ms=clock)
        if len(events):
            assert np.array_equal(g.events(), events), "%s step %d Light event words" % (name, k)
        assert np.array_equal(gpu.view(np.uint32), cpu.view(np.uint32)), "%s step %d batch results" % (name, k)
    assert np.array_equal(g.cache_state(0, n), t.cache_state(0, n)), name
    assert np.array_equal(g.cache_meta(0, n), t.cache_meta(0, n)), name
    g.close()
    return most, published


def test_contact_shadow_sun_bit_exact(vgi, orc):
    grid, _, diffuse = contact_scene()
    coords, pages = vgi.pages_of(grid)
    scene = orc.Scene()
    scene.load_pages(coords, pages)
    t = orc.Tracer(scene)
    setup = dict(coords=coords, pages=pages, diffuse=diffuse, emission=np.zeros((256, 3), np.float32),
                 lights=sun_light(70, 180), scene=scene, tracer=t)
    t.set_materials(setup["diffuse"], setup["emission"])
    tasks = exposed(grid).view(vgi.FACE_TASK).ravel()
    d = lambda x, y: np.array([[x, y, 0.0, -1.0, 1.0, 1.0, 1.0, 0, 0, 0, 0, 0]], np.float32)
    steps = [(d(-0.35, 1.0), 0, 1, 0), (d(0.35, 1.0), 1, 1, 16), (d(-0.35, 1.0), 2, 1, 100), (d(-0.35, 1.0), 3, 1, 300)]
    for s in steps:
        s[0][0, 0:3] /= np.linalg.norm(s[0][0, 0:3])
    most, published = run_steps(vgi, orc, setup, tasks, steps, "contact shadow")
    assert most > 0 and published == 2


def test_terrain_moving_sun_bit_exact(vgi, orc):
    # A 128^3 noise terrain under a sun sweeping from 60 to 25 degrees of
    # elevation: shadow borders move across thousands of faces (the first
    # sweep overflows the 2,048 records), refreshed every call, then every
    # fourth call per face.
    setup = make_scene(vgi, orc, 4, 0x5EED0005, edge=128)
    tasks = vgi.exposed_faces(setup["grid"])
    steps = [(sun_light(60, 30), 0, 1, 0), (sun_light(25, 30), 1, 1, 300), (sun_light(24, 33), 2, 1, 320),
             (sun_light(23, 36), 3, 4, 600), (sun_light(22, 39), 4, 4, 900), (sun_light(21, 42), 5, 4, 1200)]
    most, published = run_steps(vgi, orc, setup, tasks, steps, "terrain")
    assert most > 2048, "the large sweep overflows the records (%d flips)" % most
    assert published >= 3
