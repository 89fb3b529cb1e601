"""The far per-pixel tier and the view presentation on the GPU against the
oracle, frame after frame.

The reference's far-tier rig (gpu/view/pixel_tests.rs:24-330) in the
configurations its native tests use -- 2-pixel cells at rate 1 with a
one-cell pan, the 1440p defaults (3-pixel cells, rate 4) for 36 frames, a
custom light, a one-step ray budget under an emissive slab, a shortened
reach, half the view beyond the loaded page (the geometry interest) -- and a 64^3 terrain whose distant faces fall to the far tier while
near ones own view receivers, with the resident world cache exchanging
estimates and presentation: owners, receivers, estimates (with the shown
presentation), published records, counters, the material's output and the
cells bit-identical every frame, the resident state after each frame too.
"""
import numpy as np
import pytest

import far_rig as rig
from conftest import gpu_tracer, make_scene
from view_helpers import camera, prepass

pytestmark = pytest.mark.gpu


def compare(vgi, g, t, cam, depth, normals, label, clock=0):
    print("far frame: %s" % label, flush=True)
    g.view_frame(cam, depth, normals)
    pixels = cam.width * cam.height
    cown, cwords, ccount = t.view_frame(cam, depth, normals, clock_ms=clock)
    gown, gcount, overflow = g.view_owners(pixels)
    assert overflow == 0 and gcount == ccount and np.array_equal(gown, cown), "%s: owners" % label
    gw = g.view_faces(0, pixels).view(np.uint32).reshape(pixels, 28)
    bad = np.flatnonzero(np.any(gw != cwords.reshape(pixels, 28), axis=1))
    assert len(bad) == 0, "%s: %d faces differ, first pixel %d" % (label, len(bad), bad[0])
    crec, ccounters, cout, ccells = t.view_read(pixels)
    grec = g.view_pixels(0, pixels).view(np.uint32).reshape(pixels, 12)
    bad = np.flatnonzero(np.any(grec != crec, axis=1))
    assert len(bad) == 0, "%s: %d published records differ, first pixel %d" % (label, len(bad), bad[0])
    assert np.array_equal(g.view_counters(), ccounters), "%s: counters %s vs %s" % (label, g.view_counters(), ccounters)
    if len(ccells):
        gout, gcells = g.view_far_output(len(ccells))
        assert np.array_equal(gout, cout), "%s: far output differs at words %s" % (label, np.flatnonzero(gout != cout)[:8])
        assert np.array_equal(gcells, ccells), "%s: far cells differ at %s" % (label, np.flatnonzero(np.any(gcells != ccells, axis=1))[:8])
    return ccounters, cout, ccells


@pytest.mark.parametrize("case", ["pan", "defaults", "custom", "one_step", "reach", "interest"])
def test_far_rig_bit_exact(vgi, orc, case):
    kw = dict(pan=dict(cell=2, rate=1), defaults=dict(cell=3, rate=4), custom=dict(cell=3, rate=4, custom_light=(16.0, 10.0, 16.0)),
              one_step=dict(cell=3, rate=4, ceiling=5, custom_light=(16.0, 31.0, 16.0), environment=(0.0, 0.0, 0.0), steps=1),
              reach=dict(cell=3, rate=1, ceiling=5, environment=(0.0, 0.0, 0.0), reach=16.0),
              interest=dict(cell=2, rate=1))[case]
    t = rig.oracle(vgi, **kw)
    g = rig.gpu(vgi, **kw)
    depth, normals = rig.prepass()
    frames = 36 if case == "defaults" else 8
    for n in range(frames):
        # "interest": panned 16 voxels, half the view lies beyond the loaded page.
        shift = 4.0 if case == "pan" and n == frames - 1 else (16.0 if case == "interest" else 0.0)
        counters, output, _ = compare(vgi, g, t, rig.camera(vgi, n, shift), depth, normals, "%s %d" % (case, n))
        assert counters[0] > 0 and output[15] == n
    g.close()


def test_terrain_near_and_far_with_presentation_and_exchange_bit_exact(vgi, orc):
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=64)
    env = tuple(setup["env"])
    faces = vgi.exposed_faces(setup["grid"])
    t = setup["tracer"]
    t.set_transport(bounces=2, samples=4, environment=env)
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=2, samples=4, environment=env)
    slots = t.cache_register(faces)
    for n in range(17):
        g.set_clock(16 * n)
        ticket = g.submit_compact(faces, 0)
        g.complete_readback(ticket, len(faces))
        t.cache_batch(faces, slots, clock_ms=16 * n)
    w, h = 64, 48
    target = (32.0, 8.0, 32.0)
    # At 60 degrees of field of view a voxel is 41.6 / distance pixels
    # across: faces beyond ~60 voxels are far at a 0.7-pixel threshold.
    # The threshold per frame: at 0.3 (frame 3) faces cross from the far tier
    # back to receivers, seeded from the cells; at 1.5 (frame 4) receivers'
    # faces fall to the far tier, seeded from the receivers.
    thresholds = [0.7, 0.7, 0.7, 0.3, 1.5]
    eyes = [(-20.0, 50.0, -20.0), (-19.0, 50.5, -20.5), (-18.0, 51.0, -21.0), (-17.0, 51.0, -21.0), (-17.0, 51.0, -21.0)]
    for k, (eye, threshold) in enumerate(zip(eyes, thresholds)):
        far = vgi.far_settings(voxel_pixels=threshold, measure=1, cell=2, history=32.0, rate=2, steps=128)
        g.view_far(far)
        t.view_far(far)
        cfw, wfc = camera(eye, target, w, h)
        depth, normals = prepass(t, eye, cfw, wfc, w, h)
        forward = np.asarray(target) - np.asarray(eye)
        forward = forward / np.linalg.norm(forward)
        ppu = (1.0 / np.tan(np.radians(30.0))) * h * 0.5
        cam = vgi.view_camera(wfc, eye, w, h, samples=2, clip_from_world=cfw, forward=forward, pixels_per_unit=ppu, frame=100 + k)
        clock = 16 * (20 + k)
        g.set_clock(clock)
        counters, _, _ = compare(vgi, g, t, cam, depth, normals, "terrain camera %d" % k, clock)
        print("  far cells owned %d traced %d reprojected %d seeded %d; receivers seeded %d; displayed %d" % (
            counters[0], counters[1], counters[3], counters[4], counters[5], counters[11]), flush=True)
        assert counters[0] > 50 and counters[11] > 10, "both tiers in view"
        if k == 3:
            assert counters[5] > 50, "receivers seeded from far cells"
        if k == 4:
            assert counters[4] > 50, "far cells seeded from receivers"
        hw = int(max(slots)) + 1
        assert np.array_equal(g.cache_state(0, hw), t.cache_state(0, hw)), "camera %d: resident state" % k
        assert np.array_equal(g.cache_meta(0, hw), t.cache_meta(0, hw)), "camera %d: resident metadata" % k
    g.close()
