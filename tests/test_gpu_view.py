"""The view tier on the GPU against the oracle, frame after frame.

A 64^3 noise terrain with a resident world cache (half its exposed faces
resident and traced 20 batches, world-cache termination on), then view frames
from a moving camera: discovery, owners, reprojection, the exchange with the
resident cache (loading the authority on admission, promotion of a face with
more view history than a newly resident slot, storing back) and the traced
estimates. Bar: owners, receivers and estimates bit-exact every frame, and the
resident state and metadata bit-exact after each view frame.

The matched-prefix pin (gpu/view_tests.rs:185-247): a view frame's running
estimate over samples 0..n-1 of a face equals the world tier's record of the
same face and samples within 2e-5 (1 + |radiance|); known counts equal; the
direct term is the same.
"""
import numpy as np
import pytest

from conftest import gpu_tracer, make_scene
from view_helpers import camera, prepass

pytestmark = pytest.mark.gpu


def compare_frame(vgi, g, t, cam, depth, normals, clock, label):
    print("view frame: %s" % label, flush=True)
    g.view_frame(cam, depth, normals)
    pixels = cam.width * cam.height
    gown, gcount, overflow = g.view_owners(pixels)
    gfaces = g.view_faces(0, pixels)
    cown, cwords, ccount = t.view_frame(cam, depth, normals, clock_ms=clock)
    assert overflow == 0
    assert np.array_equal(gown, cown), "%s: owners differ" % label
    assert gcount == ccount
    cfaces = cwords.view(vgi.VIEW_FACE)
    gw, cw = gfaces.view(np.uint32).reshape(pixels, 28), cwords.reshape(pixels, 28)
    bad = np.flatnonzero(np.any(gw != cw, axis=1))
    assert len(bad) == 0, "%s: %d pixels differ, first %d: gpu %s cpu %s" % (label, len(bad), bad[0], gfaces[bad[0]], cfaces[bad[0]])
    return gown, gfaces


def test_view_frames_and_exchange_bit_exact(vgi, orc):
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=64)
    env = tuple(setup["env"])
    faces = vgi.exposed_faces(setup["grid"])
    t = setup["tracer"]
    t.set_transport(bounces=2, samples=4, environment=env)
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=2, samples=4, environment=env)
    half = faces[::2]
    slots = t.cache_register(half)
    for frame in range(20):
        g.set_clock(16 * frame)
        ticket = g.submit_compact(half, 0)
        g.complete_readback(ticket, len(half))
        t.cache_batch(half, slots, clock_ms=16 * frame)
    n = len(half)
    assert np.array_equal(g.cache_state(0, n), t.cache_state(0, n))
    w, h = 40, 30
    eyes = [(-6.0, 40.0, -6.0), (-5.0, 40.5, -6.5), (-4.0, 41.0, -7.0)]
    target = (32.0, 8.0, 32.0)
    for k, eye in enumerate(eyes):
        if k == 2:
            # Every exposed face resident now: faces the view already
            # measured get new slots with fewer samples (promotion).
            slots = t.cache_register(faces)
            g.set_clock(16 * (20 + k))
            ticket = g.submit_compact(faces, 0)
            g.complete_readback(ticket, len(faces))
            t.cache_batch(faces, slots, clock_ms=16 * (20 + k))
        cfw, wfc = camera(eye, target, w, h)
        depth, normals = prepass(t, eye, cfw, wfc, w, h)
        assert (depth > 0).mean() > 0.5
        cam = vgi.view_camera(wfc, eye, w, h, samples=4, measured_retention=1.0)
        clock = 16 * (21 + k)
        g.set_clock(clock)
        own, f = compare_frame(vgi, g, t, cam, depth, normals, clock, "camera %d" % k)
        owners = np.flatnonzero(own == np.arange(w * h))
        admitted = (f["identity"][owners, 3] != 0).sum()
        print("  %d owners, %d admitted to the world cache" % (len(owners), admitted), flush=True)
        assert len(owners) > 100
        resident = {(int(v[0]), int(v[1]), int(v[2]), int(fc)) for v, fc in zip(
            (faces if k == 2 else half)["voxel"], (faces if k == 2 else half)["face"])}
        in_cache = np.array([(int(v[0]), int(v[1]), int(v[2]), int(fc)) in resident
                             for v, fc in zip(f["voxel"][owners], f["face"][owners])])
        # Every visible resident face is admitted; faces on the scene's outer
        # boundary (not bound) and, before the third frame, the other half
        # stay the view's own.
        assert np.array_equal(f["identity"][owners, 3] != 0, in_cache)
        assert 0 < admitted < len(owners)
        if k == 2:
            # Faces measured by the view alone in the first two frames (8
            # samples) against a new slot's 4: the view's history wins.
            assert (f["mean"][owners, 3] == 12.0).sum() > 0, "promotion"
        hw = int(max(slots)) + 1
        assert np.array_equal(g.cache_state(0, hw), t.cache_state(0, hw)), "camera %d: resident state" % k
        assert np.array_equal(g.cache_meta(0, hw), t.cache_meta(0, hw)), "camera %d: resident metadata" % k
    # A resident batch after the view frames reads the stored authority.
    g.set_clock(16 * 30)
    ticket = g.submit_compact(faces, 0)
    gpu = g.complete_readback(ticket, len(faces))
    cpu, _ = t.cache_batch(faces, slots, clock_ms=16 * 30)
    assert np.array_equal(gpu.view(np.uint32), cpu.view(np.uint32))
    g.close()


def test_view_estimates_match_the_world_tier_prefix(vgi, orc):
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=64)
    env = tuple(setup["env"])
    t = setup["tracer"]
    samples = 16
    g = gpu_tracer(vgi, setup)
    g.set_transport(bounces=3, samples=samples, lanes=1, environment=env)
    w, h = 32, 24
    eye, target = (-6.0, 40.0, -6.0), (32.0, 8.0, 32.0)
    cfw, wfc = camera(eye, target, w, h)
    depth, normals = prepass(t, eye, cfw, wfc, w, h)
    g.view_frame(vgi.view_camera(wfc, eye, w, h, samples=samples), depth, normals)
    own, count, _ = g.view_owners(w * h)
    f = g.view_faces(0, w * h)
    owners = np.flatnonzero(own == np.arange(w * h))
    assert count == len(owners) > 50
    tasks = np.zeros(len(owners), vgi.FACE_TASK)
    tasks["voxel"] = f["voxel"][owners]
    tasks["face"] = f["face"][owners]
    ticket = g.submit_compact(tasks, 0)
    near = g.complete_readback(ticket, len(tasks))
    g.close()
    mean = f["mean"][owners]
    assert np.array_equal(mean[:, 3], near[:, 5]), "known black paths count; unknown ones do not"
    err = np.abs(mean[:, 0:3] - near[:, 0:3]).max(axis=1)
    bound = 2e-5 * (1.0 + near[:, 0:3].max(axis=1))
    print("matched prefix: worst error / bound %.3f" % float((err / bound).max()), flush=True)
    assert np.all(err <= bound)
    assert np.array_equal(f["direct"][owners, 0:3], near[:, 16:19])
    assert np.all(f["sequence"][owners] == samples)
