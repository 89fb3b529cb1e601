"""Pins of the oracle's view presentation (no GPU needed): reconstruct_view
and publish_view (voxel_view_runtime.wgsl:350-521) and the presentation's
round trip through the world cache (voxel_view_cache.wgsl:62-68,239-249).

Closed forms and invariants: with spatial strength 0 a face shows its own
mean (first frame) and then the temporal blend's exponential average of its
means; with pooling on, every shown channel is a convex combination of the
means of the face and of coplanar owners within 16 voxels and four 2x2 tiles
(so a face on another plane, or farther away, never leaks in); a face
whose relative standard error is under the threshold is shown alone;
the published records carry the shown rgb, evidence max(count, 0.001) and
no face for non-owners; an admitted face stores its presentation in the
slot's display record and loads it back.
"""
import numpy as np

import oracle_lib as orc
from test_oracle_edits import AIR, pages_of
from view_helpers import camera, prepass


def package():
    from conftest import load_package
    return load_package()


def terrace_scene(env=(0.5, 0.6, 0.7)):
    """Two terraces (y 0..3 and, for x >= 32, y 0..7) under an open sky,
    lit by the environment only, with a darker material on the upper one."""
    grid = np.full((64, 64, 64), AIR, np.uint16)
    grid[:, 0:4, :] = 1
    grid[:, 4:8, 32:] = 2
    diffuse = np.full((256, 3), 0.5, np.float32)
    diffuse[2] = (0.2, 0.3, 0.4)
    emission = np.zeros((256, 3), np.float32)
    scene = orc.Scene()
    scene.load_pages(*pages_of(grid))
    t = orc.Tracer(scene)
    t.set_materials(diffuse, emission)
    t.set_lights(np.zeros((0, 12), np.float32))
    t.set_transport(bounces=2, samples=1, environment=env)
    return grid, t


def view(vgi, t, frame=0, samples=4, eye=(20.0, 40.0, 10.0), target=(36.0, 4.0, 36.0), w=48, h=32):
    cfw, wfc = camera(eye, target, w, h)
    depth, normals = prepass(t, eye, cfw, wfc, w, h)
    cam = vgi.view_camera(wfc, eye, w, h, samples=samples, clip_from_world=cfw, frame=frame)
    owners, words, count = t.view_frame(cam, depth, normals)
    faces = words.view(vgi.VIEW_FACE)
    records, counters, _, _ = t.view_read(w * h)
    return owners, faces, count, records, counters, (w, h)


def test_without_pooling_a_face_shows_its_own_mean_then_its_blend():
    vgi = package()
    _, t = terrace_scene()
    t.view_display(0.031, 0.0, 0.25)
    owners, f0, count, _, _, _ = view(vgi, t, 0)
    own = np.flatnonzero(owners == np.arange(len(owners)))
    assert count == len(own) > 100
    sampled = own[f0["mean"][own, 3] > 0]
    assert np.array_equal(f0["shown"][sampled, :3], f0["mean"][sampled, :3]), "first frame: shown = own mean"
    assert np.all(f0["shown"][sampled, 3] == 1.0)
    _, f1, _, _, _, _ = view(vgi, t, 1)
    b = np.float32(0.25)
    want = f0["shown"][sampled, :3] * (np.float32(1.0) - b) + f1["mean"][sampled, :3] * b
    assert np.array_equal(f1["shown"][sampled, :3], want.astype(np.float32)), "second frame: the temporal blend"


def test_pooled_faces_are_convex_combinations_of_coplanar_neighbours():
    vgi = package()
    _, t = terrace_scene()
    t.view_display(0.031, 1.0, 1.0)
    owners, f, _, _, _, (w, h) = view(vgi, t, 0, samples=1)
    own = np.flatnonzero(owners == np.arange(len(owners)))
    checked = 0
    for i in own:
        if f["mean"][i, 3] <= 0:
            continue
        ti, tj = (i % w) // 2, (i // w) // 2
        axis = int(f["face"][i]) // 2
        near = [k for k in own
                if f["face"][k] == f["face"][i] and f["surface"][k, axis] == f["surface"][i, axis]
                and np.sum((f["surface"][k] - f["surface"][i]) ** 2) <= 256.0
                and abs((k % w) // 2 - ti) <= 4 and abs((k // w) // 2 - tj) <= 4 and f["mean"][k, 3] > 0]
        means = f["mean"][near, :3]
        lo, hi = means.min(axis=0), means.max(axis=0)
        s = f["shown"][i, :3]
        assert np.all(s >= lo * (1 - 1e-6) - 1e-7) and np.all(s <= hi * (1 + 1e-6) + 1e-7), (i, s, lo, hi)
        checked += 1
    assert checked > 100


def test_well_sampled_faces_are_shown_alone():
    # Relative standard error under the threshold: no neighbour is gathered,
    # and the face shows its own mean (to the rounding of m w / w).
    vgi = package()
    _, t = terrace_scene()
    t.view_display(0.5, 1.0, 1.0)  # a loose threshold every face meets at 64 samples
    owners, f, _, _, _, _ = view(vgi, t, 0, samples=64)
    own = np.flatnonzero(owners == np.arange(len(owners)))
    own = own[f["mean"][own, 3] >= 64]
    assert len(own) > 100
    assert np.allclose(f["shown"][own, :3], f["mean"][own, :3], rtol=2e-7, atol=0)


def test_published_records():
    vgi = package()
    _, t = terrace_scene()
    owners, f, count, records, counters, _ = view(vgi, t, 0, samples=2)
    px = records.view(vgi.VIEW_PIXEL).ravel()
    own = owners == np.arange(len(owners))
    assert np.all(px["face"][~own] == 0xFFFFFFFF)
    assert np.array_equal(px["face"][own], f["face"][own])
    assert np.array_equal(px["surface"][own], f["surface"][own])
    assert np.array_equal(px["irradiance"][own], f["shown"][own, :3]), "no custom light: the shown rgb"
    lit = own & ((f["shown"][:, 3] > 0) | np.any(f["shown"][:, :3] > 0, axis=1))
    assert np.array_equal(px["evidence"][lit], np.maximum(f["mean"][lit, 3], np.float32(0.001)))
    assert np.all(px["evidence"][own & ~lit] == 0.0)
    assert counters[8] == int(f["mean"][own, 3].astype(np.uint32).sum()), "history"
    assert counters[9] == int(f["mean"][own, 3].max()), "max history"
    assert counters[10] == int((f["mean"][own, 3] == 0).sum()), "uncovered"
    assert counters[11] == int((px["evidence"][own] > 0).sum()), "displayed"


def test_presentation_goes_to_the_world_cache_and_back():
    vgi = package()
    grid, t = terrace_scene()
    faces = vgi.exposed_faces(grid)
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    slots = t.cache_register(faces)
    for n in range(17):
        t.cache_batch(faces, slots, clock_ms=16 * n)
    t.view_display(0.031, 1.0, 0.5)
    owners, f, _, _, _, _ = view(vgi, t, 0, samples=1)
    own = np.flatnonzero((owners == np.arange(len(owners))) & (f["identity"][:, 3] != 0))
    assert len(own) > 100, "admitted faces"
    state = t.cache_state(0, len(faces)).reshape(len(faces), 36)
    index = {(int(a), int(b), int(c), int(d)): k for k, (a, b, c, d) in
             enumerate(zip(faces["voxel"][:, 0], faces["voxel"][:, 1], faces["voxel"][:, 2], faces["face"]))}
    for i in own:
        k = slots[index[(int(f["voxel"][i, 0]), int(f["voxel"][i, 1]), int(f["voxel"][i, 2]), int(f["face"][i]))]]
        st = state[k]
        assert np.array_equal(st[24:28].view(np.float32), f["shown"][i]), "display record"
        assert st[7] == st[24] and np.array_equal(st[13:16], st[25:28]), "copies in the result records"
    # A face leaving and re-entering the view loads its presentation back
    # (then blends toward its mean: spatial strength 0, blend 1/2).
    t.view_reset()
    t.view_display(0.031, 0.0, 0.5)
    _, f2, _, _, _, _ = view(vgi, t, 1, samples=0)
    half = np.float32(0.5)
    want = f["shown"][own, :3] * (np.float32(1.0) - half) + f2["mean"][own, :3] * half
    assert np.array_equal(f2["shown"][own, :3], want.astype(np.float32))
