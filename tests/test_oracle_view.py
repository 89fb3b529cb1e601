"""Pins of the oracle's view tier (no GPU needed).

Discovery (prepass_voxel_receiver, voxel_view_discovery.wgsl:40-104) against
closed forms: a camera over an open floor, whose receivers are the floor
voxels under the pixels' ray hits; the one-pixel-footprint neighbour rule on a
designed prepass; the inner overlap distance. Owners (index_receivers,
voxel_view_runtime.wgsl:128-142): each face's owner is the smallest pixel
index showing it. Reprojection (reproject_paths :148-217): a face keeps its
running estimate frame to frame and restarts when its material changes.
"""
import numpy as np

import oracle_lib as orc
from test_oracle_edits import floor_scene
from view_helpers import camera

VIEW_FACE_WORDS = 28


def package():
    from conftest import load_package
    return load_package()


def faces_of(words):
    vgi = package()
    return words.view(vgi.VIEW_FACE)


def floor_prepass(eye, clip_from_world, world_from_clip, width, height, edge=64):
    """The floor's top plane (y = 4) hit by each pixel's ray, in float64."""
    depth = np.zeros((height, width), np.float32)
    normals = np.zeros((height, width, 3), np.float32)
    hits = np.full((height, width, 3), np.nan)
    eye = np.asarray(eye, np.float64)
    for y in range(height):
        for x in range(width):
            p = world_from_clip @ np.array([(x + 0.5) / width * 2.0 - 1.0, 1.0 - (y + 0.5) / height * 2.0, 1.0, 1.0])
            d = p[:3] / p[3] - eye
            if d[1] >= 0.0:
                continue
            t = (4.0 - eye[1]) / d[1]
            h = eye + t * d
            if not (0.0 <= h[0] < edge and 0.0 <= h[2] < edge):
                continue
            c = clip_from_world @ np.array([h[0], h[1], h[2], 1.0])
            depth[y, x] = c[2] / c[3]
            normals[y, x] = (0.0, 1.0, 0.0)
            hits[y, x] = h
    return depth, normals, hits


def test_floor_receivers_owners_and_inner_distance():
    vgi = package()
    _, _, t, _, _ = floor_scene(edge=64)
    eye, target = (20.0, 30.0, 10.0), (32.0, 4.0, 32.0)
    w, h = 48, 32
    cfw, wfc = camera(eye, target, w, h)
    depth, normals, hits = floor_prepass(eye, cfw, wfc, w, h)
    cam = vgi.view_camera(wfc, eye, w, h, samples=0)
    owners, words, count = t.view_frame(cam, depth, normals)
    f = faces_of(words)
    valid = ~np.isnan(hits[..., 0]).ravel()
    assert valid.sum() > 0.6 * w * h
    hx, hz = hits[..., 0].ravel(), hits[..., 2].ravel()
    clear = valid & (np.abs(hx - np.round(hx)) > 1e-3) & (np.abs(hz - np.round(hz)) > 1e-3)
    want = np.stack([np.floor(hx), np.full_like(hx, 3.0), np.floor(hz)], axis=1)
    assert clear.sum() > 0.9 * valid.sum()
    assert np.array_equal(f["voxel"][clear], want[clear].astype(np.int32))
    assert np.all(f["face"][valid] == 2)
    near = valid & ~clear
    assert np.all(np.abs(f["voxel"][near] - want[near]).max(axis=1) <= 1)
    assert np.all(f["face"][~valid] == 0xFFFFFFFF)
    centre = f["voxel"][valid].astype(np.float32) + np.float32(0.5) + np.array([0, 0.5, 0], np.float32)
    assert np.array_equal(f["surface"][valid], centre)
    assert np.all(f["identity"][valid, 1] == 1), "the floor's material"
    # Owners: the smallest pixel index showing the face.
    keys = {}
    for i in np.flatnonzero(valid):
        keys.setdefault(tuple(f["voxel"][i]) + (int(f["face"][i]),), i)
    want_owner = np.full(w * h, 0xFFFFFFFF, np.uint32)
    for i in np.flatnonzero(valid):
        want_owner[i] = keys[tuple(f["voxel"][i]) + (int(f["face"][i]),)]
    assert np.array_equal(owners, want_owner)
    assert count == len(keys) and count < valid.sum(), "several pixels share faces"
    # Points nearer the camera than the inner distance own nothing.
    cam = vgi.view_camera(wfc, eye, w, h, samples=0, inner_distance=34.0)
    t.view_reset()
    owners2, words2, _ = t.view_frame(cam, depth, normals)
    dist = np.linalg.norm(np.stack([hx, np.full_like(hx, 4.0), hz], axis=1) - np.asarray(eye), axis=1)
    f2 = faces_of(words2)
    assert np.all(f2["face"][valid & (dist < 33.99)] == 0xFFFFFFFF)
    assert np.array_equal(f2["voxel"][valid & (dist > 34.01)], f["voxel"][valid & (dist > 34.01)])
    assert ((dist < 33.99) & valid).sum() > 50 and ((dist > 34.01) & valid).sum() > 50


def test_one_pixel_footprint_takes_the_exposed_neighbour():
    """A decoded point over a hole in the floor, 0.02 voxels from the
    boundary of an exposed neighbour: within a pixel footprint of 0.1 the
    neighbour owns it; within a footprint of 0.01 no face does."""
    vgi = package()
    grid, scene, t, _, _ = floor_scene(edge=64)
    scene.edit(np.array([[10, 3, 5]], np.int32), np.array([256], np.uint16))
    for step, want in ((0.1, (11, 3, 5)), (0.01, None)):
        # world = (step * x + 10.98 + step / 2, 4, 5.5) for clip x of pixel 0 = -0.5.
        m = np.zeros((4, 4))
        m[0, 0], m[0, 3] = step, 10.98 + 0.5 * step
        m[1, 3], m[2, 3], m[3, 3] = 4.0, 5.5, 1.0
        cam = vgi.view_camera(m, (10.0, 20.0, 5.0), 2, 1, samples=0)
        depth = np.full((1, 2), 0.5, np.float32)
        normals = np.tile(np.array([0.0, 1.0, 0.0], np.float32), (1, 2, 1))
        t.view_reset()
        owners, words, _ = t.view_frame(cam, depth, normals)
        f = faces_of(words)
        if want is None:
            assert f["face"][0] == 0xFFFFFFFF
        else:
            assert tuple(f["voxel"][0]) == want and f["face"][0] == 2
            assert tuple(f["voxel"][1]) == want, "pixel 1 decodes inside the neighbour itself"
            assert list(owners) == [0, 0]


def test_estimates_carry_across_frames_and_restart_on_a_new_material():
    vgi = package()
    _, scene, t, _, _ = floor_scene(edge=64)
    t.set_transport(bounces=1, samples=1, environment=(0.3, 0.5, 0.8))
    eye, target = (20.0, 30.0, 10.0), (32.0, 4.0, 32.0)
    w, h = 24, 16
    cfw, wfc = camera(eye, target, w, h)
    depth, normals, _ = floor_prepass(eye, cfw, wfc, w, h)
    cam = vgi.view_camera(wfc, eye, w, h, samples=2)
    t.view_reset()
    for frame in range(2):
        owners, words, count = t.view_frame(cam, depth, normals)
    f = faces_of(words)
    own = np.flatnonzero(owners == np.arange(w * h))
    assert len(own) == count > 10
    # Open sky above a floor: every path is known.
    assert np.all(f["mean"][own, 3] == 4.0) and np.all(f["identity"][own, 0] == 4)
    assert np.all(f["sequence"][own] == 4)
    first = f[own[0]]
    v = first["voxel"]
    scene.edit(np.array([v], np.int32), np.array([2], np.uint16))
    owners, words, _ = t.view_frame(cam, depth, normals)
    f = faces_of(words)
    i = own[0]
    assert f["identity"][i, 1] == 2 and f["mean"][i, 3] == 2.0 and f["identity"][i, 0] == 2, "a new material starts over"
    others = own[1:]
    assert np.all(f["mean"][others, 3] == 6.0)


def test_workload_prepass_agrees_with_the_oracle_traversal():
    """The benchmark's CPU prepass renderer (a plain 3D DDA over the dense
    grid, workload library) against rays through the oracle's NAADF
    traversal: the same hit faces and depths, except where a ray grazes a
    voxel edge (either side is a hit there)."""
    vgi = package()
    from view_helpers import prepass as oracle_prepass
    grid = vgi.generate(2, 0x5EED0002, 64)
    coords, pages = vgi.pages_of(grid)
    scene = orc.Scene()
    scene.load_pages(coords, pages)
    t = orc.Tracer(scene)
    eye, target = (-6.0, 40.0, -6.0), (32.0, 8.0, 32.0)
    w, h = 40, 30
    wfc, depth, normals, hits = vgi.scene_prepass(grid, eye, target, w, h)
    cfw, wfc2 = camera(eye, target, w, h)
    assert np.allclose(wfc, wfc2, rtol=1e-5, atol=1e-6)
    d2, n2 = oracle_prepass(t, eye, cfw, wfc2, w, h)
    assert hits == int((depth > 0).sum()) and hits > 0.5 * w * h
    same = np.all(normals == n2, axis=2) & ((depth > 0) == (d2 > 0))
    assert same.mean() > 0.97
    both = same & (depth > 0)
    assert np.allclose(depth[both], d2[both], rtol=1e-4)
