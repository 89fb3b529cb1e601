"""Bound frames keep each face's direct terms (results words 16-23: the sun
and point-light term of voxel_path.wgsl:283-307 and its custom channels)
while the geometry, the lights, the reach and the bound set are unchanged,
instead of tracing the face's light rays again; any call that may change one
of them makes the next frame trace them.

Bar: every record of every frame bit-identical to a tracer that traces them
each frame (the same lights set again before each frame), through an edit, a
moved sun, a shorter reach and a rebinding; the same ray count on the frames
after a change and fewer rays on the frames that keep the terms. The traced
terms themselves are pinned against the oracle by the trace tests.
"""
import numpy as np
import pytest

from conftest import gpu_tracer, make_scene

pytestmark = pytest.mark.gpu

# action before each frame; frames after None keep the terms
ACTIONS = [None, None, "edit", None, "sun", None, "reach", None, "rebind", None]


@pytest.mark.parametrize("mode", ["megakernel", "wavefront", "learned"])
def test_kept_direct_terms_match_traced_terms(vgi, orc, mode):
    seed, edge = 0x5EED0004, 96
    setup = make_scene(vgi, orc, 4, seed, edge=edge)
    lights = setup["lights"]
    sun = int(np.nonzero(lights[:, 3] < 0)[0][0])
    moved = lights.copy()
    moved[sun, 0:3] = lights[sun, 0:3] + np.array([0.35, 0.0, -0.25], np.float32)
    bounces, samples, plain = (3, 1, True) if mode == "learned" else (2, 2, False)
    env = tuple(setup["env"])
    runs = {}
    for kind in ("kept", "traced"):
        print("direct terms: mode %s, %s" % (mode, kind), flush=True)
        g = gpu_tracer(vgi, setup, resident=True)
        g.set_transport(bounces=bounces, samples=samples, environment=env, plain=plain)
        g.set_wavefront(mode == "wavefront")
        if mode == "learned":
            g.mlp_enable(seed=3, train_per_frame=0, steps=1, lr=1e-3)
        g.bind_exposed()
        current = lights
        records, rays = [], []
        for frame, action in enumerate(ACTIONS):
            if action == "edit":
                xyz, values = vgi.scene_edits(seed, 1, 200, edge)
                g.edit_voxels(xyz, values)
            elif action == "sun":
                current = moved
                g.set_lights(current)
            elif action == "reach":
                g.set_transport(bounces=bounces, samples=samples, environment=env, plain=plain, max_distance=200.0)
            elif action == "rebind":
                g.bind_faces(g.bound_tasks()[::3])
            if kind == "traced":
                g.set_lights(current)
            g.set_clock(16 * frame)
            g.run_frame(vgi.COUNT_RAYS)
            g.sync()
            rec, counters = g.read_bound()
            records.append(rec.copy())
            rays.append(int(counters[0]))
        runs[kind] = (records, rays)
        g.close()
    kept, traced = runs["kept"], runs["traced"]
    for frame, action in enumerate(ACTIONS):
        a, b = kept[0][frame], traced[0][frame]
        assert a.shape == b.shape, frame
        bad = np.nonzero(np.any(a.view(np.uint32) != b.view(np.uint32), axis=1))[0]
        assert len(bad) == 0, "frame %d (%s): %d faces differ, first %s: %s vs %s" % (
            frame, action, len(bad), bad[:3].tolist(), a[bad[0]].tolist(), b[bad[0]].tolist())
        if frame == 0 or action is not None:
            assert kept[1][frame] == traced[1][frame], (frame, action, kept[1][frame], traced[1][frame])
        else:
            assert kept[1][frame] < traced[1][frame], (frame, kept[1][frame], traced[1][frame])
    # The changes changed the terms: the kept tracer did not pass on stale ones
    # by luck of an unchanged scene.
    direct = [r[:, 16:19] for r in kept[0]]
    assert not np.array_equal(direct[3], direct[5])  # the sun moved
    print("rays kept", kept[1], "traced", traced[1], flush=True)
