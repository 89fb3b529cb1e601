"""Page loads and unloads on a resident tracer, frame after frame, against
the oracle: the upload's retention event (Load changes with their top row and
the sky's flux, Unload changes, and edit changes of a page sent again with new
content, measured on the GPU; SceneChanges::with_materials geometry.rs:110-229,
finalize_change_flux publication.rs:423-454), retention applied to every slot
(the load branch of apply_events, voxel_retention.wgsl:29-96), the bound faces
following the exposed surface, then a frame through vgi_run_frame.

Bar: event words, bound faces, hit records, 96-byte records and the resident
state bit-exact every frame.
"""
import numpy as np
import pytest

from conftest import gpu_tracer, make_scene

pytestmark = pytest.mark.gpu


def face_keys(tasks):
    return set(zip(tasks["voxel"][:, 0].tolist(), tasks["voxel"][:, 1].tolist(), tasks["voxel"][:, 2].tolist(), tasks["face"].tolist()))


def test_page_loads_and_unloads_bit_exact(vgi, orc):
    seed, edge = 0x5EED0004, 128
    setup = make_scene(vgi, orc, 4, seed, edge=edge)
    env = tuple(setup["env"])
    t = setup["tracer"]
    t.set_transport(bounces=2, samples=2, environment=env)
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=2, samples=2, environment=env, measured_retention=1.0)
    g.bind_exposed()
    grid = setup["grid"]
    # A new page beside the terrain (a copy of its neighbour raised into a
    # wall), the same page again with a carved notch, then a terrain page removed.
    wall = grid[32:64, 0:32, 96:128].copy()
    wall[:, 20:28, 8:24] = 3
    notched = wall.copy()
    notched[10:20, 0:28, 0:12] = vgi.AIR
    steps = [
        ("load", np.array([[4, 0, 1]], np.int32), wall.reshape(1, -1), None),
        ("reload", np.array([[4, 0, 1]], np.int32), notched.reshape(1, -1), None),
        ("unload", np.zeros((0, 3), np.int32), np.zeros((0, 32768), np.uint16), np.array([[0, 0, 0]], np.int32)),
        ("none", np.array([[4, 0, 1]], np.int32), notched.reshape(1, -1), None),
    ]
    kinds_seen = set()
    for frame in range(len(steps) + 1):
        words = np.zeros(0, np.uint32)
        if frame > 0:
            name, coords, pages, removed = steps[frame - 1]
            print("pages frame %d: %s" % (frame, name), flush=True)
            g.upload_pages(coords, pages, removed)
            words = t.load_event(coords, pages, removed)
            if len(words):
                t.apply_events(words, orc.event_count(words))
        tasks = g.bound_tasks()
        assert len(face_keys(tasks)) == len(tasks)
        slots = t.cache_register(tasks)
        g.set_clock(16 * frame)
        g.run_frame(vgi.RECORD_HITS)
        g.sync()
        got, _ = g.read_bound()
        hits = g.read_bound_hits()
        want, want_hits = t.cache_batch(tasks, slots, hits=True, clock_ms=16 * frame)
        if frame > 0:
            ev = g.events()
            assert np.array_equal(ev, words), "frame %d event words differ (%d vs %d words)" % (frame, len(ev), len(words))
            if len(words):
                w = words.reshape(-1, 4)
                n_events = orc.event_count(words)
                kinds_seen |= set((w[3 * n_events + 1::2, 3] & 7).tolist())
        assert np.array_equal(hits, want_hits), "frame %d hit records differ" % frame
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "frame %d records differ" % frame
    assert {1, 2}.issubset(kinds_seen), "loads and unloads were produced: %s" % sorted(kinds_seen)
    assert kinds_seen & {0, 3}, "the reload produced edit changes: %s" % sorted(kinds_seen)
    n = int(slots.max()) + 1
    assert np.array_equal(g.cache_state(0, n), t.cache_state(0, n))
    assert np.array_equal(g.cache_meta(0, n), t.cache_meta(0, n))
    g.close()
