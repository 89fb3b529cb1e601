"""The edit stream on the GPU against the oracle, frame after frame: seeded
voxel sets, clears and sphere removals on a noise terrain, the incremental
NAADF update, the edit's retention event (changed boxes, GPU-measured flux,
retention over every slot), the bound faces following the exposed surface,
then a resident batch with world-cache termination.

Bar: event words, NAADF regions, bound faces, slots, hit records and the
whole resident state bit-exact every frame.
"""
import numpy as np
import pytest

from conftest import gpu_tracer, make_scene

pytestmark = pytest.mark.gpu


def face_keys(tasks):
    return set(zip(tasks["voxel"][:, 0].tolist(), tasks["voxel"][:, 1].tolist(), tasks["voxel"][:, 2].tolist(),
                   tasks["face"].tolist()))


@pytest.mark.parametrize("edits_per_frame", [40, 400])
def test_edit_stream_bit_exact(vgi, orc, edits_per_frame):
    seed, edge, samples = 0x5EED0004, 128, 2
    setup = make_scene(vgi, orc, 4, seed, edge=edge)
    grid = setup["grid"].copy()
    t = setup["tracer"]
    t.set_transport(bounces=2, samples=samples, environment=tuple(setup["env"]))
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=2, samples=samples, environment=tuple(setup["env"]), measured_retention=1.0)
    g.bind_exposed()
    measured = 0
    for frame in range(5):
        words = np.zeros(0, np.uint32)
        if frame > 0:
            xyz, values = vgi.scene_edits(seed, frame, edits_per_frame, edge)
            print("edit frame %d: %d voxel edits" % (frame, len(values)), flush=True)
            g.edit_voxels(xyz, values)
            words = t.edit_event(xyz, values)
            if len(words):
                t.apply_events(words, orc.event_count(words))
            grid[xyz[:, 2], xyz[:, 1], xyz[:, 0]] = values  # later edits of a voxel win, as in the library
        tasks = g.bound_tasks()
        assert face_keys(tasks) == face_keys(vgi.exposed_faces(grid)), "bound faces differ from the exposed surface"
        assert len(face_keys(tasks)) == len(tasks)
        cpu_slots = t.cache_register(tasks)
        g.set_clock(16 * frame)
        ticket = g.submit_compact(tasks, vgi.RECORD_HITS)
        gpu = g.complete_readback(ticket, len(tasks))
        gh = g.read_hits(ticket, len(tasks), samples)
        assert np.array_equal(g.slots(len(tasks)), cpu_slots)
        if frame > 0:
            ev = g.events()
            assert np.array_equal(ev, words), "event words differ (%d vs %d words)" % (len(ev), len(words))
            kinds = words.reshape(-1, 4)[3 * orc.event_count(words)::2, 3] & 7 if len(words) else np.zeros(0)
            measured += int((kinds == 3).sum())
        cpu, ch = t.cache_batch(tasks, cpu_slots, hits=True, clock_ms=16 * frame)
        bad = np.nonzero(np.any(gh != ch, axis=1))[0]
        assert len(bad) == 0, "frame %d hit mismatch at %s" % (frame, bad[:3].tolist())
        assert np.array_equal(gpu.view(np.uint32), cpu.view(np.uint32)), "frame %d results not bit-identical" % frame
    assert measured > 0, "no edit was measured against resident statistics"
    n = int(cpu_slots.max()) + 1
    assert np.array_equal(g.cache_state(0, n), t.cache_state(0, n))
    b = setup["scene"].bounds()
    for rz in range(b[2], b[5] + 1):
        for ry in range(b[1], b[4] + 1):
            for rx in range(b[0], b[3] + 1):
                assert np.array_equal(g.region_canonical(rx, ry, rz), setup["scene"].region_canonical(rx, ry, rz)), (rx, ry, rz)
    g.close()
