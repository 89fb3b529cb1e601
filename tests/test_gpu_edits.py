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


def test_edits_issued_while_a_frame_runs_match_edits_after_it(vgi, orc):
    """vgi_edit_voxels issued right after vgi_run_frame, while that frame is
    still on the GPU (its host work overlapping the frame), gives the same
    event words every edit and the same bound faces, records, resident state
    and packed regions at the end as edits issued after the frame finished."""
    seed, edge = 0x5EED0004, 128
    setup = make_scene(vgi, orc, 4, seed, edge=edge)
    edits = [vgi.scene_edits(seed, f, 400, edge) for f in range(1, 6)]
    out = {}
    for mode in ("after", "overlapped"):
        print("edits %s the frame" % mode, flush=True)
        g = gpu_tracer(vgi, setup, resident=True)
        g.set_transport(bounces=2, samples=2, environment=tuple(setup["env"]), measured_retention=1.0)
        g.bind_exposed()
        g.set_clock(0)
        g.run_frame()
        g.sync()
        words = []
        if mode == "after":
            for f, (xyz, values) in enumerate(edits):
                g.edit_voxels(xyz, values)
                words.append(g.events().copy())
                g.set_clock(16 * (f + 1))
                g.run_frame()
                g.sync()
        else:
            g.edit_voxels(*edits[0])
            words.append(g.events().copy())
            for f in range(len(edits)):
                g.set_clock(16 * (f + 1))
                g.run_frame()
                if f + 1 < len(edits):
                    g.edit_voxels(*edits[f + 1])
                    words.append(g.events().copy())
                g.sync()
        g.set_clock(16 * (len(edits) + 1))
        g.run_frame()
        g.sync()
        tasks = g.bound_tasks()
        rec, _ = g.read_bound()
        n = len(tasks)
        b = setup["scene"].bounds()
        regions = [g.region_canonical(rx, ry, rz) for rz in range(b[2], b[5] + 1) for ry in range(b[1], b[4] + 1)
                   for rx in range(b[0], b[3] + 1)]
        out[mode] = (words, tasks.copy(), rec.copy(), g.cache_state(0, n), g.cache_meta(0, n), regions)
        g.close()
    a, b = out["after"], out["overlapped"]
    assert len(a[0]) == len(b[0]) == len(edits)
    for f in range(len(edits)):
        assert np.array_equal(a[0][f], b[0][f]), "edit %d: event words differ" % f
    assert np.array_equal(a[1], b[1]), "bound faces differ"
    assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32)), "records differ"
    assert np.array_equal(a[3], b[3]), "resident state differs"
    assert np.array_equal(a[4], b[4]), "slot metadata differs"
    for ra, rb in zip(a[5], b[5]):
        assert np.array_equal(ra, rb), "packed region differs"
