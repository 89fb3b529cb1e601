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
    reads as the edit issued after the frame finished: after every edit, the
    bound faces and the last frame's records for them (faces moved by the
    edit read their own record; faces bound since read zeros), and at the
    end the event words, records, resident state and packed regions."""
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
        reads, words = [], []
        g.run_frame()
        for f, (xyz, values) in enumerate(edits):
            if mode == "after":
                g.sync()
                before = g.bound_tasks()
                rec_before, _ = g.read_bound()
            g.edit_voxels(xyz, values)  # "overlapped": the frame is still on the GPU
            tasks = g.bound_tasks()
            rec, _ = g.read_bound()
            reads.append((tasks.copy(), rec.copy()))
            if mode == "after":
                # The moved faces' records are their own rows of the frame.
                key = {k: i for i, k in enumerate(zip(before["voxel"][:, 0].tolist(), before["voxel"][:, 1].tolist(),
                                                      before["voxel"][:, 2].tolist(), before["face"].tolist()))}
                rows = [key.get(k, -1) for k in zip(tasks["voxel"][:, 0].tolist(), tasks["voxel"][:, 1].tolist(),
                                                     tasks["voxel"][:, 2].tolist(), tasks["face"].tolist())]
                rows = np.array(rows)
                assert (rows >= 0).any() and (rows < 0).any(), "the edit should move some faces and add others"
                assert np.array_equal(rec[rows >= 0].view(np.uint32), rec_before[rows[rows >= 0]].view(np.uint32))
                assert not rec[rows < 0].any(), "a face bound since the frame reads zeros"
            if f > 0:
                words.append(g.events().copy())  # the previous edit's words, measured by the frame after it
            g.set_clock(16 * (f + 1))
            g.run_frame()  # the next edit overlaps this frame in the "overlapped" order
        g.sync()
        words.append(g.events().copy())
        g.set_clock(16 * (len(edits) + 1))
        g.run_frame()
        g.sync()
        tasks = g.bound_tasks()
        rec, _ = g.read_bound()
        n = len(tasks)
        b = setup["scene"].bounds()
        regions = [g.region_canonical(rx, ry, rz) for rz in range(b[2], b[5] + 1) for ry in range(b[1], b[4] + 1)
                   for rx in range(b[0], b[3] + 1)]
        out[mode] = (words, tasks.copy(), rec.copy(), g.cache_state(0, n), g.cache_meta(0, n), regions, reads)
        g.close()
    a, b = out["after"], out["overlapped"]
    assert len(a[0]) == len(b[0]) == len(edits)
    for f in range(len(edits)):
        assert np.array_equal(a[0][f], b[0][f]), "edit %d: event words differ" % f
        assert np.array_equal(a[6][f][0], b[6][f][0]), "edit %d: bound faces differ" % f
        assert np.array_equal(a[6][f][1].view(np.uint32), b[6][f][1].view(np.uint32)), "edit %d: records differ" % f
    assert np.array_equal(a[1], b[1]), "bound faces differ"
    assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32)), "records differ"
    assert np.array_equal(a[3], b[3]), "resident state differs"
    assert np.array_equal(a[4], b[4]), "slot metadata differs"
    for ra, rb in zip(a[5], b[5]):
        assert np.array_equal(ra, rb), "packed region differs"


def stone_with_cavity(vgi):
    """64^3 of stone with a 4^3 air cavity: 96 exposed faces."""
    grid = np.full((64, 64, 64), 1, np.uint16)
    grid[30:34, 30:34, 30:34] = vgi.AIR
    return grid


def carve(lo, hi, value=256):
    """Voxel edits setting the box [lo, hi) to `value`."""
    r = [np.arange(lo[a], hi[a]) for a in range(3)]
    z, y, x = np.meshgrid(r[2], r[1], r[0], indexing="ij")
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
    return xyz, np.full(len(xyz), value, np.uint16)


def test_back_to_back_edits_wrap_and_grow_the_staging_region(vgi, orc):
    """Edits issued back to back, each while the previous frame is still on
    the GPU (no sync until the end), on a scene whose first bind stages
    little: the edits' uploads outgrow the pinned staging region (it is
    allocated again, larger) and overfill it (it starts over behind copies
    the GPU may not have made yet). The same bound faces, records, resident
    state and packed regions as the same calls with a sync before every
    edit."""
    grid = stone_with_cavity(vgi)
    coords, pages = vgi.pages_of(grid)
    diffuse, emission, lights, env = vgi.optics(4, 0x5EED0004)
    edits = [carve((24, 24, 24), (40, 40, 40)), carve((8, 8, 8), (56, 56, 56)), carve((20, 20, 20), (44, 44, 44), 2),
             carve((4, 28, 4), (60, 36, 60))]
    out = {}
    for mode in ("synced", "back_to_back"):
        g = vgi.GpuTracer(0)
        g.enable_resident()
        g.upload_pages(coords, pages)
        g.set_materials(diffuse, emission)
        g.set_lights(lights)
        g.set_transport(bounces=2, samples=8, environment=tuple(env), measured_retention=1.0)
        g.bind_exposed()
        g.set_clock(0)
        g.run_frame()
        for f, (xyz, values) in enumerate(edits):
            if mode == "synced":
                g.sync()
            g.edit_voxels(xyz, values)
            g.set_clock(16 * (f + 1))
            g.run_frame()
        g.sync()
        stage = int(g.stats()[7])
        grows, wraps = stage >> 32, stage & 0xFFFFFFFF
        print("%s: staging allocated %d times, started over %d times" % (mode, grows, wraps), flush=True)
        assert grows >= 2, "an upload larger than the first region grows it"
        assert wraps >= 1, "an edit's uploads overfill the region and start it over"
        tasks = g.bound_tasks()
        rec, _ = g.read_bound()
        n = len(tasks)
        b = g.bounds()
        regions = [g.region_canonical(rx, ry, rz) for rz in range(b[2], b[5] + 1) for ry in range(b[1], b[4] + 1)
                   for rx in range(b[0], b[3] + 1)]
        out[mode] = (tasks.copy(), rec.copy(), g.cache_state(0, n), regions)
        g.close()
    a, b = out["synced"], out["back_to_back"]
    assert np.array_equal(a[0], b[0]), "bound faces differ"
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)), "records differ"
    assert np.array_equal(a[2], b[2]), "resident state differs"
    for ra, rb in zip(a[3], b[3]):
        assert np.array_equal(ra, rb), "packed region differs"


def test_bound_slots_grow_with_the_tasks_after_a_large_unresident_bind(vgi, orc):
    """A large bind without the resident cache, then the cache and a small
    exposed set (a 4^3 cavity in solid stone), then an edit that exposes
    many more faces than the small set's slot capacity while the task
    capacity of the large bind is still free: the bound slots grow with the
    bound set, and each frame stays bit-exact with the oracle."""
    grid = stone_with_cavity(vgi)
    coords, pages = vgi.pages_of(grid)
    diffuse, emission, lights, env = vgi.optics(4, 0x5EED0004)
    scene = orc.Scene()
    scene.load_pages(coords, pages)
    t = orc.Tracer(scene)
    t.set_materials(diffuse, emission)
    t.set_lights(lights)
    t.set_transport(bounces=2, samples=2, environment=tuple(env))
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    g = vgi.GpuTracer(0)
    g.upload_pages(coords, pages)
    g.set_materials(diffuse, emission)
    g.set_lights(lights)
    g.set_transport(bounces=2, samples=2, environment=tuple(env), measured_retention=1.0)
    many = np.zeros(200000, vgi.FACE_TASK)
    many["voxel"] = (29, 31, 31)
    g.bind_faces(many)
    g.enable_resident()
    g.bind_exposed()
    assert g.bound_count() == 96
    xyz, values = carve((24, 24, 24), (40, 40, 40))
    for frame in range(3):
        if frame == 1:
            g.edit_voxels(xyz, values)
            words = t.edit_event(xyz, values)
            if len(words):
                t.apply_events(words, orc.event_count(words))
            assert g.bound_count() == 6 * 16 * 16
        tasks = g.bound_tasks()
        slots = t.cache_register(tasks)
        g.set_clock(16 * frame)
        g.run_frame()
        g.sync()
        got, _ = g.read_bound()
        want, _ = t.cache_batch(tasks, slots, clock_ms=16 * frame)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "frame %d records differ" % frame
    g.close()
