"""The five benchmark workloads at their stated sizes, through the code path
the benchmark times, against the oracle on the same seeded inputs.

Every frame here runs through vgi_run_frame on bound faces (faces in Morton
order of their voxels, one lane per face when no lane would hold more than
one sample, direct terms kept across frames while nothing changed them), the
path bench.py measures; the oracle runs the reference's order (prepare ->
trace_faces_parallel with the reference's lane count -> accumulate).

Bar: hit voxel, face, status, material, cache slot and per-sample ray count
bit-exact where recorded; 96-byte records bit-identical every frame (the 1e-4
relative bound of the per-face irradiance is met with zero difference); ray
counts equal (a frame that keeps its direct terms traces exactly the
oracle's rays less its write_result direct rays); resident state, retention
event words and packed regions bit-identical.
"""
import os

import numpy as np
import pytest

from conftest import gpu_tracer, make_scene

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


def first_difference(a, b):
    bad = np.nonzero(np.any(np.ascontiguousarray(a).view(np.uint32) != np.ascontiguousarray(b).view(np.uint32), axis=1))[0]
    return bad[:3].tolist()


def compare_state(g, t, n, chunk=1 << 20):
    for first in range(0, n, chunk):
        count = min(chunk, n - first)
        assert np.array_equal(g.cache_state(first, count), t.cache_state(first, count)), "resident state differs in [%d, %d)" % (
            first, first + count)


def test_c5_cave512_plain_frames(vgi, orc):
    """512^3 worm cave, every exposed face (6.2 M), plain 4-bounce paths at 4
    rays per face: frame 1 with first-hit records, frame 2 as timed (ray
    tally, direct terms kept)."""
    seed = 0x5EED0005
    setup = make_scene(vgi, orc, 5, seed, edge=512)
    faces = vgi.exposed_faces(setup["grid"])
    print("C5: %d faces" % len(faces), flush=True)
    assert len(faces) > 6_000_000
    env = tuple(setup["env"])
    t = setup["tracer"]
    t.set_transport(bounces=4, samples=4, environment=env)  # lanes: the reference's 4
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0, plain=True)
    slots = t.cache_register(faces)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=4, samples=4, environment=env, plain=True)
    g.bind_faces(faces)
    for frame, flags in ((0, vgi.RECORD_HITS), (1, vgi.COUNT_RAYS)):
        print("C5 frame %d" % frame, flush=True)
        g.set_clock(16 * frame)
        g.run_frame(flags)
        g.sync()
        got, counters = g.read_bound()
        want, want_hits, c = t.cache_batch(faces, slots, hits=frame == 0, threads=THREADS, clock_ms=16 * frame, counters=True)
        if frame == 0:
            hits = g.read_bound_hits()
            assert np.array_equal(hits, want_hits), "hit records differ at face-samples %s" % (
                np.nonzero(np.any(hits != want_hits, axis=1))[0][:3].tolist(),)
            # The samples' rays; write_result's direct rays belong to no sample.
            assert int(hits[:, 6].astype(np.uint64).sum()) == int(c[0] - c[6]), "per-sample ray counts differ"
            assert (hits[:, 5] == 0xFFFFFFFF).all(), "plain paths report no cache slot"
        else:
            assert int(counters[0]) == int(c[0] - c[6]), "rays: gpu %d, oracle %d less %d direct" % (counters[0], c[0], c[6])
        assert same_bits(got, want), "frame %d records differ at faces %s" % (frame, first_difference(got, want))
    compare_state(g, t, len(faces))
    g.close()


def test_c2_terrain256_world_cache_frames(vgi, orc):
    """256^3 noise terrain (~35% solid, 64 emitters, sun), every exposed
    face, 4 rays per face, 2 bounces, world-cache termination on: six frames,
    so faces pass the 16-sample mark and later paths end at cached faces."""
    seed = 0x5EED0002
    setup = make_scene(vgi, orc, 2, seed, edge=256)
    faces = vgi.exposed_faces(setup["grid"])
    print("C2: %d faces" % len(faces), flush=True)
    assert len(faces) > 500_000
    env = tuple(setup["env"])
    t = setup["tracer"]
    t.set_transport(bounces=2, samples=4, environment=env)
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    slots = t.cache_register(faces)
    assert len(t.emitter_words()) > 8 * 32  # the light tree (>= 32 emissive faces)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=2, samples=4, environment=env)
    g.bind_faces(faces)
    cached = 0
    for frame in range(6):
        record = frame in (0, 5)
        g.set_clock(16 * frame)
        g.run_frame(vgi.RECORD_HITS if record else vgi.COUNT_RAYS)
        g.sync()
        got, counters = g.read_bound()
        want, want_hits, c = t.cache_batch(faces, slots, hits=record, threads=THREADS, clock_ms=16 * frame, counters=True)
        print("C2 frame %d: %d rays" % (frame, c[0]), flush=True)
        if record:
            hits = g.read_bound_hits()
            assert np.array_equal(hits, want_hits), "frame %d hit records differ" % frame
            cached += int((hits[:, 5] != 0xFFFFFFFF).sum())
        elif frame > 0:
            assert int(counters[0]) == int(c[0] - c[6])
        assert same_bits(got, want), "frame %d records differ at faces %s" % (frame, first_difference(got, want))
    assert cached > 0, "no path met a cached face"
    compare_state(g, t, len(faces))
    g.close()


def test_c1_box64_sixteen_frames_temporal(vgi, orc):
    """64^3 hollow box, every interior face, 1 ray per face, 1 bounce, 16
    frames of temporal merging with a 10-frame history (alpha 0.1 once the
    history is full)."""
    setup = make_scene(vgi, orc, 1, 0x5EED0001)
    faces = vgi.exposed_faces(setup["grid"], interior_only=True)
    assert 22000 < len(faces) < 24000
    env = tuple(setup["env"])
    t = setup["tracer"]
    t.set_transport(bounces=1, samples=1, environment=env)
    t.cache_enable(history_limit=10, recency=1.0, measured_retention=1.0)
    slots = t.cache_register(faces)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=1, samples=1, environment=env, history_limit=10)
    g.bind_faces(faces)
    for frame in range(16):
        g.set_clock(16 * frame)
        g.run_frame(vgi.RECORD_HITS if frame == 0 else 0)
        g.sync()
        got, _ = g.read_bound()
        want, want_hits = t.cache_batch(faces, slots, hits=frame == 0, threads=THREADS, clock_ms=16 * frame)
        if frame == 0:
            assert np.array_equal(g.read_bound_hits(), want_hits)
        assert same_bits(got, want), "frame %d records differ at faces %s" % (frame, first_difference(got, want))
    state = g.cache_state(0, len(faces))
    assert np.array_equal(state, t.cache_state(0, len(faces)))
    counts = state[:, 5] & 0x7FFFFFFF
    assert counts.max() == 10 and (counts == 10).mean() > 0.99, "the history is capped at 10 samples"
    g.close()


def test_c4_edit_stream512_ten_frames(vgi, orc):
    """512^3 noise terrain (~30% solid, 16 emitters, sun): ten frames of 1000
    seeded edits each (set, clear, sphere removal r 1-5; ~50k voxel edits),
    the GPU NAADF update, the retention event measured on the GPU and applied
    to every slot, then a frame of 2 rays per face, 2 bounces, world cache on,
    over a bound subset (every 8th exposed face: the oracle traces what the
    GPU traces)."""
    seed, edge = 0x5EED0004, 512
    setup = make_scene(vgi, orc, 4, seed, edge=edge)
    faces = vgi.exposed_faces(setup["grid"])[::8]
    print("C4: %d bound faces" % len(faces), flush=True)
    env = tuple(setup["env"])
    t = setup["tracer"]
    t.set_transport(bounces=2, samples=2, environment=env)
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    slots = t.cache_register(faces)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=2, samples=2, environment=env, measured_retention=1.0)
    g.bind_faces(faces)
    measured = 0
    for frame in range(11):
        words = np.zeros(0, np.uint32)
        if frame > 0:
            xyz, values = vgi.scene_edits(seed, frame, 1000, edge)
            g.edit_voxels(xyz, values)
            words = t.edit_event(xyz, values)
            if len(words):
                t.apply_events(words, orc.event_count(words))
            print("C4 frame %d: %d voxel edits, %d event words" % (frame, len(values), len(words)), flush=True)
        record = frame in (1, 10)
        g.set_clock(16 * frame)
        g.run_frame(vgi.RECORD_HITS if record else vgi.COUNT_RAYS)
        g.sync()
        got, counters = g.read_bound()
        want, want_hits, c = t.cache_batch(faces, slots, hits=record, threads=THREADS, clock_ms=16 * frame, counters=True)
        if frame > 0:
            ev = g.events()
            assert np.array_equal(ev, words), "frame %d event words differ (%d vs %d words)" % (frame, len(ev), len(words))
            kinds = words.reshape(-1, 4)[3 * orc.event_count(words)::2, 3] & 7 if len(words) else np.zeros(0)
            measured += int((kinds == 3).sum())
        if record:
            assert np.array_equal(g.read_bound_hits(), want_hits), "frame %d hit records differ" % frame
        else:
            # An edit changes geometry, so the frame after it traces its direct terms again.
            assert int(counters[0]) == int(c[0]), "frame %d rays differ" % frame
        assert same_bits(got, want), "frame %d records differ at faces %s" % (frame, first_difference(got, want))
    assert measured > 0, "no edit was measured against resident statistics"
    compare_state(g, t, len(faces))
    b = setup["scene"].bounds()
    for rz in range(b[2], b[5] + 1):
        for ry in range(b[1], b[4] + 1):
            for rx in range(b[0], b[3] + 1):
                assert np.array_equal(g.region_canonical(rx, ry, rz), setup["scene"].region_canonical(rx, ry, rz)), (rx, ry, rz)
    g.close()


def test_c3_thirty_two_batched_scenes(vgi, orc):
    """32 seeded noise scenes of edges 8*round(exp(U[ln 32, ln 512])/8), every
    exposed face, 2 rays per face, 2 bounces, in one persistent launch per
    frame inside a HIP graph with the autotuner's configuration: every scene's
    records bit-identical to its own eager frame (vgi_run_frame) and to the
    oracle on a seeded sample of 2,048 faces per scene."""
    edges = vgi.batch_edges(32, 0x5EED0003)
    rng = np.random.default_rng(3)
    scenes = []
    for k, edge in enumerate(edges):
        seed = (0x5EED0003 + 0x10000 * k) & 0xFFFFFFFF
        grid = vgi.padded(vgi.generate(3, seed, edge))
        diffuse, emission, lights, env = vgi.optics(3, seed)
        coords, pages = vgi.pages_of(grid)
        faces = vgi.exposed_faces(grid)
        del grid
        sample = np.sort(rng.choice(len(faces), size=min(2048, len(faces)), replace=False))
        scene = orc.Scene()
        scene.load_pages(coords, pages)
        tr = orc.Tracer(scene)
        tr.set_materials(diffuse, emission)
        tr.set_lights(lights)
        tr.set_transport(bounces=2, samples=2, environment=tuple(env))
        expected, _, _ = tr.trace_faces(faces[sample], threads=THREADS)
        del tr, scene
        g = vgi.GpuTracer(0)
        g.upload_pages(coords, pages)
        g.set_materials(diffuse, emission)
        g.set_lights(lights)
        g.set_transport(bounces=2, samples=2, environment=tuple(env))
        g.bind_faces(faces)
        scenes.append(dict(g=g, faces=len(faces), sample=sample, expected=expected))
        print("C3 scene %d: edge %d, %d faces" % (k, edge, len(faces)), flush=True)
    total = sum(s["faces"] for s in scenes)
    assert len(scenes) == 32 and total > 10_000_000
    batch = vgi.GpuBatch(0)
    for s in scenes:
        batch.add(s["g"])
    tuned = batch.autotune(frames=1)
    print("C3: %d faces, autotuned %s" % (total, tuned), flush=True)
    batch.set_sequence(0)
    batch.run_frame()
    batch.sync()
    batched = []
    for k, s in enumerate(scenes):
        got, _ = s["g"].read_bound()
        assert same_bits(got[s["sample"]], s["expected"]), "scene %d differs from the oracle at %s" % (
            k, first_difference(got[s["sample"]], s["expected"]))
        batched.append(got)
    batch.close()
    for k, s in enumerate(scenes):
        s["g"].run_frame()
        s["g"].sync()
        eager, _ = s["g"].read_bound()
        assert same_bits(batched[k], eager), "scene %d: batched and eager records differ at %s" % (k, first_difference(batched[k], eager))
        s["g"].close()
