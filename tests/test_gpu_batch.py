"""Batched frames: several scenes of uneven face counts in one persistent
launch per frame, captured once into a HIP graph and replayed.

Bar: every scene's 96-byte records bit-identical to its own eager launch
with the same sequence numbers, on every replay (with eager work and
synchronisation between replays), for every workgroup / lanes / tile
configuration the autotuner may pick; one scene also against the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EDGES = [40, 64, 96, 32, 128]  # 40 is padded with air to 64
SAMPLES = 2


def build_scenes(vgi, orc):
    scenes = []
    for k, edge in enumerate(EDGES):
        seed = 0x5EED0003 + k
        grid = vgi.padded(vgi.generate(3, seed, edge))
        diffuse, emission, lights, env = vgi.optics(3, seed)
        coords, pages = vgi.pages_of(grid)
        g = vgi.GpuTracer(0)
        g.upload_pages(coords, pages)
        g.set_materials(diffuse, emission)
        g.set_lights(lights)
        g.set_transport(bounces=2, samples=SAMPLES, environment=tuple(env))
        faces = vgi.exposed_faces(grid)
        g.bind_faces(faces)
        scenes.append(dict(g=g, faces=faces, grid=grid, coords=coords, pages=pages, optics=(diffuse, emission, lights, env)))
    return scenes


def eager(vgi, scene, sequence):
    """The scene's faces traced by an eager launch with sequence words
    `sequence` (the batch adds its frame base to word 0 of each task's second
    half in the same way)."""
    tasks = scene["faces"].copy()
    tasks["sequence"] = sequence
    g = scene["g"]
    ticket = g.submit_compact(tasks)
    return g.complete_readback(ticket, len(tasks))


@pytest.mark.parametrize("config", [(256, 1, 1024), (64, 2, 128), (128, 4, 64), (256, 8, 8192)])
def test_batch_replays_match_eager_frames(vgi, orc, config):
    scenes = build_scenes(vgi, orc)
    print("batch: %d scenes, faces %s, config %s" % (len(scenes), [len(s["faces"]) for s in scenes], config), flush=True)
    batch = vgi.GpuBatch(0)
    for s in scenes:
        batch.add(s["g"])
    batch.configure(*config)
    batch.set_sequence(0)
    for frame in range(4):
        batch.run_frame()
        batch.sync()
        latency, rays = batch.latency()
        assert rays > 0
        assert np.all(latency > 0) and np.all(latency <= batch.last_ms() * 1.5 + 1.0)
        for k, s in enumerate(scenes):
            got, _ = s["g"].read_bound()
            want = eager(vgi, s, SAMPLES * frame)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "frame %d scene %d differs" % (frame, k)
    batch.close()
    for s in scenes:
        s["g"].close()


def test_batch_matches_oracle_and_autotuner_keeps_results(vgi, orc):
    scenes = build_scenes(vgi, orc)
    batch = vgi.GpuBatch(0)
    for s in scenes:
        batch.add(s["g"])
    tuned = batch.autotune(frames=1)
    print("autotuned:", tuned, flush=True)
    assert tuned["candidates"] >= 12 and tuned["ms"] > 0
    batch.set_sequence(0)
    batch.run_frame()
    batch.sync()
    s = scenes[0]
    got, _ = s["g"].read_bound()
    diffuse, emission, lights, env = s["optics"]
    scene = orc.Scene()
    scene.load_pages(s["coords"], s["pages"])
    t = orc.Tracer(scene)
    t.set_materials(diffuse, emission)
    t.set_lights(lights)
    t.set_transport(bounces=2, samples=SAMPLES, environment=tuple(env))
    cpu, _, _ = t.trace_faces(s["faces"])
    assert np.array_equal(got.view(np.uint32), cpu.view(np.uint32))
    for k, sc in enumerate(scenes):
        got, _ = sc["g"].read_bound()
        assert np.array_equal(got.view(np.uint32), eager(vgi, sc, 0).view(np.uint32)), "scene %d" % k
    batch.close()
    for sc in scenes:
        sc["g"].close()
