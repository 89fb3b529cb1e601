"""Change footprints on the GPU against the oracle: the Light event each
edited box gets (where sunlight through it lands, and the light reflected
there), traced by footprint_trace_kernel in the edited scene and filled in by
footprint_finish_kernel before the events apply (gpu/resident.rs:1215-1310,
voxel_trace.wgsl:94-202), on the reference's own footprint scenes
(gpu/resident.rs:2203-2404), an oblique sun and the 64-footprint cap.

Bar: the event words (the edit's event and every footprint event) and the
next batch's results and resident state bit-exact.
"""
import math

import numpy as np
import pytest

from conftest import gpu_tracer
from test_oracle_edits import AIR
from test_oracle_footprints import EMPTY_LO, events, hole, skylight_grid

pytestmark = pytest.mark.gpu


def setup_of(vgi, orc, grid, direction):
    diffuse = np.full((256, 3), 0.5, np.float32)
    diffuse[2] = (0.3, 0.6, 0.9)
    emission = np.zeros((256, 3), np.float32)
    lights = np.zeros((1, 12), np.float32)
    lights[0, 0:3] = direction
    lights[0, 3] = -1.0
    lights[0, 4:7] = 1.0
    coords, pages = vgi.pages_of(grid)
    scene = orc.Scene()
    scene.load_pages(coords, pages)
    tracer = orc.Tracer(scene)
    tracer.set_materials(diffuse, emission)
    tracer.set_lights(lights)
    return dict(grid=grid, diffuse=diffuse, emission=emission, lights=lights, env=np.zeros(3, np.float32), coords=coords,
                pages=pages, scene=scene, tracer=tracer)


def run_edit(vgi, orc, grid, direction, xyz, values, name):
    """One traced frame, the edit on both sides, then the next frame; returns
    the event words (equal on both sides)."""
    print("footprint case %s: %d voxel edits" % (name, len(values)), flush=True)
    setup = setup_of(vgi, orc, grid, direction)
    t = setup["tracer"]
    t.set_transport(bounces=2, samples=2, environment=(0.2, 0.25, 0.3))
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=2, samples=2, environment=(0.2, 0.25, 0.3), measured_retention=1.0)
    g.bind_exposed()
    xyz = np.asarray(xyz, np.int32)
    values = np.asarray(values, np.uint16)
    words = None
    for frame in range(2):
        if frame == 1:
            g.edit_voxels(xyz, values)
            words = t.edit_event(xyz, values)
            assert len(words)
            t.apply_events(words, orc.event_count(words))
        tasks = g.bound_tasks()
        cpu_slots = t.cache_register(tasks)
        g.set_clock(16 * frame)
        ticket = g.submit_compact(tasks)
        gpu = g.complete_readback(ticket, len(tasks))
        assert np.array_equal(g.slots(len(tasks)), cpu_slots)
        if frame == 1:
            ev = g.events()
            assert np.array_equal(ev, words), "%s: event words differ at %s" % (
                name, np.nonzero(ev != words)[0][:8].tolist() if len(ev) == len(words) else (len(ev), len(words)))
        cpu, _ = t.cache_batch(tasks, cpu_slots, clock_ms=16 * frame)
        assert np.array_equal(gpu.view(np.uint32), cpu.view(np.uint32)), "%s frame %d results" % (name, frame)
    n = int(cpu_slots.max()) + 1
    assert np.array_equal(g.cache_state(0, n), t.cache_state(0, n)), name
    assert np.array_equal(g.cache_meta(0, n), t.cache_meta(0, n)), name
    g.close()
    return words


def test_skylight_footprints_bit_exact(vgi, orc):
    up = (0.0, 1.0, 0.0)
    ev = events(run_edit(vgi, orc, skylight_grid(False), up, *hole(), "skylight"))
    assert ev[1][2][0][:2] == ((8, 4, 8), (10, 4, 10)) and ev[1][2][0][3] == 4
    ev = events(run_edit(vgi, orc, skylight_grid(True), up, *hole(), "covered skylight"))
    assert ev[1][0] == (EMPTY_LO,) * 3


def test_placed_block_and_oblique_sun_bit_exact(vgi, orc):
    grid = np.full((32, 32, 32), AIR, np.uint16)
    grid[:, 4, :] = 2
    xyz = [[x, 5, z] for z in range(8, 11) for x in range(8, 11)]
    ev = events(run_edit(vgi, orc, grid, (0.0, 1.0, 0.0), xyz, [2] * 9, "placed block"))
    assert ev[1][0] == (EMPTY_LO,) * 3
    d = tuple((np.array([1.0, 1.0, 0.0], np.float32) / np.float32(math.sqrt(2.0))).tolist())
    ev = events(run_edit(vgi, orc, grid, d, [[10, 12, 10], [20, 9, 20], [21, 9, 20]], [2, 2, 2], "oblique sun"))
    assert len(ev) == 1 + len(ev[0][2]) and all(e[2][0][3] == 4 for e in ev[1:])


def test_footprint_cap_and_order_bit_exact(vgi, orc):
    # Three filled 16^3 regions (the third over the flux scan budget) and 62
    # single voxels: 65 edited boxes, 64 footprints, the unknown one left out.
    grid = np.full((128, 32, 128), AIR, np.uint16)
    grid[:, 4, :] = 2
    xyz, values = [], []
    for rz in range(3):
        for z in range(16 * rz, 16 * rz + 16):
            for y in range(16, 32):
                for x in range(16):
                    xyz.append([x, y, z]), values.append(2)
    spots = [(rx * 16 + 8, 24, rz * 16 + 8) for rz in range(8) for rx in range(1, 8)]
    spots += [(8, 24, rz * 16 + 8) for rz in range(3, 8)]
    spots += [(rx * 16 + 8, 10, rz * 16 + 8) for rz in range(8) for rx in range(8)]
    for s in spots[:62]:
        xyz.append(list(s)), values.append(2)
    ev = events(run_edit(vgi, orc, grid, (0.0, 1.0, 0.0), xyz, values, "64 footprints"))
    assert len(ev[0][2]) == 65 and len(ev) == 65
