"""The C ABI's error behaviour and empty inputs on a live tracer: invalid
arguments come back as VoxelGiError with a message (a negative code and
vgi_last_error underneath) and leave the tracer usable; empty batches, empty
bound sets and empty edit lists are ordinary calls that do nothing."""
import numpy as np
import pytest

from conftest import gpu_tracer, make_scene

pytestmark = pytest.mark.gpu


def raises(vgi, fn, *words):
    with pytest.raises(vgi.VoxelGiError) as info:
        fn()
    text = str(info.value)
    for w in words:
        assert w in text, (w, text)


def test_invalid_arguments_are_errors_and_the_tracer_stays_usable(vgi, orc):
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=64)
    g = gpu_tracer(vgi, setup)
    print("case: transport bounds", flush=True)
    raises(vgi, lambda: g.set_transport(bounces=0, samples=1), "bounces")
    raises(vgi, lambda: g.set_transport(bounces=9, samples=1), "bounces")
    raises(vgi, lambda: g.set_transport(bounces=1, samples=0), "samples")
    raises(vgi, lambda: g.set_transport(bounces=1, samples=4, lanes=3), "lanes")
    print("case: resident-only calls", flush=True)
    raises(vgi, lambda: g.select_frame(16), "resident")
    raises(vgi, lambda: g.sun_corners(0, 1), "resident")
    raises(vgi, lambda: g.bind_exposed(), "resident")
    print("case: learned cache before it is enabled", flush=True)
    q = np.zeros((16, 4), np.uint32)
    raises(vgi, lambda: g.mlp_infer(q), "not enabled")
    raises(vgi, lambda: g.mlp_table(), "not enabled")
    g.mlp_enable(seed=1)
    print("case: learned cache arguments", flush=True)
    y = np.zeros((15, 4), np.float32)
    raises(vgi, lambda: g.mlp_train(q[:15], y), "multiple of 16")
    raises(vgi, lambda: g.mlp_train_samples(0), "1..64")
    raises(vgi, lambda: g.mlp_table_lr(-1.0), "learning rate")
    raises(vgi, lambda: g.mlp_table_grad(), "no gradient")
    print("case: far tier settings and camera", flush=True)
    raises(vgi, lambda: g.view_far(vgi.far_settings(rate=9)), "rate")
    raises(vgi, lambda: g.view_far(vgi.far_settings(steps=0)), "steps")
    raises(vgi, lambda: g.view_far(vgi.far_settings(cell=5)), "cell")
    raises(vgi, lambda: g.view_display(0.031, 2.0, 0.2), "spatial")
    g.view_far(vgi.far_settings())
    depth = np.full((4, 4), 0.5, np.float32)
    normals = np.zeros((4, 4, 3), np.float32)
    normals[..., 1] = 1.0
    wfc = np.eye(4)
    raises(vgi, lambda: g.view_frame(vgi.view_camera(wfc, (0.0, 9.0, 0.0), 4, 4, forward=(0.0, -1.0, 0.0)), depth, normals),
           "pixels per unit")
    raises(vgi, lambda: g.view_frame(vgi.view_camera(wfc, (0.0, 9.0, 0.0), 4, 4, forward=(0.0, 0.0, 0.0), pixels_per_unit=2.0),
                                     depth, normals), "forward")
    g.view_far(None)
    # Still usable: a valid frame runs and its records are finite.
    print("case: a valid frame after the errors", flush=True)
    g.set_transport(bounces=2, samples=2)
    tasks = vgi.exposed_faces(setup["grid"])[:4096]
    ticket = g.submit_compact(tasks)
    out = g.complete_readback(ticket, len(tasks))
    assert np.all(np.isfinite(out))
    g.close()


def test_resident_reads_out_of_range_are_errors(vgi, orc):
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=64)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=1, samples=1)
    tasks = vgi.exposed_faces(setup["grid"])[:1000]
    g.bind_faces(tasks)
    print("case: reads beyond the resident slots", flush=True)
    raises(vgi, lambda: g.sun_corner_words(990, 20), "beyond")
    raises(vgi, lambda: g.select_frame(0), "budget")
    with pytest.raises(vgi.VoxelGiError):
        g.cache_state(990, 20)
    g.close()


def test_empty_inputs_do_nothing(vgi, orc):
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=64)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=2, samples=2)
    print("case: empty batch", flush=True)
    empty = np.zeros(0, vgi.FACE_TASK)
    ticket = g.submit_compact(empty)
    assert len(g.complete_readback(ticket, 0)) == 0
    print("case: no bound faces", flush=True)
    g.bind_faces(empty)
    assert g.bound_count() == 0
    g.run_frame()
    g.sync()
    g.sun_corners(0, 1)
    flips = g.sun_flips()
    assert flips[0] == 0 and not np.any(flips)
    print("case: no edits", flush=True)
    g.edit_voxels(np.zeros((0, 3), np.int32), np.zeros(0, np.uint16))
    # Then real work on the same tracer.
    tasks = vgi.exposed_faces(setup["grid"])[:2048]
    g.bind_faces(tasks)
    g.run_frame()
    g.sync()
    rec, _ = g.read_bound()
    assert rec.shape[0] == len(tasks) and np.all(np.isfinite(rec))
    g.close()
