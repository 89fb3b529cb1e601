"""Wavefront path tracing against the one-kernel estimator (itself checked
against the oracle): resident frames of bound faces, plain and with
world-cache termination, one to four bounces, sun and emitters (light tree
and CDF), terminal connections.

Bar: every 96-byte record bit-identical, every frame, and the same ray count.
"""
import numpy as np
import pytest

from conftest import gpu_tracer, make_scene

pytestmark = pytest.mark.gpu

CASES = [
    # workload, edge, bounces, samples, plain
    (1, 64, 1, 1, True),    # hollow box, one emitter (CDF), sun through a window
    (2, 64, 3, 4, True),    # noise terrain, 64 emitters (light tree)
    (2, 128, 4, 4, False),  # world-cache termination
    (5, 128, 4, 4, True),   # worm cave, terminal connections
    (4, 128, 2, 2, False),  # edit-stream terrain, 2 bounces
]


@pytest.mark.parametrize("workload,edge,bounces,samples,plain", CASES)
def test_wavefront_matches_megakernel(vgi, orc, workload, edge, bounces, samples, plain):
    print("wavefront: workload %d edge %d bounces %d samples %d plain %s" % (workload, edge, bounces, samples, plain), flush=True)
    setup = make_scene(vgi, orc, workload, 0x5EED0000 + workload, edge)
    faces = vgi.exposed_faces(setup["grid"], interior_only=(workload == 1))
    records, rays = {}, {}
    for mode in ("megakernel", "wavefront"):
        g = gpu_tracer(vgi, setup, resident=True)
        g.set_transport(bounces=bounces, samples=samples, environment=tuple(setup["env"]), plain=plain)
        g.set_wavefront(mode == "wavefront")
        g.bind_faces(faces)
        records[mode], rays[mode] = [], []
        for frame in range(3):
            g.set_clock(16 * frame)
            g.run_frame(vgi.COUNT_RAYS)
            g.sync()
            rec, counters = g.read_bound()
            records[mode].append(rec.copy())
            rays[mode].append(int(counters[0]))
        g.close()
    for frame in range(3):
        a, b = records["megakernel"][frame], records["wavefront"][frame]
        bad = np.nonzero(np.any(a.view(np.uint32) != b.view(np.uint32), axis=1))[0]
        assert len(bad) == 0, "frame %d: %d faces differ, first %s: %s vs %s" % (
            frame, len(bad), bad[:3].tolist(), a[bad[0]].tolist(), b[bad[0]].tolist())
    assert rays["megakernel"] == rays["wavefront"]
