"""Wavefront path tracing against the one-kernel estimator (itself checked
against the oracle): resident frames of bound faces, plain and with
world-cache termination, one to four bounces, sun and emitters (light tree
and CDF), terminal connections; with and without the per-wavefront hit-face
deduplication of the world-cache lookups.

Bar: every 96-byte record bit-identical, every frame, and the same ray count;
with deduplication, fewer cache probes than lookups wherever the cache ends
paths.
"""
import numpy as np
import pytest

from conftest import gpu_tracer, make_scene

pytestmark = pytest.mark.gpu

# Resident frames per mode: once faces hold 16 samples (the fifth frame at 4
# rays per face, the ninth at 2) the world cache ends paths, and the
# deduplicated lookups answer them.
FRAMES = 9

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
    records, rays, probes = {}, {}, {}
    for mode in ("megakernel", "wavefront", "dedup"):
        g = gpu_tracer(vgi, setup, resident=True)
        g.set_transport(bounces=bounces, samples=samples, environment=tuple(setup["env"]), plain=plain)
        g.set_wavefront(mode != "megakernel")
        g.set_dedup(mode == "dedup")
        g.bind_faces(faces)
        records[mode], rays[mode], probes[mode] = [], [], []
        for frame in range(FRAMES):
            g.set_clock(16 * frame)
            g.run_frame(vgi.COUNT_RAYS)
            g.sync()
            rec, counters = g.read_bound()
            records[mode].append(rec.copy())
            rays[mode].append(int(counters[0]))
            probes[mode].append((int(counters[5]), int(counters[6])))
        g.close()
    for mode in ("wavefront", "dedup"):
        for frame in range(FRAMES):
            a, b = records["megakernel"][frame], records[mode][frame]
            bad = np.nonzero(np.any(a.view(np.uint32) != b.view(np.uint32), axis=1))[0]
            assert len(bad) == 0, "%s frame %d: %d faces differ, first %s: %s vs %s" % (
                mode, frame, len(bad), bad[:3].tolist(), a[bad[0]].tolist(), b[bad[0]].tolist())
        assert rays["megakernel"] == rays[mode]
    print("  world-cache lookups and probes per frame with deduplication: %s" % probes["dedup"], flush=True)
    assert probes["wavefront"] == [(0, 0)] * FRAMES, "no deduplication counted when it is off"
    if not plain:
        for lookups, probed in probes["dedup"]:
            assert 0 < probed < lookups
        assert rays["megakernel"][-1] < rays["megakernel"][0], "the world cache ends paths in the last frames"
