"""vgi_trace_faces_noalloc: the non-resident trace with its tasks and records
in the library's fixed device scratch region instead of allocated buffers.

Bar: on two calls in a row (the second smaller, with faces marked
VGI_NO_FACE), records bit-identical to vgi_submit_faces + vgi_read_faces of
the same tasks; a marked face reads zeros although the first call left a
record in its place; sizes above the maximum and resident tracers are
refused with an error.
"""
import numpy as np
import pytest

from conftest import gpu_tracer, make_scene

pytestmark = pytest.mark.gpu


def test_noalloc_trace_matches_submit_on_two_calls(vgi, orc):
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=128)
    faces = vgi.exposed_faces(setup["grid"])
    env = tuple(setup["env"])
    g = gpu_tracer(vgi, setup)
    g.set_transport(bounces=2, samples=4, environment=env)
    g.complete_readback(g.submit_compact(faces[:64]), 64)  # set-up: builds the emitter table (allocates)
    first = faces[:vgi.NOALLOC_MAX_FACES]
    second = faces[1000:3000].copy()
    second["face"][::7] = vgi.NO_FACE
    print("noalloc: %d then %d faces" % (len(first), len(second)), flush=True)
    got_first = g.trace_faces_noalloc(first)
    got_second = g.trace_faces_noalloc(second)
    want_first = g.complete_readback(g.submit_compact(first), len(first))
    want_second = g.complete_readback(g.submit_compact(second), len(second))
    assert np.array_equal(got_first.view(np.uint32), want_first.view(np.uint32)), "first call differs"
    assert np.array_equal(got_second.view(np.uint32), want_second.view(np.uint32)), "second call differs"
    assert not got_second[::7].any(), "a face marked VGI_NO_FACE reads zeros"
    assert got_first[1000:3000][::7].any(), "the first call wrote records where the second call's marked faces sit"
    assert (got_second[1::7, 5] > 0).all(), "traced faces have known samples"
    with pytest.raises(vgi.VoxelGiError, match="at most"):
        g.trace_faces_noalloc(np.zeros(vgi.NOALLOC_MAX_FACES + 1, vgi.FACE_TASK))
    g.close()
    r = gpu_tracer(vgi, setup, resident=True)
    with pytest.raises(vgi.VoxelGiError, match="resident"):
        r.trace_faces_noalloc(faces[:16])
    r.close()
