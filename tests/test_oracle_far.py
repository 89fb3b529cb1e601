"""Pins of the oracle's far per-pixel tier (no GPU needed).

The reference's own native far-tier tests (gpu/view/pixel_tests.rs:344-561)
restated on the oracle's classify_pixels / trace_pixel_queue
(voxel_view_pixels.wgsl:129-216) over the same rig: faces under the pixel
threshold own no receiver and are lit per cell; a cell's count grows by one
a frame at rate 1 and follows the world point through a one-cell pan; at the
1440p defaults every cell visits all nine of its pixels and spends one path
sequence per trace; custom direct light is shown outside the history, also
with a one-step ray budget (a ray out of steps is a miss); the transport
reach holds. Plus the packing of the material's output (pack2x16float,
round-to-nearest-even against numpy's binary16) and the per-pixel handoff
between the tiers.
"""
import numpy as np

import far_rig as rig
import oracle_lib as orc


def package():
    from conftest import load_package
    return load_package()


def frame(t, vgi, n, shift=0.0):
    depth, normals = rig.prepass()
    cam = rig.camera(vgi, n, shift)
    owners, _, count = t.view_frame(cam, depth, normals)
    return t.view_read(rig.SIZE * rig.SIZE), count


def test_half_packing_rounds_to_nearest_even():
    rng = np.random.default_rng(7)
    values = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(100.0),
        rng.uniform(0, 1e-4, 4000).astype(np.float32),                      # binary16 subnormals
        rng.uniform(6.0e4, 7.0e4, 2000).astype(np.float32),                 # overflow to infinity
        np.array([0.0, -0.0, 1.0, 2049.0, 2051.0, 65504.0, 65520.0, 65519.996, 5.9604645e-08, 2.9802322e-08,
                  2.9802326e-08, np.inf, -np.inf], np.float32),
    ])
    want = values.astype(np.float16).view(np.uint16)
    got = np.array([orc.lib().orc_pack_half2(float(v), 0.0) & 0xFFFF for v in values], np.uint16)
    assert np.array_equal(got, want)
    assert orc.lib().orc_pack_half2(1.0, -2.0) == 0x3C00 | (0xC000 << 16)


def test_small_faces_are_lit_per_pixel_without_receivers():
    # pixel_tests.rs:344-376
    vgi = package()
    off = rig.oracle(vgi, cell=2, rate=1, far=False)
    (_, _, _, _), receivers = frame(off, vgi, 0)
    assert receivers > 0, "without the far tier the floor gets view receivers"
    t = rig.oracle(vgi, cell=2, rate=1)
    for n in range(6):
        (_, counters, output, cells), receivers = frame(t, vgi, n)
        assert receivers == 0, "per-pixel faces own no receiver"
        assert output[8:12].tolist() == [8, 8, 2, 1], "header grid"
        assert output[15] == n, "header stamp"
        c = rig.counts(output, 64)
        assert np.all(np.abs(c - (n + 1)) < 0.01), "frame %d: %s" % (n, c)
        assert rig.f16(output[16] & 0xFFFF) > 0.0, "the floor sees the sky"
        assert counters[0] == 64 and counters[1] == 64, "every cell owned and traced (rate 1)"
        assert counters[3] == (64 if n > 0 else 0), "reprojected from the previous frame"
    # A one-cell pan (four voxels): history follows the world point. (The
    # last column of cells now samples x = 33..35, beyond the loaded page:
    # outside the interest, those cells are not written.)
    (_, counters, output, _), _ = frame(t, vgi, 6, shift=4.0)
    words = output[16:].reshape(64, 4)
    stamped = np.flatnonzero(words[:, 2] >> 3 == 6)
    assert len(stamped) == counters[0] >= 56
    kept = int((rig.counts(output, 64, stamped) > 1.0).sum())
    assert kept >= 48, "reprojected history: %d of %d" % (kept, len(stamped))


def test_cells_at_the_default_settings_sample_every_pixel_and_index():
    # pixel_tests.rs:378-437: the 1440p defaults, 3-pixel cells, one trace in
    # four frames after four warm-up traces.
    vgi = package()
    cell = 3  # FarTierSettings::cell_for(1440)
    t = rig.oracle(vgi, cell=cell, rate=4)
    cells_x = rig.cells_of(cell)
    inside = rig.interior(cells_x)
    visited = {i: set() for i in inside}
    sequences = {i: [] for i in inside}
    for n in range(36):
        (_, _, output, history), _ = frame(t, vgi, n)
        c = rig.counts(output, cells_x * cells_x, inside)
        assert np.all(c > 0.0), "frame %d: every cell holds light" % n
        for i in inside:
            x = history[i, 4:5].view(np.float32)[0]
            z = history[i, 6:7].view(np.float32)[0]
            visited[i].add((round((x - 1.0) / 2.0), round((z - 1.0) / 2.0)))
            sequences[i].append(int(history[i, 8]))
    for i in inside:
        assert len(visited[i]) == 9, "cell %d samples all nine pixels: %s" % (i, visited[i])
        s = sequences[i]
        assert all(b == a or b == a + 1 for a, b in zip(s, s[1:])), "cell %d: %s" % (i, s)
        assert 12 <= s[-1] <= 13, "cell %d: %s" % (i, s)


def test_custom_direct_light_is_shown_outside_the_history():
    # pixel_tests.rs:439-464
    vgi = package()
    t = rig.oracle(vgi, cell=3, rate=4, custom_light=(16.0, 10.0, 16.0))
    cells_x = rig.cells_of(3)
    for n in range(8):
        (_, _, output, history), _ = frame(t, vgi, n)
        for i in rig.interior(cells_x):
            shown = rig.f16(output[16 + 4 * i] & 0xFFFF)
            mean = history[i, 0:1].view(np.float32)[0]
            assert shown > mean + 0.05, "frame %d cell %d: shown %s over mean %s" % (n, i, shown, mean)


def test_custom_light_with_a_one_step_budget_stays_lit():
    # pixel_tests.rs:466-511: a custom light above the emissive slab, which
    # shadows the floor's middle; with one traversal step every visibility
    # ray runs out of steps, a miss under the far tier's rule, so the
    # shadowed cells are lit, whether they trace this frame or only reproject.
    vgi = package()
    t = rig.oracle(vgi, cell=3, rate=4, ceiling=5, custom_light=(16.0, 31.0, 16.0), environment=(0.0, 0.0, 0.0), steps=1)
    cells_x = rig.cells_of(3)
    terms = {i: [] for i in rig.interior(cells_x)}
    for n in range(12):
        (_, _, output, history), _ = frame(t, vgi, n)
        if n >= 4:
            for i in terms:
                terms[i].append(rig.f16(output[16 + 4 * i] & 0xFFFF) - history[i, 0:1].view(np.float32)[0])
    for i, v in terms.items():
        assert all(x > 0.03 for x in v), "cell %d: custom light %s" % (i, v)
    # With the full budget the slab shadows the middle cells instead.
    full = rig.oracle(vgi, cell=3, rate=4, ceiling=5, custom_light=(16.0, 31.0, 16.0), environment=(0.0, 0.0, 0.0))
    (_, _, output, history), _ = frame(full, vgi, 0)
    middle = 2 * cells_x + 2
    assert abs(rig.f16(output[16 + 4 * middle] & 0xFFFF) - history[middle, 0:1].view(np.float32)[0]) < 0.01


def test_far_tier_honours_the_transport_reach():
    # pixel_tests.rs:513-544: an emissive slab 24 voxels above the floor, no
    # sky: with the reach at 16 the far tier sees none of it; at full reach it does.
    vgi = package()

    def light_at(reach):
        t = rig.oracle(vgi, cell=3, rate=1, ceiling=5, environment=(0.0, 0.0, 0.0), reach=reach)
        for n in range(4):
            (_, _, output, _), _ = frame(t, vgi, n)
        middle = 16 + (2 * rig.cells_of(3) + 2) * 4
        return rig.f16(output[middle] & 0xFFFF)

    assert light_at(None) > 0.01, "the slab lights the floor at full reach"
    assert light_at(16.0) == 0.0, "beyond the reach the slab contributes nothing"


def test_cells_hand_off_to_and_from_view_receivers():
    # voxel_view_pixels.wgsl:163-178 and voxel_view_cache.wgsl:201-218: a
    # face crossing into the far band starts from its receiver's estimate
    # (count and sequence); a face entering receiver range starts from the far
    # cell at its centre with one unit of borrowed weight.
    vgi = package()
    near = rig.oracle(vgi, cell=2, rate=1, far=False)
    depth, normals = rig.prepass()
    for n in range(3):
        near.view_frame(rig.camera(vgi, n, samples=2), depth, normals)
    near.view_far(rig.settings(vgi, 2, 1))
    (_, counters, output, history), receivers = frame(near, vgi, 3)
    assert receivers == 0
    assert counters[4] == 64, "every cell seeded from its face's receiver"
    c = rig.counts(output, 64)
    assert np.all(c == 7.0), "6 receiver samples + 1 path: %s" % c
    assert np.all(history[:, 8] == 7), "the receiver's next sequence + 1"
    # Back to receivers (the tier still on, its threshold now under a
    # voxel's 0.04 pixels): a face new to the view takes the cell at its centre.
    near.view_far(rig.settings(vgi, 2, 1, voxel_pixels=0.01))
    _, faces, count = near.view_frame(rig.camera(vgi, 4, samples=0), depth, normals)
    _, counters, _, _ = near.view_read(rig.SIZE * rig.SIZE)
    assert count > 0
    f = faces.view(vgi.VIEW_FACE)
    seeded = f[(f["face"] == 2) & (f["mean"][:, 3] == 1.0)]
    assert counters[5] > 0 and len(seeded) == counters[5]
    assert np.all(seeded["shown"][:, 3] == 1.0) and np.all(seeded["moment"][:, 3] == 0.0)


def test_far_cells_keep_to_the_interest():
    # voxel_view_runtime.wgsl:26-30, voxel_view_pixels.wgsl:160: a pixel whose
    # surface lies outside the loaded regions (the host's geometry interest)
    # owns no far cell. Panned 16 voxels, the prepass's right half shows
    # x = 33..47, beyond the 32-voxel page.
    vgi = package()
    t = rig.oracle(vgi, cell=2, rate=1)
    (_, counters, output, history), _ = frame(t, vgi, 0, shift=16.0)
    stamped = (history[:, 7] >> 3) == 0
    stamped &= (history[:, 7] & 7) == 2
    x = history[:, 4].view(np.float32)
    assert counters[0] == int(stamped.sum()) and 16 <= counters[0] < 64
    assert np.all(x[stamped] < 32.0), "no cell outside the loaded page"
    columns = {int(c) % 8 for c in np.flatnonzero(stamped)}
    assert columns <= set(range(8)) and max(columns) <= 7

