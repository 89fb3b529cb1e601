"""Pins of the oracle's change footprints (no GPU needed): the Light event
each edited box gets, whose bound is where sunlight through the box lands
and whose flux is the light reflected there (gpu/resident.rs:1215-1310,
voxel_trace.wgsl:94-202), against the reference's own tests of them
(gpu/resident.rs:2203-2404) and against the geometry of the tubes worked
out by hand.

Encoding: the edit's event first (three headers, then its boxes), then one
event per footprint (three headers, one bound); a footprint that lands
nowhere keeps the empty bound (min 2^29, max -2^29) out of every receiver's
reach.
"""
import math

import numpy as np

import oracle_lib as orc
from test_oracle_edits import AIR, lum, pages_of

f32 = np.float32
EMPTY_LO, EMPTY_HI = 1 << 29, -(1 << 29)


def sun_scene(grid, direction=(0.0, 1.0, 0.0), radiance=1.0, footprints=True):
    diffuse = np.full((256, 3), 0.5, np.float32)
    diffuse[2] = (0.3, 0.6, 0.9)
    emission = np.zeros((256, 3), np.float32)
    scene = orc.Scene()
    scene.load_pages(*pages_of(grid))
    tracer = orc.Tracer(scene)
    tracer.set_materials(diffuse, emission)
    sun = np.zeros(12, np.float32)
    sun[0:3] = direction
    sun[3] = -1.0
    sun[4:7] = radiance
    tracer.set_lights(sun.reshape(1, 12))
    tracer.set_transport(bounces=1, samples=1)
    tracer.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    tracer.set_footprints(footprints)
    return tracer, diffuse


def events(words):
    """Per event: (union min, union max, [(min, max, flux, kind)])."""
    w = words.reshape(-1, 4)
    e = orc.event_count(words)
    out = []
    for k in range(e):
        first, n = int(w[3 * k + 2, 2]), int(w[3 * k + 2, 3])
        boxes = []
        for b in range(n):
            lo, hi = w[3 * e + 2 * (first + b)], w[3 * e + 2 * (first + b) + 1]
            boxes.append((tuple(lo[:3].view(np.int32).tolist()), tuple(hi[:3].view(np.int32).tolist()),
                          float(lo[3:4].view(np.float32)[0]), int(hi[3] & 7)))
        out.append((tuple(w[3 * k, :3].view(np.int32).tolist()), tuple(w[3 * k + 1, :3].view(np.int32).tolist()), boxes))
    return out


def skylight_grid(roof):
    # gpu/resident.rs:2205-2306: a floor (y 4) under a ceiling (y 20) whose
    # 3x3 hole (x, z 8..10) the edit opens, with or without a roof (y 28).
    grid = np.full((32, 32, 32), AIR, np.uint16)
    grid[:, 4, :] = 2
    grid[:, 20, :] = 2
    if roof:
        grid[:, 28, :] = 2
    return grid


def hole():
    return [[x, 20, z] for z in range(8, 11) for x in range(8, 11)], [AIR] * 9


def tube_flux(spacing, diffuse_rgb, radiance=1.0):
    """One tube's reflected flux x 64, rounded (voxel_trace.wgsl:176-187)."""
    c = lum([radiance] * 3)
    f = f32(f32(f32(c * lum(diffuse_rgb)) * spacing) * spacing)
    return int(f32(f * f32(64.0)) + f32(0.5))


def test_skylight_footprint_lands_on_the_floor_below():
    tracer, diffuse = sun_scene(skylight_grid(roof=False))
    words = tracer.edit_event(*hole())
    ev = events(words)
    assert len(ev) == 2, "the edit's event and its footprint's"
    (emin, emax, eboxes), (fmin, fmax, fboxes) = ev
    assert eboxes[0][:2] == ((8, 20, 8), (10, 20, 10))
    # The grid spans the box diagonal |(3, 1, 3)| = sqrt(19): 5 x 5 tubes at
    # spacing sqrt(19)/5; with the sun straight up the outer ring (offsets
    # +-1.74 from the centre 9.5) misses the 8..11 box and the inner 3 x 3
    # tubes (x, z = 9.5 + {-0.87, 0, 0.87}) land on floor voxels 8, 9, 10.
    assert (fmin, fmax) == ((8, 4, 8), (10, 4, 10))
    (bmin, bmax, flux, kind), = fboxes
    assert (bmin, bmax, kind) == ((8, 4, 8), (10, 4, 10), 4)
    spacing = f32(f32(math.sqrt(19.0)) / f32(5.0))
    assert flux == f32(9 * tube_flux(spacing, diffuse[2])) / f32(64.0)
    # The reference's bound: nine voxels of white sunlight times the albedo.
    albedo = float(lum(diffuse[2]))
    assert abs(flux - 9.0 * albedo) <= 0.25 * 9.0 * albedo


def test_covered_skylight_footprint_lands_nowhere():
    tracer, _ = sun_scene(skylight_grid(roof=True))
    ev = events(tracer.edit_event(*hole()))
    assert len(ev) == 2
    _, (fmin, fmax, fboxes) = ev
    assert fmin == (EMPTY_LO,) * 3 and fmax == (EMPTY_HI,) * 3, "a hole under a roof changes no sunlight"
    assert fboxes == [((EMPTY_LO,) * 3, (EMPTY_HI,) * 3, 0.0, 4)]


def test_placed_block_footprint_does_not_land_inside_its_support():
    # gpu/resident.rs:2333-2404: a 3x3 block placed on the floor; its tubes
    # start inside the floor under it and land nowhere.
    grid = np.full((32, 32, 32), AIR, np.uint16)
    grid[:, 4, :] = 2
    tracer, _ = sun_scene(grid)
    xyz = [[x, 5, z] for z in range(8, 11) for x in range(8, 11)]
    ev = events(tracer.edit_event(xyz, [2] * 9))
    assert len(ev) == 2
    assert ev[1][2][0][0] == (EMPTY_LO,) * 3, "the edit's own faces are the edit event's"


def test_oblique_sun_footprint_is_the_shifted_shadow():
    # A 1x1x1 block floating at (10, 12, 10) over the floor (top at y 5), sun
    # towards (1, 1, 0)/sqrt(2): the shadow falls 7 voxels towards -x.
    grid = np.full((32, 32, 32), AIR, np.uint16)
    grid[:, 4, :] = 2
    d = np.array([1.0, 1.0, 0.0], np.float32) / np.float32(math.sqrt(2.0))
    tracer, _ = sun_scene(grid, direction=tuple(d))
    ev = events(tracer.edit_event([[10, 12, 10]], [2]))
    assert len(ev) == 2
    (fmin, fmax, flux, kind), = ev[1][2]
    assert kind == 4 and flux > 0
    # Rays through the cube's section (x 10..11 at y 12..13, z 10..11) leave
    # it downwards along -(1, 1, 0) and meet y = 5 after 7..8 voxels of drop.
    assert fmin[1] == fmax[1] == 4
    assert 10 - 8 <= fmin[0] <= fmax[0] <= 11 - 7 + 1 and fmin[2] >= 10 and fmax[2] <= 11


def test_footprints_off_plans_the_edit_alone():
    # gpu.rs:419-425 (without_footprints).
    tracer, _ = sun_scene(skylight_grid(roof=False), footprints=False)
    words = tracer.edit_event(*hole())
    assert orc.event_count(words) == 1 and len(words) == 4 * (3 + 2)


def test_footprints_take_edits_first_and_at_most_64():
    # gpu/resident.rs:1230-1273: one footprint per edited box, edits before
    # boxes of unknown flux, 64 at most. Sorted boxes go through the scan
    # budget in (min, max) order: two whole 16^3 regions (18^3 = 5832
    # positions each) fit the 16384 budget, the third does not and is
    # unknown; single voxels (27 positions) after them all fit.
    def layout(small):
        grid = np.full((128, 32, 128), AIR, np.uint16)  # z, y, x
        grid[:, 4, :] = 2
        tracer, _ = sun_scene(grid)
        xyz, values = [], []
        for rz in range(3):  # regions x 0..15, y 16..31, z 16 rz..: filled
            for z in range(16 * rz, 16 * rz + 16):
                for y in range(16, 32):
                    for x in range(16):
                        xyz.append([x, y, z]), values.append(2)
        spots = [(rx * 16 + 8, 24, rz * 16 + 8) for rz in range(8) for rx in range(1, 8)]
        spots += [(8, 24, rz * 16 + 8) for rz in range(3, 8)]
        spots += [(rx * 16 + 8, 10, rz * 16 + 8) for rz in range(8) for rx in range(8)]
        for s in spots[:small]:
            xyz.append(list(s)), values.append(2)
        return tracer, np.array(xyz, np.int32), np.array(values, np.uint16)

    for small, footprints in [(61, 64), (62, 64), (10, 13)]:
        tracer, xyz, values = layout(small)
        ev = events(tracer.edit_event(xyz, values))
        boxes = ev[0][2]
        assert len(boxes) == 3 + small
        # The third region in box order is over the budget (no face observed
        # any of them, so the measured ones end as unknown flux too).
        assert boxes[2] == ((0, 16, 32), (15, 31, 47), 0.0, 0)
        assert len(ev) == 1 + footprints
        lands = [e[2][0][:2] for e in ev[1:]]
        under_unknown = [i for i, (mn, mx) in enumerate(lands) if mn[0] < 16 and 32 <= mn[2] <= 47]
        if 3 + small <= 64:
            # Every box has a footprint; the unknown one's comes last and
            # covers the floor under it.
            assert under_unknown == [footprints - 1]
            mn, mx = lands[-1]
            assert mn[1] == mx[1] == 4 and mn[0] >= 0 and mx[0] <= 15 and mn[2] >= 32 and mx[2] <= 47
        else:
            assert under_unknown == [], "over 64 edited boxes: the unknown one gets no footprint"
        # The measured regions' footprints come first, in box order.
        assert lands[0][0][2] < 16 and 16 <= lands[1][0][2] < 32
