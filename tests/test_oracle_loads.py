"""Pins of the oracle's page-load and page-unload retention (no GPU needed).

A region that loads changes as its whole box with kind Load and its top
solid row; one that unloads as kind Unload (SceneChanges::with_materials,
geometry.rs:181-215); a load's flux is the sky's luminance over the region's
width (finalize_change_flux, publication.rs:428-438); encoded as kind 2 |
top << 3 and kind 1 (resident.rs:1319-1382). Retention (voxel_retention.wgsl:
29-96) weighs a load by the height of its top above the receiver; unloads
predict nothing -- the KATs of retention.rs:276-320.
"""
import math

import numpy as np

import oracle_lib as orc
from test_oracle_edits import floor_scene, lum, pages_of

f32 = np.float32
AIR = 256


def encode(boxes):
    """One event of (min, max, flux, kind word) boxes, as encode_events lays it out."""
    words = []
    mn = np.min([b[0] for b in boxes], axis=0)
    mx = np.max([b[1] for b in boxes], axis=0)
    words += [*np.asarray(mn, np.int32).view(np.uint32), 0, *np.asarray(mx, np.int32).view(np.uint32), 0, 0, 0, 0, len(boxes)]
    for lo, hi, flux, kind in boxes:
        words += [*np.asarray(lo, np.int32).view(np.uint32), f32(flux).view(np.uint32)]
        words += [*np.asarray(hi, np.int32).view(np.uint32), kind]
    return np.array(words, np.uint32)


def receiver_tracer(receivers, retention=1.0):
    """An oracle tracer whose resident slots are +y faces of `receivers`."""
    grid, scene, t, _, _ = floor_scene()
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=retention)
    tasks = np.zeros(len(receivers), [("voxel", "<i4", (3,)), ("face", "<u4"), ("sequence", "<u4"), ("padding", "<u4", (3,))])
    tasks["voxel"] = receivers
    tasks["face"] = 2
    slots = t.cache_register(tasks)
    st = np.zeros((len(receivers), 36), np.uint32)
    st[:, 5] = 1000  # a long history, unflagged
    st[:, 0:3] = np.float32(1.0).view(np.uint32)
    st[:, 8:11] = np.float32(1.0).view(np.uint32)
    t.set_state(0, st)
    return t, slots


def pending_change(t, slot):
    return float(t.cache_meta(slot, 1)[0, 4:5].view(np.float32)[0])


def test_load_height_and_unload_kats():
    """retention.rs:276-320: a load at (0, 5, 0) with top 5 and flux 2 adds
    2 * height / (pi d^2) where height = clamp(top + 1 - y, 0, 16): nothing at
    a receiver above its top (y 12), 2 * 6 / (pi * 41) at (4, 0, 0); an
    air-only load (top saturated from i32::MIN) and an unload add nothing;
    measured mode caps no history for any of them."""
    receivers = [(4, 12, 0), (4, 0, 0), (1, 1, 1)]
    cases = [
        (2 | (5 << 3), 2.0, [0.0, 2.0 * 6.0 / (math.pi * 41.0), None]),
        (2 | ((-(1 << 27) & 0xFFFFFFFF) << 3 & 0xFFFFFFFF), 2.0, [0.0, 0.0, 0.0]),
        (1, 0.0, [0.0, 0.0, 0.0]),
    ]
    for kind, flux, expect in cases:
        t, slots = receiver_tracer(receivers)
        t.apply_events(encode([((0, 5, 0), (0, 5, 0), flux, kind)]), 1)
        for k, want in enumerate(expect):
            if want is None:
                continue
            got = pending_change(t, slots[k])
            assert abs(got - want) <= 1e-6 * max(1.0, want), (kind, k, got, want)
            assert int(t.cache_state(slots[k], 1)[0][5] & 0x7FFFFFFF) == 1000, "measured mode caps nothing"


def test_unknown_weight_falls_back_to_the_distance_curve():
    """retention.rs:317-319: an unknown change at distance 64 caps the history
    at 809 samples ((64^2)^2 + 10368) / 20736 rounded down)."""
    t, slots = receiver_tracer([(64, 0, 0)])
    t.apply_events(encode([((0, 0, 0), (0, 0, 0), 0.0, 0)]), 1)
    assert int(t.cache_state(slots[0], 1)[0][5] & 0x7FFFFFFF) == 809


def event_boxes(words):
    w = words.reshape(-1, 4)
    out = []
    for b in range(int(w[2, 3])):
        lo, hi = w[3 + 2 * b], w[4 + 2 * b]
        out.append((tuple(lo[:3].view(np.int32).tolist()), tuple(hi[:3].view(np.int32).tolist()), float(lo[3:4].view(np.float32)[0]),
                    int(hi[3])))
    return out


def test_page_load_and_unload_events():
    """A page loaded beside the floor: its regions holding solid voxels change
    as whole-region Loads with their top row and flux sky x 16 (the brighter
    of the environment's luminance and half the sun's); its air-only regions
    outside the floor's occupied bounds are no change (geometry.rs:181-189).
    Removing it: Unloads, flux 0, of its regions within the then-occupied
    bounds. Re-sending a page unchanged: no event."""
    grid, scene, t, _, _ = floor_scene(edge=64)
    t.set_transport(bounces=2, samples=16, environment=(0.2, 0.4, 0.1))
    tasks = np.zeros(4, [("voxel", "<i4", (3,)), ("face", "<u4"), ("sequence", "<u4"), ("padding", "<u4", (3,))])
    tasks["voxel"] = [(10, 3, 10), (40, 3, 40), (60, 3, 2), (2, 3, 60)]
    tasks["face"] = 2
    t.cache_register(tasks)
    sky = max(float(lum((0.2, 0.4, 0.1))), 0.5 * float(lum((3.0, 3.0, 3.0))))
    # A new page at page (2, 0, 0): a pillar up to y = 20 in its lower-left corner.
    page = np.full((32, 32, 32), AIR, np.uint16)
    page[0:4, 0:21, 0:4] = 2
    loads = event_boxes(t.load_event(np.array([[2, 0, 0]], np.int32), page.reshape(1, -1)))
    assert sorted(b[0] for b in loads) == [(64, 0, 0), (64, 16, 0)]
    for lo, hi, flux, kind in loads:
        assert tuple(np.subtract(hi, lo)) == (15, 15, 15), "a load is its whole region"
        assert kind & 7 == 2
        top = np.array([kind], np.uint32).view(np.int32)[0] >> 3
        assert top == (15 if lo[1] == 0 else 20), "each region's own top solid row"
        assert abs(flux - sky * 16.0) <= 1e-5 * sky * 16
    # The same page again: nothing changed.
    assert len(t.load_event(np.array([[2, 0, 0]], np.int32), page.reshape(1, -1))) == 0
    # Removing it: unloads of its regions within the occupied bounds (x = 4, y, z <= 1).
    unloads = event_boxes(t.load_event(np.zeros((0, 3), np.int32), np.zeros((0, 32768), np.uint16), removed=np.array([[2, 0, 0]], np.int32)))
    assert sorted(b[0] for b in unloads) == [(64, 0, 0), (64, 0, 16), (64, 16, 0), (64, 16, 16)]
    for lo, hi, flux, kind in unloads:
        assert kind == 1 and flux == 0.0
