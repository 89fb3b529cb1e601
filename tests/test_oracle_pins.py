"""Pins of the CPU oracle against what the reference's own tests hold and what
the mathematics fixes (no GPU needed).

Each test names the reference test or closed form it follows. A mistake of
the usual kinds in the oracle (a dropped term, a wrong sign or index, a
transposed operand) fails at least one of these.
"""
import ctypes
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as orc
from conftest import load_package

f32 = np.float32


# ------------------------------------------------------------------ NAADF
def lcg_occupancy(size, divisor, state):
    """naadf/construction_tests.rs:97-108,122-142: the seeded LCG occupancy."""
    sx, sy, sz = size
    out = np.zeros((sz, sy, sx), np.uint8)
    for i in range(sx * sy * sz):
        state = (state * 747796405 + 2891336453) & 0xFFFFFFFF
        x, y, z = i % sx, (i // sx) % sy, i // (sx * sy)
        out[z, y, x] = 1 if divisor != 0 and state % divisor == 0 else 0
    return out, state


def test_build_matches_frozen_full_sweep():
    """naadf/construction_tests.rs:110-145: every 2^3 topology, then seeded sizes."""
    for bits in range(256):
        blocked = np.array([(bits >> i) & 1 for i in range(8)], np.uint8).reshape(2, 2, 2)
        assert np.array_equal(orc.build_aadf(blocked), orc.build_aadf(blocked, full_sweep=True))
    state = 0x12345678
    for size in [(0, 0, 0), (1, 1, 1), (4, 4, 4), (9, 7, 8), (17, 3, 11), (255, 1, 1), (1, 255, 1), (1, 1, 255)]:
        for divisor in [0, 1, 2, 7, 29]:
            blocked, state = lcg_occupancy(size, divisor, state)
            if blocked.size == 0:
                continue
            assert np.array_equal(orc.build_aadf(blocked), orc.build_aadf(blocked, full_sweep=True)), (size, divisor)


def test_empty_grid_closed_form():
    """naadf.rs:201-208: the empty 4^3 grid expands to every edge."""
    f = orc.build_aadf(np.zeros((4, 4, 4), np.uint8))
    at = f[1, 2, 1]  # (x, y, z) = (1, 2, 1)
    assert (at[0], at[2], at[4]) == (1, 2, 1)  # negative x, y, z
    assert (at[1], at[3], at[5]) == (2, 1, 2)  # positive x, y, z


def test_boxes_hold_only_empty_cells():
    """naadf.rs:166-199: every box holds only empty cells. (The merging rule's
    boxes are not maximal boxes in general; the rule itself is pinned in
    test_oracle_kats.py by an independent full sweep and by the closed form
    of runs along one axis.)"""
    blocked, _ = lcg_occupancy((9, 7, 8), 7, 0x12345678)
    f = orc.build_aadf(blocked).astype(np.int64)
    sz, sy, sx = blocked.shape
    for z in range(sz):
        for y in range(sy):
            for x in range(sx):
                b = f[z, y, x]
                if blocked[z, y, x]:
                    assert not b.any()
                    continue
                box = blocked[z - b[4]:z + b[5] + 1, y - b[2]:y + b[3] + 1, x - b[0]:x + b[1] + 1]
                assert box.size == (b[0] + b[1] + 1) * (b[2] + b[3] + 1) * (b[4] + b[5] + 1)
                assert not box.any()


# ------------------------------------------------------------------ hashes and samplers
def test_pcg_is_a_bijection_with_avalanche():
    """PCG RXS-M-XS 32 (sampling.rs:5-9) is a permutation of u32: its LCG step
    and output xorshift-multiply are each invertible. A dropped or wrong shift
    typically shows as collisions or poor avalanche."""
    L = orc.lib()
    xs = np.arange(1 << 18, dtype=np.uint64)
    ys = np.array([L.orc_pcg(int(x)) for x in xs], np.uint64)
    assert len(np.unique(ys)) == len(xs)
    flips = []
    for x in range(2000):
        for bit in (0, 7, 15, 31):
            flips.append(bin(L.orc_pcg(x) ^ L.orc_pcg(x ^ (1 << bit))).count("1"))
    assert 14.5 < np.mean(flips) < 17.5


def test_halton_dimensions_are_rotated_radical_inverses():
    """voxel_path.wgsl:188-200: dims < 12 are the base-p radical inverse of
    sequence+1 plus a Cranley-Patterson rotation, mod 1 (exact rational check)."""
    L = orc.lib()
    primes = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]
    for dim, p in enumerate(primes):
        for stream in (0, 12345, 0xDEADBEEF):
            rot = Fraction(L.orc_pcg(stream ^ L.orc_pcg(dim)) >> 9, 1 << 23)
            for seq in (0, 1, 2, 5, 63, 1000, 77777, (1 << 24) - 2):
                i, ri, scale = seq + 1, Fraction(0), Fraction(1)
                while i:
                    scale /= p
                    ri += (i % p) * scale
                    i //= p
                expect = float((ri + rot) % 1)
                got = L.orc_sample_1d(stream, seq, dim)
                d = abs(got - expect)
                assert min(d, 1 - d) < 4e-7, (dim, stream, seq, got, expect)


def test_sobol_pairs_form_scrambled_nets():
    """voxel_path.wgsl:202-212: dims >= 12 are Owen-scrambled Sobol pairs, so
    the first 2^m points of each pair are a (0, m, 2)-net: every elementary
    interval of area 2^-m holds exactly one point."""
    L = orc.lib()
    m = 8
    n = 1 << m
    for stream in (7, 0xABCDEF):
        for dim in (12, 14, 16):
            pts = np.array([[L.orc_sample_1d(stream, s, dim), L.orc_sample_1d(stream, s, dim + 1)] for s in range(n)])
            for a in range(m + 1):
                b = m - a
                cells = (np.floor(pts[:, 0] * (1 << a)).astype(int) << b) | np.floor(pts[:, 1] * (1 << b)).astype(int)
                assert len(np.unique(cells)) == n, (stream, dim, a)


def test_own_sin_cos_exp_accuracy():
    """The project's sin/cos/exp against float64 references: within 2 ulp-scale."""
    L = orc.lib()
    xs = np.linspace(0.0, 2 * math.pi, 20001, dtype=np.float64).astype(np.float32)
    s = np.array([L.orc_sinf(float(x)) for x in xs])
    c = np.array([L.orc_cosf(float(x)) for x in xs])
    assert np.max(np.abs(s - np.sin(xs.astype(np.float64)))) < 2.5e-7
    assert np.max(np.abs(c - np.cos(xs.astype(np.float64)))) < 2.5e-7
    es = np.linspace(-20.0, 5.0, 5001, dtype=np.float32)
    e = np.array([L.orc_expf(float(x)) for x in es])
    ref = np.exp(es.astype(np.float64))
    assert np.max(np.abs(e - ref) / ref) < 3e-7


def test_hemisphere_is_cosine_weighted():
    """voxel_path.wgsl:231-238: unit directions in the normal's hemisphere with
    pdf cos/pi: E[cos] = 2/3, E[cos^2] = 1/2, and tangential means vanish."""
    L = orc.lib()
    out = np.zeros(3, np.float32)
    for face in range(6):
        n = np.zeros(3, np.float32)
        n[face // 2] = -1.0 if face & 1 else 1.0
        cos, cos2, tang = [], [], []
        for s in range(4096):
            u = L.orc_sample_1d(99, s, 0)
            v = L.orc_sample_1d(99, s, 1)
            L.orc_hemisphere(n.ctypes.data_as(ctypes.c_void_p), u, v, out.ctypes.data_as(ctypes.c_void_p))
            assert abs(np.linalg.norm(out) - 1.0) < 1e-6
            cos.append(float(out @ n))
            cos2.append(float(out @ n) ** 2)
            tang.append(out - (out @ n) * n)
        assert min(cos) >= 0.0
        assert abs(np.mean(cos) - 2.0 / 3.0) < 2e-3
        assert abs(np.mean(cos2) - 0.5) < 2e-3
        assert np.abs(np.mean(tang, axis=0)).max() < 5e-3


# ------------------------------------------------------------------ traversal
def scalar_box_step(position, direction, cell, negative, positive, size):
    """gpu/traversal_tests.rs:15-39, the reference's scalar XYZ oracle of
    advance_box, in float32."""
    size = f32(size)
    mn = np.array([f32(c - n) * size for c, n in zip(cell, negative)], np.float32)
    mx = np.array([f32(c + p + 1) * size for c, p in zip(cell, positive)], np.float32)
    distance = f32(1e30)
    normal = [0.0, 0.0, 0.0]
    for axis in range(3):
        d = f32(direction[axis])
        if abs(d) <= f32(1e-8):
            continue
        plane = mx[axis] if d > 0 else mn[axis]
        travel = f32(f32(plane - f32(position[axis])) / d)
        if travel >= f32(-0.001) and travel < distance:
            distance = max(travel, f32(0.0))
            normal = [0.0, 0.0, 0.0]
            normal[axis] = -1.0 if d > 0 else 1.0
    return np.array(normal, np.float32), max(f32(distance + f32(0.001)), f32(0.001))


def test_advance_box_matches_scalar_boundary_policy():
    """gpu/traversal_tests.rs:41-208: normal exact, distance within 1e-5 (1 + d),
    on the reference's grid of sizes, bases, fractions, extents and tie-prone
    direction components (subsampled), plus its unequal-negative-crossing case."""
    L = orc.lib()
    comps = [-1.0, -1.01e-8, -1e-8, -0.99e-8, -0.0, 0.0, 0.99e-8, 1e-8, 1.01e-8, 1.0]
    out = np.zeros(4, np.float32)
    cases = []
    for size in (1.0, 4.0, 16.0):
        for base in (-4096, -16, -1, 0, 1, 16, 4096):
            cell = (base, -base, int(base / 2))
            for frac in (-0.0005, 0.0, 0.5, 1.0, 1.0005):
                for ext in ((0, 0, 0), (3, 2, 1)):
                    pos = [f32((f32(c) + f32(frac)) * f32(size)) for c in cell]
                    for d in itertools.product(comps, repeat=3):
                        cases.append((pos, d, cell, ext, ext, size))
    rng = np.random.default_rng(3)
    picked = [cases[i] for i in rng.choice(len(cases), 40000, replace=False)]
    picked.append(([1.0009, 1.0001, 1.0002], (1.0, 1.0, 1.0), (0, 0, 0), (0, 0, 0), (0, 0, 0), 1.0))
    for pos, d, cell, neg, posx, size in picked:
        p = np.array(pos, np.float32)
        dd = np.array(d, np.float32)
        c = np.array(cell, np.int32)
        n = np.array(neg, np.int32)
        q = np.array(posx, np.int32)
        L.orc_advance_box(*(a.ctypes.data_as(ctypes.c_void_p) for a in (p, dd, c, n, q)), size,
                          out.ctypes.data_as(ctypes.c_void_p))
        rn, rd = scalar_box_step(p, dd, cell, neg, posx, size)
        assert np.array_equal(out[:3] + 0.0, rn + 0.0), (pos, d, cell, out, rn)
        assert abs(out[3] - rd) <= 1e-5 * (1.0 + rd)


def brute_force_first_solid(grid, origin, direction, max_t=400.0):
    """Voxel-by-voxel DDA (Amanatides & Woo) in float64 over a dense grid."""
    pos = np.array(origin, np.float64)
    d = np.array(direction, np.float64)
    v = np.floor(pos).astype(int)
    step = np.sign(d).astype(int)
    with np.errstate(divide="ignore"):
        tmax = np.where(d != 0, ((v + (step > 0)) - pos) / d, np.inf)
        tdelta = np.where(d != 0, np.abs(1.0 / d), np.inf)
    n = grid.shape[0]
    t = 0.0
    while t < max_t:
        if np.any(v < 0) or np.any(v >= n):
            return None, None
        if grid[v[2], v[1], v[0]] < 256:
            return tuple(int(a) for a in v), t
        a = int(np.argmin(tmax))
        t = tmax[a]
        v[a] += step[a]
        tmax[a] += tdelta[a]
    return None, None


def test_traversal_agrees_with_brute_force_walk():
    """Brute force on small inputs: the NAADF traversal's first hit is the first
    solid voxel a voxel-by-voxel walk meets (ties between nearly equal plane
    crossings excepted), and the hit voxel is solid."""
    vgi = load_package()
    grid = vgi.generate(3, 0x5EED0003, 64)
    coords, pages = vgi.pages_of(grid)
    scene = orc.Scene()
    scene.load_pages(coords, pages)
    tr = orc.Tracer(scene)
    rng = np.random.default_rng(11)
    agree = checked = 0
    air = np.argwhere(grid == 256)
    for _ in range(1200):
        z, y, x = air[rng.integers(len(air))]
        o = np.array([x + 0.5, y + 0.5, z + 0.5], np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        oi, of = tr.trace_ray(o, d)
        ref, t = brute_force_first_solid(grid, o, d)
        if oi[4] == 1:
            assert grid[oi[2], oi[1], oi[0]] < 256  # the hit voxel is solid
            prev = np.floor(of[:3] - 0.01 * d).astype(int)  # just before the hit: air or outside
            if np.all(prev >= 0) and np.all(prev < 64):
                assert grid[prev[2], prev[1], prev[0]] == 256 or tuple(prev) == tuple(oi[:3])
        if ref is None:
            assert oi[4] == 0
            continue
        checked += 1
        agree += int(oi[4] == 1 and tuple(oi[:3]) == ref)
    assert checked > 300 and agree >= checked - 3


# ------------------------------------------------------------------ emitters, retention
def test_emitter_distribution_kat():
    """emission.rs:401-432: two adjacent emissive voxels expose 10 faces,
    probabilities sum to 1, select(0) has cumulative 0.1, the last is 1."""
    page = np.full(32 ** 3, 256, np.uint16)
    page[4 + 32 * (4 + 32 * 4)] = 10
    page[5 + 32 * (4 + 32 * 4)] = 10
    scene = orc.Scene()
    scene.load_pages(np.zeros((1, 3), np.int32), page.reshape(1, -1))
    tr = orc.Tracer(scene)
    diffuse = np.full((256, 3), 0.45, np.float32)
    emission = np.zeros((256, 3), np.float32)
    emission[10] = 5.0
    tr.set_materials(diffuse, emission)
    w = tr.emitter_words()
    assert len(w) == 10 * 8  # 10 faces, no tree below 32 faces
    rec = w.reshape(10, 8)
    cum = rec[:, 4].view(np.float32)
    prob = rec[:, 5].view(np.float32)
    assert abs(prob.sum() - 1.0) < 1e-6
    assert cum[0] == np.float32(0.1)
    assert cum[-1] == 1.0
    # select(0.999) lands on the last face (cumulative 1.0)
    assert int(np.searchsorted(cum, np.float32(0.999), side="right")) == 9


def test_light_tree_probabilities_normalise():
    """tree.rs:28-74: a binary tree of 2F-1 nodes whose root holds the whole
    power (probabilities sum to 1) and whose leaves name every face once."""
    vgi = load_package()
    grid = vgi.generate(2, 0x5EED0002, 64)
    diffuse, emission, lights, env = vgi.optics(2, 0x5EED0002)
    coords, pages = vgi.pages_of(grid)
    scene = orc.Scene()
    scene.load_pages(coords, pages)
    tr = orc.Tracer(scene)
    tr.set_materials(diffuse, emission)
    words = tr.emitter_words()
    # faces*8 words of records, then nodes*16 words with nodes = 2*faces - 1
    nfaces = (len(words) + 16) // 40
    assert nfaces * 8 + (2 * nfaces - 1) * 16 == len(words)
    nodes = words[nfaces * 8:].reshape(-1, 16)
    power = nodes[:, 3].view(np.float32)
    assert abs(power[0] - 1.0) < 1e-5  # root holds all probability mass
    leaves = nodes[nodes[:, 12] == 0xFFFFFFFF]
    assert len(leaves) == nfaces
    assert sorted(leaves[:, 13].tolist()) == list(range(nfaces))


def test_retention_curve_kat():
    """retention.rs:346-376: H(d) = round((d/12)^4) with the history ceiling."""
    L = orc.lib()
    for d, expected in [(0, 0), (8, 0), (16, 3), (24, 16), (32, 51), (48, 256), (64, 809), (96, 4096), (128, 12945),
                        (192, 0xFFFFFFFF), (384, 0xFFFFFFFF)]:
        assert L.orc_curve_limit(d * d) == expected
    for d2 in range(0, 36864, 7):
        expected = round((math.sqrt(d2) / 12.0) ** 4)
        assert L.orc_curve_limit(d2) == (0xFFFFFFFF if expected >= 65536 else expected)


# ------------------------------------------------------------------ estimator closed forms
@pytest.mark.parametrize("bounces", [1, 2, 3])
def test_finite_depth_furnace(bounces):
    """gpu/view_tests.rs:387-459 / research.rs:71-115: in the closed furnace of
    emission E and albedo r, a depth-d estimate converges to
    E (1 - r^(d+1)) / (1 - r) (including the terminal emission connection);
    within 1.5% per channel."""
    vgi = load_package()
    grid = vgi.generate(6, 0)
    diffuse, emission, lights, env = vgi.optics(6, 0)
    coords, pages = vgi.pages_of(grid)
    scene = orc.Scene()
    scene.load_pages(coords, pages)
    tr = orc.Tracer(scene)
    tr.set_materials(diffuse, emission)
    tr.set_lights(lights)
    tr.set_transport(bounces=bounces, samples=64)
    faces = vgi.exposed_faces(grid)
    inside = np.all((faces["voxel"] >= 8) & (faces["voxel"] <= 23), axis=1)
    normals = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
    nb = faces["voxel"] + normals[faces["face"]]
    interior = inside & np.all((nb >= 9) & (nb <= 22), axis=1)
    tasks = faces[interior]
    assert len(tasks) == 6 * 14 * 14
    res, _, _ = tr.trace_faces(tasks)
    mean = res[:, :3].mean(axis=0).astype(np.float64)
    E = emission[4].astype(np.float64)
    r = diffuse[4].astype(np.float64)
    expected = E * (1 - r ** (bounces + 1)) / (1 - r)
    assert np.all(np.abs(mean / expected - 1) < 0.015), (mean, expected)


def test_cached_furnace_converges_to_infinite_series():
    """gpu/cache_tests.rs:43-101: with world-cache termination and a short
    history the furnace converges to E / (1 - r) (within 2%)."""
    vgi = load_package()
    grid = vgi.generate(6, 0)
    diffuse, emission, lights, env = vgi.optics(6, 0)
    coords, pages = vgi.pages_of(grid)
    scene = orc.Scene()
    scene.load_pages(coords, pages)
    tr = orc.Tracer(scene)
    tr.set_materials(diffuse, emission)
    tr.set_transport(bounces=3, samples=16)
    tr.cache_enable(history_limit=256)
    faces = vgi.exposed_faces(grid)
    normals = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
    nb = faces["voxel"] + normals[faces["face"]]
    inside = np.all((faces["voxel"] >= 8) & (faces["voxel"] <= 23), axis=1) & np.all((nb >= 9) & (nb <= 22), axis=1)
    tasks = faces[inside]
    slots = tr.cache_register(tasks)
    avg = np.zeros(3)
    for batch in range(400):
        res, _ = tr.cache_batch(tasks, slots, clock_ms=batch * 16)
        if batch in (349, 399):
            avg += res[:, :3].mean(axis=0) / 2
    expected = emission[4] / (1 - diffuse[4])
    assert np.all(np.abs(avg / expected - 1) < 0.02), (avg, expected)


# ------------------------------------------------------------------ top level
def assert_valid_fixed_point(scene):
    """world.rs:596-684: every box holds only non-barrier cells inside the
    bounds and no box can grow by the merging rule."""
    top, size = scene.top_level()
    b = scene.bounds()
    sx, sy, sz = size
    blocked = np.zeros((sz, sy, sx), np.uint8)
    L = orc.lib()
    for z in range(sz):
        for y in range(sy):
            for x in range(sx):
                s = L.orc_scene_region_solid(scene.h, b[0] + x, b[1] + y, b[2] + z)
                blocked[z, y, x] = 1 if s != 0 else 0  # unloaded (-1) or solid
    steps = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
    for z in range(sz):
        for y in range(sy):
            for x in range(sx):
                f = top[z, y, x]
                if blocked[z, y, x]:
                    assert not f.any()
                    continue
                lo = (x - f[0], y - f[2], z - f[4])
                hi = (x + f[1], y + f[3], z + f[5])
                assert min(lo) >= 0 and hi[0] < sx and hi[1] < sy and hi[2] < sz
                assert not blocked[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1].any()
                for dirn, (dx, dy, dz) in enumerate(steps):
                    nx, ny, nz = x + dx, y + dy, z + dz
                    if not (0 <= nx < sx and 0 <= ny < sy and 0 <= nz < sz) or blocked[nz, ny, nx]:
                        continue
                    if f[dirn] == 255:
                        continue
                    nbr = top[nz, ny, nx]
                    opp = dirn ^ 1
                    grows = all(nbr[k] >= f[k] for k in range(6) if k != opp)
                    assert not grows, ((x, y, z), dirn)


def test_top_level_fixed_point_build_and_updates():
    """The FIFO growth (world.rs:838-912) reaches a valid fixed point on a full
    build and after edits that flip regions solid and empty (cut_boxes)."""
    vgi = load_package()
    grid = np.full((128, 128, 128), 256, np.uint16)
    rng = np.random.default_rng(5)
    for _ in range(40):  # scattered solid regions
        x, y, z = rng.integers(0, 8, 3)
        grid[z * 16 + 3, y * 16 + 5, x * 16 + 7] = 1
    coords, pages = vgi.pages_of(grid)
    scene = orc.Scene()
    scene.load_pages(coords, pages)
    assert_valid_fixed_point(scene)
    for _ in range(3):
        xyz = rng.integers(0, 128, (30, 3)).astype(np.int32)
        vals = np.where(rng.random(30) < 0.5, 1, 256).astype(np.uint16)
        scene.edit(xyz, vals)
        assert_valid_fixed_point(scene)
