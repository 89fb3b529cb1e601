"""Pins of the oracle's learned-cache network (no GPU needed): the hash-grid
encoding against its definition (Mueller et al. 2022: geometric level
resolutions, trilinear interpolation from the cell's corners, dense indexing
where the corners fit, the spatial hash x ^ 2654435761 y ^ 805459861 z
elsewhere) and the backward pass (weights and hash-table entries) against
central finite differences of the double-precision forward.

The oracle's exact mode drops the BF16 rounding (which makes the forward
piecewise constant) so that the gradient it computes is the derivative of the
function it evaluates; the relative L2 loss holds its denominator at the
unperturbed prediction, as the training does.
"""
import math

import numpy as np
import pytest

import oracle_lib as orc

SIZE = 1 << 22


def scene_for(extent):
    return np.array([0.0, 0.0, 0.0, 1.0 / extent], np.float32)


def queries_in(rng, n, extent):
    q = np.zeros((n, 4), np.uint32)
    q[:, 0:3] = rng.integers(0, extent, (n, 3)).astype(np.uint32)
    q[:, 3] = rng.integers(0, 6, n).astype(np.uint32) | (rng.integers(1, 8, n).astype(np.uint32) << 8)
    return q


def test_level_resolutions_are_geometric():
    for extent in (16, 64, 128, 512):
        want = [int(math.floor(16.0 * (extent / 16.0) ** (l / 7.0) + 1e-3)) for l in range(8)]
        assert orc.mlp_levels(scene_for(extent)).tolist() == want
        assert want[0] == 16 and want[-1] == extent
    assert orc.mlp_levels(scene_for(8)).tolist() == [16] * 8


def position(q, extent):
    n = np.zeros(3, np.float32)
    f = int(q[3] & 0xFF)
    n[f // 2] = 1.0 if f % 2 == 0 else -1.0
    return ((q[0:3].astype(np.int32).astype(np.float32) + np.float32(0.5) + n * np.float32(0.5)) * np.float32(1.0 / extent))


def test_corners_weights_and_indices_follow_the_definition():
    extent = 512
    rng = np.random.default_rng(1)
    q = queries_in(rng, 64, extent)
    table = np.zeros(orc.MLP_TABLE, np.float32)
    _, index, weight = orc.mlp_hash(table, q, scene_for(extent))
    res = orc.mlp_levels(scene_for(extent))
    for i in range(len(q)):
        u = position(q[i], extent)
        for level in range(8):
            r = int(res[level])
            p = u * np.float32(r)
            p0 = np.floor(p).astype(np.int64)
            w = p - np.floor(p)
            side = r + 2
            for c in range(8):
                o = np.array([c & 1, (c >> 1) & 1, c >> 2])
                x, y, z = (p0 + o).tolist()
                if side ** 3 <= SIZE:
                    want = (x + side * (y + side * z)) % SIZE
                else:
                    want = (x ^ ((y * 2654435761) & 0xFFFFFFFF) ^ ((z * 805459861) & 0xFFFFFFFF)) % SIZE
                assert index[i, level, c] == want, (i, level, c)
                ww = np.prod(np.where(o == 1, w, 1.0 - w))
                assert abs(float(weight[i, level, c]) - float(ww)) <= 1e-6
            assert abs(float(weight[i, level].sum()) - 1.0) <= 1e-5 and np.all(weight[i, level] >= 0)


def test_trilinear_interpolation_reproduces_linear_fields():
    # A dense level holding f(x, y, z) = a + b x + c y + d z at its corner
    # points (in cells) interpolates to f at the query point exactly (up to
    # FP32 rounding); feature 1 holds another plane.
    extent = 512
    res = orc.mlp_levels(scene_for(extent))
    table = np.zeros(orc.MLP_TABLE, np.float32)
    planes = [(0.25, 0.01, -0.02, 0.005), (-0.1, 0.003, 0.002, -0.004)]
    dense = [l for l in range(8) if (int(res[l]) + 2) ** 3 <= SIZE]
    assert dense == [0, 1, 2, 3, 4]
    for level in dense:
        side = int(res[level]) + 2
        z, y, x = np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij")
        flat = (x + side * (y + side * z)).ravel()
        for f, (a, b, c, d) in enumerate(planes):
            vals = (a + b * x + c * y + d * z).ravel().astype(np.float32)
            table[(level * SIZE + flat) * 2 + f] = vals
    rng = np.random.default_rng(2)
    q = queries_in(rng, 200, extent)
    feats, _, _ = orc.mlp_hash(table, q, scene_for(extent))
    for i in range(len(q)):
        p = position(q[i], extent).astype(np.float64)
        for level in dense:
            pc = p * float(np.float32(int(res[level])))
            for f, (a, b, c, d) in enumerate(planes):
                want = a + b * pc[0] + c * pc[1] + d * pc[2]
                assert abs(float(feats[i, 2 * level + f]) - want) <= 2e-5 * (1.0 + abs(want)), (i, level, f)


@pytest.fixture
def exact():
    orc.mlp_exact(True)
    yield
    orc.mlp_exact(False)


def test_gradients_match_finite_differences(exact):
    extent = 64
    sc = scene_for(extent)
    rng = np.random.default_rng(3)
    n = 24
    q = queries_in(rng, n, extent)
    diffuse = rng.uniform(0.2, 0.8, (256, 3)).astype(np.float32)
    w = np.zeros(14528, np.float32)
    offs = [0, 2048, 6144, 10240, 14336, 14528]
    ins = [32, 64, 64, 64, 64]
    for l in range(5):
        a = math.sqrt(6.0 / ins[l])
        w[offs[l]:offs[l + 1]] = rng.uniform(-a, a, offs[l + 1] - offs[l]).astype(np.float32)
    table = rng.uniform(-0.5, 0.5, orc.MLP_TABLE).astype(np.float32)
    y = np.zeros((n, 4), np.float32)
    y[:, :3] = rng.uniform(0.0, 2.0, (n, 3))
    y[:, 3] = rng.uniform(0.5, 1.0, n)
    grad, table_grad, _ = orc.mlp_gradient(w, table, q, y, diffuse, sc)
    p0 = orc.mlp_forward(w, table, q, diffuse, sc)
    den = p0 ** 2 + 0.01

    def loss(wv, tv):
        p = orc.mlp_forward(wv, tv, q, diffuse, sc)
        return float(((y[:, 3:4] / n) * (p - y[:, :3]) ** 2 / den).sum())

    checked = 0
    for l in range(5):
        for k in rng.choice(np.arange(offs[l], offs[l + 1]), 4, replace=False):
            if grad[k] == 0.0:
                continue
            hi, lo = w.copy(), w.copy()
            hi[k] += np.float32(1e-3)
            lo[k] -= np.float32(1e-3)
            fd = (loss(hi, table) - loss(lo, table)) / float(hi[k] - lo[k])
            assert abs(fd - grad[k]) <= 1e-6 * abs(grad[k]) + 1e-12, (l, k, fd, grad[k])
            checked += 1
    assert checked >= 10
    # Table entries the queries touch, a few per level.
    _, index, _ = orc.mlp_hash(table, q, sc)
    checked = 0
    for level in range(8):
        for i in range(3):
            for f in range(2):
                e = (level * SIZE + int(index[i, level, (i * 3) % 8])) * 2 + f
                hi, lo = table.copy(), table.copy()
                hi[e] += np.float32(1e-3)
                lo[e] -= np.float32(1e-3)
                fd = (loss(w, hi) - loss(w, lo)) / float(hi[e] - lo[e])
                assert abs(fd - table_grad[e]) <= 1e-6 * abs(table_grad[e]) + 1e-12, (level, i, f, fd, table_grad[e])
                checked += int(table_grad[e] != 0.0)
    assert checked >= 30
    # Entries no query touches have no gradient.
    touched = np.zeros(orc.MLP_TABLE, bool)
    for i in range(n):
        for level in range(8):
            for c in range(8):
                touched[(level * SIZE + index[i, level, c]) * 2 + np.arange(2)] = True
    assert np.all(table_grad[~touched] == 0.0)


def test_learned_paths_stop_at_the_plain_first_hit():
    """The oracle's learned path (the learned frame's: a path ends at its
    first explicit hit into the network) against its plain path on the same
    samples: the query is the plain path's first-hit voxel, face and material;
    the factor is that material's albedo (throughput 1 at the first hit); a
    sample that leaves to the sky is the plain sample exactly; a sample that
    hits gathers no more than the plain path (every omitted term is
    non-negative). A learned path that queried the wrong vertex, dropped the
    albedo or kept tracing would fail one of these."""
    from test_oracle_edits import pages_of
    edge = 32
    grid = np.full((edge, edge, edge), 256, np.uint16)
    grid[:, 0:2, :] = 1
    grid[10:13, 2:13, 10:13] = 2
    grid[28:32, 2:24, :] = 1  # two walls
    grid[:, 2:24, 28:32] = 2
    grid[20, 2, 5] = 7
    diffuse = np.full((256, 3), 0.5, np.float32)
    diffuse[2] = (0.3, 0.6, 0.9)
    emission = np.zeros((256, 3), np.float32)
    emission[7] = (4.0, 2.0, 1.0)
    scene = orc.Scene()
    scene.load_pages(*pages_of(grid))
    t = orc.Tracer(scene)
    t.set_materials(diffuse, emission)
    sun = np.zeros(12, np.float32)
    sun[0:3] = np.array([0.3, 0.8, -0.5], np.float32) / np.linalg.norm([0.3, 0.8, -0.5])
    sun[3] = -1.0
    sun[4:7] = 3.0
    t.set_lights(sun.reshape(1, 12))
    t.set_transport(bounces=3, samples=1, lanes=1, environment=(0.2, 0.3, 0.4))
    voxels = [(x, 1, z, 2) for x in range(1, 27, 2) for z in range(1, 27, 2) if not (9 <= x <= 13 and 9 <= z <= 13)]
    voxels += [(12, y, 11, 0) for y in range(2, 13)] + [(11, y, 10, 5) for y in range(2, 13)]
    tasks = np.zeros(len(voxels) * 4, [("voxel", "<i4", (3,)), ("face", "<u4"), ("sequence", "<u4"), ("padding", "<u4", (3,))])
    for k, (x, y, z, f) in enumerate(voxels * 4):
        tasks[k] = ((x, y, z), f, 17 + 5 * k, (0, 0, 0))
    res, hits, _ = t.trace_faces(tasks, hits=True)
    partial, factor, q = t.trace_learned(tasks)
    status = hits[:, 3] >> 8
    known = partial[:, 3] > 0
    assert np.array_equal(known, hits[:, 7] == 1)
    hit = (status == 1) & known
    sky = (status == 0) & known
    assert hit.sum() > 100 and sky.sum() > 100
    assert np.array_equal(q[hit, 0:3], hits[hit, 0:3])
    assert np.array_equal(q[hit, 3] & 0xFF, hits[hit, 3] & 0xFF)
    assert np.array_equal(q[hit, 3] >> 8, hits[hit, 4])
    assert np.array_equal(factor[hit], diffuse[hits[hit, 4]])
    assert np.all(q[~hit, 3] == 0xFFFFFFFF) and np.all(factor[~hit] == 0.0)
    assert np.array_equal(partial[sky, 0:3], res[sky, 0:3])
    assert np.all(res[hit, 0:3] >= partial[hit, 0:3] - 1e-6)
    assert np.any(res[hit, 0:3] > partial[hit, 0:3] + 1e-3)


@pytest.fixture
def frequency():
    orc.mlp_encoding(1)
    yield
    orc.mlp_encoding(0)


def test_encoding_layouts(frequency):
    """Both encodings' 32 features against their definitions, computed here in
    float64 numpy: hash mode -- the grid's 16 features (orc_mlp_hash), sin and
    cos of pi u per axis, the normal, the albedo, 1, zeros; frequency mode --
    sin and cos of 2^k pi u for k = 1, 2 (axis-major pairs per band) in 0..11,
    k = 3 for x, y in 12..15 and for z in 29..30, the rest as in hash mode.
    u is the face centre over the scene's largest extent."""
    rng = np.random.default_rng(8)
    extent = 256
    sc = np.array([-32.0, 0.0, 16.0, 1.0 / extent], np.float32)
    q = queries_in(rng, 300, 200)
    q[:, 0:3] = (q[:, 0:3].astype(np.int64) + [-32, 0, 16]).astype(np.uint32)
    diffuse = rng.random((256, 3)).astype(np.float32)
    table = rng.uniform(-1, 1, SIZE * 16).astype(np.float32)
    normals = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float64)
    face = q[:, 3] & 0xFF
    mat = (q[:, 3] >> 8) & 0xFF
    centre = q[:, 0:3].view(np.int32).astype(np.float64) + 0.5 + 0.5 * normals[face]
    u = (centre - sc[0:3].astype(np.float64)) / extent
    tail = np.concatenate([np.stack([np.sin(np.pi * u[:, a]) if c == 0 else np.cos(np.pi * u[:, a])
                                     for a in range(3) for c in range(2)], axis=1),
                           normals[face], diffuse[mat].astype(np.float64), np.ones((len(q), 1))], axis=1)
    freq = np.stack([np.sin(2.0 ** k * np.pi * u[:, a]) if c == 0 else np.cos(2.0 ** k * np.pi * u[:, a])
                     for k in (1, 2) for a in range(3) for c in range(2)], axis=1)
    band3 = lambda a, c: np.sin(8 * np.pi * u[:, a]) if c == 0 else np.cos(8 * np.pi * u[:, a])
    got = orc.mlp_encode(table, q, diffuse, sc)
    assert np.allclose(got[:, 0:12], freq, atol=2e-6)
    assert np.allclose(got[:, 12:16], np.stack([band3(0, 0), band3(0, 1), band3(1, 0), band3(1, 1)], axis=1), atol=2e-6)
    assert np.allclose(got[:, 16:29], tail, atol=2e-6)
    assert np.allclose(got[:, 29:31], np.stack([band3(2, 0), band3(2, 1)], axis=1), atol=2e-6)
    assert np.all(got[:, 31] == 0.0)
    orc.mlp_encoding(0)
    got = orc.mlp_encode(table, q, diffuse, sc)
    feats, _, _ = orc.mlp_hash(table, q, sc)
    assert np.array_equal(got[:, 0:16].astype(np.float32), feats)
    assert np.allclose(got[:, 16:29], tail, atol=2e-6)
    assert np.all(got[:, 29:32] == 0.0)
