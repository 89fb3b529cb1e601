"""Known-answer pins of the CPU oracle taken from the reference's own tests,
and closed forms written independently of the oracle's code (no GPU needed).

- pcg: an independently written inverse of PCG-RXS-M-XS-32 with the
  reference's constants (sampling.rs:5-9) recovers every input.
- packing: the KATs of gpu_scene.rs:672-781,942-956 (interned brick sizes,
  uniform-brick regions without payload, the single-voxel brick word, the
  disconnected world's 16 headers) and the decode oracle of
  gpu_scene.rs:604-632 over whole pages.
- regions: the reference's frozen region constructor
  (world/construction_tests.rs:8-107) restated here against the oracle's
  regions in their packed canonical form.
- face index: face_index.rs:231-261 (40,000 seeded inserts and removals
  against a map, longest probe below 64).
- accumulate (voxel_accumulate.wgsl:135-227): the running mean up to the
  history cap and the exponential tail past it, recency 2's closed-form
  weights, the measured-retention keep c*sigma^2/D^2 on lit and dark faces,
  and the display blend's dependence on elapsed time.
- NAADF: a numpy restatement of the frozen full sweep
  (naadf/construction_tests.rs:6-94) against the oracle's build, and the
  closed form of runs along one axis.
"""
import numpy as np
import pytest

import oracle_lib as orc

AIR = 256
M32 = 0xFFFFFFFF


# ------------------------------------------------------------------ pcg
PCG_MUL, PCG_INC, PCG_WORD = 747796405, 2891336453, 277803737  # sampling.rs:6-7


def pcg_inverse(h):
    """Undo PCG-RXS-M-XS-32 step by step: the output xorshift by 22 (one
    application inverts a shift of at least half the width), the multiply by
    the word constant (its inverse mod 2^32), the input-dependent xorshift
    (its top 4 bits pass through unchanged and give the shift), then the LCG."""
    word = (h ^ (h >> 22)) & M32
    t = (word * pow(PCG_WORD, -1, 1 << 32)) & M32
    shift = (t >> 28) + 4
    state, acc = t, t
    while True:
        acc >>= shift
        if acc == 0:
            break
        state ^= acc
    return ((state - PCG_INC) * pow(PCG_MUL, -1, 1 << 32)) & M32


def test_pcg_round_trips_through_an_independent_inverse():
    L = orc.lib()
    rng = np.random.default_rng(5)
    values = [0, 1, 2, 0xFFFFFFFF, 0x80000000, 0x7FFFFFFF, 0xA341316C] + rng.integers(0, 1 << 32, 20000, dtype=np.uint64).tolist()
    for v in values:
        assert pcg_inverse(L.orc_pcg(int(v))) == int(v), hex(int(v))
    # hash_voxel = pcg(pcg(pcg(x) ^ y) ^ z) (voxel_traversal.wgsl:29-31), peeled from the outside in.
    for x, y, z in [(0, 0, 0), (5, -3, 17), (-200, 511, 3), (1 << 20, -(1 << 20), 7)]:
        h = L.orc_hash_voxel(x, y, z)
        inner = pcg_inverse(h) ^ (z & M32)
        assert pcg_inverse(inner) ^ (y & M32) == L.orc_pcg(x & M32)
        # stream_id = hash_voxel ^ pcg(face + 1) (sampling.rs:11-13)
        for face in range(6):
            assert L.orc_stream_id(x, y, z, face) ^ L.orc_pcg(face + 1) == h


# ------------------------------------------------------------------ packing
def page_tracer(pages):
    """An oracle tracer over {page xyz: 32^3 u16 (z, y, x)} (no lights)."""
    coords = np.array(list(pages.keys()), np.int32).reshape(-1, 3)
    data = np.stack([np.ascontiguousarray(p, np.uint16).reshape(-1) for p in pages.values()]) if pages else np.zeros((0, 32768), np.uint16)
    scene = orc.Scene()
    scene.load_pages(coords, data)
    t = orc.Tracer(scene)
    t.set_materials(np.full((256, 3), 0.5, np.float32), np.zeros((256, 3), np.float32))
    t.set_lights(np.zeros((0, 12), np.float32))
    return scene, t


def decoded_page(t, origin=(0, 0, 0)):
    out = np.zeros((32, 32, 32), np.uint16)
    for z in range(32):
        for y in range(32):
            for x in range(32):
                out[z, y, x] = t.packed_material(origin[0] + x, origin[1] + y, origin[2] + z)
    return out


def test_repeated_bricks_share_storage():
    """gpu_scene.rs:672-739: a plane of material 255 at y = 4 across a page is
    one mixed brick shared by 64 table entries: four region tables (256
    words) and one brick (32 words) are 1152 bytes; decoding the packed tables
    gives back every voxel, also after an edit of one reference."""
    page = np.full((32, 32, 32), AIR, np.uint16)
    page[:, 4, :] = 255
    scene, t = page_tracer({(0, 0, 0): page})
    cells, regions, _, unique, refs = t.packed_info()
    assert cells * 4 == 1152 and unique == 1 and refs == 64 and regions == 8
    assert np.array_equal(decoded_page(t), page)
    page[4, 4, 4] = AIR
    scene.edit(np.array([[4, 4, 4]], np.int32), np.array([AIR], np.uint16))
    assert t.packed_material(4, 4, 4) == AIR
    assert t.packed_material(20, 4, 20) == 255, "editing one reference must not modify another"
    assert np.array_equal(decoded_page(t), page)


def test_region_of_uniform_bricks_has_no_voxel_payload():
    """gpu_scene.rs:741-765: x < 8 material 0, else 255, in the first region:
    one brick table and no brick payload."""
    page = np.full((32, 32, 32), AIR, np.uint16)
    page[:16, :16, :8] = 0
    page[:16, :16, 8:16] = 255
    _, t = page_tracer({(0, 0, 0): page})
    cells, _, _, unique, _ = t.packed_info()
    assert cells == 64 and unique == 0
    assert t.packed_material(3, 3, 3) == 0 and t.packed_material(9, 3, 3) == 255


def test_single_voxel_brick_word():
    """gpu_scene.rs:767-781: one voxel (4, 4, 4) of 255: a table and one
    brick, whose first cell is material 255 ((cell & 65535) >> 1)."""
    page = np.full((32, 32, 32), AIR, np.uint16)
    page[4, 4, 4] = 255
    _, t = page_tracer({(0, 0, 0): page})
    cells = t.packed_words(2)
    assert len(cells) == 64 + 32
    assert (cells[64] & 65535) >> 1 == 255
    # The brick's table entry points at it with tag 2 (gpu_scene.rs:228-233).
    assert cells[1 + 4 * 1 + 16 * 1] == (64 << 2) | 2
    assert np.array_equal(decoded_page(t), page)


def test_disconnected_world_does_not_allocate_its_bounding_box():
    """gpu_scene.rs:942-956: two empty pages 10,000 pages apart: 16 headers,
    no cells."""
    empty = np.full((32, 32, 32), AIR, np.uint16)
    _, t = page_tracer({(0, 0, 0): empty, (10000, 10000, 10000): empty})
    cells, regions, _, _, _ = t.packed_info()
    assert cells == 0 and regions == 16


def test_every_live_region_is_found_within_the_directory_probe_limit():
    """gpu_scene.rs:866-880: each region's directory entry is reached within
    32 probes from its hash and names its header."""
    rng = np.random.default_rng(11)
    pages = {}
    for k in range(12):
        p = np.where(rng.random((32, 32, 32)) < 0.2, rng.integers(0, 8, (32, 32, 32)), AIR).astype(np.uint16)
        pages[(int(rng.integers(-6, 6)), int(rng.integers(-2, 2)), int(rng.integers(-6, 6)))] = p
    _, t = page_tracer(pages)
    directory = t.packed_words(0).view(np.int32).reshape(-1, 4)
    mask = len(directory) - 1
    L = orc.lib()
    live = [tuple(e[:3]) for e in directory if e[3] != -1]
    assert len(live) == 8 * len(pages)
    for key in live:
        slot = L.orc_hash_voxel(*map(int, key)) & mask
        for _ in range(32):
            if tuple(directory[slot][:3]) == key:
                break
            assert directory[slot][3] != -1
            slot = (slot + 1) & mask
        else:
            pytest.fail("region %s beyond 32 probes" % (key,))
    for (px, py, pz), p in pages.items():
        assert np.array_equal(decoded_page(t, (px * 32, py * 32, pz * 32)), p)


# ------------------------------------------------------------------ regions
def pack_extents(b):
    """gpu_scene.rs:462-467: six 2-bit extents, NEG_X first."""
    bits = 0
    for i in range(6):
        assert b[i] < 4
        bits |= int(b[i]) << (i * 2)
    return bits


def frozen_region_canonical(materials):
    """The reference's frozen region constructor (world/construction_tests.rs:
    8-107: per-brick voxel AADFs, brick occupancy, the brick AADF), written in
    the packed canonical form (kind, material, table words with tag 2 for
    mixed bricks, 32 cell words per mixed brick; gpu_scene.rs:199-236)."""
    m = np.asarray(materials, np.uint16).reshape(16, 16, 16)  # (z, y, x)
    out = np.zeros(4 + 64 + 64 * 32, np.uint32)
    if (m == AIR).all():
        out[0], out[1] = 1, AIR
        return out
    if (m == m.flat[0]).all():
        out[0], out[1] = 2, int(m.flat[0])
        return out
    out[0] = 3
    occupied = np.zeros((4, 4, 4), np.uint8)
    voxel_aadf = {}
    for bz in range(4):
        for by in range(4):
            for bx in range(4):
                brick = m[bz * 4:bz * 4 + 4, by * 4:by * 4 + 4, bx * 4:bx * 4 + 4]
                blocked = (brick != AIR).astype(np.uint8)
                occupied[bz, by, bx] = blocked.any()
                voxel_aadf[(bx, by, bz)] = orc.build_aadf(blocked)
    brick_aadf = orc.build_aadf(occupied)
    for index in range(64):
        bx, by, bz = index & 3, (index >> 2) & 3, index >> 4
        brick = m[bz * 4:bz * 4 + 4, by * 4:by * 4 + 4, bx * 4:bx * 4 + 4]
        first = int(brick[0, 0, 0])
        if (brick == first).all():
            out[4 + index] = (pack_extents(brick_aadf[bz, by, bx]) << 2) if first == AIR else ((first << 2) | 1)
            continue
        out[4 + index] = 2
        for i in range(64):
            lx, ly, lz = i & 3, (i >> 2) & 3, i >> 4
            mat = int(brick[lz, ly, lx])
            cell = (pack_extents(voxel_aadf[(bx, by, bz)][lz, ly, lx]) << 1) if mat == AIR else ((mat << 1) | 1)
            out[68 + index * 32 + i // 2] |= cell << ((i & 1) * 16)
    return out


def region_cases(rng):
    yield np.full((16, 16, 16), AIR, np.uint16)
    yield np.full((16, 16, 16), 7, np.uint16)
    bricks = np.where(rng.random((4, 4, 4)) < 0.5, rng.integers(0, 5, (4, 4, 4)), AIR).astype(np.uint16)
    yield np.kron(bricks, np.ones((4, 4, 4), np.uint16)).astype(np.uint16)  # uniform bricks only
    for density in (0.02, 0.2, 0.5, 0.9):
        yield np.where(rng.random((16, 16, 16)) < density, rng.integers(0, 255, (16, 16, 16)), AIR).astype(np.uint16)
    one = np.full((16, 16, 16), 3, np.uint16)
    one[9, 2, 13] = AIR  # a single air voxel in stone
    yield one
    hole = np.full((16, 16, 16), AIR, np.uint16)
    hole[5, 6, 7] = 1  # a single solid voxel in air
    yield hole


def test_regions_match_the_frozen_region_constructor():
    rng = np.random.default_rng(21)
    cases = list(region_cases(rng))
    page = np.full((32, 32, 32), AIR, np.uint16)
    coords = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)]
    for start in range(0, len(cases), 8):
        chunk = cases[start:start + 8]
        for k, region in enumerate(chunk):
            rx, ry, rz = coords[k]
            page[rz * 16:rz * 16 + 16, ry * 16:ry * 16 + 16, rx * 16:rx * 16 + 16] = region
        scene, _ = page_tracer({(0, 0, 0): page})
        for k, region in enumerate(chunk):
            got = scene.region_canonical(*coords[k])
            want = frozen_region_canonical(region)
            assert np.array_equal(got[:2], want[:2]), ("kind", start + k, got[:2], want[:2])
            assert np.array_equal(got[4:], want[4:]), ("payload", start + k)


# ------------------------------------------------------------------ face index
def test_face_index_matches_a_map_through_insertions_and_removals():
    """face_index.rs:231-261."""
    table = orc.FaceIndex(16)
    reference = {}
    L = orc.lib()
    state = 12345
    for step in range(40000):
        state = L.orc_pcg((state ^ step) & M32)
        voxel = ((state % 23) - 11, (state >> 8) % 7, -((state >> 16) % 19))
        face = (state >> 24) % 6
        if state & 3 == 0:
            table.remove(voxel, face)
            reference.pop((voxel, face), None)
        else:
            table.insert(voxel, face, step)
            reference[(voxel, face)] = step
        if step % 997 == 0:
            for (v, f), slot in reference.items():
                assert table.get(v, f) == slot
            assert len(table) == len(reference)
    for (v, f), slot in reference.items():
        assert table.get(v, f) == slot
    # Removed keys are gone.
    gone = [(v, f) for v in [(0, 0, 0), (3, 2, -5), (-11, 6, -18)] for f in range(6) if (v, f) not in reference]
    for v, f in gone:
        assert table.get(v, f) is None
    assert table.max_probe() < 64, table.max_probe()


# ------------------------------------------------------------------ accumulate
def resident_slot(history_limit=0, recency=1.0, retention=1.0):
    """An oracle tracer with one resident slot (slot 0)."""
    page = np.full((32, 32, 32), AIR, np.uint16)
    page[0, 0, 0] = 1
    scene, t = page_tracer({(0, 0, 0): page})
    t.set_transport(bounces=1, samples=1)
    t.cache_enable(history_limit=history_limit, recency=recency, measured_retention=retention)
    task = np.zeros(1, [("voxel", "<i4", (3,)), ("face", "<u4"), ("sequence", "<u4"), ("padding", "<u4", (3,))])
    task["face"] = 2
    slot = t.cache_register(task)
    assert slot[0] == 0
    return scene, t


def fresh(mean, second, known=1.0):
    r = np.zeros(24, np.float32)
    r[0:3] = mean
    r[5] = known
    r[8:11] = second
    return r


def test_running_mean_up_to_the_cap_then_exponential_tail():
    """history cap L = 4, one sample per batch: the first four merges are the
    running mean; past the cap each merge moves the mean by 1/L of the
    difference, so x_j weighs (1/L)(1 - 1/L)^(n-j) (a closed form)."""
    _, t = resident_slot(history_limit=4)
    xs = [1.0, 5.0, 2.0, 8.0, 3.0, 9.0, 4.0, 0.5, 6.0]
    L = 4
    for n, x in enumerate(xs, start=1):
        t.accumulate(0, fresh([x, 2 * x, 0.0], [x * x, 4 * x * x, 0.0]), 1, clock_ms=16 * n)
        st = t.cache_state(0, 1)[0]
        mean = st[0:3].view(np.float32)
        count = int(st[5] & 0x7FFFFFFF)
        assert count == min(n, L)
        if n <= L:
            expect = np.mean(xs[:n])
        else:
            expect = np.mean(xs[:L]) * (1 - 1 / L) ** (n - L) + sum((1 / L) * (1 - 1 / L) ** (n - j) * xs[j - 1] for j in range(L + 1, n + 1))
        assert abs(mean[0] - expect) <= 1e-5 * max(1.0, abs(expect)), (n, mean[0], expect)
        assert abs(mean[1] - 2 * expect) <= 2e-5 * max(1.0, abs(expect))
        assert int(st[6]) == n  # the sequence advances by the batch's samples


def test_recency_two_weights_samples_by_their_rank():
    """recency k = 2, no cap: alpha = min(2 fresh / count, 1), so after n
    single-sample batches sample i weighs 2 (i - 1) / (n (n - 1)) (the
    product of (1 - 2/j) telescopes)."""
    _, t = resident_slot(recency=2.0)
    rng = np.random.default_rng(9)
    xs = rng.uniform(0.0, 4.0, 40)
    for n, x in enumerate(xs, start=1):
        t.accumulate(0, fresh([x, 0.0, 0.0], [x * x, 0.0, 0.0]), 1, clock_ms=16 * n)
        if n >= 2:
            mean = t.cache_state(0, 1)[0][0:3].view(np.float32)[0]
            w = np.array([2.0 * (i - 1) / (n * (n - 1)) for i in range(1, n + 1)])
            assert abs(mean - float(w @ xs[:n])) <= 2e-5 * max(1.0, float(w @ xs[:n])), n


def prepared_state(mean, second, count, changed):
    st = np.zeros(36, np.uint32)
    st[0:3] = np.asarray(mean, np.float32).view(np.uint32)
    st[8:11] = np.asarray(second, np.float32).view(np.uint32)
    st[5] = count | (0x80000000 if changed else 0)
    return st


@pytest.mark.parametrize("lit", [True, False])
def test_measured_retention_keeps_c_sigma2_over_d2(lit):
    """voxel_accumulate.wgsl:151-172: a face flagged changed by D keeps
    floor(c sigma^2 / D^2) samples of its history before the fresh batch
    merges; sigma^2 is its own retained variance when it was lit relative to
    the change, else the fresh batch's (luminance weights; only green varies
    here, so sigma^2 = 0.7152 var). Unflagged, the history is kept whole."""
    c, D = 3.0, 0.5
    var_old, var_new = 0.64, 0.36
    if lit:
        mean, second = [0.0, 2.0, 0.0], [0.0, 4.0 + var_old, 0.0]
        sigma2 = 0.7152 * var_old
    else:
        mean, second = [0.0, 0.0, 0.0], [0.0, var_old, 0.0]
        sigma2 = 0.7152 * var_new
    expect_keep = int(np.floor(c * sigma2 / (D * D)))
    assert 2 <= expect_keep <= 20
    scene, t = resident_slot(retention=c)
    meta = t.cache_meta(0, 1)
    meta[0, 4] = np.float32(D).view(np.uint32)
    t.set_meta(0, meta)
    t.set_state(0, prepared_state(mean, second, 100, changed=True))
    t.accumulate(0, fresh([0.0, 1.0, 0.0], [0.0, 1.0 + var_new, 0.0]), 1)
    st = t.cache_state(0, 1)[0]
    assert int(st[5] & 0x7FFFFFFF) == expect_keep + 1
    assert t.cache_meta(0, 1)[0, 4] == 0, "a batch with known samples consumes the pending change"
    # Without the changed flag the pending change does not bind.
    t.set_meta(0, meta)
    t.set_state(0, prepared_state(mean, second, 100, changed=False))
    t.accumulate(0, fresh([0.0, 1.0, 0.0], [0.0, 1.0 + var_new, 0.0]), 1)
    assert int(t.cache_state(0, 1)[0][5] & 0x7FFFFFFF) == 101


def test_display_blend_follows_elapsed_time():
    """voxel_accumulate.wgsl:187-215: with a settled, noiseless face the
    display moves towards the mean by min(1 - exp(-dt / 75 ms), n / (n + 32))
    of the way: unchanged when no time passed, nearly all the way after a
    long wait once the count is large."""
    _, t = resident_slot()
    st = prepared_state([2.0, 2.0, 2.0], [4.0, 4.0, 4.0], 100000, changed=False)
    st[24:27] = np.array([0.5, 0.5, 0.5], np.float32).view(np.uint32)
    st[29], st[30] = 1000, 0xFFFFFFFF  # valid clock at 1000 ms
    t.set_state(0, st)
    t.accumulate(0, fresh([2.0, 2.0, 2.0], [4.0, 4.0, 4.0]), 1, clock_ms=1000)
    disp = t.cache_state(0, 1)[0][24:27].view(np.float32)
    assert np.allclose(disp, 0.5, atol=1e-7), "no elapsed time: display unchanged"
    t.accumulate(0, fresh([2.0, 2.0, 2.0], [4.0, 4.0, 4.0]), 1, clock_ms=1000 + 75)
    disp2 = t.cache_state(0, 1)[0][24:27].view(np.float32)
    # One response time: blend 1 - e^-1; the weight ratio favours the brighter new value less:
    # alpha = min(blend, (blend / 2.25) / ((1 - blend) / 0.75 + blend / 2.25)).
    b = 1 - np.exp(-1.0)
    alpha = min(b, (b / 2.25) / ((1 - b) / 0.75 + b / 2.25))
    assert np.allclose(disp2, 0.5 + (2.0 - 0.5) * alpha, rtol=2e-6)
    t.accumulate(0, fresh([2.0, 2.0, 2.0], [4.0, 4.0, 4.0]), 1, clock_ms=1000 + 75 + 60000)
    disp3 = t.cache_state(0, 1)[0][24:27].view(np.float32)
    assert np.allclose(disp3, 2.0, rtol=1e-3), "a long wait at a large count shows the mean"


# ------------------------------------------------------------------ NAADF
def numpy_full_sweep(blocked):
    """naadf/construction_tests.rs:6-94 as whole-array numpy sweeps (blocked:
    (z, y, x)); field (z, y, x, 6) with NEG_X, POS_X, NEG_Y, POS_Y, NEG_Z, POS_Z."""
    blocked = np.asarray(blocked, bool)
    shape = blocked.shape
    field = np.zeros(shape + (6,), np.int32)
    rounds = max(max(shape) - 1, 0)
    for _ in range(rounds):
        for axis in range(3):  # x, y, z
            np_axis = 2 - axis
            nxt = field.copy()
            for direction, step in ((axis * 2, -1), (axis * 2 + 1, 1)):
                # neighbour = position + step along the axis
                nb = np.roll(field, -step, axis=np_axis)
                nb_blocked = np.roll(blocked, -step, axis=np_axis)
                inside = np.ones(shape, bool)
                idx = [slice(None)] * 3
                idx[np_axis] = 0 if step < 0 else shape[np_axis] - 1
                inside[tuple(idx)] = False
                # try_grow: the cell's box grows by one towards an empty
                # neighbour whose box covers it in every other direction.
                ok = inside & ~nb_blocked & (field[..., direction] < 255)
                opposite = direction ^ 1
                for d in range(6):
                    if d != opposite:
                        ok &= nb[..., d] >= field[..., d]
                nxt[..., direction] += ok
            nxt[blocked] = 0
            field = nxt
    return field


def test_build_matches_an_independent_full_sweep():
    """The oracle's build_aadf against the numpy full sweep: every 2^3
    topology, then seeded volumes (naadf/construction_tests.rs:110-145)."""
    from test_oracle_pins import lcg_occupancy

    for bits in range(256):
        blocked = np.array([(bits >> i) & 1 for i in range(8)], np.uint8).reshape(2, 2, 2)
        assert np.array_equal(orc.build_aadf(blocked), numpy_full_sweep(blocked)), bits
    state = 0x12345678
    for size in [(1, 1, 1), (4, 4, 4), (9, 7, 8), (17, 3, 11), (64, 1, 1), (1, 64, 1), (1, 1, 64)]:
        for divisor in [0, 1, 2, 7, 29]:
            blocked, state = lcg_occupancy(size, divisor, state)
            assert np.array_equal(orc.build_aadf(blocked), numpy_full_sweep(blocked)), (size, divisor)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_runs_along_one_axis_reach_their_ends(axis):
    """On a 1-voxel-thick line the boxes are the empty runs: each empty cell
    extends to the last empty cell before a blocked one or the edge."""
    rng = np.random.default_rng(axis)
    n = 40
    line = rng.random(n) < 0.3
    shape = [1, 1, 1]
    shape[2 - axis] = n
    blocked = line.reshape(shape).astype(np.uint8)
    f = orc.build_aadf(blocked).reshape(n, 6)
    for i in range(n):
        if line[i]:
            assert not f[i].any()
            continue
        lo = i
        while lo > 0 and not line[lo - 1]:
            lo -= 1
        hi = i
        while hi < n - 1 and not line[hi + 1]:
            hi += 1
        want = np.zeros(6, np.uint8)
        want[2 * axis], want[2 * axis + 1] = i - lo, hi - i
        assert np.array_equal(f[i], want), (i, f[i], want)
