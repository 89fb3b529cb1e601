"""The oracle's restatement of GPU work selection (voxel_accumulate.wgsl:99-130,
448-695), pinned on the CPU:

- the bucket logarithm against the mathematics (exact at powers of two; a
  bucket differs from floor(8 log2(d^2 + 1)) nowhere outside a 1e-5 band
  around a bucket edge);
- select_threshold's quotas, lending and thresholds on histograms whose
  answer is derived by hand from the rules (comments give the arithmetic);
- candidate classification of hand-built slot records, one per rule;
- record_stability on faces whose statistics are known in closed form (black
  faces: zero mean and variance), through their stable-batch count to the
  settled flag.
"""
import math

import numpy as np

import oracle_lib

URGENT = 0x40000000
LIVE = 8


def test_log2_matches_mathematics():
    for k in range(-30, 31):
        assert oracle_lib.select_log2(2.0 ** k) == float(k)
    rng = np.random.default_rng(5)
    xs = np.concatenate([np.arange(1, 5000, dtype=np.float64), rng.uniform(1e-6, 1e6, 4000),
                         np.float64(np.float32(2.0 ** rng.uniform(-20, 30, 2000)))])
    for x in xs:
        x = float(np.float32(x))
        got = oracle_lib.select_log2(x)
        want = math.log2(x)
        assert abs(got - want) <= 3e-7 + 2.0 * np.spacing(np.float32(abs(want))), (x, got, want)
    assert oracle_lib.select_log2(0.0) == -math.inf
    # Coverage buckets floor(8 log2(d^2 + 1)) over integer d^2.
    mismatches = 0
    for d2 in range(0, 200000, 7):
        got = min(int(np.float32(oracle_lib.select_log2(float(d2 + 1))) * np.float32(8.0)), 255)
        exact = 8.0 * math.log2(d2 + 1)
        want = min(int(exact), 255)
        if got != want and abs(exact - round(exact)) > 1e-5:
            mismatches += 1
    assert mismatches == 0


def hist(entries):
    h = np.zeros((4, 256), np.uint32)
    for cls, bucket, n in entries:
        h[cls, bucket] += n
    return h


def test_threshold_quotas_and_lending():
    # 100 coverage keys at bucket 10, 10 refinement keys at bucket 20, budget 64:
    # quotas coverage 64/4 = 16, refinement 64/2 = 32 -> 10, sweep 16 -> 0;
    # spare 38 goes to refinement (none left), then coverage: 16 + 38 = 54.
    thr, rem = oracle_lib.select_threshold(hist([(0, 10, 100), (1, 20, 10)]), 64)
    assert thr.tolist() == [10, 20, 256, 256] and rem.tolist() == [54, 10, 0, 0]
    # Urgent first, at most 3/4 of 64 = 48; rest 16: coverage 4 of 5; spare
    # 12: coverage +1, then urgent +11 = 59.
    thr, rem = oracle_lib.select_threshold(hist([(3, 3, 100), (0, 0, 5)]), 64)
    assert thr.tolist() == [0, 256, 256, 3] and rem.tolist() == [5, 0, 0, 59]
    # Sweep keys 3 each at buckets 0, 1, 2 and budget 7: quota 7 - 1 - 3 = 3,
    # spare 4 lent back to sweep: 7 = 3 + 3 + 1, threshold bucket 2, one tie.
    thr, rem = oracle_lib.select_threshold(hist([(2, 0, 3), (2, 1, 3), (2, 2, 3)]), 7)
    assert thr.tolist() == [256, 256, 2, 256] and rem.tolist() == [0, 0, 1, 0]
    # Budget 1: only sweep gets a share (1 - 0 - 0); coverage keeps no
    # threshold bucket (256) although it has keys.
    thr, rem = oracle_lib.select_threshold(hist([(0, 0, 5), (2, 0, 5)]), 1)
    assert thr.tolist() == [256, 256, 0, 256] and rem.tolist() == [0, 0, 1, 0]
    # No keys: no thresholds.
    thr, rem = oracle_lib.select_threshold(hist([]), 64)
    assert thr.tolist() == [256] * 4 and rem.tolist() == [0] * 4


def meta_record(voxel, face=0, live=True, schedule=0, revision_tag=0, transport=0, noise=0.0, merge_ms=0, brightest=0.0,
                rank=0):
    m = np.zeros(12, np.uint32)
    m[0:3] = np.asarray(voxel, np.int32).view(np.uint32)
    m[3] = face | (LIVE if live else 0)
    m[5] = schedule
    m[6] = revision_tag
    m[7] = transport
    m[8] = np.float32(noise).view(np.uint32)
    m[9] = merge_ms
    m[10] = np.float32(brightest).view(np.uint32)
    m[11] = rank
    return m


def test_candidate_classes(orc):
    scene = orc.Scene()
    scene.load_pages(np.zeros((1, 3), np.int32), np.zeros((1, 32 ** 3), np.uint16))
    t = orc.Tracer(scene)
    t.cache_enable()
    focus = (40, 40, 40)
    t.set_schedule(focus=focus, transport_revision=0, stable_error=0.12, stable_samples=32, audit_ms=1000, mirror_budget=64)
    cases = []
    # Coverage (never measured, count < 8): bucket floor(8 log2(d^2 + 1)).
    cases.append((meta_record((40, 40, 40)), 0x000 | 0))                    # d^2 = 0
    cases.append((meta_record((41, 40, 40)), 0x000 | 8))                    # log2(2) = 1
    cases.append((meta_record((41, 41, 41)), 0x000 | 16))                   # log2(4) = 2
    cases.append((meta_record((50, 40, 40)), 0x000 | 53))                   # 8 log2(101) = 53.27
    # Urgent: class 3, same distance buckets, even when measured and stable.
    cases.append((meta_record((41, 41, 41), schedule=12, rank=0x80000000 | URGENT | 500, brightest=1.0), 0x300 | 16))
    # Refinement (measured, 100 samples, within one page): error 1/sqrt(100)
    # = 0.1, age 1 s -> 0.1, +0.25 near the focus: priority 0.45;
    # 16 (log2(0.45) + 8) = 109.57 -> bucket 255 - 109 = 146.
    cases.append((meta_record((45, 40, 40), noise=1.0, merge_ms=9000, rank=0x80000000 | 100), 0x100 | 146))
    # Sweep (three pages away): age 1023 ms, 8 log2(1024) = 80 -> 255 - 80 = 175.
    cases.append((meta_record((140, 40, 40), noise=1.0, merge_ms=10000 - 1023, rank=0x80000000 | 100), 0x200 | 175))
    # Stable (4 batches, fresh in this revision), bright, enough samples, audited
    # recently: not due.
    cases.append((meta_record((45, 40, 40), schedule=8 | 4, noise=1.0, merge_ms=9990, brightest=1.0,
                              rank=0x80000000 | 64), 0xFFFFFFFF))
    # The same face once its audit interval has passed: due again (refinement).
    cases.append((meta_record((45, 40, 40), schedule=8 | 4, noise=1.0, merge_ms=10000 - 1000, brightest=1.0,
                              rank=0x80000000 | 64), 0x100 | (255 - int(16 * (math.log2(1.0 / 8.0 + 1.0 * 0.1 + 0.25) + 8)))))
    # Dark stable faces need 256 samples: 64 is not enough, still due.
    cases.append((meta_record((45, 40, 40), schedule=8 | 4, noise=1.0, merge_ms=9990, brightest=0.01,
                              rank=0x80000000 | 64), 0x100 | (255 - int(16 * (math.log2(1.0 / 8.0 + 0.01 * 0.1 + 0.25) + 8)))))
    # A host job of batch 98 owns the face through batch 101; free at 102.
    cases.append((meta_record((40, 40, 40), schedule=98 << 4), 0xFFFFFFFF))
    # Retired slot.
    cases.append((meta_record((40, 40, 40), live=False), 0xFFFFFFFF))
    tasks = np.zeros((len(cases), 8), np.uint32)
    for i in range(len(cases)):
        tasks[i, 0:3] = [i, 0, 0]  # distinct faces; the records below replace their identity
    slots = t.cache_register(tasks)
    assert slots.tolist() == list(range(len(cases)))
    t.set_meta(0, np.stack([c[0] for c in cases]))
    for i, (_, want) in enumerate(cases):
        got = t.select_key(i, 100, 10000)
        assert got == want, "case %d: %#x vs %#x" % (i, got, want)
    assert t.select_key(len(cases) - 2, 102, 10000) == 0x000 | 0


def test_record_stability_on_black_faces(orc):
    # A closed box with no light: every sample is 0, so mean, variance, drift
    # and error are 0 and the display settles at once. Dark faces need 256
    # samples: at 4 per batch the stable count starts at batch 64, reaches 4
    # at batch 67 and the settled flag (schedule & 15 == 12) rides on the
    # records from then on.
    scene = orc.Scene()
    page = np.full(32 ** 3, 1, np.uint16)
    page.reshape(32, 32, 32)[1:31, 1:31, 1:31] = 256  # air
    scene.load_pages(np.zeros((1, 3), np.int32), page[None, :])
    t = orc.Tracer(scene)
    t.set_materials(np.full((256, 3), 0.5, np.float32), np.zeros((256, 3), np.float32))
    t.set_lights(np.zeros((0, 12), np.float32))
    t.set_transport(bounces=1, samples=4)
    t.cache_enable()
    t.set_schedule(stable_error=0.12, stable_samples=32)
    tasks = np.zeros((3, 8), np.uint32)
    tasks[:, 0:3] = [[0, 5, 5], [5, 0, 5], [5, 5, 0]]
    tasks[:, 3] = [0, 2, 4]  # wall faces into the interior
    slots = t.cache_register(tasks)
    for batch in range(1, 71):
        res, _ = t.cache_batch(tasks, slots, clock_ms=16 * batch)
        meta = t.cache_meta(0, 3)
        state = t.cache_state(0, 3)
        count = state[:, 5] & 0x7FFFFFFF
        assert np.all(count == 4 * batch)
        stable = meta[:, 5] & 7
        want = 0 if 4 * batch < 256 else min(batch - 63, 4)
        assert np.all(stable == want), (batch, stable)
        assert np.all(meta[:, 5] & 8) and np.all(meta[:, 7] == 0) and np.all(meta[:, 9] == 16 * batch)
        assert np.all(meta[:, 11] == (0x80000000 | count))
        settled = (res.view(np.uint32)[:, 5] >> 31) == 1
        assert np.all(settled == (want == 4)), batch
