"""GPU work selection against the oracle, and the behaviour the reference's
own native test asks of it (resident.rs:2066-2190).

Bar: the selection scratch (four 256-bucket histograms, thresholds,
remainders, tie and pick totals, mirror count) and every slot's key
bit-identical; the pick list identical (the keys below each class threshold
plus the first `remainder` ties at it, in slot order); the mirror records
the oracle's set; after the picks are traced and merged, every slot's state
and metadata bit-identical.
"""
import numpy as np
import pytest

import oracle_lib
from conftest import gpu_tracer, make_scene

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF


def check_picks(picks, keys_before, selection, budget):
    thr = selection[1024:1028]
    rem = selection[1028:1032]
    picked = picks[picks != NONE]
    assert len(np.unique(picked)) == len(picked)
    k = keys_before
    live = k != NONE
    cls = np.minimum(k >> 8, 3)
    bucket = k & 255
    accepted_strict = live & ((bucket < thr[cls]) | (thr[cls] == 256))
    ties = live & (bucket == thr[cls])
    n_accept = int(accepted_strict.sum()) + int(np.minimum(rem, np.bincount(cls[ties], minlength=4)).sum())
    assert int(selection[1036]) == n_accept
    assert len(picked) == min(budget, n_accept)
    in_pick = np.zeros(len(k), bool)
    in_pick[picked] = True
    assert np.all(live[picked])
    if n_accept <= budget:
        assert np.all(in_pick[accepted_strict])
        for c in range(4):
            if thr[c] < 256:
                assert int((in_pick & ties & (cls == c)).sum()) == int(rem[c])
        assert not np.any(in_pick & ~(accepted_strict | ties))
    else:
        assert np.all((accepted_strict | ties)[picked])


def setup_pair(vgi, orc, edge=64, stride=2):
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=edge)
    tasks = vgi.exposed_faces(setup["grid"])[::stride]
    t = setup["tracer"]
    t.set_transport(bounces=2, samples=2, environment=tuple(setup["env"]))
    t.cache_enable(history_limit=0, recency=1.0, measured_retention=1.0)
    g = gpu_tracer(vgi, setup, resident=True)
    g.set_transport(bounces=2, samples=2, environment=tuple(setup["env"]))
    return setup, tasks, t, g


@pytest.mark.parametrize("budget", [256, 1024])
def test_selection_matches_oracle(vgi, orc, budget):
    setup, tasks, t, g = setup_pair(vgi, orc)
    vox = tasks["voxel"]
    focus = tuple(int(v) for v in vox[len(vox) // 2])
    sched = dict(focus=focus, transport_revision=3, stable_error=0.12, stable_samples=32, audit_ms=1_000_000, mirror_budget=1 << 16)
    g.set_schedule(**sched)
    t.set_schedule(**sched)
    # Every face resident (the same slots on both sides); a first host batch
    # traces a third of them, which GPU picks then leave for four batches.
    cpu_slots = t.cache_register(tasks)
    g.bind_faces(tasks)
    hw = len(tasks)
    g.set_clock(0)
    ticket = g.submit_compact(tasks[::3])
    g.complete_readback(ticket, len(tasks[::3]))
    assert np.array_equal(g.slots(len(tasks[::3])), cpu_slots[::3])
    t.cache_batch(tasks[::3], cpu_slots[::3], clock_ms=0)
    for rnd in range(10):
        clock = 40 * (rnd + 1)
        print("selection round %d budget %d" % (rnd, budget), flush=True)
        state_before = t.cache_state(0, hw)
        g.set_clock(clock)
        g.select_frame(budget)
        sel_cpu, mirrored, picks_cpu = t.select(budget, clock_ms=clock)
        keys_before = t.cache_keys(0, hw)
        sel_gpu = g.selection()
        assert np.array_equal(sel_gpu, sel_cpu), "round %d: selection scratch differs at %s" % (
            rnd, np.nonzero(sel_gpu != sel_cpu)[0][:8].tolist())
        picks = g.picks()
        check_picks(picks, keys_before, sel_gpu, budget)
        assert np.array_equal(picks, picks_cpu), "round %d: pick lists differ" % rnd
        mirror = g.mirror()
        assert sorted(mirror[:, 0].tolist()) == sorted(mirrored.tolist())
        for rec in mirror:
            s = int(rec[0])
            words = state_before[s, :24]
            rec_words = rec[4:]
            # The settled flag rides on bit 31 of the count word.
            same = rec_words == words
            same[5] = (rec_words[5] & 0x7FFFFFFF) == (words[5] & 0x7FFFFFFF)
            assert np.all(same), "mirror record of slot %d" % s
        t.cache_picks(picks, clock_ms=clock)
        assert np.array_equal(g.keys(0, hw), t.cache_keys(0, hw)), "round %d keys" % rnd
        assert np.array_equal(g.cache_state(0, hw), t.cache_state(0, hw)), "round %d state" % rnd
        assert np.array_equal(g.cache_meta(0, hw), t.cache_meta(0, hw)), "round %d metadata" % rnd
    ms = g.select_timings()
    assert ms[0] > 0
    g.close()


def test_picks_cover_nearest_first_and_leave_host_jobs(vgi, orc):
    # resident.rs:2066-2190 on this build's API: every face resident and
    # unmeasured; one host batch claims one face; the first four pick batches
    # (budget 1) rank the four quarters while picking at most one face and
    # never the claimed one; at batch 4 coverage goes nearest the focus first
    # (log-distance buckets); after enough rounds every face is covered and
    # every picked face reached the statistics mirror.
    setup, tasks, t, g = setup_pair(vgi, orc, edge=64, stride=16)
    vox = tasks["voxel"].astype(np.int64)
    focus = vox[len(vox) // 3]
    g.set_schedule(focus=tuple(int(v) for v in focus), transport_revision=1, stable_error=0.12, stable_samples=32,
                   audit_ms=1_000_000, mirror_budget=1024)
    hw = len(tasks)
    g.set_clock(0)
    # The host traces one face (claiming it); binding the whole set makes
    # every other face resident without tracing it.
    ticket = g.submit_compact(tasks[:1])
    g.complete_readback(ticket, 1)
    owned = int(g.slots(1)[0])
    assert owned == 0
    g.bind_faces(tasks)
    assert g.bound_count() == hw
    mirrored = set()
    # Coverage buckets as the selection forms them (float32, its log2).
    dv = (vox - focus).astype(np.float32)
    d2 = dv[:, 0] * dv[:, 0] + dv[:, 1] * dv[:, 1] + dv[:, 2] * dv[:, 2] + np.float32(1.0)
    buckets = np.array([min(int(np.float32(oracle_lib.select_log2(float(x))) * np.float32(8.0)), 255) for x in d2])
    for rnd in range(400):
        budget = 1 if rnd < 4 else (64 if rnd == 4 else 256)
        before = g.cache_state(0, hw)[:, 5] & 0x3FFFFFFF
        g.set_clock(16 * (rnd + 1))
        g.select_frame(budget)
        picks = g.picks()
        mirrored.update(g.mirror()[:, 0].tolist())
        after = g.cache_state(0, hw)[:, 5] & 0x3FFFFFFF
        picked = picks[picks != NONE]
        assert len(picked) <= budget
        if 1 <= rnd < 4:
            assert owned not in picked.tolist(), "a face with a host job in flight is never picked"
        if rnd == 4:
            keys = g.keys(0, hw)
            left = np.array([s for s in range(hw) if after[s] == 0 and s not in picked and keys[s] != NONE])
            assert len(picked) == budget and len(left)
            assert buckets[picked].max() <= buckets[left].min()
        if np.all(after >= 16):
            break
    assert np.all(after >= 16), "every face covered by GPU picks (%d left after %d rounds)" % (int((after < 16).sum()), rnd)
    # Every picked face reached the mirror (the host face may have come back
    # with its host job instead).
    final_mirror = set(g.mirror()[:, 0].tolist()) | mirrored
    g.select_frame(1)
    final_mirror |= set(g.mirror()[:, 0].tolist())
    missing = [s for s in range(hw) if s != owned and s not in final_mirror]
    assert not missing, "%d picked faces never mirrored" % len(missing)
    g.close()
