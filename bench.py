"""Benchmark of the lighting hot path on one MI355X.

A step is one frame of the resident world-space pass over every exposed face
of the scene: prepare (sequences) -> trace (cosine-hemisphere paths with
sun, emitter NEE and world-cache termination) -> accumulate (temporal merge).
Inputs (scene, faces, optics) are resident in HBM before the timed region.

Default workload: the 512^3 worm-cave scene (seeded noise rock, 200 worm
tunnels of radius 3-8, 3 sun shafts, 300 emitters), plain 4-bounce tracing at
4 rays per face per frame, FP32.

Prints one JSON line. `value` is Mrays/s over the timed frames (every
traversal call counted on the GPU); the roofline object prices the trace
kernel's algorithmic gather bytes against HBM peak; the cpu_baseline object
times the CPU oracle (the same algorithm, scalar C++) on a bounded sample of
the same scene's faces on the host cores.
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0

# Learned-cache configuration: training targets per frame, Adam steps per
# frame, paths averaged into each target; rays per face of a learned frame.
LEARNED_TRAIN = (81920, 8, 16)
LEARNED_SPP = 1
# Adam rate of the network (the hash table's is ten times it).
LEARNED_LR = 3e-3
# History recency of the learned form's temporal merge (the transport's
# history_recency, resident.rs:721-725): sample i of n weighs about
# (i/n)^0.25, so the frames of the network's warm-up fade faster than in
# the exact mean.
LEARNED_RECENCY = 1.25
# The caller's training schedule (vgi_mlp_training): after 800 learned frames
# the network has settled and half the targets in half the Adam steps keep
# it there, at 1.1 ms less per frame.
LEARNED_SCHEDULE = {800: (40960, 4, LEARNED_LR)}
# The network as the 512^3 cave configuration states it: a 32-dim frequency
# encoding (no hash grid), 64k one-path samples per frame, Adam at 1e-3 (the
# steps per frame are not stated: 8, as above).
STATED_TRAIN = (65536, 8, 1)
STATED_LR = 1e-3
# The error judge: every JUDGE_STRIDE-th face against FP64 sums of
# REFERENCE_SPP-ray launches.
JUDGE_STRIDE = 512
REFERENCE_SPP = 32

WORKLOADS = {
    # name: (generator id, seed, edge, bounces, spp, description)
    "cave512": (5, 0x5EED0005, 512, 4, 4, "512^3 worm cave, 200 worms r3-8, 300 emitters, 3 sun shafts; plain 4-bounce, 4 rays/face"),
    "terrain256": (2, 0x5EED0002, 256, 2, 4, "256^3 noise terrain ~35% solid, 64 emitters, sun; 2 bounces, 4 rays/face, world cache"),
    "box64": (1, 0x5EED0001, 64, 1, 1, "64^3 hollow box, 4x4 window, 1 emitter, sun; 1 bounce, 1 ray/face"),
}


def load_package():
    path = os.path.join(ROOT, "global-illumination_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location("vgi_amd", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vgi_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(vgi, grid, coords, pages, optics, faces, bounces, spp, budget_s):
    """Time the CPU oracle on a bounded sample of the same faces."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc

    diffuse, emission, lights, env = optics
    t0 = time.time()
    scene = orc.Scene()
    scene.load_pages(coords, pages)
    tracer = orc.Tracer(scene)
    tracer.set_materials(diffuse, emission)
    tracer.set_lights(lights)
    tracer.set_transport(bounces=bounces, samples=spp, environment=tuple(env))
    lib = orc.lib()
    lib.orc_tracer_sync(tracer.h)
    build_s = time.time() - t0
    threads = max(1, min(16, os.cpu_count() or 1))
    rng = np.random.default_rng(7)
    order = rng.permutation(len(faces))
    n = 2048
    rays = 0
    elapsed = 0.0
    done = 0
    while elapsed < budget_s and done < len(order):
        chunk = faces[order[done:done + n]]
        t1 = time.time()
        _, _, c = tracer.trace_faces(chunk, counters=True, threads=threads)
        elapsed += time.time() - t1
        rays += int(c[0])
        done += len(chunk)
        n = min(n * 2, 65536)
    return {
        "value": rays / elapsed / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d of %d faces (random), %d rays, %.1f s of tracing; oracle scene build %.1f s" % (
            done, len(faces), rays, elapsed, build_s),
    }


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA
MLP_FLOP_PER_QUERY = 2 * 14528     # 32-64-64-64-64-3: 14,528 multiply-adds per forward


def mlp_roofline(timings):
    """The learned cache's MFMA fraction from the library's GPU timestamps of
    learned frames (medians): inference at 29,056 FLOP per query; a training
    sample's forward, backward-data and weight-gradient GEMMs at 3 x 29,056."""
    med = {k: float(np.median([t[k] for t in timings])) for k in timings[0]}
    # Rates frame by frame, then their median: a schedule that changes the
    # training load part-way keeps each frame's samples with its own time.
    infer_tflops = float(np.median([t["queries"] * MLP_FLOP_PER_QUERY / (t["infer_ms"] / 1000.0) / 1e12
                                    for t in timings]))
    train_tflops = float(np.median([t["train_samples"] * 3 * MLP_FLOP_PER_QUERY / (t["train_mfma_ms"] / 1000.0) / 1e12
                                    for t in timings if t["train_mfma_ms"] > 0.0] or [0.0]))
    return {
        "bound": "mfma", "unit": "TFLOP/s", "peak": MFMA_BF16_PEAK_TFLOPS,
        "inference": {"achieved": round(infer_tflops, 2), "frac": round(infer_tflops / MFMA_BF16_PEAK_TFLOPS, 5),
                      "ms": round(med["infer_ms"], 4), "queries": int(med["queries"])},
        "training": {"achieved": round(train_tflops, 2), "frac": round(train_tflops / MFMA_BF16_PEAK_TFLOPS, 5),
                     "mfma_ms": round(med["train_mfma_ms"], 4), "step_ms": round(med["train_ms"], 4),
                     "samples": int(med["train_samples"]), "steps": int(med["steps"])},
        "note": "algorithmic FLOP (no padding) over the kernels' event-timed duration; inference includes the hash-grid "
                "gathers and sin/cos encoding, training the table's FP32 gradient atomics",
    }


def run_to_target(vgi, g, judged, reference, var, max_frames, target, budget_ms, every=8, schedule=None, mlp=None):
    """Frames on the bound faces (the judged ones bound first) until the
    judged faces' relative RMSE against the reference reaches `target` or the
    GPU time (from `budget_ms`'s origin) runs out; `schedule` maps a frame to
    the learned cache's (targets, steps, rate) from then on; a list `mlp`
    receives the learned cache's kernel timings of every evaluated frame.
    Returns the (frame, cumulative GPU ms, rmse with the reference's variance
    removed, raw rmse) points evaluated."""
    points = []
    gpu_ms = 0.0
    for f in range(1, max_frames + 1):
        if schedule and f in schedule:
            g.mlp_training(*schedule[f])
        g.run_frame()
        g.sync()
        gpu_ms += g.last_frame_ms()
        if f <= 8 or f % every == 0 or f == max_frames or gpu_ms >= budget_ms:
            if mlp is not None:
                mlp.append(g.mlp_timings())
            rec, _ = g.read_bound(judged)
            est = rec[:, 0:3] + rec[:, 16:19]
            points.append((f, gpu_ms, vgi.relative_rmse(est, reference, var), vgi.relative_rmse(est, reference)))
            if points[-1][2] <= target or gpu_ms >= budget_ms:
                break
    return points


def at_budget(points, budget_ms):
    """RMSE of the last evaluated frame within a GPU-time budget."""
    best = None
    for f, ms, r, _ in points:
        if ms <= budget_ms:
            best = r
    return None if best is None else round(best, 4)


def convergence(vgi, coords, pages, optics, faces, bounces, spp, args):
    """GPU time to reach the target face-irradiance RMSE: plain 4-bounce
    paths, the learned cache, and the learned cache with the network the
    workload states, judged on every JUDGE_STRIDE-th face against an FP64
    long-horizon reference."""
    diffuse, emission, lights, env = optics
    judged_idx = np.arange(0, len(faces), JUDGE_STRIDE)
    rest = np.setdiff1d(np.arange(len(faces)), judged_idx)
    ordered = np.concatenate([faces[judged_idx], faces[rest]])
    judged = len(judged_idx)

    def tracer(samples=spp, recency=1.0):
        g = vgi.GpuTracer(0)
        g.enable_resident()
        g.upload_pages(coords, pages)
        g.set_materials(diffuse, emission)
        g.set_lights(lights)
        g.set_transport(bounces=bounces, samples=samples, environment=tuple(env), plain=True, history_recency=recency)
        return g

    t0 = time.time()
    g = tracer(REFERENCE_SPP)
    g.bind_faces(faces[judged_idx])
    sums = g.converged(args.reference_launches)
    g.run_frame()
    rec, _ = g.read_bound()
    reference, var = vgi.converged_irradiance(sums, rec[:, 16:19])
    g.close()
    noise = float(np.sqrt(var.mean()) / reference.mean())
    samples = args.reference_launches * REFERENCE_SPP
    print("reference: %d faces x %d samples, noise %.4f, %.1f s" % (judged, samples, noise, time.time() - t0), file=sys.stderr,
          flush=True)
    out = {
        "target_rmse": args.target_rmse,
        "judged_faces": judged,
        "reference": "FP64 sums of %d plain launches x %d rays per judged face (every %dth face; relative noise %.4f, its variance "
                     "subtracted from rmse; rmse_raw keeps it)" % (args.reference_launches, REFERENCE_SPP, JUDGE_STRIDE, noise),
        "reference_noise": round(noise, 5),
        "gpu_ms_budget": args.convergence_budget_ms,
    }
    legs = {}
    g = tracer()
    g.bind_faces(ordered)
    pts = run_to_target(vgi, g, judged, reference, var, args.plain_frames, args.target_rmse, args.convergence_budget_ms)
    g.close()
    legs["plain"] = (pts, 0.0, spp, None)
    print("plain: %d frames, %.0f GPU ms, rmse %.4f (raw %.4f)" % pts[-1], file=sys.stderr, flush=True)
    # The learned form at 1 ray per face per frame; then the network as the
    # workload states it (frequency encoding, one-path targets), which
    # measures what the hash grid and the averaged targets are for.
    for name, encoding, train, lr, schedule in (
            ("learned", vgi.MLP_ENCODING_HASH, LEARNED_TRAIN, LEARNED_LR, LEARNED_SCHEDULE),
            ("learned_stated_network", vgi.MLP_ENCODING_FREQUENCY, STATED_TRAIN, STATED_LR, None)):
        g = tracer(LEARNED_SPP, LEARNED_RECENCY)
        g.bind_faces(ordered)
        g.mlp_enable(seed=1, train_per_frame=train[0], steps=train[1], lr=lr)
        g.mlp_encoding(encoding)
        g.mlp_train_samples(train[2])
        t1 = time.perf_counter()
        g.mlp_pretrain(args.pretrain_frames)
        g.sync()
        pretrain_ms = (time.perf_counter() - t1) * 1000.0
        timings = []
        pts = run_to_target(vgi, g, judged, reference, var, args.learned_frames, args.target_rmse,
                            args.convergence_budget_ms - pretrain_ms, schedule=schedule, mlp=timings)
        g.close()
        legs[name] = (pts, pretrain_ms, LEARNED_SPP, (encoding, train, lr, schedule, timings))
        print("%s: %d frames, %.0f GPU ms, rmse %.4f (raw %.4f) (pretrain %.0f ms)" % ((name,) + pts[-1] + (pretrain_ms,)),
              file=sys.stderr, flush=True)
    budgets = [1000.0, 3000.0, 10000.0, 20000.0, 30000.0]
    for name, (pts, pre_ms, rays, net) in legs.items():
        reached = pts[-1][2] <= args.target_rmse
        frames = pts[-1][0]
        entry = {
            "gpu_ms_per_frame": round(pts[-1][1] / frames, 3),
            "frames_to_target": frames if reached else None,
            "gpu_ms_to_target": round(pts[-1][1] + pre_ms, 1) if reached else None,
            "frames_run": frames,
            "gpu_ms_run": round(pts[-1][1] + pre_ms, 1),
            "rmse_last": round(pts[-1][2], 4),
            "rmse_raw_last": round(pts[-1][3], 4),
            "rmse_curve": [[f, round(ms + pre_ms, 1), round(r, 4), round(raw, 4)] for f, ms, r, raw in pts[:: max(1, len(pts) // 12)]],
            "rmse_at_gpu_ms": {str(int(b)): at_budget([(f, ms + pre_ms, r, raw) for f, ms, r, raw in pts], b) for b in budgets},
            "rays_per_face_per_frame": rays,
        }
        if net is not None:
            encoding, train, lr, schedule, timings = net
            entry["mlp"] = mlp_roofline(timings)
            entry["pretrain_frames"] = args.pretrain_frames
            entry["pretrain_ms"] = round(pre_ms, 1)
            entry["history_recency"] = LEARNED_RECENCY
            if schedule:
                entry["schedule"] = {str(k): {"targets": v[0], "steps": v[1], "lr": v[2]} for k, v in schedule.items()}
            entry["training"] = ("%d targets per frame of %d path%s each (3-bounce), %d Adam steps (lr %g%s), BF16 MFMA; encoding: %s" % (
                train[0], train[2], "s" if train[2] > 1 else "", train[1], lr,
                ", hash table %g, lazy" % (10 * lr) if encoding == vgi.MLP_ENCODING_HASH else "",
                "8-level hash grid (2^%d entries per level) + sin/cos(pi u), normal, albedo" % int(np.log2(vgi.MLP_TABLE // 16))
                if encoding == vgi.MLP_ENCODING_HASH else "sin/cos(2^k pi u), k = 0..3, normal, albedo (as stated)"))
        out[name] = entry
    # The target's terms: the learned form at <= target RMSE, its GPU time per
    # frame against plain 4-bounce tracing's, and total GPU time to the target.
    out["plain_over_learned_ms_per_frame"] = round(out["plain"]["gpu_ms_per_frame"] / out["learned"]["gpu_ms_per_frame"], 2)
    out["learned_reaches_target"] = out["learned"]["frames_to_target"] is not None
    if out["learned_reaches_target"] and out["plain"]["gpu_ms_to_target"]:
        out["plain_over_learned_ms_to_target"] = round(out["plain"]["gpu_ms_to_target"] / out["learned"]["gpu_ms_to_target"], 3)
    return out


def edit_stream(vgi, args):
    """The edit stream: a 512^3 noise terrain (~30% solid, 16 emitters, sun),
    each frame 1000 seeded edits (voxel set / clear / sphere removal r 1-5),
    the incremental NAADF update and retention event, the bound faces
    following the exposed surface, then one resident frame of 2 rays per face
    (2 bounces, world cache on, temporal merge)."""
    gen, seed, edge = 4, 0x5EED0004, 512
    t0 = time.time()
    grid = vgi.generate(gen, seed, edge)
    diffuse, emission, lights, env = vgi.optics(gen, seed)
    coords, pages = vgi.pages_of(grid)
    g = vgi.GpuTracer(0)
    g.enable_resident()
    g.upload_pages(coords, pages)
    g.set_materials(diffuse, emission)
    g.set_lights(lights)
    g.set_transport(bounces=2, samples=2, environment=tuple(env), measured_retention=1.0)
    g.bind_exposed()
    print("edit stream: %d faces bound, set up in %.1f s" % (g.bound_count(), time.time() - t0), file=sys.stderr, flush=True)
    edits = [vgi.scene_edits(seed, f + 1, 1000, edge) for f in range(args.edit_frames)]
    for _ in range(2):
        g.run_frame()
    g.sync()
    edit_ms, build_ms, frame_ms, trace_ms, rays, voxels, phases = [], [], [], [], 0, 0, []
    # First half: each edit after the previous frame has finished (the host
    # edit's own time and phases); second half: each edit issued while the
    # previous frame runs, so its host work overlaps the GPU.
    half = args.edit_frames // 2
    t_start = time.perf_counter()
    for f in range(half):
        xyz, values = edits[f]
        t1 = time.perf_counter()
        g.edit_voxels(xyz, values)
        edit_ms.append((time.perf_counter() - t1) * 1000.0)
        tm = g.timings()
        build_ms.append(float(tm[3]))
        phases.append(tm[4:8])
        g.run_frame(vgi.COUNT_RAYS)
        g.sync()
        frame_ms.append(g.last_frame_ms())
        trace_ms.append(g.last_trace_ms())
        rays += int(g.read_counters_bound()[0])
        voxels += len(values)
    sequential = time.perf_counter() - t_start
    t_start = time.perf_counter()
    g.edit_voxels(*edits[half])
    for f in range(half, args.edit_frames):
        voxels += len(edits[f][1])
        g.run_frame(vgi.COUNT_RAYS)
        if f + 1 < args.edit_frames:
            g.edit_voxels(*edits[f + 1])
        g.sync()
        frame_ms.append(g.last_frame_ms())
        trace_ms.append(g.last_trace_ms())
        rays += int(g.read_counters_bound()[0])
    overlapped = time.perf_counter() - t_start
    faces = g.bound_count()
    g.close()
    return {
        "workload": "512^3 noise terrain ~30%% solid, 16 emitters, sun; %d frames x 1000 edits (set/clear/sphere r1-5); "
                    "2 rays/face, 2 bounces, world cache, measured retention" % args.edit_frames,
        "frames": args.edit_frames,
        "voxel_edits_per_frame": round(voxels / args.edit_frames, 1),
        "faces_bound_at_end": int(faces),
        "ms_per_frame_wall": round(overlapped * 1000.0 / (args.edit_frames - half), 3),
        "ms_per_frame_wall_edit_after_frame": round(sequential * 1000.0 / half, 3),
        "edit_ms_host_median": round(float(np.median(edit_ms)), 3),
        "naadf_pack_gpu_ms_median": round(float(np.median(build_ms)), 3),
        "edit_host_ms_median_by_phase": dict(zip(["regions_and_repack", "top_level", "retention_event", "exposure"],
                                                 [round(float(v), 3) for v in np.median(np.array(phases), axis=0)])),
        "frame_gpu_ms_median": round(float(np.median(frame_ms)), 3),
        "trace_gpu_ms_median": round(float(np.median(trace_ms)), 3),
        "mrays_per_s_trace": round(rays / (float(np.sum(trace_ms)) / 1000.0) / 1e6, 1),
    }


def work_selection(vgi, g, faces, args):
    """GPU work selection over the cave's resident faces (one slot per bound
    face): pick batches of `--select-picks` faces ranked by a focus at the
    scene's centre. The first eight batches re-rank every quarter twice (the
    timed frames' host claims expire after four); the next `--select-batches`
    are timed. Algorithmic bytes of the selection passes per batch: the
    classified quarter's 48-byte metadata read and 4-byte key written, every
    slot's key read by the histogram, the count and the emit pass, each
    pick's cleared place (36 bytes) and its metadata, state word, task, slot
    and cleared key, each mirror record's 112 bytes written and its 96 + 48
    bytes read."""
    slots = int(len(faces))
    vox = faces["voxel"]
    centre = tuple(int(v) for v in (vox.min(axis=0) + vox.max(axis=0)) // 2)
    g.set_schedule(focus=centre, transport_revision=1, stable_error=0.12, stable_samples=32, audit_ms=1000,
                   mirror_budget=65536)
    picks = args.select_picks
    timings, picked, mirrors = [], [], []
    for i in range(8 + args.select_batches):
        g.select_frame(picks)
        ms = g.select_timings()
        if i >= 8:
            sel = g.selection()
            timings.append(ms)
            picked.append(int(min(sel[1036], picks)))
            mirrors.append(int(min(sel[1037], 65536)))
    timings = np.array(timings)
    select_ms = float(np.median(timings[:, 0]))
    quarter = (slots + 3) // 4
    n_pick = float(np.mean(picked))
    n_mirror = float(np.mean(mirrors))
    algorithmic = (quarter * (48 + 4) + slots * 4 * 3 + picks * 36 + n_pick * (48 + 4 + 32 + 4 + 4)
                   + n_mirror * (112 + 96 + 48))
    achieved = algorithmic / (select_ms / 1000.0) / 1e9
    print("work selection: %d slots, %d picks, passes %.3f ms, trace %.3f ms, accumulate %.3f ms" % (
        slots, picks, select_ms, float(np.median(timings[:, 1])), float(np.median(timings[:, 2]))), file=sys.stderr, flush=True)
    return {
        "workload": "512^3 worm cave, every bound face resident; focus at the scene centre",
        "slots": slots,
        "picks_budget": picks,
        "picks_per_batch": round(n_pick, 1),
        "mirror_records_per_batch": round(n_mirror, 1),
        "batches_timed": args.select_batches,
        "selection_passes_ms": round(select_ms, 4),
        "pick_trace_ms": round(float(np.median(timings[:, 1])), 3),
        "pick_accumulate_ms": round(float(np.median(timings[:, 2])), 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "algorithmic_bytes": int(algorithmic)},
    }


def sun_corner_words(vgi, g, faces, lights, args):
    """Sun corner words of the cave's bound resident faces: the first call
    traces every face turned towards the sun; then the sun moves by one
    degree of azimuth per call, `--sun-calls` calls re-tracing every face and
    as many re-tracing a quarter (the reference's refresh every fourth frame).
    Each lit face traces four corner rays (visible-geometry walks to the
    scene's end or a hit)."""
    base = np.array(lights, np.float32).copy()
    d = base[0, 0:3].astype(np.float64)
    elevation = math.asin(d[1] / np.linalg.norm(d))
    azimuth = math.atan2(d[2], d[0])
    normals = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)], np.float64)
    out = {}
    g.sun_corners(0, 1)
    for refresh in (1, 4):
        gpu_ms, host_ms, flips, rays = [], [], [], []
        for k in range(args.sun_calls):
            a = azimuth + math.radians(1.0 + k + (args.sun_calls if refresh == 4 else 0))
            light = base.copy()
            light[0, 0:3] = (math.cos(elevation) * math.cos(a), math.sin(elevation), math.cos(elevation) * math.sin(a))
            g.set_lights(light)
            g.sun_corners(k + 1, refresh)
            ms = g.sun_timings()
            gpu_ms.append(float(ms[0]))
            host_ms.append(float(ms[1]))
            flips.append(int(g.sun_flips()[0]))
            lit = (normals[faces["face"]] @ light[0, 0:3].astype(np.float64)) > 0
            due = ((k + 1 + np.arange(len(faces))) % refresh == 0) if refresh > 1 else np.ones(len(faces), bool)
            rays.append(4 * int((lit & due).sum()))
        med = float(np.median(gpu_ms))
        out["refresh_%d" % refresh] = {
            "gpu_ms_median": round(med, 3),
            "host_ms_median": round(float(np.median(host_ms)), 3),
            "corner_rays_per_call": int(np.median(rays)),
            "mrays_per_s": round(float(np.median(rays)) / (med / 1000.0) / 1e6, 1),
            "flips_per_call_median": int(np.median(flips)),
        }
    g.set_lights(base)
    print("sun corners: %s" % out, file=sys.stderr, flush=True)
    return dict({"workload": "512^3 worm cave, every bound face resident; the sun's azimuth moves 1 degree per call",
                 "faces": int(len(faces)), "calls_timed": args.sun_calls}, **out)


def tunnel_camera(grid):
    """An eye inside a worm tunnel near the scene centre (an air voxel with air
    all around it within 2 voxels) and a target along its longest free line
    of sight among the 26 lattice directions."""
    from scipy import ndimage
    c = grid.shape[0] // 2
    lo = c - 64
    air = grid[lo:lo + 128, lo:lo + 128, lo:lo + 128] == 256
    inner = ndimage.binary_erosion(air, structure=np.ones((5, 5, 5), bool))
    zyx = np.argwhere(inner)
    if len(zyx) == 0:
        return None
    pick = zyx[np.argmin(((zyx - 64) ** 2).sum(axis=1))] + lo
    eye = np.array([pick[2], pick[1], pick[0]], np.float64) + 0.5
    best, target = -1, None
    for d in [np.array(v, np.float64) for v in np.ndindex(3, 3, 3)]:
        d = d - 1.0
        if not d.any():
            continue
        d /= np.linalg.norm(d)
        free = 0
        while free < 256:
            p = np.floor(eye + d * (free + 1)).astype(int)
            if (p < 0).any() or (p >= grid.shape[0]).any() or grid[p[2], p[1], p[0]] != 256:
                break
            free += 1
        # A line of sight near the vertical has no defined camera roll.
        if free > best and abs(d[1]) < 0.9:
            best, target = free, eye + d * 16.0
    return eye, target


def view_frames(vgi, g, grid, args):
    """View frames inside the cave on the bound resident tracer: a 640x360
    depth/normal prepass from a camera in a worm tunnel (CPU ray cast of the
    voxel grid), `--view-frames` frames of discovery, face hash, reprojection,
    world-cache exchange, 4 paths per visible face, presentation
    (reconstruction and publication) and the far per-pixel tier at the
    reference's default settings (faces under 8 pixels across, 1-pixel cells
    at this height, one path per cell every fourth frame after four, 128
    steps per ray); the camera moves 0.1 voxel per frame along its line of
    sight."""
    found = tunnel_camera(grid)
    if found is None:
        return None
    eye, target = found
    w, h = 640, 360
    fov_y = np.radians(60.0)
    d = (target - eye) / np.linalg.norm(target - eye)
    g.view_far(vgi.far_settings())
    g.view_display()
    frame_ms, trace_ms, far_ms, owners, cells, traced, displayed = [], [], [], [], [], [], []
    for k in range(args.view_frames + 1):
        e = eye + d * 0.1 * k
        wfc, depth, normals, hits = vgi.scene_prepass(grid, e, e + d * 16.0, w, h, fov_y)
        cam = vgi.view_camera(wfc, e, w, h, samples=4, forward=d, pixels_per_unit=h * 0.5 / np.tan(fov_y * 0.5), frame=k)
        g.view_frame(cam, depth, normals)
        g.sync()
        if k == 0:
            continue  # first frame: nothing to reproject yet
        ms = g.view_timings()
        frame_ms.append(float(ms[0]))
        trace_ms.append(float(ms[1]))
        far_ms.append(float(ms[2]))
        owners.append(g.view_owners(0)[1])
        counters = g.view_counters()
        cells.append(int(counters[0]))
        traced.append(int(counters[1]))
        displayed.append(int(counters[11]))
    g.view_far(None)
    med = float(np.median(frame_ms))
    faces = int(np.median(owners))
    out = {"workload": "512^3 worm cave, 640x360 prepass from inside a tunnel, every bound face resident; 4 paths per visible face; "
                       "far tier at its defaults (8 px, 1-px cells, rate 4, 128 steps)",
           "pixels": w * h, "pixels_hit": int(hits), "visible_faces": faces, "far_cells": int(np.median(cells)),
           "far_cells_traced": int(np.median(traced)), "faces_displayed": int(np.median(displayed)), "frames_timed": args.view_frames,
           "frame_gpu_ms_median": round(med, 3), "trace_gpu_ms_median": round(float(np.median(trace_ms)), 3),
           "far_gpu_ms_median": round(float(np.median(far_ms)), 3),
           "faces_per_s": round(faces / (med / 1000.0), 1)}
    print("view frames: %s" % out, file=sys.stderr, flush=True)
    return out


def batched_scenes(vgi, args):
    """32 seeded noise scenes of log-uniform edges in [32, 512] (multiples of
    8), 2 rays per face, 2 bounces, traced in one persistent launch per frame
    inside a HIP graph with the autotuned workgroup / lanes / tile; compared
    with one launch per scene."""
    t0 = time.time()
    edges = vgi.batch_edges(32, 0x5EED0003)
    tracers, faces = [], []
    for k, edge in enumerate(edges):
        seed = (0x5EED0003 + 0x10000 * k) & 0xFFFFFFFF
        grid = vgi.padded(vgi.generate(3, seed, edge))
        diffuse, emission, lights, env = vgi.optics(3, seed)
        coords, pages = vgi.pages_of(grid)
        g = vgi.GpuTracer(0)
        g.upload_pages(coords, pages)
        g.set_materials(diffuse, emission)
        g.set_lights(lights)
        g.set_transport(bounces=2, samples=2, environment=tuple(env))
        f = vgi.exposed_faces(grid)
        g.bind_faces(f)
        tracers.append(g)
        faces.append(len(f))
        del grid, coords, pages
    setup_s = time.time() - t0
    print("batched scenes: %d scenes, %d faces, set up in %.1f s" % (len(tracers), sum(faces), setup_s), file=sys.stderr, flush=True)
    # One launch per scene, back to back (the unbatched way).
    single_ms = []
    for g in tracers:
        g.run_frame()
        g.sync()
        g.run_frame()
        g.sync()
        single_ms.append(g.last_frame_ms())
    batch = vgi.GpuBatch(0)
    for g in tracers:
        batch.add(g)
    t1 = time.time()
    tuned = batch.autotune(frames=1)
    tune_s = time.time() - t1
    batch.set_sequence(0)
    batch.run_frame()
    batch.sync()
    ms, rays = [], 0
    for _ in range(args.batch_frames):
        batch.run_frame()
        batch.sync()
        ms.append(batch.last_ms())
        rays += batch.latency()[1]
    latency, _ = batch.latency()
    batch.close()
    for g in tracers:
        g.close()
    total = float(np.sum(ms))
    order = np.argsort(faces)
    return {
        "workload": "32 noise scenes, edges 8*round(exp(U[ln32,ln512])/8) (%d..%d), 2 rays/face, 2 bounces; "
                    "one persistent launch per frame in a HIP graph" % (min(edges), max(edges)),
        "scenes": len(tracers),
        "faces": int(sum(faces)),
        "faces_min_max": [int(min(faces)), int(max(faces))],
        "frames": args.batch_frames,
        "autotuned": tuned,
        "autotune_s": round(tune_s, 1),
        "gpu_ms_per_frame": round(total / len(ms), 3),
        "faces_per_s": round(sum(faces) * len(ms) / (total / 1000.0), 1),
        "mrays_per_s": round(rays / (total / 1000.0) / 1e6, 1),
        "one_launch_per_scene_gpu_ms": round(float(np.sum(single_ms)), 3),
        "latency_ms_by_faces": [[int(faces[i]), round(float(latency[i]), 3)] for i in order],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cave512", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU oracle tracing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, help="MiB copied by a plain copy kernel (PMC calibration)")
    ap.add_argument("--no-convergence", action="store_true", help="skip the time-to-RMSE comparison")
    ap.add_argument("--target-rmse", type=float, default=0.02)
    ap.add_argument("--reference-launches", type=int, default=2560)
    ap.add_argument("--convergence-budget-ms", type=float, default=40000.0, help="GPU ms per convergence leg")
    ap.add_argument("--plain-frames", type=int, default=900)
    ap.add_argument("--learned-frames", type=int, default=4000)
    ap.add_argument("--pretrain-frames", type=int, default=400)
    ap.add_argument("--no-edit-stream", action="store_true", help="skip the edit-stream leg")
    ap.add_argument("--edit-frames", type=int, default=120)
    ap.add_argument("--learned-only", action="store_true", help="time learned frames only (profiling)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes leg")
    ap.add_argument("--batch-frames", type=int, default=20)
    ap.add_argument("--no-select", action="store_true", help="skip the GPU work-selection leg")
    ap.add_argument("--select-picks", type=int, default=262144)
    ap.add_argument("--select-batches", type=int, default=8)
    ap.add_argument("--no-sun", action="store_true", help="skip the sun corner-word leg")
    ap.add_argument("--sun-calls", type=int, default=8)
    ap.add_argument("--no-view", action="store_true", help="skip the view-tier leg")
    ap.add_argument("--view-frames", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs the hot path on a single MI355X "
              "(multi-GPU sharding is a design note, not an implemented path).", file=sys.stderr)
        sys.exit(2)

    vgi = load_package()
    if args.calibrate_copy:
        g = vgi.GpuTracer(0)
        nbytes = args.calibrate_copy << 20
        ms = g.probe_copy(nbytes, 5)
        g.close()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "ms_per_copy": ms, "GBps": 2 * nbytes / ms / 1e6}))
        return

    gen, seed, edge, bounces, spp, desc = WORKLOADS[args.workload]
    t0 = time.time()
    grid = vgi.generate(gen, seed, edge)
    optics = vgi.optics(gen, seed)
    faces = vgi.exposed_faces(grid, interior_only=(gen == 1))
    coords, pages = vgi.pages_of(grid)
    gen_s = time.time() - t0
    print("scene %s: %d faces, generated in %.1f s" % (args.workload, len(faces), gen_s), file=sys.stderr, flush=True)

    g = vgi.GpuTracer(0)
    g.enable_resident()
    t1 = time.time()
    g.upload_pages(coords, pages)
    build_ms = float(g.timings()[3])
    upload_s = time.time() - t1
    diffuse, emission, lights, env = optics
    g.set_materials(diffuse, emission)
    g.set_lights(lights)
    g.set_transport(bounces=bounces, samples=spp, environment=tuple(env), plain=True)
    g.bind_faces(faces)
    g.emitter_words()  # build the emitter table before the stats line
    stats = g.stats()
    print("uploaded: %d regions, %d cell words, GPU NAADF build+pack %.2f ms, %d emitter faces, upload %.1f s"
          % (stats[0], stats[2], build_ms, stats[5], upload_s), file=sys.stderr, flush=True)

    if args.learned_only:
        # Learned frames alone (for profiling): the cache warmed up on
        # `pretrain_frames` training-only frames, then `steps` timed frames.
        g.set_transport(bounces=bounces, samples=LEARNED_SPP, environment=tuple(optics[3]), plain=True,
                        history_recency=LEARNED_RECENCY)
        g.mlp_enable(seed=1, train_per_frame=LEARNED_TRAIN[0], steps=LEARNED_TRAIN[1], lr=LEARNED_LR)
        g.mlp_train_samples(LEARNED_TRAIN[2])
        g.mlp_pretrain(args.pretrain_frames)
        for _ in range(args.warmup):
            g.run_frame()
            g.sync()
        frame_ms, trace_ms, rays, mlp = [], [], 0, []
        for _ in range(args.steps):
            g.run_frame(vgi.COUNT_RAYS)
            g.sync()
            frame_ms.append(g.last_frame_ms())
            trace_ms.append(g.last_trace_ms())
            rays += int(g.read_counters_bound()[0])
            mlp.append(g.mlp_timings())
        g.close()
        print(json.dumps({"mode": "learned", "workload": desc, "rays_per_face_per_frame": LEARNED_SPP, "steps": args.steps,
                          "pretrain_frames": args.pretrain_frames,
                          "frame_gpu_ms": round(float(np.median(frame_ms)), 3),
                          "trace_infer_combine_gpu_ms": round(float(np.median(trace_ms)), 3),
                          "rays_per_frame": rays // args.steps,
                          "queries_per_frame": int(len(faces) * LEARNED_SPP),
                          "mlp": mlp_roofline(mlp)}))
        return

    # The library owns its HIP stream; vgi_sync drains it on both sides of the
    # timed region (one process, one GPU: no barrier to take).
    for _ in range(args.warmup):
        g.run_frame(vgi.COUNT_RAYS)
        g.sync()
    # One untimed frame with every counter: the gather mix per ray that prices
    # the roofline's algorithmic bytes.
    g.run_frame(vgi.COUNT_GATHERS)
    g.sync()
    gather_counts = g.read_counters_bound().astype(np.float64)
    counters = np.zeros(vgi.BOUND_COUNTERS, np.float64)
    trace_ms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        g.run_frame(vgi.COUNT_RAYS)
        g.sync()
        trace_ms.append(g.last_trace_ms())
        counters += g.read_counters_bound().astype(np.float64)
    g.sync()
    elapsed = time.perf_counter() - t_start
    print("frames: %s ms (trace kernel)" % ["%.2f" % t for t in trace_ms], file=sys.stderr, flush=True)

    rays = counters[0]
    mrays = rays / elapsed / 1e6
    ms_per_step = elapsed * 1000.0 / args.steps
    # Algorithmic bytes of one trace launch: the traversal's dependent
    # gathers (16 B region header, 4 B brick word, 4 B cell word; per-ray mix
    # from the counting frame, times the timed frames' rays) plus the face
    # task (32 B) and result record (96 B).
    gathers_per_ray = (16.0 * gather_counts[2] + 4.0 * gather_counts[3] + 4.0 * gather_counts[4]) / gather_counts[0]
    per_launch = gathers_per_ray * rays / args.steps + 128.0 * len(faces)
    avg_trace_s = float(np.mean(trace_ms)) / 1000.0
    achieved = per_launch / avg_trace_s / 1e9
    line = {
        "metric": "Mrays/s + GPU ms/frame to 2% face-irradiance RMSE, learned cache vs plain",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator)",
        "config": {
            "workload": desc,
            "faces": int(len(faces)),
            "rays_per_frame": int(rays / args.steps),
            "per_ray": {"steps": round(gather_counts[1] / gather_counts[0], 3),
                        "header_loads": round(gather_counts[2] / gather_counts[0], 3),
                        "brick_word_loads": round(gather_counts[3] / gather_counts[0], 3),
                        "cell_word_loads": round(gather_counts[4] / gather_counts[0], 3)},
            "faces_per_s": round(len(faces) * args.steps / elapsed, 1),
            "trace_ms_per_frame": round(float(np.mean(trace_ms)), 3),
            "method": "plain: paths of up to 4 bounces (Russian roulette after the second hit, survival clamp(max throughput, 0.1, 0.95); no cache termination); temporal merge on",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None,
        },
    }
    # HBM bytes per launch of this kernel on this workload, from the committed
    # rocprofv3 PMC passes (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel).
    traffic_file = os.path.join(ROOT, "profiles", "trace_faces_%s_traffic.json" % args.workload)
    if os.path.exists(traffic_file):
        line["roofline"]["traffic"] = round(json.load(open(traffic_file))["traffic_bytes_per_launch"])
    if not args.no_select and args.workload == "cave512":
        line["work_selection"] = work_selection(vgi, g, faces, args)
    if not args.no_sun and args.workload == "cave512":
        line["sun_corner_words"] = sun_corner_words(vgi, g, faces, lights, args)
    if not args.no_view and args.workload == "cave512":
        line["view_frames"] = view_frames(vgi, g, grid, args)
    g.close()
    if not args.no_edit_stream and args.workload == "cave512":
        line["edit_stream"] = edit_stream(vgi, args)
    if not args.no_batch and args.workload == "cave512":
        line["batched_scenes"] = batched_scenes(vgi, args)
    if not args.no_convergence and args.workload == "cave512":
        line["convergence"] = convergence(vgi, coords, pages, optics, faces, bounces, spp, args)
        # The learned cache's BF16 MFMA kernels (inference, training GEMMs)
        # against the dense BF16 peak, from the learned leg's frames.
        line["roofline_mlp"] = line["convergence"]["learned"]["mlp"]
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(vgi, grid, coords, pages, optics, faces, bounces, spp, args.cpu_budget)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
