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
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cave512", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU oracle tracing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, help="MiB copied by a plain copy kernel (PMC calibration)")
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
    counters = np.zeros(5, np.float64)
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
            "faces_per_s": round(len(faces) * args.steps / elapsed, 1),
            "trace_ms_per_frame": round(float(np.mean(trace_ms)), 3),
            "method": "plain: every path runs its 4 bounces (no cache termination); temporal merge on",
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
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(vgi, grid, coords, pages, optics, faces, bounces, spp, args.cpu_budget)
    g.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
