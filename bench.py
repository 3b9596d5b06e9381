"""Benchmark of the HIP rasterization pipeline on one MI355X.

Main workload (default, ``--config batch``): 256 scenes with square canvases of side
log-U[256, 4096] px and log-U[10, 20 000] closed paths each (scenegen.batch_scenes: the
4096^2 workload's blob generator with radii log-U[4, min(512, side/2)]).  A step rasterizes
every scene from outlines resident in HBM to its RGBA8 framebuffer in HBM:

* batched: all 256 scenes in one pass (flatten, bin, propagate, sort, fill, composite),
  buffers sized by the learned predictor, launch schedule chosen by the autotuner (once,
  before timing), replayed from one HIP graph -> ``value``;
* plain: one scene at a time, stage by stage, exact sizes read back between stages (the
  reference renderer's structure) -> ``plain_paths_per_s``, ``batched_over_plain``.

The ``roofline`` object covers the slower of the two HBM-bound kernels (k_fill_masks,
k_composite) in the timed batched steps: HIP events recorded as nodes of the replayed graph
on the library's stream bracket each kernel; achieved = algorithmic bytes / that time.

Also reported: ``single_4k`` (one 4096x4096 scene of 10 000 paths, scenegen.single_scene_4k,
the same two modes and roofline), ``serving`` (Poisson request streams over the resident
scenes through the continuous-batching request loop: p50/p99 latency and throughput), and
``cpu_baseline`` (the reference's own hybrid-mode tiler, oracle/_ref compiled from its
sources, on the host's cores over a bounded sample of the same scenes; tiling only: the
reference's raster half needs a graphics pipeline CDNA lacks).

``--config single`` runs the 4096^2 scene as the main workload instead; ``--config stress_clips`` the
stress shape (4096^2, 2000 star polygons under 200 single-level clip paths; its cpu_baseline tiles
every path without the clip resolution).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pathfinder-cpp_amd"))

import pfh  # noqa: E402
import predictor  # noqa: E402
import scenegen  # noqa: E402

METRIC = "paths/sec + Mpix/sec composited (batched vs plain), ≤1/255 alpha err"
HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "batch": "256 scenes, square canvases of side log-U[256,4096] px, 10-20000 closed paths each (log-uniform), "
             "4-64 line/quad/cubic segments per path, radii log-U[4,min(512,side/2)] px, alpha U(0.2,1), "
             "nonzero + even-odd",
    "stress_clips": "4096x4096 scene, 2000 star polygons of 100-2000 vertices (windings far above 1, both fill "
                    "rules) under 200 blob clip paths (single level), half of the stars clipped",
    "single": "4096x4096 scene, 10000 closed paths of 4-64 line/quad/cubic segments, radii log-U[4,512] px, "
              "alpha U(0.2,1), nonzero + even-odd",
}


def fill_bytes(st):
    # 12 B per fill read (reference Fill layout), 8 B alpha-tile index + 4 B backdrop/ctrl
    # per alpha tile, 256 B of u8 coverage written per 16x16 alpha tile; over the alpha tiles
    # the fill kernel rasterizes (those draw entries sample) and their fills.
    return 12 * st["fills_masked"] + 8 * st["alpha_masked"] + 4 * st["alpha_masked"] + 256 * st["alpha_masked"]


def composite_bytes(st, pixels, n_paths):
    # 4 B RGBA8 written per pixel, 16 B per draw entry, 256 B of mask per alpha entry,
    # 4 B colour per paint.
    return 4 * pixels + 16 * st["entries"] + 256 * st["alpha_entries"] + 4 * n_paths


def roofline(st, fill_ms, comp_ms, pixels, n_paths):
    fb, cb = fill_bytes(st), composite_bytes(st, pixels, n_paths)
    fill_gbs = fb / (fill_ms * 1e-3) / 1e9
    comp_gbs = cb / (comp_ms * 1e-3) / 1e9
    if comp_ms >= fill_ms:
        kern, nbytes, ms, gbs = "k_composite", cb, comp_ms, comp_gbs
    else:
        kern, nbytes, ms, gbs = "k_fill_masks", fb, fill_ms, fill_gbs
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kern,
            "algorithmic_bytes": int(nbytes), "kernel_ms": round(ms, 4),
            "other": {"k_fill_masks_gbs": round(fill_gbs, 1), "k_fill_masks_ms": round(fill_ms, 4),
                      "k_composite_gbs": round(comp_gbs, 1), "k_composite_ms": round(comp_ms, 4)}}


def measure(ctx, ids, flags, steps, warmup, plain_steps, schedule=None):
    """Plain and batched timings of one workload; batched kernel times from in-graph events.
    The launch schedule is autotuned for the workload, or ``schedule`` when given."""
    for _ in range(max(1, warmup)):
        for sid in ids:
            ctx.rasterize([sid], pfh.PFH_PLAIN)
    stage_ms = np.zeros(8)
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for sid in ids:
            ctx.rasterize([sid], pfh.PFH_PLAIN)
            stage_ms += np.array(ctx.stats()["ms_stage"])
    plain_s = (time.perf_counter() - t0) / plain_steps
    stage_ms /= plain_steps
    # launch schedule for this workload (once, outside the timed region)
    if schedule is None:
        sched, ms0, ms1 = ctx.autotune(ids, flags & pfh.PFH_PREDICT, reps=3)
        tuning = {"schedule": sched, "default_ms": round(ms0, 4), "tuned_ms": round(ms1, 4)}
    else:
        ctx.set_schedule(**schedule)
        tuning = {"schedule": ctx.schedule(), "tuned": False}
    for _ in range(warmup):
        ctx.rasterize(ids, flags)
    fill_ms = comp_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.rasterize(ids, flags)  # returns after the launch has completed on the device
        ms = ctx.stats()["ms_stage"]
        fill_ms += ms[6]
        comp_ms += ms[7]
    batched_s = (time.perf_counter() - t0) / steps
    return plain_s, stage_ms, batched_s, fill_ms / steps, comp_ms / steps, ctx.stats(), tuning


def serve(ctx, ids, qps, n_requests, seed, slots=2, budget=64 << 30, max_batch=32):
    """Poisson stream of ``n_requests`` requests for scenes drawn uniformly from ``ids`` at ``qps``
    through the request loop; returns its stats and the achieved request rate."""
    rng = np.random.Generator(np.random.PCG64(seed))
    order = rng.integers(0, len(ids), n_requests)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, n_requests))
    ctx.server_start(slots=slots, memory_budget=budget, max_batch=max_batch,
                     flags=pfh.PFH_PREDICT | pfh.PFH_NO_GRAPH)
    done = 0
    t0 = time.perf_counter()
    for k in range(n_requests):
        while time.perf_counter() - t0 < arrivals[k]:
            done += len(ctx.poll(wait_ms=0.0))
        ctx.submit(ids[int(order[k])])
    while done < n_requests:
        done += len(ctx.poll(wait_ms=5.0))
    wall = time.perf_counter() - t0
    st = ctx.server_stats()
    ctx.server_stop()
    keep = ("completed", "launches", "retries", "max_concurrent_launches", "mean_batch", "latency_ms_p50",
            "latency_ms_p99", "latency_ms_mean", "latency_ms_max", "peak_reserved_bytes", "device_bytes")
    out = {k: (round(st[k], 3) if isinstance(st[k], float) else st[k]) for k in keep}
    out.update({"offered_rps": round(qps, 1), "achieved_rps": round(n_requests / wall, 1), "slots": slots,
                "max_batch": max_batch})
    return out


def cpu_baseline(scenes, budget_s=15.0, threads=4):
    """The reference's tiler over the workload's scenes, in order, until about budget_s of CPU work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # bench-only checker leg: the reference's tiler built from its sources

    if not oracle.ref_available():
        return None
    oracle.time_ref(scenes[0], threads=threads, repeats=1)  # warm-up
    paths, secs, used = 0, 0.0, 0
    for sc in scenes:
        if secs >= budget_s:
            break
        one = oracle.time_ref(sc, threads=threads, repeats=1)
        reps = max(1, min(50, int(min(budget_s - secs, 2.0) / max(one, 1e-4))))
        secs += oracle.time_ref(sc, threads=threads, repeats=reps)
        paths += reps * sc.n_paths
        used += 1
    return {"value": round(paths / secs, 1), "unit": "paths/s", "cores": threads, "kind": "reference",
            "sample": f"reference hybrid-mode Tiler (core/d3d9/tiler.cpp, compiled from the reference sources) "
                      f"over the first {used} of the {len(scenes)} scenes of this workload, {paths} paths in "
                      f"{secs:.1f} s, {threads} threads (PATHFINDER_THREADS); tiling only",
            "host_nproc": os.cpu_count()}


def copy_calibration(mbytes):
    """One device-to-device copy of a known size through the HIP runtime, so that rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE of the runtime's copy kernel calibrate those counters to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = int(mbytes) << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] copied {n} bytes device to device", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", choices=["batch", "single", "stress_clips"], default="batch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="main workload only (no single_4k / serving)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--copy-mb", type=int, default=0, help="also run one D2D copy of this size (counter calibration)")
    ap.add_argument("--schedule", default=None,
                    help="launch schedule instead of autotuning, e.g. fill_grid=32768,composite_block=64")
    args = ap.parse_args()
    schedule = None
    if args.schedule:
        schedule = {k: int(v) for k, v in (kv.split("=") for kv in args.schedule.split(","))}
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: scenes would shard across GPUs with no exchange (one process "
              f"per GPU, each rasterizing its own share of the scenes); --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    scenes = (scenegen.batch_scenes() if args.config == "batch" else
              [scenegen.single_scene_4k()] if args.config == "single" else [scenegen.stress_clip_scene()])
    n_paths = sum(s.n_paths for s in scenes)
    pixels = sum(s.width * s.height for s in scenes)
    print(f"[bench] {args.config}: {len(scenes)} scene(s), {n_paths} paths, {pixels / 1e6:.1f} Mpix", flush=True)

    ctx = pfh.Context(0)
    flags = pfh.PFH_BATCHED
    weights = predictor.load()
    if weights is not None:
        ctx.set_predictor(weights, safety=1.25)
        flags |= pfh.PFH_PREDICT
    ids = [ctx.upload(s) for s in scenes]
    print("[bench] uploaded", flush=True)

    plain_steps = max(1, args.steps // 4) if args.config == "batch" else args.steps
    plain_s, stage_ms, batched_s, fill_ms, comp_ms, st, tuning = measure(ctx, ids, flags, args.steps, args.warmup,
                                                                         plain_steps, schedule)
    print(f"[bench] plain {plain_s * 1e3:.3f} ms/step, batched {batched_s * 1e3:.3f} ms/step", flush=True)
    names = ["flatten", "bin_count", "bin_emit", "propagate", "entries", "sort", "fill", "composite"]
    out = {
        "metric": METRIC, "value": round(n_paths / batched_s, 1), "unit": "paths/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(batched_s * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": WORKLOADS[args.config], "mode": "batched, graph replay" + (
            ", predictor-sized" if flags & pfh.PFH_PREDICT else ""), "scenes": len(scenes), "paths": n_paths,
                   "pixels": pixels},
        "mpix_per_s": round(pixels / batched_s / 1e6, 1),
        "plain_paths_per_s": round(n_paths / plain_s, 1), "plain_ms_per_step": round(plain_s * 1e3, 4),
        "batched_over_plain": round(plain_s / batched_s, 3),
        "plain_stage_ms": {k: round(v, 4) for k, v in zip(names, stage_ms)},
        "counts": {k: st[k] for k in ("segments", "lines", "tiles", "fills", "alpha_tiles", "entries",
                                      "alpha_entries", "fills_masked", "alpha_masked", "fb_tiles", "retries",
                                      "host_syncs")},
        "roofline": roofline(st, fill_ms, comp_ms, pixels, n_paths),
        "autotune": tuning,
    }

    if not args.no_extra and args.config == "batch":
        one = scenegen.single_scene_4k()
        sid = ctx.upload(one)
        p_s, s_ms, b_s, f_ms, c_ms, st1, tune1 = measure(ctx, [sid], flags, 20, 3, 20, schedule)
        out["single_4k"] = {"workload": WORKLOADS["single"], "paths_per_s": round(one.n_paths / b_s, 1),
                            "ms_per_step": round(b_s * 1e3, 4), "plain_ms_per_step": round(p_s * 1e3, 4),
                            "mpix_per_s": round(one.width * one.height / b_s / 1e6, 1),
                            "plain_stage_ms": {k: round(v, 4) for k, v in zip(names, s_ms)},
                            "roofline": roofline(st1, f_ms, c_ms, one.width * one.height, one.n_paths),
                            "autotune": tune1}
        ctx.release(sid)
        print(f"[bench] single_4k batched {b_s * 1e3:.3f} ms/step", flush=True)
        # stress shape with clip paths: two-phase propagate and the clip combine kernel
        stress = scenegen.stress_clip_scene()
        sid = ctx.upload(stress)
        p_s, s_ms, b_s, f_ms, c_ms, st2, _ = measure(ctx, [sid], flags, 10, 2, 3, ctx.schedule())
        out["stress_clips"] = {"workload": WORKLOADS["stress_clips"], "paths_per_s": round(stress.n_paths / b_s, 1),
                               "ms_per_step": round(b_s * 1e3, 4), "plain_ms_per_step": round(p_s * 1e3, 4),
                               "fill_ms": round(f_ms, 4), "composite_ms": round(c_ms, 4),
                               "counts": {k: st2[k] for k in ("lines", "fills", "alpha_tiles", "entries",
                                                              "alpha_entries", "retries")}}
        ctx.release(sid)
        print(f"[bench] stress_clips batched {b_s * 1e3:.3f} ms/step", flush=True)
        # serving: request rate as a fraction of the batched throughput in scenes/s
        capacity = len(ids) / batched_s
        out["serving"] = {"requests": "Poisson arrivals, each a uniformly drawn resident scene of the workload",
                          "capacity_scenes_per_s": round(capacity, 1),
                          "runs": [serve(ctx, ids, frac * capacity, 512, seed=100 + k)
                                   for k, frac in enumerate((0.5, 1.0))]}
        print("[bench] serving done", flush=True)

    out["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline(scenes, budget_s=args.cpu_budget)
    if args.copy_mb:
        copy_calibration(args.copy_mb)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
