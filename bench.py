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

The ``roofline`` object covers the slower of k_fill_masks and k_composite in the timed batched
steps: HIP events recorded as nodes of the replayed graph on the library's stream bracket each
kernel; achieved = algorithmic bytes / that time.  For the batch workload it also carries the
kernel's HBM bytes per launch (``traffic``) and the share of the chip's VALU issue rate it uses
(``valu_issue_share``), both from the rocprofv3 counter runs summarised in profiles/.

Also reported: ``single_4k`` (one 4096x4096 scene of 10 000 paths, scenegen.single_scene_4k,
the same two modes and roofline), ``stress`` (scenegen.stress_scene: 4096x4096, 2 000 star
polygons, 500 stroked cubic polylines, 200 clip paths nesting up to 4 deep), ``serving``
(Poisson request streams over the resident scenes through the continuous-batching request loop:
p50/p99 latency and throughput), and ``cpu_baseline`` (the reference's hybrid path on the host's
cores over a stratified sample of the same scenes: its own Tiler, oracle/_ref compiled from its
sources, and a restatement of its raster half; see cpu_baseline()).

``--config single`` runs the 4096^2 scene as the main workload instead; ``--config stress`` the
stress scene (its cpu_baseline is null: the CPU raster restatement does not resolve clip paths).
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
    "stress": "4096x4096 scene, 2000 star polygons of 100-2000 vertices (windings up to 50, both fill rules), "
              "500 open polylines of 4-16 cubics stroked at width U(1,20) (caps and joins of every kind), 200 "
              "clip paths (circles and rects) in nesting chains of depth 1-4, half of the draw paths clipped",
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


PROFILES = os.path.join(ROOT, "profiles")


def profiled(file, key, kernel):
    """A per-launch figure of `kernel` on the 256-scene batch from a committed rocprofv3 summary
    (its `key: kernel=value ...` line); None without it."""
    try:
        for line in open(os.path.join(PROFILES, file)):
            if line.startswith(key + ":"):
                vals = dict(kv.split("=") for kv in line.split(":", 1)[1].split())
                return float(vals[kernel]) if kernel in vals else None
    except OSError:
        pass
    return None


def roofline(st, fill_ms, comp_ms, pixels, n_paths, traffic_known=False):
    fb, cb = fill_bytes(st), composite_bytes(st, pixels, n_paths)
    fill_gbs = fb / (fill_ms * 1e-3) / 1e9
    comp_gbs = cb / (comp_ms * 1e-3) / 1e9
    if comp_ms >= fill_ms:
        kern, nbytes, ms, gbs = "k_composite", cb, comp_ms, comp_gbs
    else:
        kern, nbytes, ms, gbs = "k_fill_masks", fb, fill_ms, fill_gbs
    # HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated) and the share of the chip's VALU
    # issue rate the kernel uses, both from the batch workload's counter runs
    traffic = profiled("batch256_hbm_traffic.md", "traffic_bytes_per_launch", kern) if traffic_known else None
    valu = profiled("batch256_sq_counters.md", "valu_issue_share", kern) if traffic_known else None
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": int(traffic) if traffic is not None else None,
            "valu_issue_share": valu,
            "counters": "profiles/batch256_hbm_traffic.md, profiles/batch256_sq_counters.md" if traffic_known else None,
            "kernel": kern,
            "algorithmic_bytes": int(nbytes), "kernel_ms": round(ms, 4),
            "other": {"k_fill_masks_gbs": round(fill_gbs, 1), "k_fill_masks_ms": round(fill_ms, 4),
                      "k_composite_gbs": round(comp_gbs, 1), "k_composite_ms": round(comp_ms, 4)}}


def measure(ctx, ids, flags, steps, warmup, plain_steps, schedule=None):
    """Plain and batched timings of one workload; batched kernel times from in-graph events.
    The launch schedule is autotuned for the workload, or ``schedule`` when given.  With
    plain_steps = 0 the plain mode is not run (plain_s is None)."""
    stage_ms = np.zeros(8)
    plain_s = None
    if plain_steps:
        for _ in range(max(1, warmup)):
            for sid in ids:
                ctx.rasterize([sid], pfh.PFH_PLAIN)
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


def serve(ctx, ids, qps, n_requests, seed, slots=3, budget=64 << 30, max_batch=64):
    """Poisson stream of ``n_requests`` requests for scenes drawn uniformly from ``ids`` at ``qps``
    through the request loop; returns its stats and the achieved request rate (requests over the
    wall time from the first arrival to the last completion, the final drain included: a stream
    of several thousand requests keeps that drain a small part of it)."""
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
    out.update({"offered_rps": round(qps, 1), "achieved_rps": round(n_requests / wall, 1),
                "achieved_fraction": round(n_requests / wall / qps, 3), "requests": n_requests, "slots": slots,
                "max_batch": max_batch})
    return out


def cpu_baseline(scenes, reps=5, n_sample=48, threads_ref=4):
    """The reference's hybrid path on the host's cores over a stratified sample of the workload's
    scenes (n_sample scenes at evenly spaced ranks of path count, the largest included), each leg
    the median of `reps` timed runs after one warm-up:
      (i)   tiling: the reference's own Tiler (oracle/_ref, compiled from its sources) on
            threads_ref threads (PATHFINDER_THREADS, core/d3d9/scene_builder.cpp:13);
      (ii)  the same on every host thread of this GPU's share (at most 16);
      (iii) end to end: (ii) plus the raster half on as many threads, a restatement of the hybrid
            fill and per-layer composite (oracle.raster_cpu; the reference runs it as GPU raster
            shaders, which CDNA lacks), into RGBA8.
    `value` is leg (iii) in paths/s; Mpix/s beside it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # bench-only checker leg: the reference's tiler built from its sources

    if not oracle.ref_available():
        return None
    cores = max(1, min(16, os.cpu_count() or 1))
    if len(scenes) == 1 and (scenes[0].n_clip_paths or getattr(scenes[0], "path_stroke", None) is not None):
        return cpu_baseline_stroked(scenes[0], reps, threads_ref, cores)
    order = np.argsort([s.n_paths for s in scenes], kind="stable")
    pick = sorted(set(int(order[int(round(q * (len(order) - 1)))]) for q in np.linspace(0, 1, n_sample)))
    sample = [scenes[k] for k in pick if not scenes[k].n_clip_paths]
    if not sample:
        return None
    paths = sum(sc.n_paths for sc in sample)
    pixels = sum(sc.width * sc.height for sc in sample)

    def median_time(fn):
        fn()
        return float(np.median([fn() for _ in range(reps)]))

    t_i = sum(median_time(lambda sc=sc: oracle.time_ref(sc, threads=threads_ref)) for sc in sample)
    t_ii = sum(median_time(lambda sc=sc: oracle.time_ref(sc, threads=cores)) for sc in sample)
    t_r = sum(median_time(lambda sc=sc: oracle.raster_cpu(sc, threads=cores)[1]) for sc in sample)
    t_iii = t_ii + t_r
    legs = {
        "i_tiling_ref_threads": {"threads": threads_ref, "paths_per_s": round(paths / t_i, 1), "s": round(t_i, 4),
                                 "kind": "reference"},
        "ii_tiling_all_cores": {"threads": cores, "paths_per_s": round(paths / t_ii, 1), "s": round(t_ii, 4),
                                "kind": "reference"},
        "iii_end_to_end_all_cores": {"threads": cores, "paths_per_s": round(paths / t_iii, 1),
                                     "mpix_per_s": round(pixels / t_iii / 1e6, 2), "s": round(t_iii, 4),
                                     "raster_s": round(t_r, 4), "kind": "reference tiling + port raster"},
    }
    return {"value": legs["iii_end_to_end_all_cores"]["paths_per_s"], "unit": "paths/s", "cores": cores,
            "kind": "port", "mpix_per_s": legs["iii_end_to_end_all_cores"]["mpix_per_s"],
            "sample": f"{len(sample)} of the {len(scenes)} scenes (evenly spaced ranks of path count, the "
                      f"largest included; {paths} paths, {pixels / 1e6:.1f} Mpix), median of {reps} runs per scene "
                      f"and leg; value = leg iii (reference Tiler on {cores} threads + the hybrid raster half "
                      f"restated on {cores} threads, to RGBA8)",
            "legs": legs, "host_nproc": os.cpu_count()}


def cpu_baseline_stroked(scene, reps, threads_ref, cores):
    """The stress scene on the host: its strokes expanded (the stroker restated in C++, bit-identical to
    the reference's OutlineStrokeToFill, one thread), then the reference's own Tiler over every path of
    the stroked scene, clip paths included, on threads_ref and on all threads.  The tiler's clip
    resolution step and the raster half are not timed (the harness tiles paths without resolving
    clips, and the CPU raster restatement does not resolve them): value = stroke + tiling on all
    threads, a lower bound on the CPU's time."""
    import oracle

    def median_time(fn):
        fn()
        return float(np.median([fn() for _ in range(reps)]))

    def stroke_once():
        t0 = time.perf_counter()
        oracle.apply_strokes(scene)
        return time.perf_counter() - t0

    t_s = median_time(stroke_once)
    stroked = oracle.apply_strokes(scene)
    t_i = median_time(lambda: oracle.time_ref(stroked, threads=threads_ref))
    t_ii = median_time(lambda: oracle.time_ref(stroked, threads=cores))
    n = scene.n_paths
    return {"value": round(n / (t_s + t_ii), 1), "unit": "paths/s", "cores": cores, "kind": "reference",
            "sample": f"the whole scene ({n} paths, {scene.n_clip_paths} of them clip paths), median of {reps} runs "
                      f"per leg; value = strokes expanded (1 thread) + the reference Tiler on {cores} threads; clip "
                      f"resolution and the raster half not timed",
            "legs": {"stroke_1_thread_s": round(t_s, 4),
                     "i_tiling_ref_threads": {"threads": threads_ref, "paths_per_s": round(n / t_i, 1),
                                              "s": round(t_i, 4), "kind": "reference"},
                     "ii_tiling_all_cores": {"threads": cores, "paths_per_s": round(n / t_ii, 1), "s": round(t_ii, 4),
                                             "kind": "reference"}},
            "host_nproc": os.cpu_count()}


def copy_calibration(mbytes):
    """One device-to-device copy of a known size through the HIP runtime, so that rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE of the runtime's copy kernel calibrate those counters to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime libpfh.so is linked with
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
    ap.add_argument("--config", choices=["batch", "single", "stress"], default="batch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="main workload only (no single_4k / serving)")
    ap.add_argument("--cpu-reps", type=int, default=5, help="timed runs per scene and leg of the CPU baseline")
    ap.add_argument("--copy-mb", type=int, default=0, help="also run one D2D copy of this size (counter calibration)")
    ap.add_argument("--batched-only", action="store_true",
                    help="skip the plain mode (profiling: every dispatch of the run is a batched one)")
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
              [scenegen.single_scene_4k()] if args.config == "single" else [scenegen.stress_scene()])
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

    plain_steps = 0 if args.batched_only else max(1, args.steps // 4) if args.config == "batch" else args.steps
    plain_s, stage_ms, batched_s, fill_ms, comp_ms, st, tuning = measure(ctx, ids, flags, args.steps, args.warmup,
                                                                         plain_steps, schedule)
    print(f"[bench] plain {plain_s * 1e3 if plain_s else float('nan'):.3f} ms/step, "
          f"batched {batched_s * 1e3:.3f} ms/step", flush=True)
    names = ["flatten", "bin_count", "bin_emit", "propagate", "entries", "sort", "fill", "composite"]
    out = {
        "metric": METRIC, "value": round(n_paths / batched_s, 1), "unit": "paths/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(batched_s * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": WORKLOADS[args.config], "mode": "batched, graph replay" + (
            ", predictor-sized" if flags & pfh.PFH_PREDICT else ""), "scenes": len(scenes), "paths": n_paths,
                   "pixels": pixels},
        "mpix_per_s": round(pixels / batched_s / 1e6, 1),
        "plain_paths_per_s": round(n_paths / plain_s, 1) if plain_s else None,
        "plain_ms_per_step": round(plain_s * 1e3, 4) if plain_s else None,
        "batched_over_plain": round(plain_s / batched_s, 3) if plain_s else None,
        "plain_stage_ms": {k: round(v, 4) for k, v in zip(names, stage_ms)} if plain_s else None,
        "counts": {k: st[k] for k in ("segments", "lines", "tiles", "fills", "alpha_tiles", "entries",
                                      "alpha_entries", "fills_masked", "alpha_masked", "fb_tiles", "retries",
                                      "host_syncs")},
        "roofline": roofline(st, fill_ms, comp_ms, pixels, n_paths, traffic_known=args.config == "batch"),
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
        # the stress scene: star polygons, stroked polylines, nested clip paths (its own autotune)
        stress = scenegen.stress_scene()
        sid = ctx.upload(stress)
        p_s, s_ms, b_s, f_ms, c_ms, st2, tune2 = measure(ctx, [sid], flags, 10, 2, 3, schedule)
        out["stress"] = {"workload": WORKLOADS["stress"], "paths_per_s": round(stress.n_paths / b_s, 1),
                         "ms_per_step": round(b_s * 1e3, 4), "plain_ms_per_step": round(p_s * 1e3, 4),
                         "fill_ms": round(f_ms, 4), "composite_ms": round(c_ms, 4), "autotune": tune2,
                         "plain_stage_ms": {k: round(v, 4) for k, v in zip(names, s_ms)},
                         "counts": {k: st2[k] for k in ("segments", "lines", "fills", "alpha_tiles", "entries",
                                                        "alpha_entries", "retries")}}
        ctx.release(sid)
        print(f"[bench] stress batched {b_s * 1e3:.3f} ms/step", flush=True)
        # serving: request rate as a fraction of the batched throughput in scenes/s, with the
        # schedule tuned for these scenes (the single and stress legs tuned their own)
        ctx.set_schedule(**tuning["schedule"])
        capacity = len(ids) / batched_s
        out["serving"] = {"requests": "Poisson arrivals, each a uniformly drawn resident scene of the workload",
                          "capacity_scenes_per_s": round(capacity, 1),
                          "runs": [serve(ctx, ids, frac * capacity, 8192, seed=100 + k)
                                   for k, frac in enumerate((0.5, 1.0))]}
        print("[bench] serving done", flush=True)

    out["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline(scenes, reps=args.cpu_reps)
    if args.copy_mb:
        copy_calibration(args.copy_mb)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
