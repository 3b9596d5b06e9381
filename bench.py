"""Benchmark of the HIP rasterization pipeline on one MI355X.

Workload (default): one 4096x4096 scene of 10 000 closed paths (4-64
line/quad/cubic segments each, centres uniform, radii log-uniform in
[4, 512] px, alpha U(0.2, 1), nonzero and even-odd), scenegen.single_scene_4k.
A step rasterizes the whole scene (flatten, bin, propagate, sort, fill,
composite) from outlines resident in HBM to the RGBA8 framebuffer, batched
mode (no host round trip; replayed from a HIP graph).  ``value`` is paths/s
of that mode; the plain stage-by-stage mode (exact sizes read back between
stages, the reference renderer's structure) is timed beside it.

--config batch: 256 scenes with sides log-U[256, 4096] and path counts
log-U[10, 20 000] (scenegen.batch_scenes); plain = one scene at a time,
batched = all scenes in one graph launch.

The roofline object covers the dominant HBM-bound kernel (fill or composite,
whichever takes longer), with achieved bandwidth from algorithmic bytes over
its HIP-event duration on the library's stream.  cpu_baseline is the
reference's own hybrid-mode tiler (oracle/_ref, compiled from the reference
sources) run on this host's cores over the same scene.
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


def fill_bytes(st):
    # 12 B per fill read (reference Fill layout), 8 B alpha-tile index + 4 B backdrop/ctrl
    # per alpha tile, 256 B of u8 coverage written per 16x16 alpha tile.
    return 12 * st["fills"] + 8 * st["alpha_tiles"] + 4 * st["alpha_tiles"] + 256 * st["alpha_tiles"]


def composite_bytes(st, pixels, n_paths):
    # 4 B RGBA8 written per pixel, 16 B per draw entry, 256 B of mask per alpha entry,
    # 4 B colour per paint.
    return 4 * pixels + 16 * st["entries"] + 256 * st["alpha_entries"] + 4 * n_paths


def cpu_baseline(scene, budget_s=15.0, threads=4):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # test/bench-only: the reference's tiler built from its sources

    if not oracle.ref_available():
        return None
    oracle.time_ref(scene, threads=threads, repeats=1)  # warm-up
    one = oracle.time_ref(scene, threads=threads, repeats=1)
    runs = max(1, int(budget_s / max(one, 1e-3)))
    total = oracle.time_ref(scene, threads=threads, repeats=runs)
    return {"value": runs * scene.n_paths / total, "unit": "paths/s", "cores": threads, "kind": "reference",
            "sample": f"reference hybrid-mode Tiler (core/d3d9/tiler.cpp) over the same {scene.n_paths}-path "
                      f"{scene.width}x{scene.height} scene, {runs} runs, {threads} threads "
                      f"(PATHFINDER_THREADS), tiling only: the reference's raster half needs a graphics "
                      f"pipeline CDNA lacks", "host_nproc": os.cpu_count()}


def time_steps(ctx, ids, flags, steps):
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.rasterize(ids, flags)
    return (time.perf_counter() - t0) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", choices=["single", "batch"], default="single")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: scenes shard across GPUs with no exchange (see DESIGN.md); "
              f"--gpus {args.gpus} is not supported in this build", file=sys.stderr)
        sys.exit(2)

    if args.config == "single":
        scenes = [scenegen.single_scene_4k()]
        workload = ("4096x4096 scene, 10000 closed paths of 4-64 line/quad/cubic segments, radii log-U[4,512] px, "
                    "alpha U(0.2,1), nonzero + even-odd")
    else:
        scenes = scenegen.batch_scenes()
        workload = ("256 scenes, sides log-U[256,4096] px, 10-20000 paths each (log-uniform), "
                    "blob generator of the 4096^2 workload")
    n_paths = sum(s.n_paths for s in scenes)
    pixels = sum(s.width * s.height for s in scenes)
    print(f"[bench] {args.config}: {len(scenes)} scene(s), {n_paths} paths, {pixels / 1e6:.1f} Mpix", flush=True)

    ctx = pfh.Context(0)
    batched_flags = pfh.PFH_BATCHED
    weights = predictor.load()
    if weights is not None:
        ctx.set_predictor(weights, safety=1.25)
        batched_flags |= pfh.PFH_PREDICT
    ids = [ctx.upload(s) for s in scenes]
    print("[bench] uploaded", flush=True)

    # plain: per scene, stage by stage, sizes read back between stages
    for _ in range(max(1, args.warmup)):
        for sid in ids:
            ctx.rasterize([sid], pfh.PFH_PLAIN)
    stage_ms = np.zeros(8)
    plain_steps = max(1, args.steps // 2 if args.config == "batch" else args.steps)
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for sid in ids:
            ctx.rasterize([sid], pfh.PFH_PLAIN)
            st_plain = ctx.stats()
            stage_ms += np.array(st_plain["ms_stage"])
    plain_s = (time.perf_counter() - t0) / plain_steps
    stage_ms /= plain_steps
    print(f"[bench] plain {plain_s * 1e3:.3f} ms/step", flush=True)

    # batched: all scenes in one pass, replayed from a HIP graph
    for _ in range(args.warmup):
        ctx.rasterize(ids, batched_flags)
    batched_s = time_steps(ctx, ids, batched_flags, args.steps)
    st = ctx.stats()
    print(f"[bench] batched {batched_s * 1e3:.3f} ms/step", flush=True)

    names = ["count", "scan", "emit", "propagate", "entries", "sort", "fill", "composite"]
    if args.config == "single":
        fb, cb = fill_bytes(st_plain), composite_bytes(st_plain, pixels, n_paths)
        fill_gbs = fb / (stage_ms[6] * 1e-3) / 1e9
        comp_gbs = cb / (stage_ms[7] * 1e-3) / 1e9
        if stage_ms[7] >= stage_ms[6]:
            roof = {"kernel": "k_composite", "bytes": cb, "ms": stage_ms[7], "gbs": comp_gbs}
        else:
            roof = {"kernel": "k_fill_masks", "bytes": fb, "ms": stage_ms[6], "gbs": fill_gbs}
        roofline = {"bound": "hbm", "achieved": round(roof["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(roof["gbs"] / HBM_PEAK_GBS, 4), "traffic": None, "kernel": roof["kernel"],
                    "algorithmic_bytes": int(roof["bytes"]), "kernel_ms": round(roof["ms"], 4),
                    "other": {"k_fill_masks_gbs": round(fill_gbs, 1), "k_composite_gbs": round(comp_gbs, 1)}}
    else:
        roofline = None

    cpu = None
    if not args.no_cpu_baseline and args.config == "single":
        cpu = cpu_baseline(scenes[0], budget_s=args.cpu_budget)

    out = {
        "metric": METRIC, "value": round(n_paths / batched_s, 1), "unit": "paths/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(batched_s * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "mode": "batched, graph replay" + (
            ", predictor-sized" if batched_flags & pfh.PFH_PREDICT else ""), "scenes": len(scenes), "paths": n_paths,
                   "pixels": pixels},
        "mpix_per_s": round(pixels / batched_s / 1e6, 1),
        "plain_paths_per_s": round(n_paths / plain_s, 1), "plain_ms_per_step": round(plain_s * 1e3, 4),
        "batched_over_plain": round(plain_s / batched_s, 3),
        "plain_stage_ms": {k: round(v, 4) for k, v in zip(names, stage_ms)},
        "counts": {k: st[k] for k in ("segments", "lines", "tiles", "fills", "alpha_tiles", "entries",
                                      "alpha_entries", "fb_tiles", "retries", "host_syncs")},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
