"""Learned buffer sizing for batched rasterization.

Per path, an FP32 MLP 12 -> 64 -> 64 -> 3 predicts log1p of the number of
flattened chords, fills and alpha tiles the flatten and bin stages will emit;
summed over a scene (times a safety factor) these size the device buffers of a
batched launch, so the launch needs no count-and-read-back round trip.  This
replaces the reference's sizing loop, which counts on the GPU, reads the count
back, grows the buffer to the next power of two and redoes the stage
(core/d3d11/renderer.cpp:537-577, 707-724, 834-845).  An under-prediction is
detected on the device and costs a retry, never a wrong pixel.

Training data are (features, exact counts) pairs that the pipeline itself
emits (pfh_scene_features, pfh_path_counts); training runs on the GPU with the
project's own forward/backward/Adam kernels (pfh_mlp_train).

    python predictor.py --paths 1000000 --out predictor_weights.npy
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import pfh  # noqa: E402
import scenegen  # noqa: E402

WEIGHTS_PATH = os.path.join(HERE, "predictor_weights.npy")


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised MLP parameters in the pfh_mlp layout."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for fan_in, fan_out in ((12, 64), (64, 64), (64, pfh.MLP_OUTPUTS)):
        parts.append(rng.normal(0.0, math.sqrt(2.0 / fan_in), (fan_in, fan_out)).reshape(-1))
        parts.append(np.zeros(fan_out))
    p = np.concatenate(parts).astype(np.float32)
    assert p.shape[0] == pfh.MLP_PARAMS
    return p


def collect(ctx: "pfh.Context", scenes, verbose: bool = False):
    """(features [n, 12], exact counts [n, 3]) for every path of ``scenes``."""
    feats, counts = [], []
    for k, sc in enumerate(scenes):
        sid = ctx.upload(sc)
        ctx.rasterize([sid], pfh.PFH_PLAIN | pfh.PFH_DUMP)
        feats.append(ctx.features(sid))
        counts.append(ctx.path_counts(sid))
        ctx.release(sid)
        if verbose and k % 20 == 0:
            print(f"[predictor] collected {k + 1}/{len(scenes)} scenes", flush=True)
    return np.concatenate(feats), np.concatenate(counts)


def training_scenes(n_paths: int, seed: int = 7):
    """Scenes from the benchmark generators (4096^2 blobs and the mixed-size batch
    generator) until about ``n_paths`` paths."""
    rng = np.random.Generator(np.random.PCG64(seed))
    scenes, total, k = [], 0, 0
    while total < n_paths:
        side = int(round(math.exp(rng.uniform(math.log(256), math.log(4096)))))
        count = int(min(20000, max(10, round(math.exp(rng.uniform(math.log(10), math.log(20000)))))))
        if k % 3 == 0:
            side, count = 4096, 10000
        sc = scenegen.blob_scene(seed * 100003 + k, side, side, count, r_max=min(512.0, side / 2.0))
        scenes.append(sc)
        total += sc.n_paths
        k += 1
    return scenes


def normaliser(x: np.ndarray):
    mean = x.mean(0).astype(np.float32)
    std = x.std(0).astype(np.float32)
    return mean, (1.0 / np.maximum(std, 1e-6)).astype(np.float32)


def pack(mean, inv_std, params) -> np.ndarray:
    return np.concatenate([mean, inv_std, params]).astype(np.float32)


def train(ctx, feats, counts, epochs=20, batch=4096, lr=2e-3, seed=1, params=None):
    mean, istd = normaliser(feats)
    x = ((feats - mean) * istd).astype(np.float32)
    y = np.log1p(counts.astype(np.float64)).astype(np.float32)
    p0 = init_params(seed) if params is None else params
    params, hist = ctx.mlp_train(p0, x, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return pack(mean, istd, params), hist


def predict_totals(ctx, weights, feats) -> np.ndarray:
    """Predicted (chords, fills, alpha tiles) summed over the given paths."""
    mean, istd, params = weights[:12], weights[12:24], weights[24:]
    x = ((feats - mean) * istd).astype(np.float32)
    out = ctx.mlp_eval(params, x)
    return np.expm1(np.maximum(out.astype(np.float64), 0.0)).sum(0)


def save(weights, path=WEIGHTS_PATH):
    np.save(path, np.asarray(weights, np.float32), allow_pickle=False)


def load(path=WEIGHTS_PATH):
    if not os.path.exists(path):
        return None
    w = np.load(path, allow_pickle=False)
    return w if w.shape == (pfh.PREDICTOR_FLOATS,) else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--paths", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--out", default=WEIGHTS_PATH)
    args = ap.parse_args()
    t0 = time.time()
    scenes = training_scenes(args.paths)
    print(f"[predictor] {len(scenes)} scenes, {sum(s.n_paths for s in scenes)} paths generated "
          f"in {time.time() - t0:.1f}s", flush=True)
    ctx = pfh.Context(0)
    feats, counts = collect(ctx, scenes, verbose=True)
    print(f"[predictor] collected {feats.shape[0]} samples in {time.time() - t0:.1f}s", flush=True)
    t1 = time.time()
    weights, hist = train(ctx, feats, counts, epochs=args.epochs)
    print(f"[predictor] trained {args.epochs} epochs in {time.time() - t1:.1f}s; loss per epoch:",
          " ".join(f"{v:.4f}" for v in hist), flush=True)
    save(weights, args.out)
    print(f"[predictor] wrote {args.out}", flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
