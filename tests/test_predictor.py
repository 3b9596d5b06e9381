"""Sizing predictor on the GPU: kernels against the FP64 oracle, exact training
targets against the oracle, training, and what the predictor gains."""
import time

import numpy as np
import pytest

import mlp
import oracle
import pfh
import predictor
import scenegen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = pfh.Context(0)
    yield c
    c.close()


def test_mlp_kernels_match_fp64_oracle(ctx):
    rng = np.random.default_rng(11)
    params = (rng.normal(0, 0.2, pfh.MLP_PARAMS)).astype(np.float32)
    x = rng.normal(0, 1, (3000, pfh.N_FEATURES)).astype(np.float32)
    t = rng.normal(1, 1, (3000, pfh.MLP_OUTPUTS)).astype(np.float32)
    out = ctx.mlp_eval(params, x)
    ref, _ = mlp.forward(params, x)
    np.testing.assert_allclose(out, ref, rtol=1e-4, atol=1e-4)
    grad, loss = ctx.mlp_grad(params, x, t)
    rloss, rgrad = mlp.loss_and_grad(params, x, t)
    assert abs(loss - rloss) <= 1e-4 * abs(rloss)
    scale = np.abs(rgrad).max()
    assert np.abs(grad - rgrad).max() <= 1e-4 * scale, np.abs(grad - rgrad).max() / scale


def test_training_targets_are_exact(ctx):
    sc = scenegen.blob_scene(21, 1024, 1024, 400, r_max=300.0)
    sid = ctx.upload(sc)
    ctx.rasterize([sid], pfh.PFH_PLAIN | pfh.PFH_DUMP)
    got = ctx.path_counts(sid)
    ctx.release(sid)
    canon = oracle.run_oracle(sc)
    np.testing.assert_array_equal(got[:, 0], canon["path_lines"])
    np.testing.assert_array_equal(got[:, 1], np.diff(canon["fill_off"]))
    alpha = np.add.reduceat(canon["tile_alpha"].astype(np.int64), canon["tile_off"][:-1]) \
        if canon["tile_alpha"].size else np.zeros(sc.n_paths)
    alpha = np.where(np.diff(canon["tile_off"]) > 0, alpha, 0)
    np.testing.assert_array_equal(got[:, 2], alpha)


def _scenes(seed, n):
    return [scenegen.blob_scene(seed + k, s, s, p, r_max=min(512.0, s / 2))
            for k, (s, p) in enumerate([(2048, 3000), (1024, 1500), (512, 800), (4096, 4000), (256, 300)][:n])]


@pytest.fixture(scope="module")
def trained(ctx):
    feats, counts = predictor.collect(ctx, _scenes(500, 5) + _scenes(600, 5))
    weights, hist = predictor.train(ctx, feats, counts, epochs=25, batch=1024, lr=2e-3)
    return weights, hist


def test_training_learns_scene_totals(ctx, trained):
    weights, hist = trained
    assert hist[-1] < 0.5 * hist[0], hist
    held = _scenes(900, 4)
    feats, counts = predictor.collect(ctx, held)
    pred = predictor.predict_totals(ctx, weights, feats)
    exact = counts.sum(0)
    rel = np.abs(pred - exact) / exact
    print("[predictor] held-out totals exact", exact, "predicted", np.round(pred), "rel err", np.round(rel, 3))
    assert (rel < 0.35).all(), rel


def test_shipped_predictor_on_held_out_scenes(ctx):
    """The shipped weights (trained by predictor.main on scenes of seed 7) against the exact counts of
    scenes it never saw (the same generators, seed 11: 34 scenes, 156 k paths): the scene totals of
    chords, fills and alpha tiles within 3 % (measured 0.8 / 0.4 / 0.2 %)."""
    w = predictor.load()
    assert w is not None, "predictor weights missing"
    held = predictor.training_scenes(150_000, seed=11)
    feats, counts = predictor.collect(ctx, held)
    pred = predictor.predict_totals(ctx, w, feats)
    exact = counts.sum(0).astype(np.float64)
    rel = np.abs(pred - exact) / exact
    print(f"[predictor] shipped weights on {len(held)} held-out scenes ({feats.shape[0]} paths): exact {exact}, "
          f"predicted {np.round(pred)}, rel err {np.round(rel, 4)}", flush=True)
    assert (rel < 0.03).all(), rel


def test_predictor_gain_over_count_round_trip(ctx, trained):
    """Fresh scenes in a batch: without the predictor the batch needs an extra count
    pass and two host round trips before the launch; with it, one graph launch and
    one read back of the overflow counters, and the same pixels."""
    weights, _ = trained
    scenes = _scenes(1200, 5)
    results = {}
    for use in (False, True):
        c = pfh.Context(0)
        ids = [c.upload(s) for s in scenes]
        flags = pfh.PFH_BATCHED
        if use:
            c.set_predictor(weights, safety=2.0)
            flags |= pfh.PFH_PREDICT
        t0 = time.perf_counter()
        c.rasterize(ids, flags)
        dt = time.perf_counter() - t0
        st = c.stats()
        results[use] = (dt, st, [c.readback(i) for i in ids])
        print(f"[predictor] predictor={use}: first batched launch {dt * 1e3:.2f} ms, host syncs "
              f"{st['host_syncs']}, retries {st['retries']}, capacities fills {st['capacity_fills']} "
              f"(exact {st['fills']})", flush=True)
        c.close()
    assert results[False][1]["host_syncs"] == 3
    assert results[True][1]["host_syncs"] == 1 + results[True][1]["retries"]
    assert results[True][1]["retries"] == 0
    for a, b in zip(results[False][2], results[True][2]):
        np.testing.assert_array_equal(a, b)
