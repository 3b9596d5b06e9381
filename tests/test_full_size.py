"""The benchmarked workloads at their full sizes, through the C ABI.

* The 4096², 10 000-path scene (scenegen.single_scene_4k): plain and batched against the
  oracle (tiles, fills, backdrops, order bit-exact; masks within 1/255; RGBA8 within 1 LSB).
* The 256-scene batch (scenegen.batch_scenes): one batched, predictor-sized, graph-replayed launch
  of all 256 scenes draws every scene bit for bit as the plain per-scene pipeline does; a batched
  launch with dumps matches the oracle on 8 scenes (evenly spaced ranks of path count, the largest
  included).
* Serving: a Poisson stream of requests for those scenes through the request loop, at the batched
  capacity, in graph mode: every completed frame equals the scene's plain frame, 0 LSB.
"""
import hashlib
import time

import numpy as np
import pytest

import oracle
import pfh
import predictor
import scenegen
from test_gpu_parity import COV_TOL, RGBA_TOL, check_scene

pytestmark = pytest.mark.gpu


def digest(img):
    return hashlib.sha1(np.ascontiguousarray(img).tobytes()).hexdigest()


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED])
def test_single_4k_scene_full_size(flags):
    ctx = pfh.Context(0)
    check_scene(ctx, scenegen.single_scene_4k(), flags=flags, name="single 4096^2, 10000 paths")
    ctx.close()


@pytest.fixture(scope="module")
def batch256():
    ctx = pfh.Context(0)
    w = predictor.load()
    assert w is not None, "predictor weights missing"
    ctx.set_predictor(w, safety=1.25)
    scenes = scenegen.batch_scenes()
    ids = [ctx.upload(s) for s in scenes]
    plain = {}
    for sid in ids:
        ctx.rasterize([sid], pfh.PFH_PLAIN)
        plain[sid] = digest(ctx.readback(sid))
    print(f"[full] 256 scenes uploaded and drawn plain", flush=True)
    yield ctx, scenes, ids, plain
    ctx.close()


def test_batch256_batched_equals_plain(batch256):
    ctx, scenes, ids, plain = batch256
    flags = pfh.PFH_BATCHED | pfh.PFH_PREDICT
    for rep in range(3):  # capture, then replays
        ctx.rasterize(ids, flags)
        st = ctx.stats()
        print(f"[full] batched rep {rep}: {st['ms_total']:.3f} ms, retries {st['retries']}", flush=True)
        bad = [sid for sid in ids if digest(ctx.readback(sid)) != plain[sid]]
        assert not bad, f"scenes differing from their plain frames: {bad[:10]}"


def test_batch256_dump_matches_oracle_on_a_sample(batch256):
    ctx, scenes, ids, _ = batch256
    ctx.rasterize(ids, pfh.PFH_BATCHED | pfh.PFH_PREDICT | pfh.PFH_DUMP)
    order = np.argsort([s.n_paths for s in scenes], kind="stable")
    sample = sorted(set(int(order[int(round(q * 255))]) for q in np.linspace(0, 1, 8)))
    assert int(order[-1]) in sample
    for k in sample:
        sc, sid = scenes[k], ids[k]
        print(f"[full] scene {k}: {sc.n_paths} paths, {sc.width}^2", flush=True)
        got = ctx.dump(sid)
        img = ctx.readback(sid)
        model = oracle.full_model(sc, zmode=1)
        for key in oracle.CANONICAL_KEYS:
            np.testing.assert_array_equal(got[key], model["canon"][key], err_msg=f"scene {k}: {key}")
        if got["masks"].size:
            assert np.abs(got["masks"].astype(np.float64) / 255.0 - model["coverage"]).max() <= COV_TOL
        want = np.rint(np.clip(model["image"], 0.0, 1.0) * 255.0).astype(np.int64)
        assert np.abs(img.astype(np.int64) - want).max() <= RGBA_TOL


def test_serving_poisson_stream_of_batch_scenes(batch256):
    ctx, scenes, ids, plain = batch256
    ctx.rasterize(ids, pfh.PFH_BATCHED | pfh.PFH_PREDICT)
    capacity = len(ids) / (ctx.stats()["ms_total"] * 1e-3)  # scenes/s of the batched launch
    rng = np.random.default_rng(4)
    n = 400
    order = rng.integers(0, len(ids), n)
    arrivals = np.cumsum(rng.exponential(1.0 / capacity, n))
    ctx.server_start(slots=2, memory_budget=64 << 30, max_batch=32, flags=pfh.PFH_PREDICT)
    rid_scene, checked = {}, 0
    t0 = time.perf_counter()

    def retire(done):
        nonlocal checked
        for rid in done:
            sid = rid_scene.pop(rid)
            assert digest(ctx.readback(sid)) == plain[sid], f"request {rid}: scene {sid} differs"
            checked += 1

    for k in range(n):
        while time.perf_counter() - t0 < arrivals[k]:
            retire(ctx.poll(wait_ms=0.0))
        rid_scene[ctx.submit(ids[int(order[k])])] = ids[int(order[k])]
    t_end = time.perf_counter() + 120.0
    while rid_scene:
        assert time.perf_counter() < t_end, "requests still pending"
        retire(ctx.poll(wait_ms=5.0))
    st = ctx.server_stats()
    ctx.server_stop()
    print(f"[full] served {checked} requests at {capacity:.0f}/s offered: {st}", flush=True)
    assert checked == n and st["completed"] == n
