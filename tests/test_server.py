"""Request loop (continuous batching) through the C ABI.

Each request's framebuffer must be the same as the scene rasterized alone: the
pipeline is deterministic (fixed-point coverage), so the tolerance is 0 LSB.
"""
import time

import numpy as np
import pytest

import pfh
import predictor
import scenegen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served():
    ctx = pfh.Context(0)
    w = predictor.load()
    assert w is not None, "predictor weights missing"
    ctx.set_predictor(w, safety=1.25)
    scenes = scenegen.batch_scenes(n_scenes=18, seed=4242, side_max=1536, paths_max=2500)
    ids = [ctx.upload(s) for s in scenes]
    alone = {}
    for sid in ids:
        ctx.rasterize([sid], pfh.PFH_PLAIN)
        alone[sid] = ctx.readback(sid)
    yield ctx, ids, alone
    ctx.close()


def _drain(ctx, rid_scene, alone, timeout_s=60.0):
    left = dict(rid_scene)
    t_end = time.perf_counter() + timeout_s
    while left:
        assert time.perf_counter() < t_end, f"requests still pending: {sorted(left)}"
        for rid in ctx.poll(wait_ms=50.0):
            sid = left.pop(rid)
            np.testing.assert_array_equal(ctx.readback(sid), alone[sid], err_msg=f"request {rid} scene {sid}")


def test_burst_overlaps_and_matches_alone(served):
    ctx, ids, alone = served
    ctx.server_start(slots=3, memory_budget=16 << 30, max_batch=4)
    try:
        rid_scene = {}
        for sid in ids[:12]:
            rid_scene[ctx.submit(sid)] = sid
        assert ctx.server_stats()["launches"] == 0  # submit only queues
        ctx.poll(wait_ms=0.0, capacity=0)  # one step: all three slots take a batch of 4
        st = ctx.server_stats()
        print("[server] burst after one step", st, flush=True)
        assert st["max_concurrent_launches"] == 3 and st["launches"] == 3
        _drain(ctx, rid_scene, alone)
        st = ctx.server_stats()
        print("[server] burst", st, flush=True)
        assert st["completed"] == 12 and st["queued"] == 0 and st["in_flight"] == 0
        assert st["launches"] >= 3 and st["mean_batch"] > 1.0
    finally:
        ctx.server_stop()


def test_poisson_stream_matches_alone(served):
    ctx, ids, alone = served
    ctx.server_start(slots=2, memory_budget=16 << 30, max_batch=8)
    try:
        rng = np.random.Generator(np.random.PCG64(7))
        order = list(rng.permutation(ids))
        gaps = rng.exponential(1.0 / 400.0, len(order))  # 400 requests/s
        rid_scene, left = {}, {}
        t = time.perf_counter()
        for sid, gap in zip(order, gaps):
            t += gap
            while time.perf_counter() < t:
                for rid in ctx.poll(wait_ms=0.0):
                    s = left.pop(rid)
                    np.testing.assert_array_equal(ctx.readback(s), alone[s], err_msg=f"request {rid}")
            rid = ctx.submit(int(sid))
            rid_scene[rid] = int(sid)
            left[rid] = int(sid)
        _drain(ctx, left, alone)
        st = ctx.server_stats()
        print("[server] poisson", st, flush=True)
        assert st["completed"] == len(order)
        assert 0 < st["latency_ms_p50"] <= st["latency_ms_p99"] <= st["latency_ms_max"]
    finally:
        ctx.server_stop()


def test_memory_budget_serialises_launches(served):
    ctx, ids, alone = served
    # a budget below any one request: each runs alone, one launch at a time
    ctx.server_start(slots=3, memory_budget=1 << 20, max_batch=8)
    try:
        rid_scene = {ctx.submit(sid): sid for sid in ids[:6]}
        _drain(ctx, rid_scene, alone)
        st = ctx.server_stats()
        print("[server] budget", st, flush=True)
        assert st["max_concurrent_launches"] == 1 and st["launches"] == 6 and st["mean_batch"] == 1.0
    finally:
        ctx.server_stop()


def test_same_scene_twice_waits_for_itself(served):
    ctx, ids, alone = served
    ctx.server_start(slots=2, memory_budget=16 << 30, max_batch=8)
    try:
        sid = ids[0]
        rid_scene = {ctx.submit(sid): sid, ctx.submit(sid): sid}
        with pytest.raises(pfh.PfhError):
            ctx.release(sid)  # refused while requested
        _drain(ctx, rid_scene, alone)
        assert ctx.server_stats()["launches"] == 2
    finally:
        ctx.server_stop()


def test_clip_scenes_through_the_loop(served):
    """Scenes with clip paths mixed with clip-free ones in the request loop (predicted capacities,
    eager launches): every result equals the scene rasterized alone."""
    ctx, ids, alone = served
    clip_ids = []
    for k in range(3):
        sid = ctx.upload(scenegen.clip_scene(600 + k, size=640 + 128 * k, n_draw=250, n_clip=6))
        ctx.rasterize([sid], pfh.PFH_PLAIN)
        alone[sid] = ctx.readback(sid)
        clip_ids.append(sid)
    ctx.server_start(slots=2, memory_budget=16 << 30, max_batch=4,
                     flags=pfh.PFH_PREDICT | pfh.PFH_NO_GRAPH)
    try:
        rid_scene = {}
        for sid in [clip_ids[0], ids[0], clip_ids[1], ids[1], clip_ids[2], ids[2]]:
            rid_scene[ctx.submit(sid)] = sid
        _drain(ctx, rid_scene, alone)
        assert ctx.server_stats()["completed"] == 6
    finally:
        ctx.server_stop()
    for sid in clip_ids:
        ctx.release(sid)
