"""Caller-owned targets and blending over the destination (Canvas::set_dst_texture,
core/canvas.cpp:538-540; Renderer::draw(..., clear_dst = false), core/renderer.h:83), through the
C ABI, against the oracle's FP64 composite:

* a target in caller device memory with a pitch wider than the row (its padding bytes untouched);
* PFH_LOAD_DST: the scene blended over the target's previous contents, which show through the
  transmittance the layers leave: expected = C + (1 - A) * B with (C, A) the FP64 premultiplied
  composite over transparent and B the background, within 1 LSB;
* both in plain and batched (graph-replayed) modes.
"""
import numpy as np
import pytest

import oracle
import pfh
import scenegen
from hip_mem import DeviceBuffer

pytestmark = pytest.mark.gpu


def background(h, w, seed):
    """A premultiplied RGBA8 background (colour <= alpha)."""
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 256, (h, w, 1))
    rgb = (rng.random((h, w, 3)) * (a + 1)).astype(np.int64).clip(0, a)
    return np.concatenate([rgb, a], axis=2).astype(np.uint8)


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED, pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH])
def test_caller_target_with_pitch_and_load_dst(flags):
    sc = scenegen.blob_scene(701, 600, 344, 250, r_max=160.0, alpha_min=0.2)
    model = oracle.full_model(sc, zmode=1)
    img = np.clip(model["image"], 0.0, 1.0)  # premultiplied FP64 over transparent
    h, w = sc.height, sc.width
    pitch = w * 4 + 192
    ctx = pfh.Context(0)
    buf = DeviceBuffer(h * pitch)
    buf.fill(0xAB)
    sid = ctx.upload(sc)
    ctx.set_target(sid, buf.ptr, pitch)
    assert ctx.target(sid) == (buf.ptr, pitch)
    print(f"[targets] flags={flags} clear", flush=True)
    # cleared target (clear_dst = true)
    for _ in range(2):  # the second batched call replays its graph
        ctx.rasterize([sid], flags)
    got = ctx.readback(sid)
    want = np.rint(img * 255.0).astype(np.int64)
    assert np.abs(got.astype(np.int64) - want).max() <= 1
    host = buf.download().reshape(h, pitch)
    np.testing.assert_array_equal(host[:, :w * 4].reshape(h, w, 4), got)
    assert (host[:, w * 4:] == 0xAB).all(), "bytes past the row were written"
    # blend over a background
    bg = background(h, w, 702)
    rows = np.full((h, pitch), 0xCD, np.uint8)
    rows[:, :w * 4] = bg.reshape(h, w * 4)
    for rep in range(2):
        print(f"[targets] flags={flags} load_dst rep={rep}", flush=True)
        buf.upload(rows)
        ctx.rasterize([sid], flags | pfh.PFH_LOAD_DST)
        got = ctx.readback(sid).astype(np.float64)
        want = img * 255.0 + (1.0 - img[..., 3:4]) * bg.astype(np.float64)
        err = np.abs(got - np.rint(want)).max()
        assert err <= 1, f"load_dst rgba8 error {err}"
        host = buf.download().reshape(h, pitch)
        assert (host[:, w * 4:] == 0xCD).all(), "bytes past the row were written"
    # back to the context-owned image
    ctx.set_target(sid, None, 0)
    ctx.rasterize([sid], flags)
    assert np.abs(ctx.readback(sid).astype(np.int64) - np.rint(img * 255.0).astype(np.int64)).max() <= 1
    ctx.close()
    buf.free()


def test_rejects_bad_fill_rule_and_pitch():
    ctx = pfh.Context(0)
    sc = scenegen.cubic_scene(703, size=256, n_paths=4)
    sc.path_fill_rule = sc.path_fill_rule.copy()
    sc.path_fill_rule[2] = 2
    with pytest.raises(pfh.PfhError, match="path_fill_rule"):
        ctx.upload(sc)
    sc.path_fill_rule[2] = 1
    sid = ctx.upload(sc)
    buf = DeviceBuffer(256 * 1024)
    with pytest.raises(pfh.PfhError, match="pitch"):
        ctx.set_target(sid, buf.ptr, 256 * 4 - 4)
    with pytest.raises(pfh.PfhError, match="pitch"):
        ctx.set_target(sid, buf.ptr, 256 * 4 + 2)
    ctx.close()
    buf.free()
