"""pfh_rasterize_noalloc (the plain pipeline with its work buffers in the library's static
scratch array) against pfh_rasterize(PFH_PLAIN), through the C ABI: identical RGBA8 on two
calls in a row, the second with a smaller scene, then a scene too large for the scratch is
refused with PFH_E_LIMIT and the entry point keeps working."""
import numpy as np
import pytest

import pfh
import scenegen

pytestmark = pytest.mark.gpu


def plain_image(ctx, sid, flags=pfh.PFH_PLAIN):
    ctx.rasterize([sid], flags)
    return ctx.readback(sid)


def test_noalloc_matches_plain_on_two_calls():
    ctx = pfh.Context(0)
    big = scenegen.blob_scene(601, 1024, 1024, 900, r_max=300.0)
    small = scenegen.cubic_scene(602, size=512, n_paths=64)
    ids = [ctx.upload(big), ctx.upload(small)]
    want = [plain_image(ctx, sid) for sid in ids]
    for sid, img in zip(ids, want):  # a larger input first, then a smaller one
        print(f"[noalloc] scene {sid}", flush=True)
        ctx.rasterize_noalloc([sid])
        np.testing.assert_array_equal(ctx.readback(sid), img)
    # both scenes in one call (laid out in the scratch one after the other)
    ctx.rasterize_noalloc(ids)
    for sid, img in zip(ids, want):
        np.testing.assert_array_equal(ctx.readback(sid), img)
    ctx.close()


def test_noalloc_load_dst_matches_plain():
    ctx = pfh.Context(0)
    sc = scenegen.blob_scene(603, 640, 480, 300, r_max=200.0, alpha_min=0.3)
    sid = ctx.upload(sc)
    ctx.rasterize([sid])  # some contents to blend over
    first = ctx.readback(sid)
    want = plain_image(ctx, sid, pfh.PFH_PLAIN | pfh.PFH_LOAD_DST)
    ctx.rasterize([sid])
    np.testing.assert_array_equal(ctx.readback(sid), first)
    ctx.rasterize_noalloc([sid], pfh.PFH_LOAD_DST)
    np.testing.assert_array_equal(ctx.readback(sid), want)
    ctx.close()


def test_noalloc_refuses_a_scene_beyond_the_scratch():
    ctx = pfh.Context(0)
    huge = scenegen.single_scene_4k()  # ~4.5 M tiles: well past 128 MiB of work buffers
    small = scenegen.cubic_scene(604, size=512, n_paths=32)
    hid, sid = ctx.upload(huge), ctx.upload(small)
    want = plain_image(ctx, sid)
    with pytest.raises(pfh.PfhError, match="static scratch"):
        ctx.rasterize_noalloc([hid])
    ctx.rasterize_noalloc([sid])
    np.testing.assert_array_equal(ctx.readback(sid), want)
    with pytest.raises(pfh.PfhError, match="PFH_PLAIN"):
        ctx.rasterize_noalloc([sid], pfh.PFH_BATCHED)
    ctx.close()
