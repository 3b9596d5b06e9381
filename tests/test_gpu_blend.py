"""Blend modes on the device (pfh_scene_desc::path_blend) against the oracle: tiles, fills, backdrops
and order bit-exact (destructive modes tile the view box, only src-over and clear occlude), masks
within 1/255, RGBA8 within 2 LSB of the oracle's FP64 composite of the device's own coverage bytes.
Against the FP64 model with exact coverage the colour burn / dodge and HSL ops amplify the masks'
1/255 quantisation (burn divides by the source, dodge by 1 - source: a source channel of 2/255 scales
a difference in the destination 127 times), so there 99.9 % of the pixels are within 2 LSB.

* The golden blend_* fixtures of the reference's own tiler run in tests/test_gpu_parity.py.
* Every composite op as a layer over an opaque backdrop on pixel-aligned rects (the closed forms of
  tests/test_blend.py), and scenes of random blend modes over solid and gradient paints, plain and
  batched.
* Scenes with and without blend modes in one batched launch: each equals its plain run.
"""
import numpy as np
import pytest

import oracle
import pfh
import scenegen
from test_blend import _rects
from test_gpu_parity import check_scene

pytestmark = pytest.mark.gpu
TOL = 2


@pytest.fixture(scope="module")
def ctx():
    c = pfh.Context(0)
    yield c
    c.close()


def check_blend(ctx, sc, flags, name):
    got, img = check_scene(ctx, sc, flags=flags, name=name, rgba_tol=255)  # tiles, fills, order, masks
    model = oracle.full_model(sc, zmode=1)
    cov = got["masks"].astype(np.float64).reshape(model["coverage"].shape) / 255.0
    quant = oracle.composite_fp64(model["canon"], model["alpha"], cov, sc)
    err_q = np.abs(img.astype(np.int64) - np.rint(np.clip(quant, 0.0, 1.0) * 255.0).astype(np.int64)).max()
    err = np.abs(img.astype(np.int64) - np.rint(np.clip(model["image"], 0.0, 1.0) * 255.0).astype(np.int64)).max(-1)
    print(f"[blend] {name}: {err_q} LSB from the composite of the device's masks; against exact coverage "
          f"max {err.max()}, {100.0 * (err <= TOL).mean():.4f} % within {TOL}", flush=True)
    assert err_q <= TOL, f"{name}: rgba8 error {err_q} against the composite of the device's masks"
    assert (err <= TOL).mean() >= 0.999


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH])
def test_every_blend_mode_over_a_backdrop(ctx, flags):
    for k in range(27):
        sc = _rects(64, 48, [(0, 0, 40, 48), (4.5, 4.25, 44, 28.5), (20, 8, 60, 40)], [0xff3c8cc8, 0xcc40b4dc, 0x99d0a020],
                    [4, k, 4])
        check_scene(ctx, sc, flags=flags, name=f"blend {pfh.BLEND_MODES[k]}", rgba_tol=TOL)


@pytest.mark.parametrize("seed,size,n", [(91, 512, 80), (92, 1024, 300)])
@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED])
def test_blend_scene_parity(ctx, seed, size, n, flags):
    check_blend(ctx, scenegen.blend_scene(seed, size=size, n=n), flags=flags, name=f"blend scene {size}")


def test_blend_and_plain_scenes_in_one_batch(ctx):
    scenes = [scenegen.blend_scene(93, size=768, n=150), scenegen.blob_scene(94, 512, 384, 200, r_max=128.0),
              scenegen.blend_scene(95, size=512, n=60, painted=0.0)]
    ids = [ctx.upload(s) for s in scenes]
    alone = []
    for sid in ids:
        ctx.rasterize([sid], pfh.PFH_PLAIN)
        alone.append(ctx.readback(sid))
    for flags in (pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH, pfh.PFH_BATCHED, pfh.PFH_BATCHED):
        ctx.rasterize(ids, flags)
        for sid, img in zip(ids, alone):
            np.testing.assert_array_equal(ctx.readback(sid), img)
    for sid in ids:
        ctx.release(sid)


def test_bad_blend_mode_is_refused(ctx):
    sc = scenegen.blend_scene(96, size=256, n=10)
    sc.path_blend = np.full(sc.n_paths, 27, np.uint8)
    with pytest.raises(pfh.PfhError):
        ctx.upload(sc)
