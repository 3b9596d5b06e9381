"""Gradient paints on the device (the composite's per-pixel colour texture, pfh_raster.hip) against
the oracle's FP64 model of the tile shader (oracle/paint_oracle.cpp, pinned to the reference's
ramps, half texture and transform inverse in tests/test_paint.py).

Tiles, fills, backdrops, order and the z-buffer (gradient opacity) bit-exact; masks within 1/255;
RGBA8 within 2 LSB of round(FP64 composite * 255): the device evaluates the colour texture
coordinate, the radial filter and the ramp's linear sample in fp32 (the shader's precision) where
the model samples in FP64, so a pixel on a texel boundary of a steep ramp can land one ramp step
apart -- and at least 99.9 % of the pixels within 1 LSB.
Cases: linear and radial gradients (focal points, start radii, rotated and scaled gradient spaces),
clamp and repeat, 1-5 stops (equal offsets, out of order), opaque and translucent stops, both fill
rules, plain and batched; a batch mixing gradient and solid scenes."""
import numpy as np
import pytest

import oracle
import pfh
import scenegen
from test_gpu_parity import check_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = pfh.Context(0)
    yield c
    c.close()


def within_one_lsb_share(img, scene):
    want = np.rint(np.clip(oracle.full_model(scene, zmode=1)["image"], 0.0, 1.0) * 255.0).astype(np.int64)
    return float((np.abs(img.astype(np.int64) - want) <= 1).mean())


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED])
@pytest.mark.parametrize("seed,size,n", [(21, 1024, 200), (22, 2048, 600)])
def test_gradient_scene_parity(ctx, flags, seed, size, n):
    scene = scenegen.gradient_scene(seed, size, n)
    _, img = check_scene(ctx, scene, flags=flags, name=f"gradients{size}", rgba_tol=2)
    share = within_one_lsb_share(img, scene)
    print(f"[paint] gradients{size} flags={flags}: {share * 100:.4f} % of pixels within 1 LSB", flush=True)
    assert share >= 0.999


def test_gradient_and_solid_scenes_in_one_batch(ctx):
    scenes = [scenegen.gradient_scene(31, 512, 80), scenegen.blob_scene(32, 512, 512, 80, r_max=90.0),
              scenegen.gradient_scene(33, 768, 120, painted=1.0)]
    ids = [ctx.upload(s) for s in scenes]
    ctx.rasterize(ids, pfh.PFH_BATCHED)
    imgs = [ctx.readback(i) for i in ids]
    for i, sid in enumerate(ids):
        ctx.rasterize([sid], pfh.PFH_PLAIN)
        np.testing.assert_array_equal(ctx.readback(sid), imgs[i], err_msg=f"scene {i}")
        want = np.rint(np.clip(oracle.full_model(scenes[i], zmode=1)["image"], 0, 1) * 255).astype(np.int64)
        assert np.abs(imgs[i].astype(np.int64) - want).max() <= 2, f"scene {i}"
    for sid in ids:
        ctx.release(sid)


def test_bad_gradient_is_refused(ctx):
    scene = scenegen.gradient_scene(41, 256, 10, painted=1.0)
    scene.gradients = scene.gradients.copy()
    scene.gradients[0]["kind"] = 7
    with pytest.raises(pfh.PfhError, match="gradient"):
        ctx.upload(scene)
    scene = scenegen.gradient_scene(41, 256, 10, painted=1.0)
    scene.path_gradient = scene.path_gradient.copy()
    scene.path_gradient[0] = len(scene.gradients)
    with pytest.raises(pfh.PfhError, match="path_gradient"):
        ctx.upload(scene)
