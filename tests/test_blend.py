"""Blend modes (DrawPath::blend_mode) in the oracle, without a GPU.

* The tiling side (destructive modes tile the view box, only src-over and clear occlude) is pinned by
  the reference's own tiler: the blend_* scenes of tests/golden/hybrid_fixtures.npz
  (tests/test_oracle_golden.py).
* The 15 composite ops (shaders/d3d9/tile.frag:444-583) by their special values and identities:
  neutral and absorbing colours, symmetric cases, the HSL conversions on the primaries.
* The FP64 composite of an op layer over an opaque backdrop on pixel-aligned rects: a B(dest, src) +
  (1 - a) dest in covered pixels, alpha 1 over the layer's whole tiles (tile.frag:585-598).
Runs without a GPU; tests/test_gpu_blend.py checks the device against this model."""
import math

import numpy as np
import pytest

import oracle
import scenegen

MULTIPLY, SCREEN, OVERLAY, DARKEN, LIGHTEN, DODGE, BURN, HARD, SOFT, DIFF, EXCL = range(1, 12)
HUE, SAT, COLOR, LUM = 12, 13, 14, 15
RNG = np.random.default_rng(515)
COLS = [RNG.uniform(0.02, 0.98, 3) for _ in range(20)]


def op(d, s, k):
    return oracle.composite_op(d, s, k)


def test_separable_ops_special_values():
    white, black, grey = np.ones(3), np.zeros(3), np.full(3, 0.5)
    for d in COLS:
        np.testing.assert_allclose(op(d, white, MULTIPLY), d, atol=1e-15)
        np.testing.assert_allclose(op(d, black, MULTIPLY), 0.0, atol=1e-15)
        np.testing.assert_allclose(op(d, black, SCREEN), d, atol=1e-15)
        np.testing.assert_allclose(op(d, white, SCREEN), 1.0, atol=1e-15)
        np.testing.assert_allclose(op(d, d, DIFF), 0.0, atol=1e-15)
        np.testing.assert_allclose(op(d, black, EXCL), d, atol=1e-15)
        np.testing.assert_allclose(op(d, white, EXCL), 1.0 - d, atol=1e-15)
        np.testing.assert_allclose(op(d, grey, HARD), d, atol=1e-15)   # hard light at 0.5: the backdrop
        np.testing.assert_allclose(op(d, grey, SOFT), d, atol=1e-15)   # soft light at 0.5: the backdrop
        np.testing.assert_allclose(op(grey, d, OVERLAY), d, atol=1e-15)  # overlay on mid grey: the source
        np.testing.assert_allclose(op(d, black, DODGE), d, atol=1e-15)
        np.testing.assert_allclose(op(d, white, DODGE), 1.0, atol=1e-15)
        np.testing.assert_allclose(op(black, d, DODGE), 0.0, atol=1e-15)
        np.testing.assert_allclose(op(d, white, BURN), d, atol=1e-15)
        np.testing.assert_allclose(op(d, black, BURN), 0.0, atol=1e-15)
        np.testing.assert_allclose(op(white, d, BURN), 1.0, atol=1e-15)
        for s in COLS[:5]:
            dk, lt = op(d, s, DARKEN), op(d, s, LIGHTEN)
            assert (dk <= np.minimum(d, s) + 1e-15).all() and (dk >= np.minimum(d, s) - 1e-15).all()
            assert (lt >= np.maximum(d, s) - 1e-15).all() and (lt <= np.maximum(d, s) + 1e-15).all()
            np.testing.assert_allclose(op(d, s, MULTIPLY), op(s, d, MULTIPLY), atol=1e-15)  # symmetric
            np.testing.assert_allclose(op(d, s, SCREEN) + op(d, s, MULTIPLY), d + s, atol=1e-14)
            # overlay is hard light with the layers swapped (W3C compositing)
            np.testing.assert_allclose(op(d, s, OVERLAY), op(s, d, HARD), atol=1e-15)
            # soft light stays between the backdrop and the light/dark direction of the source
            sl = op(d, s, SOFT)
            assert ((s <= 0.5) <= (sl <= d + 1e-15)).all() and ((s > 0.5) <= (sl >= d - 1e-15)).all()


def _hsl_shader(c):
    """Hue (radians), the shader's saturation c / v (the HSV one, tile.frag:476), lightness: the
    standard definitions, computed by their own formulas."""
    v, mn = max(c), min(c)
    ch = v - mn
    if ch == 0:
        h = 0.0
    elif v == c[0]:
        h = (math.pi / 3) * (((c[1] - c[2]) / ch) % 6)
    elif v == c[1]:
        h = (math.pi / 3) * ((c[2] - c[0]) / ch + 2)
    else:
        h = (math.pi / 3) * ((c[0] - c[1]) / ch + 4)
    return h, (ch / v if v else 0.0), (v + mn) / 2


def _rgb_from_hsl(h, sat, light):
    """HSL -> RGB by sectors (chroma (1 - |2l - 1|) s), a method other than the shader's."""
    c = (1 - abs(2 * light - 1)) * sat
    hp = (h / (math.pi / 3)) % 6
    x = c * (1 - abs(hp % 2 - 1))
    r, g, b = [(c, x, 0), (x, c, 0), (0, c, x), (0, x, c), (x, 0, c), (c, 0, x)][int(hp) % 6]
    m = light - c / 2
    return np.array([r + m, g + m, b + m])


def test_hsl_ops_follow_the_shader():
    """Hue / saturation / color / luminosity take (h, s, l) of the backdrop and the source as the
    shader's compositeRGBToHSL defines them, whose saturation is c / v (HSV's), and convert back
    as HSL: so a colour composited with itself comes back only where c / v equals HSL's saturation
    (pure hues, greys).  Checked against the standard conversions fed that saturation."""
    red, green, blue = np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0])
    for d in COLS:
        for s in COLS[:4]:
            hd, hs = _hsl_shader(d), _hsl_shader(s)
            np.testing.assert_allclose(op(d, s, HUE), _rgb_from_hsl(hs[0], hd[1], hd[2]), atol=1e-12)
            np.testing.assert_allclose(op(d, s, SAT), _rgb_from_hsl(hd[0], hs[1], hd[2]), atol=1e-12)
            np.testing.assert_allclose(op(d, s, COLOR), _rgb_from_hsl(hs[0], hs[1], hd[2]), atol=1e-12)
            np.testing.assert_allclose(op(d, s, LUM), _rgb_from_hsl(hd[0], hd[1], hs[2]), atol=1e-12)
        # not a round trip for a colour of c / v above HSL's saturation
        assert np.abs(op(d, d, SAT) - d).max() > 1e-4 or abs(max(d) - min(d)) < 1e-3
        g = np.full(3, 0.3)
        lum = 0.5 * (d.max() + d.min())
        np.testing.assert_allclose(op(d, g, SAT), np.full(3, lum), atol=1e-12)  # a grey source: no saturation
    for p in (red, green, blue, np.array([1.0, 1.0, 0.0]), np.array([0.0, 1.0, 1.0])):
        for k in (HUE, SAT, COLOR, LUM):  # pure hues round-trip (c / v = 1 = HSL's saturation at l = 1/2)
            np.testing.assert_allclose(op(p, p, k), p, atol=1e-12)
        np.testing.assert_allclose(op(np.full(3, 0.5), p, COLOR), p, atol=1e-12)


def _rects(w, h, rects, colours, blend):
    pts, cst, pcs = [], [0], [0]
    for x0, y0, x1, y1 in rects:
        pts.append(np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], np.float32))
        cst.append(cst[-1] + 4)
        pcs.append(len(cst) - 1)
    n = len(rects)
    sc = scenegen.Scene(w, h, np.concatenate(pts), np.zeros(4 * n, np.uint8), np.array(cst), np.ones(n, np.uint8),
                        np.array(pcs), np.zeros(n, np.uint8), np.array(colours, np.uint32))
    sc.path_blend = np.array(blend, np.uint8)
    return sc


def _rgb(c):
    return np.array([(c >> 8 * i) & 255 for i in range(4)], np.float64) / 255.0


@pytest.mark.parametrize("k", range(12, 27))
def test_op_layer_over_opaque_backdrop(k):
    """An opaque backdrop over the canvas, an op layer (alpha 0.8) on a pixel-aligned rect: covered
    pixels = a B(dest, src) + (1 - a) dest, pixels of the layer's tiles outside it keep their colour
    with alpha 1 (here already 1), pixels outside its tiles are the backdrop; over a transparent
    region the layer's uncovered tile pixels turn opaque black (tile.frag's alpha 1)."""
    bg, src = 0xff3c8cc8, 0xcc40b4dc
    sc = _rects(64, 48, [(0, 0, 40, 48), (4, 4, 36, 28)], [bg, src], [4, k])
    img = oracle.full_model(sc)["image"]
    d, s = _rgb(bg), _rgb(src)
    a = s[3]
    want = a * op(d[:3], s[:3], int(oracle.COMPOSITE_OP[k])) + (1 - a) * d[:3]
    np.testing.assert_allclose(img[10, 10, :3], want, atol=1e-6)
    assert img[10, 10, 3] == 1.0
    np.testing.assert_allclose(img[40, 20], d, atol=1e-6)        # outside the layer's tiles
    np.testing.assert_allclose(img[2, 2], d, atol=1e-6)          # in its tiles, uncovered: the backdrop
    # the layer's tiles reach x = 48 > 40: uncovered pixels over nothing become (0, 0, 0, 1)
    sc2 = _rects(64, 48, [(0, 0, 40, 48), (4, 4, 44, 28)], [bg, src], [4, k])
    img2 = oracle.full_model(sc2)["image"]
    np.testing.assert_allclose(img2[2, 45], [0.0, 0.0, 0.0, 1.0], atol=1e-12)
    np.testing.assert_allclose(img2[2, 50], [0.0, 0.0, 0.0, 0.0], atol=1e-12)


@pytest.mark.parametrize("k", [0, 1, 2, 3, 5, 6, 7, 8, 9, 10, 11])
def test_porter_duff_modes_draw_over(k):
    """The Porter-Duff modes draw as src-over (the tile pass's blend state, d3d9/renderer.cpp:209)."""
    bg, src = 0xff3c8cc8, 0x9940b4dc
    over = oracle.full_model(_rects(64, 48, [(0, 0, 40, 48), (4, 4, 36, 28)], [bg, src], [4, 4]))["image"]
    got = oracle.full_model(_rects(64, 48, [(0, 0, 40, 48), (4, 4, 36, 28)], [bg, src], [4, k]))["image"]
    np.testing.assert_array_equal(got, over)


def test_blend_scene_model_runs():
    sc = scenegen.blend_scene(90, size=256, n=40)
    m = oracle.full_model(sc)
    assert np.isfinite(m["image"]).all() and (m["image"][..., 3] <= 1.0 + 1e-12).all()
    assert math.isclose(float(oracle.COMPOSITE_OP[14]), 1.0)
