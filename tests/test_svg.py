"""Real-world scenes: the reference's SVG assets through its own SVG ingest
(tests/golden/svg_fixtures.npz, written by nanosvg + SvgScene's steps compiled from the reference).

* Every stroke of tiger.svg (78, widths 0.18-3.5 px) and features.svg (12, every cap and join, a
  dash pattern) dashed and stroked by the oracle equals the reference's outline bit for bit: the
  checker of the device stroker holds on real-world input, not only on the synthetic fixtures.
* The oracle's stroking of the source scene (apply_strokes) gives exactly the reference's scene.
* The documents' shape: tiger is 304 draw paths of solid paint (226 fills, 78 strokes) on
  900 x 900; features has linear and radial gradient paints, an even-odd fill and a dashed stroke.
Runs without a GPU; tests/test_gpu_svg.py renders the scenes on the device."""
import numpy as np
import pytest

import oracle
import svg_scenes

KEYS = ("points", "point_flags", "contour_start", "contour_closed")


def same(got, want, what):
    for k, g, r in zip(KEYS, got, want):
        g, r = np.asarray(g), np.asarray(r)
        assert g.shape == r.shape, f"{what}: {k} {g.shape} vs {r.shape}"
        np.testing.assert_array_equal(g.view(np.uint8), r.view(np.uint8), err_msg=f"{what}: {k}")


@pytest.mark.parametrize("name", ["tiger", "features"])
def test_oracle_strokes_match_reference_ingest(name):
    for i, (src, style, pat, out) in enumerate(svg_scenes.strokes(name)):
        o = src if pat is None else oracle.dash(*src, pat, 0.0)
        same(oracle.stroke(*o, *style), out, f"{name} stroke {i}")


@pytest.mark.parametrize("name", ["tiger", "features"])
def test_apply_strokes_gives_the_reference_scene(name):
    a = oracle.apply_strokes(svg_scenes.scene_src(name))
    b = svg_scenes.scene_out(name)
    for k in ("points", "point_flags", "contour_start", "contour_closed", "path_contour_start", "path_fill_rule",
              "path_rgba"):
        np.testing.assert_array_equal(getattr(a, k).view(np.uint8), getattr(b, k).view(np.uint8), err_msg=k)


def test_document_shapes():
    t = svg_scenes.load("tiger")
    pi = t["path_i32"]
    assert tuple(t["size"]) == (900.0, 900.0)
    assert pi.shape[0] == 304 and (pi[:, 0] == 0).sum() == 226 and (pi[:, 0] == 1).sum() == 78
    assert (pi[:, 2] == svg_scenes.PAINT_COLOR).all()
    f = svg_scenes.load("features")
    pf = f["path_i32"]
    assert set(pf[:, 2].tolist()) == {1, 2, 3}
    assert (pf[:, 1] == 1).any() and (pf[:, 7] > 0).any()
    assert set(pf[pf[:, 0] == 1, 4].tolist()) == {0, 1, 2} and set(pf[pf[:, 0] == 1, 5].tolist()) == {0, 1, 2}


def test_apply_strokes_keeps_gradient_paints():
    src = svg_scenes.scene_src("features", solid_only=False)
    out = oracle.apply_strokes(src)
    np.testing.assert_array_equal(out.path_gradient, src.path_gradient)
    np.testing.assert_array_equal(out.path_opaque, src.path_opaque)
    assert out.gradients is src.gradients
