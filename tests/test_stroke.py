"""The oracle's stroke-to-fill (oracle/stroke_oracle.cpp), the checker of the GPU stroker, against
the reference.

* Bit for bit against tests/golden/stroke_fixtures.npz, written by the reference's own compiled
  OutlineStrokeToFill (core/stroke.cpp): caps butt / square / round, joins miter / bevel / round,
  the miter limit, closed contours (closing line skipped when the first point repeats), lines,
  quadratics, cubics with loops, cusps and coincident controls, zero-length and sub-tolerance
  segments, single points, several contours, widths above the curvature radius.
* Closed forms the geometry fixes (independent of the reference): a butt-capped segment strokes
  to the rectangle length x width; square caps add width/2 at each end; round caps add a disc
  (within the cubic arc's 2.7e-4 radial error); a closed square of side s stroked at width w with
  miter joins fills the frame (s + w)^2 - (s - w)^2 = 4 s w.
Runs without a GPU; tests/test_gpu_stroke.py holds the GPU stroker to the same fixtures."""
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "stroke_fixtures.npz")
KEYS = ("points", "point_flags", "contour_start", "contour_closed")
BUTT, SQUARE, ROUND = 0, 1, 2  # LineCap (core/data/data.h:73-82)
MITER = 0                      # LineJoin::Miter


def load_cases():
    z = np.load(FIX, allow_pickle=False)
    out = {}
    for name in [str(n) for n in z["names"]]:
        w, cap, join, ml = [float(v) for v in z[f"{name}/style"]]
        out[name] = ([z[f"{name}/in/{k}"] for k in KEYS], (w, int(cap), int(join), ml),
                     [z[f"{name}/ref/{k}"] for k in KEYS])
    return out


CASES = load_cases()


@pytest.mark.parametrize("name", list(CASES))
def test_stroke_matches_reference_fixture(name):
    inp, style, ref = CASES[name]
    got = oracle.stroke(*inp, *style)
    for k, g, r in zip(KEYS, got, ref):
        assert g.dtype == r.dtype and g.shape == r.shape, f"{name}: {k} {g.shape} vs {r.shape}"
        np.testing.assert_array_equal(g.view(np.uint8), r.view(np.uint8), err_msg=f"{name}: {k}")


def signed_area(points, flags, contour_start):
    """Sum of the contours' shoelace areas of the on-curve/control polygon after flattening every
    curve finely (de Casteljau at 256 steps)."""
    total = 0.0
    for c in range(len(contour_start) - 1):
        p = points[contour_start[c]:contour_start[c + 1]].astype(np.float64)
        f = flags[contour_start[c]:contour_start[c + 1]]
        poly, i = [p[0]], 0
        while i + 1 < len(p):
            if f[i + 1] == 0:
                poly.append(p[i + 1])
                i += 1
            elif f[i + 1] == 1 and i + 2 < len(p) and f[i + 2] == 0:
                t = np.linspace(0, 1, 257)[1:, None]
                poly.extend((1 - t) ** 2 * p[i] + 2 * t * (1 - t) * p[i + 1] + t ** 2 * p[i + 2])
                i += 2
            else:
                t = np.linspace(0, 1, 257)[1:, None]
                poly.extend((1 - t) ** 3 * p[i] + 3 * t * (1 - t) ** 2 * p[i + 1] + 3 * t ** 2 * (1 - t) * p[i + 2]
                            + t ** 3 * p[i + 3])
                i += 3
        q = np.array(poly)
        total += 0.5 * np.sum(q[:, 0] * np.roll(q[:, 1], -1) - np.roll(q[:, 0], -1) * q[:, 1])
    return total


def stroke_area(pts, width, cap, join=0, closed=False):
    o = oracle.stroke(np.asarray(pts, np.float32), np.zeros(len(pts), np.uint8), np.array([0, len(pts)]),
                      np.array([1 if closed else 0], np.uint8), width, cap, join, 10.0)
    return o, abs(signed_area(o[0], o[1], o[2]))


@pytest.mark.parametrize("length,width", [(100.0, 8.0), (37.5, 2.0), (250.0, 19.0)])
def test_segment_stroke_closed_forms(length, width):
    (pts, flg, cst, ccl), a = stroke_area([(10, 20), (10 + length, 20)], width, BUTT)
    r = width / 2
    np.testing.assert_array_equal(pts, np.array([(10, 20 - r), (10 + length, 20 - r), (10 + length, 20 + r),
                                                 (10, 20 + r)], np.float32))
    assert list(cst) == [0, 4] and list(ccl) == [1]
    assert a == pytest.approx(length * width, rel=1e-6)
    _, a = stroke_area([(10, 20), (10 + length, 20)], width, SQUARE)
    assert a == pytest.approx((length + width) * width, rel=1e-6)
    _, a = stroke_area([(10, 20), (10 + length, 20)], width, ROUND)
    assert a == pytest.approx(length * width + np.pi * r * r, rel=1e-3)


@pytest.mark.parametrize("side,width", [(80.0, 10.0), (40.0, 3.0)])
def test_closed_square_frame_area(side, width):
    sq = [(10, 10), (10 + side, 10), (10 + side, 10 + side), (10, 10 + side)]
    (pts, flg, cst, ccl), _ = stroke_area(sq, width, BUTT, MITER, closed=True)
    assert len(cst) == 3 and list(ccl) == [1, 1]  # an outer and an inner contour
    a_out = signed_area(pts, flg, cst[:2])
    a_in = signed_area(pts[cst[1]:], flg[cst[1]:], cst[1:] - cst[1])
    assert np.sign(a_out) == -np.sign(a_in)  # opposite orientations: their nonzero fill is the frame
    assert abs(a_out + a_in) == pytest.approx(4 * side * width, rel=1e-6)
    assert max(abs(a_out), abs(a_in)) == pytest.approx((side + width) ** 2, rel=1e-6)


def test_stroke_rejects_bad_input():
    pts = np.zeros((2, 2), np.float32)
    with pytest.raises(ValueError):
        oracle.stroke(pts, np.zeros(2, np.uint8), np.array([0, 2]), np.zeros(1, np.uint8), 1.0, 7, 0, 10.0)
    with pytest.raises(ValueError):
        oracle.stroke(pts, np.zeros(2, np.uint8), np.array([0, 3]), np.zeros(1, np.uint8), 1.0, 0, 0, 10.0)
