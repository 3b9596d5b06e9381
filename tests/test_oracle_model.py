"""Closed-form pins of the oracle's FP64 coverage and composite model and of the
backdrop propagation: exact polygon/pixel intersection areas, brute-force
winding numbers, and n-layer src-over."""
import math

import numpy as np
import pytest

import oracle
import scenegen
from golden.gen_golden import make_scene, rect, star


def clip_area(poly, x0, y0, x1, y1):
    """Signed area of polygon ∩ [x0,x1]x[y0,y1] (Sutherland–Hodgman against a convex window)."""
    def clip(pts, inside, inter):
        out = []
        n = len(pts)
        for i in range(n):
            a, b = pts[i], pts[(i + 1) % n]
            ia, ib = inside(a), inside(b)
            if ia:
                out.append(a)
            if ia != ib:
                out.append(inter(a, b))
        return out

    def ix(c):
        return lambda a, b: (c, a[1] + (b[1] - a[1]) * (c - a[0]) / (b[0] - a[0]))

    def iy(c):
        return lambda a, b: (a[0] + (b[0] - a[0]) * (c - a[1]) / (b[1] - a[1]), c)

    pts = [tuple(map(float, p)) for p in poly]
    for inside, inter in ((lambda p: p[0] >= x0, ix(x0)), (lambda p: p[0] <= x1, ix(x1)),
                          (lambda p: p[1] >= y0, iy(y0)), (lambda p: p[1] <= y1, iy(y1))):
        if not pts:
            return 0.0
        pts = clip(pts, inside, inter)
    if len(pts) < 3:
        return 0.0
    s = 0.0
    for i in range(len(pts)):
        a, b = pts[i], pts[(i + 1) % len(pts)]
        s += a[0] * b[1] - b[0] * a[1]
    return 0.5 * s


def exact_coverage(poly, w, h):
    img = np.zeros((h, w))
    xs, ys = [p[0] for p in poly], [p[1] for p in poly]
    for y in range(max(0, int(math.floor(min(ys)))), min(h, int(math.ceil(max(ys))))):
        for x in range(max(0, int(math.floor(min(xs)))), min(w, int(math.ceil(max(xs))))):
            img[y, x] = abs(clip_area(poly, x, y, x + 1, y + 1))
    return img


# Vertices on the 1/256 grid, edge slopes powers of two: every tile-boundary
# crossing is exactly representable, so the 8.8 fills carry the exact polygon.
POLYS = {
    "rect_grid": [(3.5, 4.25), (40.75, 4.25), (40.75, 37.5), (3.5, 37.5)],
    "triangle_slope1": [(5, 5), (45, 45), (5, 45)],
    "triangle_slope2": [(40, 3), (60, 43), (20, 43)],
    "quad_slope_half": [(2, 20), (34, 4), (62, 18), (30, 34)],
    "notch_nonconvex": [(4, 4), (60, 4), (40, 24), (60, 44), (60, 60), (4, 60)],
}


@pytest.mark.parametrize("name", list(POLYS))
@pytest.mark.parametrize("rule", [0, 1])
def test_fp64_coverage_equals_exact_area(name, rule):
    poly = POLYS[name]
    sc = make_scene(64, 64, [([(poly, None, True)], rule, 0xFF000000)])
    model = oracle.full_model(sc)
    got = oracle.path_coverage_image(sc, model, 0)
    want = exact_coverage(poly, 64, 64)
    # The reference clamps tile-local fill coordinates to 4095/256 px
    # (core/d3d9/object_builder.cpp:33-35): a fill ending on a tile's right or
    # bottom edge is pulled in by 1/256 px, which tilts it by at most that much.
    # For axis-aligned edges only the tile's last column/row can feel it.
    diff = np.abs(got - want)
    assert diff.max() <= 1.0 / 256 + 1e-9
    if name == "rect_grid":
        ys, xs = np.nonzero(diff > 1e-9)
        assert ((xs % 16 == 15) | (ys % 16 == 15)).all()


def test_half_pixel_edges_are_half_covered():
    sc = make_scene(64, 64, [([rect(10.5, 10.5, 30.5, 20.5)], 0, 0xFF000000)])
    got = oracle.path_coverage_image(sc, oracle.full_model(sc), 0)
    # (pixels away from a tile's last row/column, which the 4095 clamp touches)
    assert got[12, 10] == pytest.approx(0.5)
    assert got[10, 12] == pytest.approx(0.5)
    assert got[10, 10] == pytest.approx(0.25)
    assert got[12, 12] == pytest.approx(1.0)
    assert got[20, 20] == pytest.approx(0.5)
    assert got[15, 31] == 0.0


def test_coverage_sum_is_polygon_area_for_disjoint_rects():
    # each inside one tile, clear of the right/bottom tile edges the 4095 clamp touches
    rects = [(3.25, 2.5, 13.75, 9.0), (20.0, 20.0, 30.5, 31.25), (40.5, 40.5, 47.0, 47.0), (50.0, 3.0, 63.5, 15.5)]
    sc = make_scene(64, 64, [([rect(*r) for r in rects], 1, 0xFF000000)])
    got = oracle.path_coverage_image(sc, oracle.full_model(sc), 0)
    area = sum((r[2] - r[0]) * (r[3] - r[1]) for r in rects)
    assert got.sum() == pytest.approx(area, rel=1e-12)


def winding(poly, x, y):
    """Winding number of (x, y) by signed crossings of a rightward ray."""
    w = 0
    n = len(poly)
    for i in range(n):
        (x0, y0), (x1, y1) = poly[i], poly[(i + 1) % n]
        if (y0 <= y) != (y1 <= y):
            xc = x0 + (y - y0) * (x1 - x0) / (y1 - y0)
            if xc > x:
                w += 1 if y1 > y0 else -1
    return w


@pytest.mark.parametrize("n,m", [(7, 3), (11, 4), (9, 2), (5, 2)])
def test_solid_tile_backdrop_is_winding_number(n, m):
    pts, _, _ = star(256, 256, 250, n, m)
    poly = [tuple(map(float, p)) for p in np.asarray(pts, np.float32)]
    sc = make_scene(512, 512, [([(pts, None, True)], 0, 0xFF000000)])
    canon = oracle.run_oracle(sc)
    l, t, r, b = canon["path_rect"][0]
    w = r - l
    signs = set()
    checked = 0
    for i in range(canon["tile_alpha"].size):
        if canon["tile_alpha"][i]:
            continue
        tx, ty = l + i % w, t + i // w
        wn = winding(poly, tx * 16 + 8.0, ty * 16 + 8.0)
        bd = int(canon["tile_backdrop"][i])
        assert abs(bd) == abs(wn)
        if wn:
            signs.add(bd // wn)
            checked += 1
    assert len(signs) == 1 and checked > 0
    # centre winding of {n/m} is m
    ci = (256 // 16 - t) * w + (256 // 16 - l)
    if not canon["tile_alpha"][ci]:
        assert abs(int(canon["tile_backdrop"][ci])) == m


def test_composite_single_opaque_layer_is_colour():
    sc = make_scene(64, 64, [([rect(16, 16, 48, 48)], 0, 0xFF336699)])
    img = oracle.full_model(sc)["image"]
    np.testing.assert_allclose(img[20, 20], [0x99 / 255, 0x66 / 255, 0x33 / 255, 1.0])
    np.testing.assert_allclose(img[5, 5], [0, 0, 0, 0])


@pytest.mark.parametrize("n", [1, 3, 8])
def test_composite_n_identical_layers(n):
    a = 0x66 / 255
    sc = make_scene(64, 64, [([rect(16, 16, 48, 48)], 0, 0x66FF8040)] * n)
    img = oracle.full_model(sc)["image"]
    cov = 1 - (1 - a) ** n
    np.testing.assert_allclose(img[30, 30], [0x40 / 255 * cov, 0x80 / 255 * cov, 1.0 * cov, cov], rtol=1e-12)


def test_evenodd_rule_on_integers():
    for c, want in [(0, 0), (1, 1), (2, 0), (3, 1), (-1, 1), (-2, 0), (0.25, 0.25), (1.75, 0.25), (-0.5, 0.5)]:
        assert oracle.load_oracle().pfo_apply_rule(c, 1) == pytest.approx(want)
    for c, want in [(0, 0), (1, 1), (-3, 1), (0.3, 0.3), (-0.3, 0.3)]:
        assert oracle.load_oracle().pfo_apply_rule(c, 0) == pytest.approx(want)


def test_blob_scene_canonical_invariants():
    sc = scenegen.blob_scene(5, 512, 512, 60, r_max=128.0)
    canon = oracle.run_oracle(sc)
    # closed contours inside the view box: every column sums to zero winding below the path
    inside = 0
    for p in range(sc.n_paths):
        a, b = sc.contour_start[sc.path_contour_start[p]], sc.contour_start[sc.path_contour_start[p + 1]]
        pts = sc.points[a:b]
        if pts.min() >= 0 and pts.max() <= 512:
            inside += 1
            assert (canon["col_sum"][canon["col_off"][p]:canon["col_off"][p + 1]] == 0).all()
    assert inside > 10
    # fills always lie in alpha tiles of their own path
    fp = np.repeat(np.arange(sc.n_paths), np.diff(canon["fill_off"]))
    gt = canon["tile_off"][fp] + canon["fill_tile"]
    assert canon["tile_alpha"][gt].all()
    assert (canon["fill_seg"][:, 0] != canon["fill_seg"][:, 2]).all()
