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


def _rect_path(x0, y0, x1, y1):
    return np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], np.float32)


def _rect_scene(w, h, rects, rules, n_clip, path_clip):
    """One path per list of rects (its contours); the last n_clip paths are clip paths."""
    pts, cst, pcs = [], [0], [0]
    for contours in rects:
        for r in contours:
            pts.append(_rect_path(*r))
            cst.append(cst[-1] + 4)
        pcs.append(len(cst) - 1)
    k = len(rects)
    sc = scenegen.Scene(w, h, np.concatenate(pts), np.zeros(4 * (len(cst) - 1), np.uint8), np.array(cst),
                        np.ones(len(cst) - 1, np.uint8), np.array(pcs), np.array(rules, np.uint8),
                        np.full(k, 0xff00ff00, np.uint32))
    sc.n_clip_paths = n_clip
    sc.path_clip = np.array(path_clip, np.int32)
    return sc


def _box_area(w, h, r):
    """Exact per-pixel area of pixel [x, x+1) x [y, y+1) inside rect r (the box-filter coverage)."""
    x = np.arange(w, dtype=np.float64)
    y = np.arange(h, dtype=np.float64)
    ox = np.clip(np.minimum(x + 1, r[2]) - np.maximum(x, r[0]), 0, 1)
    oy = np.clip(np.minimum(y + 1, r[3]) - np.maximum(y, r[1]), 0, 1)
    return oy[:, None] * ox[None, :]


def test_clip_model_is_min_of_box_coverages():
    """A rect clipped by a rect: the clip combine gives min(draw area, clip area) per pixel, a
    closed form in the box-filter areas (shaders/d3d9/tile_clip_combine.frag:25-27).  Edges off
    the pixel grid so both masks are fractional where they meet; tiles outside the clip's rect
    are culled, solid tiles inside take the clip's mask.  Tolerance 1/255: a fill's tile-local x
    is clamped to 4095/256 (core/d3d9/object_builder.cpp:19-64), so an edge across a whole tile
    leaves its last pixel column 1/256 short."""
    d, c = (3.5, 4.25, 70.75, 50.5), (20.25, 10.5, 90.5, 40.75)
    sc = _rect_scene(96, 64, [[d], [c]], [0, 0], 1, [0, -1])
    m = oracle.full_model(sc)
    got = oracle.path_coverage_image(sc, m, 0)
    want = np.minimum(_box_area(96, 64, d), _box_area(96, 64, c))
    assert np.abs(got - want).max() <= 1.0 / 255.0
    assert np.abs(got - want).mean() < 1e-3
    # the clip path draws nothing itself
    assert (m["canon"]["entry_path"] == 0).all()


def test_clip_model_windings_above_one():
    """Clip winding 2 (two overlapping clip contours) under an even-odd draw path of winding 2, on
    pixel-aligned rects.  Where the clip tile has a mask the combine gives
    evenodd(min(|w_draw|, |w_clip|)); under a solid clip tile the draw tile passes unchanged,
    evenodd(w_draw) (core/d3d9/tiler.cpp:385-431 combines only two masks).  The clip's own fill
    rule does not matter."""
    d = [(8, 8, 56, 56), (24, 24, 72, 72)]        # winding 2 in [24, 56)^2
    c = [(16, 0, 64, 64), (32, 16, 80, 48)]       # winding 2 in [32, 64) x [16, 48)
    eo = np.ones(1, np.uint8)
    for crule in (0, 1):
        sc = _rect_scene(96, 96, [d, c], [1, crule], 1, [0, -1])
        m = oracle.full_model(sc)
        got = oracle.path_coverage_image(sc, m, 0)
        wd = sum(_box_area(96, 96, r) for r in d)
        wc = sum(_box_area(96, 96, r) for r in c)
        canon = m["canon"]
        l, t, r, b = [int(v) for v in canon["path_rect"][1]]
        t0 = int(canon["tile_off"][1])
        clip_mask = np.zeros((96, 96), bool)
        for i in range((r - l) * (b - t)):
            tx, ty = l + i % (r - l), t + i // (r - l)
            clip_mask[ty * 16:ty * 16 + 16, tx * 16:tx * 16 + 16] = canon["tile_alpha"][t0 + i] == 1
        want = np.where(clip_mask, oracle.apply_rule(np.minimum(wd, wc), eo), oracle.apply_rule(wd, eo))
        want[wc == 0] = 0.0
        assert np.abs(got - want).max() <= 1.0 / 255.0, crule  # 8.8 clamp, as above
        both = (wd == 2) & (wc == 2) & clip_mask
        assert both.any() and (got[both] <= 1.0 / 255.0).all()
        assert ((wd == 2) & (wc == 1) & clip_mask).any()


def test_nested_clip_model_is_min_over_the_chain():
    """A rect clipped by a chain of 4 nested clip rects (each clip path clipped by the next one out,
    the level structure of core/d3d11/scene_builder.cpp:92-133): per pixel the coverage is the min of
    the five box-filter areas, whatever the tile cases at each level (a solid tile over a mask takes
    the mask, a mask over a solid tile keeps its own, a blank or missing tile culls).  Edges off the
    pixel grid; tolerance 1/255 for the 8.8 clamp, as above."""
    d = (2.5, 3.25, 120.75, 100.5)
    chain = [(10.25, 6.5, 110.5, 90.75), (14.75, 12.25, 100.25, 80.5), (30.5, 20.75, 94.25, 70.25),
             (22.25, 18.5, 70.75, 60.5)]
    # paths: draw d, then clips c1..c4 where c1 is the outermost; d is clipped by c4
    sc = _rect_scene(128, 112, [[d]] + [[c] for c in chain], [0, 0, 1, 0, 1], 4, [3, -1, 0, 1, 2])
    m = oracle.full_model(sc)
    got = oracle.path_coverage_image(sc, m, 0)
    want = _box_area(128, 112, d)
    for c in chain:
        want = np.minimum(want, _box_area(128, 112, c))
    assert np.abs(got - want).max() <= 1.0 / 255.0
    assert np.abs(got - want).mean() < 1e-3
    assert (m["canon"]["entry_path"] == 0).all()  # clip paths draw nothing


def test_nested_clip_solid_draw_takes_the_chain():
    """A draw rect covering the canvas (solid tiles everywhere inside) under a chain of 2 clips: its
    tiles over the inner clip's masks take that clip's K, itself min'd with its parent where both
    have masks: the coverage is min(box(c1), box(c2))."""
    d = (0, 0, 96, 96)
    c1, c2 = (8.5, 9.25, 80.75, 70.5), (20.25, 4.5, 90.5, 60.25)
    sc = _rect_scene(96, 96, [[d], [c1], [c2]], [0, 0, 0], 2, [1, -1, 0])
    m = oracle.full_model(sc)
    got = oracle.path_coverage_image(sc, m, 0)
    want = np.minimum(_box_area(96, 96, c1), _box_area(96, 96, c2))
    assert np.abs(got - want).max() <= 1.0 / 255.0


def test_nested_clip_windings_above_one():
    """Windings above 1 at every level, pixel-aligned: where the draw tile and both clip levels have
    masks the result is evenodd(min(|w_draw|, |w_c2|, |w_c1|)); outside the outer clip's rect nothing."""
    d = [(8, 8, 88, 88), (24, 24, 72, 72)]        # winding 2 in [24, 72)^2
    c1 = [(4, 4, 84, 84), (20, 20, 60, 92)]       # winding 2 in [20, 60) x [20, 84)
    c2 = [(12, 0, 76, 76), (28, 14, 90, 50)]      # winding 2 in [28, 76) x [14, 50)
    sc = _rect_scene(96, 96, [d, c1, c2], [1, 0, 1], 2, [1, -1, 0])
    m = oracle.full_model(sc)
    got = oracle.path_coverage_image(sc, m, 0)
    wd = sum(_box_area(96, 96, r) for r in d)
    w1 = sum(_box_area(96, 96, r) for r in c1)
    w2 = sum(_box_area(96, 96, r) for r in c2)
    canon = m["canon"]

    def alpha_mask(p):
        l, t, r, b = [int(v) for v in canon["path_rect"][p]]
        t0 = int(canon["tile_off"][p])
        out = np.zeros((96, 96), bool)
        for i in range((r - l) * (b - t)):
            tx, ty = l + i % (r - l), t + i // (r - l)
            out[ty * 16:ty * 16 + 16, tx * 16:tx * 16 + 16] = canon["tile_alpha"][t0 + i] == 1
        return out

    all_masks = alpha_mask(0) & alpha_mask(1) & alpha_mask(2)
    want = oracle.apply_rule(np.minimum(np.minimum(wd, w1), w2), np.ones(1, np.uint8))
    assert all_masks.any()
    assert np.abs(got - want)[all_masks].max() <= 1.0 / 255.0
    assert (got[w1 == 0] == 0).all()
    three = all_masks & (wd == 2) & (w1 == 2) & (w2 == 2)
    assert three.any() and (got[three] <= 1.0 / 255.0).all()  # evenodd(2) covers nothing


def test_clip_nesting_deeper_than_four_is_refused():
    rects = [[(0, 0, 32, 32)]] + [[(1 + i, 1 + i, 30 - i, 30 - i)] for i in range(5)]
    sc = _rect_scene(32, 32, rects, [0] * 6, 5, [4, -1, 0, 1, 2, 3])
    with pytest.raises(ValueError):
        oracle.full_model(sc)
    sc.path_clip = np.array([4, -1, 0, 1, 2, 4], np.int32)  # a cycle
    with pytest.raises(ValueError):
        oracle.full_model(sc)


def test_clip_fill_rule_is_ignored_where_windings_reach_two():
    """The golden clip_ccw_evenodd_rule scene: its clip path is two overlapping counter-clockwise rects
    (winding -2 where they overlap) marked even-odd.  The hybrid path takes |winding| and never applies
    the clip's own rule, so the nonzero draw rect under it stays fully covered inside the overlap (the
    even-odd rule, honoured, would have clipped it away there) and in the rest of the clip."""
    from conftest import load_golden

    sc, _ = load_golden()["clip_ccw_evenodd_rule"]
    m = oracle.full_model(sc, zmode=0)
    cov = oracle.path_coverage_image(sc, m, 0)
    overlap = cov[32:99, 52:117]      # inside both clip rects (and the draw rect), off their edges
    single = cov[110:118, 30:45]      # inside one clip rect and the draw rect only
    assert np.abs(overlap - 1.0).max() <= 1.0 / 255.0
    assert np.abs(single - 1.0).max() <= 1.0 / 255.0
    assert cov[11:12, 60:100].max() == 0.0  # inside the draw rect, above both clip rects: clipped away


def _rule(w, even_odd):
    """The fill rule of integer windings (pixel-aligned rects): nonzero |w| >= 1, even-odd w odd."""
    w = np.rint(w)
    return (w % 2 != 0).astype(np.float64) if even_odd else (w != 0).astype(np.float64)


# Draw path, outer clip, inner clip: two overlapping rects each (winding 2 where they overlap, and the
# overlaps span whole 16 x 16 tiles, so solid tiles with backdrop 2 meet every case of the
# resolution: over a clip mask, over a solid clip tile, under a solid draw tile).
OWN_D = [(8, 8, 88, 88), (24, 24, 72, 72)]        # winding 2 in [24, 72)^2
OWN_C1 = [(4, 4, 92, 84), (16, 30, 64, 92)]       # winding 2 in [16, 64) x [30, 84)
OWN_C2 = [(12, 0, 80, 76), (28, 14, 90, 50)]      # winding 2 in [28, 80) x [14, 50)


@pytest.mark.parametrize("rd,rc", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_own_clip_rules_are_the_min_of_each_rule(rd, rc):
    """clip_rule 1 (ClipPath::fill_rule honoured): the coverage is min(rule_draw(w_draw),
    rule_clip(w_clip)) at every pixel, with no dependence on which tiles carry masks; pixel-aligned
    rects make both exact 0/1 (tolerance 1/255 for the 8.8 clamp at tile edges, as above).  The
    reference's |winding| form differs from it where an even winding meets a mask: checked to
    differ on this scene whenever either rule is even-odd."""
    sc = _rect_scene(96, 96, [OWN_D, OWN_C1], [rd, rc], 1, [0, -1])
    sc.clip_rule = 1
    got = oracle.path_coverage_image(sc, oracle.full_model(sc), 0)
    wd = sum(_box_area(96, 96, r) for r in OWN_D)
    wc = sum(_box_area(96, 96, r) for r in OWN_C1)
    want = np.minimum(_rule(wd, rd), _rule(wc, rc))
    assert np.abs(got - want).max() <= 1.0 / 255.0
    sc.clip_rule = 0
    ref = oracle.path_coverage_image(sc, oracle.full_model(sc), 0)
    differ = np.abs(ref - want).max() > 0.5
    assert differ == bool(rd or rc)


@pytest.mark.parametrize("rules", [(1, 1, 1), (0, 1, 0), (1, 0, 1), (0, 0, 1)])
def test_own_clip_rules_nested_chain(rules):
    """A draw path under a chain of two clip paths, each with windings up to 2 and its own rule: the
    coverage is the min of the three rules' values at every pixel (clip_rule 1), whatever the tile
    cases at each level; outside the outer clip nothing.  Tolerance 2/256: the outer clip has two
    horizontal edges across the same tile row, each clamped 1/256 short in a tile's last pixel
    column (the 8.8 clamp, as above), which the even-odd rule shows at winding 2."""
    sc = _rect_scene(96, 96, [OWN_D, OWN_C1, OWN_C2], list(rules), 2, [1, -1, 0])
    sc.clip_rule = 1
    got = oracle.path_coverage_image(sc, oracle.full_model(sc), 0)
    want = np.ones((96, 96))
    for rects, r in zip((OWN_D, OWN_C1, OWN_C2), rules):
        want = np.minimum(want, _rule(sum(_box_area(96, 96, x) for x in rects), r))
    assert np.abs(got - want).max() <= 2.0 / 256.0
    assert np.abs(got - want).mean() < 1e-3
    assert want.max() == 1.0 and (want == 0).any()


def test_own_clip_rules_solid_tiles_of_even_windings_cover_nothing():
    """A canvas-wide draw rect under an even-odd clip of two nested rects (winding 2 over whole
    tiles inside): with clip_rule 1 the clip's even tiles cull the draw tiles under them, and a solid
    draw tile of even winding under an even-odd draw rule takes no clip mask (coverage 0 there);
    with clip_rule 0 the reference's |winding| keeps the inner square covered."""
    clip = [(8, 8, 88, 88), (32, 32, 72, 72)]  # no two horizontal edges in one tile row (the 8.8 clamp)
    sc = _rect_scene(96, 96, [[(0, 0, 96, 96)], clip], [0, 1], 1, [0, -1])
    outer, inner = _box_area(96, 96, clip[0]), _box_area(96, 96, clip[1])
    sc.clip_rule = 1
    m = oracle.full_model(sc)
    assert np.abs(oracle.path_coverage_image(sc, m, 0) - (outer - inner)).max() <= 1.0 / 255.0
    canon = m["canon"]

    def tile(p, tx, ty):
        l, t, r, b = [int(v) for v in canon["path_rect"][p]]
        return int(canon["tile_off"][p]) + (ty - t) * (r - l) + (tx - l)

    # the clip's solid tiles of winding 2 (no mask) cull the draw tiles over them
    even = [(tx, ty) for ty in range(6) for tx in range(6)
            if not canon["tile_alpha"][tile(1, tx, ty)] and canon["tile_backdrop"][tile(1, tx, ty)] in (-2, 2)]
    assert even
    for tx, ty in even:
        assert canon["tile_backdrop"][tile(0, tx, ty)] == 0 and not canon["tile_alpha"][tile(0, tx, ty)]
    sc.clip_rule = 0
    assert np.abs(oracle.path_coverage_image(sc, oracle.full_model(sc), 0) - outer).max() <= 1.0 / 255.0
    # solid draw tiles of winding 2 under the even-odd rule over a nonzero clip's masks (its bottom
    # edge crosses them): the reference's tiles take the clip mask under their rule (covered), here
    # they cover nothing
    d = [(0, 0, 96, 96), (16, 16, 80, 80)]
    c = (4.5, 4.5, 91.5, 50.5)
    sc = _rect_scene(96, 96, [d, [c]], [1, 0], 1, [0, -1])
    want = np.minimum(_rule(sum(_box_area(96, 96, r) for r in d), 1), _box_area(96, 96, c))
    for clip_rule in (0, 1):
        sc.clip_rule = clip_rule
        got = oracle.path_coverage_image(sc, oracle.full_model(sc), 0)
        err = np.abs(got - want)
        if clip_rule:
            assert err.max() <= 1.0 / 255.0
        else:
            assert np.abs(got[48:50, 16:80] - 1.0).max() <= 1.0 / 255.0  # the reference's quirk


def test_own_clip_rules_golden_evenodd_clip():
    """The golden clip_ccw_evenodd_rule scene with clip_rule 1: the clip path's even-odd rule is
    honoured, so the nonzero draw rect under it is clipped away where the two clip rects overlap
    (winding -2) and stays covered where one covers."""
    from conftest import load_golden

    sc, _ = load_golden()["clip_ccw_evenodd_rule"]
    sc.clip_rule = 1
    cov = oracle.path_coverage_image(sc, oracle.full_model(sc), 0)
    assert cov[32:99, 52:117].max() <= 1.0 / 255.0
    assert np.abs(cov[110:118, 30:45] - 1.0).max() <= 1.0 / 255.0
    assert cov[11:12, 60:100].max() == 0.0
