"""HIP pipeline vs the oracle, through the C ABI.

Bit-exact: tile rects, per-tile backdrops and alpha flags, column sums, fills
(tile-local 8.8 segments) in emission order per path, z-buffer and per
framebuffer-tile draw order.  Tolerances (stated): coverage masks within 1/255
of the FP64 model (u8 quantisation 0.5/255 + fp32 arithmetic); RGBA8 within 1 LSB
of round(FP64 composite * 255).
"""
import numpy as np
import pytest

import oracle
import pfh
import scenegen
from conftest import load_golden

pytestmark = pytest.mark.gpu

COV_TOL = 1.0 / 255 + 1e-6
RGBA_TOL = 1

GOLDEN = load_golden()


@pytest.fixture(scope="module")
def ctx():
    c = pfh.Context(0)
    yield c
    c.close()


def check_scene(ctx, scene, flags=pfh.PFH_PLAIN, name="", guard_hits=0, rgba_tol=RGBA_TOL):
    print(f"[parity] {name} paths={scene.n_paths} canvas={scene.width}x{scene.height} flags={flags}", flush=True)
    sid = ctx.upload(scene)
    ctx.rasterize([sid], flags | pfh.PFH_DUMP)
    got = ctx.dump(sid)
    img = ctx.readback(sid)
    model = oracle.full_model(scene, zmode=1)
    canon = model["canon"]
    assert got["guard_hits"] == canon["guard_hits"] == guard_hits
    for k in oracle.CANONICAL_KEYS:
        np.testing.assert_array_equal(got[k], canon[k], err_msg=f"{name}: {k}")
    if got["masks"].size:
        err = np.abs(got["masks"].astype(np.float64) / 255.0 - model["coverage"]).max()
        assert err <= COV_TOL, f"{name}: coverage error {err * 255:.3f}/255"
    want = np.rint(np.clip(model["image"], 0.0, 1.0) * 255.0).astype(np.int64)
    px = np.abs(img.astype(np.int64) - want).max() if img.size else 0
    assert px <= rgba_tol, f"{name}: rgba8 error {px}"
    ctx.release(sid)
    return got, img


@pytest.mark.parametrize("name", list(GOLDEN))
def test_golden_scene_parity(ctx, name):
    scene, _ = GOLDEN[name]
    check_scene(ctx, scene, name=name)


@pytest.mark.parametrize("seed,size,paths", [(11, 1024, 300), (12, 2048, 1500), (13, 4096, 2500)])
def test_blob_scene_parity(ctx, seed, size, paths):
    check_scene(ctx, scenegen.blob_scene(seed, size, size, paths, r_max=min(512.0, size / 2)),
                name=f"blob{size}")


def test_star_scene_parity(ctx):
    check_scene(ctx, scenegen.star_scene(seed=5, size=1024, n_stars=40, n_max=600, m_max=30, r_max=400.0),
                name="stars")


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH])
def test_negative_zero_direction_chords(ctx, flags):
    """Chords whose direction has a -0 component: process_line_segment's crossing time is
    -inf there (x / -0), its sample point NaN, and the walk ends at the loop guard.  The
    kernels take such chords through the full walk (not the one-tile shortcut) and must
    reproduce the reference's fills and guard count exactly."""
    pts = np.array([[0.0, 10.0], [-0.0, 50.0], [30.0, 50.0], [30.0, 0.0], [10.0, -0.0]], np.float32)
    scene = scenegen.Scene(64, 64, pts, np.zeros(5, np.uint8), np.array([0, 5]), np.ones(1, np.uint8),
                           np.array([0, 1]), np.zeros(1, np.uint8), np.array([0xff3366cc], np.uint32))
    check_scene(ctx, scene, flags=flags, name="-0 chords", guard_hits=2)


def polygon_scene(size, polys):
    """One path per (cx, cy, r, n_vertices, rgba): a closed n-gon of lines."""
    pts, cstart, pstart = [], [0], [0]
    for cx, cy, r, n, _ in polys:
        a = np.arange(n) * (2 * np.pi / n)
        pts.append(np.stack([cx + r * np.cos(a), cy + r * np.sin(a)], 1))
        cstart.append(cstart[-1] + n)
        pstart.append(pstart[-1] + 1)
    pts = np.concatenate(pts).astype(np.float32)
    k = len(polys)
    return scenegen.Scene(size, size, pts, np.zeros(len(pts), np.uint8), np.array(cstart), np.ones(k, np.uint8),
                          np.array(pstart), np.array([i % 2 for i in range(k)], np.uint8),
                          np.array([p[4] for p in polys], np.uint32))


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH])
def test_paths_past_the_lds_limits(ctx, flags):
    """Paths the group-class kernels cannot bin in LDS: a rect of more than 4 600 tiles (count
    and emit through global atomics) and a path of more than 2^15 chords (its backdrop sums
    could overflow the packed 16-bit LDS deltas, so its count pass uses global atomics)."""
    scene = polygon_scene(2048, [(1024.0, 1024.0, 900.0, 64, 0xc0ff8000),      # ~113 x 113 tiles
                                 (700.0, 800.0, 300.0, 40000, 0x80336699),      # 40 000 chords
                                 (1300.0, 1200.0, 200.0, 9, 0xff20c020)])
    check_scene(ctx, scene, flags=flags, name="past LDS limits")


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH])
def test_non_finite_points_and_chord_chains(ctx, flags):
    """Segments with a NaN or infinite point emit nothing (process_line_segment drops every
    piece), so the chord chain of their contour restarts after them; a contour that starts
    where the previous one ended continues the chain.  Fills must match the reference's."""
    nan, inf = np.nan, np.inf
    pts = np.array([[10, 10], [60, 12], [nan, 40], [70, 70], [15, 60],   # a NaN mid-contour
                    [15, 60], [40, 90], [5, 100],                        # starts where the last ended
                    [80, 20], [120, 10], [130, 60], [90, 70],             # a cubic
                    [100, 100], [120, 110], [90, 120],
                    [20, 110], [inf, 115], [40, 125]], np.float32)       # an infinite point
    flags_pt = np.zeros(len(pts), np.uint8)
    flags_pt[9], flags_pt[10] = 1, 2
    scene = scenegen.Scene(128, 128, pts, flags_pt, np.array([0, 5, 8, 12, 15, 18]), np.ones(5, np.uint8),
                           np.array([0, 2, 4, 5]), np.array([0, 1, 0], np.uint8),
                           np.array([0xff2040e0, 0xa0e04020, 0xc020c040], np.uint32))
    check_scene(ctx, scene, flags=flags, name="non-finite points")


def test_empty_and_degenerate_scenes(ctx):
    empty = scenegen.Scene(64, 48, np.zeros((0, 2), np.float32), np.zeros(0, np.uint8), np.array([0]),
                           np.zeros(0, np.uint8), np.array([0]), np.zeros(0, np.uint8), np.zeros(0, np.uint32))
    sid = ctx.upload(empty)
    ctx.rasterize([sid])
    assert (ctx.readback(sid) == 0).all()
    ctx.release(sid)
    # a path with an empty contour and one fully outside the view box
    sc = GOLDEN["fully_outside"][0]
    check_scene(ctx, sc, name="fully_outside")


def test_batched_equals_plain(ctx):
    scenes = scenegen.batch_scenes(n_scenes=6, seed=123, side_max=1024, paths_max=800)
    ids = [ctx.upload(s) for s in scenes]
    plain = []
    for sid in ids:
        ctx.rasterize([sid], pfh.PFH_PLAIN | pfh.PFH_DUMP)
        plain.append((ctx.readback(sid), ctx.dump(sid)))
    for flags in (pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH, pfh.PFH_BATCHED, pfh.PFH_BATCHED):
        ctx.rasterize(ids, flags | pfh.PFH_DUMP)
        for sid, (img, d) in zip(ids, plain):
            np.testing.assert_array_equal(ctx.readback(sid), img)
            got = ctx.dump(sid)
            for k in oracle.CANONICAL_KEYS + ("masks",):
                np.testing.assert_array_equal(got[k], d[k], err_msg=k)
    for sid in ids:
        ctx.release(sid)


def sampled_alpha(model) -> tuple:
    """(alpha tiles that draw entries sample, their fills) of the oracle's model."""
    at = model["alpha"]
    ai = at["alpha_index"][model["canon"]["entry_tile"]]
    ai = ai[ai >= 0]
    return int(ai.shape[0]), int((at["fill_off"][ai + 1] - at["fill_off"][ai]).sum())


def test_fill_rasterizes_only_sampled_alpha_tiles(ctx):
    """Without PFH_DUMP the coverage pass rasterizes only the alpha tiles some draw entry
    samples (a mask under an opaque solid tile of a later path is never read): its counts
    match the oracle's entries, and the pixels equal the dump run's, which rasterizes every
    alpha tile, and the oracle's."""
    scenes = [scenegen.cubic_scene(41, size=512, n_paths=64),
              scenegen.blob_scene(42, 1024, 1024, 1500, alpha_min=0.95),
              scenegen.blob_scene(43, 768, 768, 600, alpha_min=0.95)]
    models = [oracle.full_model(s, zmode=1) for s in scenes]
    sampled = [sampled_alpha(m) for m in models]
    n_alpha = sum(int(m["alpha"]["tile"].shape[0]) for m in models)
    n_fills = sum(int(m["canon"]["fill_seg"].shape[0]) for m in models)
    assert sum(a for a, _ in sampled) < n_alpha  # the scenes do hide some alpha tiles
    ids = [ctx.upload(s) for s in scenes]
    for flags in (pfh.PFH_PLAIN, pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH, pfh.PFH_BATCHED):
        print(f"[sampled alpha] flags={flags}", flush=True)
        runs = [[sid] for sid in ids] if flags == pfh.PFH_PLAIN else [ids]
        listed_imgs = {}  # the run without dumps, which rasterizes the sampled tiles only
        for dump in (False, True):
            got_a = got_f = 0
            for run in runs:
                ctx.rasterize(run, flags | (pfh.PFH_DUMP if dump else 0))
                st = ctx.stats()
                got_a += st["alpha_masked"]
                got_f += st["fills_masked"]
            if dump:
                assert (got_a, got_f) == (n_alpha, n_fills)
            else:
                assert (got_a, got_f) == (sum(a for a, _ in sampled), sum(f for _, f in sampled))
            for sid, m in zip(ids, models):
                img = ctx.readback(sid)
                want = np.rint(np.clip(m["image"], 0.0, 1.0) * 255.0).astype(np.int64)
                assert np.abs(img.astype(np.int64) - want).max() <= RGBA_TOL
                if dump:
                    np.testing.assert_array_equal(img, listed_imgs[sid])
                else:
                    listed_imgs[sid] = img
    for sid in ids:
        ctx.release(sid)


def test_undersized_prediction_retries_without_wrong_pixels(ctx):
    scenes = scenegen.batch_scenes(n_scenes=3, seed=321, side_max=512, paths_max=300)
    ids = [ctx.upload(s) for s in scenes]
    ref = []
    for sid in ids:
        ctx.rasterize([sid], pfh.PFH_PLAIN)
        ref.append(ctx.readback(sid))
    for sid in ids:
        ctx.release(sid)
    # A fresh context (no buffers grown yet) and a predictor whose outputs are ~0:
    # every fill / alpha-tile capacity is too small, so the batch must be retried.
    fresh = pfh.Context(0)
    ids = [fresh.upload(s) for s in scenes]
    w = np.zeros(pfh.PREDICTOR_FLOATS, np.float32)
    w[12:24] = 1.0
    fresh.set_predictor(w, safety=1.0)
    fresh.rasterize(ids, pfh.PFH_BATCHED | pfh.PFH_PREDICT)
    assert fresh.stats()["retries"] > 0
    for sid, img in zip(ids, ref):
        np.testing.assert_array_equal(fresh.readback(sid), img)
    fresh.close()


def test_plain_fresh_context_more_chords_than_tiles():
    # Small canvas, many small paths: the chord count exceeds the tile count, so the
    # plain pipeline grows its chord buffer (and the scan scratch sized by it) after
    # the first read-back.  Every later scan must use the regrown scratch.
    scene = scenegen.blob_scene(77, 384, 384, 900, r_min=4.0, r_max=40.0)
    o = oracle.run_oracle(scene)
    assert int(o["path_lines"].sum()) > 2 * int(o["tile_off"][-1])
    fresh = pfh.Context(0)
    check_scene(fresh, scene, name="chords>tiles fresh")
    fresh.close()


def test_graph_replay_after_buffers_regrow():
    # A captured graph holds raw buffer addresses; a larger rasterize in between
    # regrows the buffers, and the replay of the first scene set must not use the old ones.
    fresh = pfh.Context(0)
    small = scenegen.batch_scenes(n_scenes=2, seed=9, side_max=384, paths_max=120)
    big = scenegen.blob_scene(10, 2048, 2048, 3000, r_max=512.0)
    ids = [fresh.upload(s) for s in small]
    fresh.rasterize(ids, pfh.PFH_BATCHED)
    first = [fresh.readback(i) for i in ids]
    bid = fresh.upload(big)
    fresh.rasterize([bid], pfh.PFH_BATCHED)
    big_img = fresh.readback(bid)
    for _ in range(2):
        fresh.rasterize(ids, pfh.PFH_BATCHED)
        for i, img in zip(ids, first):
            np.testing.assert_array_equal(fresh.readback(i), img)
    fresh.rasterize([bid], pfh.PFH_PLAIN)
    np.testing.assert_array_equal(fresh.readback(bid), big_img)
    fresh.close()


def test_predicted_capacities_above_allocation_replay_graph():
    # Plain first (buffers sized to the exact counts, rounded to 2^k), then a predictor
    # whose totals exceed those allocations: the first batched call regrows the buffers
    # and captures a graph; repeat calls must ask for nothing new and replay it.
    scene = scenegen.blob_scene(31, 768, 768, 600, r_max=300.0)
    fresh = pfh.Context(0)
    sid = fresh.upload(scene)
    fresh.rasterize([sid], pfh.PFH_PLAIN)
    want = fresh.readback(sid)
    st = fresh.stats()
    totals = np.array([st["lines"], st["fills"], st["alpha_tiles"]], np.float64)
    w = np.zeros(pfh.PREDICTOR_FLOATS, np.float32)
    w[12:24] = 1.0
    w[-3:] = np.log1p(2.5 * totals / scene.n_paths)  # output bias only: 2.5x the exact totals
    fresh.set_predictor(w, safety=1.0)
    for _ in range(4):
        fresh.rasterize([sid], pfh.PFH_BATCHED | pfh.PFH_PREDICT)
        st = fresh.stats()
        assert st["retries"] == 0 and st["host_syncs"] == 1
        assert st["capacity_lines"] >= 2 * totals[0]
        np.testing.assert_array_equal(fresh.readback(sid), want)
    fresh.close()


CLIP_GOLDEN = [n for n in GOLDEN if n.startswith("clip_")]


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH, pfh.PFH_BATCHED])
@pytest.mark.parametrize("name", CLIP_GOLDEN)
def test_clip_paths_golden(ctx, name, flags):
    """Clip paths (core/d3d9/tiler.cpp:385-431, shaders/d3d9/tile_clip_combine.frag:25-27): the
    resolved tiles and backdrops bit-exact, the combined masks and the image within tolerance of
    the FP64 model, on scenes whose tiling the reference's own Tiler wrote."""
    scene, _ = GOLDEN[name]
    check_scene(ctx, scene, flags=flags, name=name)


@pytest.mark.parametrize("seed,size,n_draw,n_clip", [(31, 1024, 400, 12), (32, 2048, 1500, 40)])
def test_clip_scene_parity(ctx, seed, size, n_draw, n_clip):
    scene = scenegen.clip_scene(seed, size=size, n_draw=n_draw, n_clip=n_clip, r_max=min(512.0, size / 3))
    check_scene(ctx, scene, flags=pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH, name=f"clips{size}")


def test_clip_scenes_in_a_batch(ctx):
    """Scenes with and without clip paths in one batched launch (two-phase propagate for the whole
    batch): each scene's image and dump equal its plain run, over graph replays too, and the
    listed (non-dump) launch draws the same image."""
    scenes = [scenegen.clip_scene(41, size=768, n_draw=300, n_clip=9),
              scenegen.blob_scene(42, 512, 384, 200, r_max=128.0),
              scenegen.clip_scene(43, size=1024, n_draw=500, n_clip=3, clipped=0.9)]
    ids = [ctx.upload(s) for s in scenes]
    plain = []
    for sid in ids:
        ctx.rasterize([sid], pfh.PFH_PLAIN | pfh.PFH_DUMP)
        plain.append((ctx.readback(sid), ctx.dump(sid)))
    for flags in (pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH | pfh.PFH_DUMP, pfh.PFH_BATCHED | pfh.PFH_DUMP,
                  pfh.PFH_BATCHED | pfh.PFH_DUMP, pfh.PFH_BATCHED, pfh.PFH_BATCHED):
        ctx.rasterize(ids, flags)
        for sid, (img, d) in zip(ids, plain):
            np.testing.assert_array_equal(ctx.readback(sid), img)
            if flags & pfh.PFH_DUMP:
                got = ctx.dump(sid)
                for k in oracle.CANONICAL_KEYS + ("masks",):
                    np.testing.assert_array_equal(got[k], d[k], err_msg=k)
    for sid in ids:
        ctx.release(sid)


def test_clip_nesting_limits(ctx):
    """Clip paths may nest in chains of at most 4 (core/d3d11/scene_builder.cpp:92-133 levels); a
    deeper chain, a cycle or a clip index out of range fails the upload with PFH_E_ARG."""
    scene = scenegen.clip_scene(51, size=256, n_draw=10, n_clip=5)
    n_draw = scene.n_paths - 5
    scene.path_clip = scene.path_clip.copy()
    scene.path_clip[n_draw:] = [-1, 0, 1, 2, 3]  # a chain of 5
    with pytest.raises(pfh.PfhError, match="nest deeper"):
        ctx.upload(scene)
    scene.path_clip[n_draw:] = [4, 0, 1, 2, 3]  # a cycle
    with pytest.raises(pfh.PfhError, match="nest deeper"):
        ctx.upload(scene)
    scene.path_clip[n_draw:] = [-1, 0, 1, 2, -1]
    scene.path_clip[0] = 5  # no clip path 5
    with pytest.raises(pfh.PfhError, match="path_clip"):
        ctx.upload(scene)
    scene.path_clip[0] = 3
    check_scene(ctx, scene, name="chain of 4")


def nested_rect_scene(size, seed):
    """Draw blobs under chains of nested clip rects and circles (scenegen.nested_clips)."""
    draws = scenegen.blob_scene(seed, size, size, 300, r_max=size / 4)
    clips, parent, _ = scenegen.nested_clips(seed ^ 0x77, size=size, n_clip=24)
    cs = scenegen.outlines_scene(size, size, clips, np.arange(24) % 2, np.full(24, 0xff000000, np.uint32))
    sc = scenegen.concat_paths(draws, cs)
    rng = np.random.default_rng(seed)
    pc = rng.integers(0, 24, 300).astype(np.int32)
    pc[rng.random(300) < 0.3] = -1
    sc.n_clip_paths = 24
    sc.path_clip = np.concatenate([pc, parent])
    return sc


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH, pfh.PFH_BATCHED])
def test_nested_clip_paths(ctx, flags):
    """Clip paths nested up to 4 deep: each clip path is resolved against its own clip path like a
    draw path (tiles bit-exact), its coverage the min over the chain; masks within 1/255 of the
    FP64 model."""
    check_scene(ctx, nested_rect_scene(1024, 61), flags=flags, name="nested clips")


def test_stroked_polylines(ctx):
    """Stroked cubic polylines (the library's stroke-to-fill, bit-exact with the reference's) through
    the pipeline: tiles bit-exact, masks and pixels within tolerance."""
    check_scene(ctx, scenegen.stroke_scene(62, size=1024, n=120), flags=pfh.PFH_BATCHED, name="strokes")


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED])
def test_dashed_stroked_polylines(ctx, flags):
    """Stroked cubic polylines, half of them dashed before stroking (the library's dasher and
    stroker, bit-exact with the reference's) through the pipeline: tiles bit-exact, masks and pixels
    within tolerance."""
    check_scene(ctx, scenegen.dashed_stroke_scene(63, size=1024, n=120), flags=flags, name="dashed strokes")


def _deep_curve_scene(seed, size=1024, n=40):
    """Closed contours of 3 cubics (or quads) with controls up to a canvas beyond the canvas: subdivision
    depths up to 7; and one cubic reaching ~1e9 px, which hits the depth cap (MAX_FLATTEN_DEPTH
    levels, every leaf there a guard hit)."""
    rng = np.random.default_rng(seed)
    pts, flags, cst = [], [], [0]
    for i in range(n):
        quad = i % 4 == 3
        p = rng.uniform(0, size, (3, 2))
        for k in range(3):
            pts.append(p[k])
            flags.append(0)
            for c in range(1 if quad else 2):
                pts.append(rng.uniform(-size, 2 * size, 2))
                flags.append(1 + c)
        cst.append(len(pts))
    big = [(10, 10), (5e8, 1e9), (-5e8, 1e9), (1000, 12), (600, 900)]
    pts += [np.array(q, np.float64) for q in big]
    flags += [0, 1, 2, 0, 0]
    cst.append(len(pts))
    P = len(cst) - 1
    rgba = [int(c) for c in rng.integers(0, 2**32, P, dtype=np.uint64) | 0xff000000]
    return scenegen.Scene(size, size, np.array(pts, np.float32), np.array(flags, np.uint8), np.array(cst),
                          np.ones(P, np.uint8), np.arange(P + 1), rng.integers(0, 2, P).astype(np.uint8),
                          np.array(rgba, np.uint32))


@pytest.mark.parametrize("deep", [1, 4, 6])
def test_deep_curves_split_over_lanes(ctx, deep):
    """Schedule::flatten_deep routes curves of an estimated subdivision depth of `deep` or more to
    16 lanes each (k_flatten_deep): the chords, fills and pixels are those of the oracle's
    depth-first walk, including the capped depth's guard hits, on curves of every depth, the
    golden curve scenes and stroked outlines (dashed too), plain and batched."""
    before = ctx.schedule()
    try:
        ctx.set_schedule(flatten_deep=deep)
        sc = _deep_curve_scene(70 + deep)
        guard = int(oracle.full_model(sc)["canon"]["guard_hits"])
        assert guard > 1000  # the capped cubic's leaves at depth 16
        for flags in (pfh.PFH_PLAIN, pfh.PFH_BATCHED):
            check_scene(ctx, sc, flags=flags, name=f"deep curves ({deep})", guard_hits=guard)
        for name in ("cubic_512_64x8", "tangent_cubic", "collinear_cubic"):
            if name in GOLDEN:
                check_scene(ctx, GOLDEN[name][0], name=f"{name} ({deep})")
        check_scene(ctx, scenegen.blob_scene(71, 1024, 1024, 300, r_max=512.0), flags=pfh.PFH_BATCHED,
                    name=f"blobs ({deep})")
        check_scene(ctx, scenegen.dashed_stroke_scene(72, size=1024, n=60), flags=pfh.PFH_BATCHED,
                    name=f"dashed strokes ({deep})")
    finally:
        ctx.set_schedule(**before)


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED])
def test_stress_scene_full_size(ctx, flags):
    """The stress workload at its full size: 4096², 2 000 star polygons of 100-2 000 vertices, 500
    stroked cubic polylines of width 1-20, 200 clip paths nesting up to 4 deep, half of the draw
    paths clipped.  Tiles, fills, backdrops and order bit-exact; masks within 1/255; RGBA8 within
    1 LSB; the listed launch (no dumps) draws the same image as the dump launch."""
    scene = scenegen.stress_scene()
    got, img = check_scene(ctx, scene, flags=flags, name="stress full")
    sid = ctx.upload(scene)
    ctx.rasterize([sid], flags)
    np.testing.assert_array_equal(ctx.readback(sid), img)
    ctx.release(sid)


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED])
def test_stress_stars_under_clip_paths(ctx, flags):
    """The stress workload's shape at 2048²: heavily overlapping star polygons (windings far above
    1, both fill rules) with half of them clipped by 40 blob clip paths."""
    scene = scenegen.stress_clip_scene(7, size=2048, n_stars=300, n_max=600, n_clip=40)
    check_scene(ctx, scene, flags=flags, name="stress stars + clips")
