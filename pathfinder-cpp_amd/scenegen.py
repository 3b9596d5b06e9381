"""Seeded synthetic scene generator for the rasterization workloads.

A scene is held as flat outline arrays, the post-``Path2d`` form the tiler
consumes (core/data/path.h:12-57 ``Outline``/``Contour``; core/data/data.h:150-160
``PointFlag``):

    points           float32 [N, 2]
    point_flags      uint8   [N]     0 on-curve, 1 first control, 2 second control
    contour_start    int32   [C + 1] point range of each contour
    contour_closed   uint8   [C]
    path_contour_start int32 [P + 1] contour range of each draw path
    path_fill_rule   uint8   [P]     0 nonzero (FillRule::Winding), 1 even-odd
    path_rgba        uint32  [P]     solid paint, r | g << 8 | b << 16 | a << 24
    view_box         float32 [4]     left, top, right, bottom

Workloads (the scene families the benchmark and parity tests use):
  * ``cubic_scene``: 512x512, 64 closed paths of 8 cubic Béziers with all
    control points uniform in the canvas, nonzero rule, opaque colours.
  * ``blob_scene``: W x H canvas, closed star-shaped paths of 4..64 segments
    (line/quad/cubic uniform), centres uniform, radii log-uniform, alpha
    U(0.2, 1), fill rule chosen by a coin.  With 4096x4096 and 10 000 paths
    this is the 4096^2 single-scene workload.
  * ``batch_scenes``: many blob scenes with side log-uniform in [256, 4096]
    and path count log-uniform in [10, 20 000].
  * ``star_scene``: star polygons {n/m} with winding up to m.

Points are not deduplicated the way ``Path2d::line_to`` does
(core/path2d.cpp:21-29); the arrays are outlines, so every consumer (the HIP
path, the oracle and the reference harness) tiles exactly the same geometry.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

BASE_SEED = 0x9E3779B97F4A7C15


@dataclasses.dataclass
class Scene:
    width: int
    height: int
    points: np.ndarray
    point_flags: np.ndarray
    contour_start: np.ndarray
    contour_closed: np.ndarray
    path_contour_start: np.ndarray
    path_fill_rule: np.ndarray
    path_rgba: np.ndarray
    view_box: np.ndarray = None
    # Clip paths (Scene::clip_paths, core/scene.h): the last n_clip_paths paths; path_clip[p] is the
    # clip path of draw path p as an index among them, or -1 (None: no path is clipped).
    n_clip_paths: int = 0
    path_clip: np.ndarray = None
    # How a clip path's coverage enters: 0 its |winding sum| whatever its fill rule (the reference),
    # 1 under its own path_fill_rule (ClipPath::fill_rule honoured).
    clip_rule: int = 0
    # Blend modes (DrawPath::blend_mode): None (src-over), or per path a BlendMode value
    # (core/paint/effects.h:56-123 order; clip paths' are not used).
    path_blend: np.ndarray = None
    # Strokes (Canvas::stroke_path): None, or per path (line_width, line_cap, line_join, miter_limit)
    # and optionally (dash_start, dash_count, dash_offset); a path with line_width > 0 is stroked at
    # upload (dashed first with stroke_dashes[dash_start:dash_start + dash_count] when dash_count > 0)
    # and its stroke filled with the nonzero rule.
    path_stroke: np.ndarray = None
    stroke_dashes: np.ndarray = None
    # Gradient paints (Paint::from_gradient): None, or per path an index into `gradients` (GRADIENT
    # records; -1: the solid path_rgba), with the stops of every gradient in the two stop arrays.
    path_gradient: np.ndarray = None
    gradients: np.ndarray = None
    gradient_stop_offsets: np.ndarray = None
    gradient_stop_rgba: np.ndarray = None
    gradient_ramps: np.ndarray = None  # [n_gradients, 256] for gradients with stop_count < 0

    def __post_init__(self):
        if self.view_box is None:
            self.view_box = np.array([0, 0, self.width, self.height], np.float32)
        if self.path_clip is not None:
            self.path_clip = np.ascontiguousarray(self.path_clip, np.int32)
        if self.path_stroke is not None:
            ps = np.asarray(self.path_stroke, np.float32)
            self.path_stroke = np.ascontiguousarray(ps.reshape(-1, ps.shape[-1] if ps.ndim == 2 else 4))
        if self.stroke_dashes is not None:
            self.stroke_dashes = np.ascontiguousarray(self.stroke_dashes, np.float32)
        self.points = np.ascontiguousarray(self.points, np.float32).reshape(-1, 2)
        self.point_flags = np.ascontiguousarray(self.point_flags, np.uint8)
        self.contour_start = np.ascontiguousarray(self.contour_start, np.int32)
        self.contour_closed = np.ascontiguousarray(self.contour_closed, np.uint8)
        self.path_contour_start = np.ascontiguousarray(self.path_contour_start, np.int32)
        self.path_fill_rule = np.ascontiguousarray(self.path_fill_rule, np.uint8)
        self.path_rgba = np.ascontiguousarray(self.path_rgba, np.uint32)
        self.view_box = np.ascontiguousarray(self.view_box, np.float32)

    @property
    def n_paths(self) -> int:
        return int(self.path_contour_start.shape[0] - 1)

    @property
    def path_opaque(self) -> np.ndarray:
        # ColorU::is_opaque (common/color.cpp:42-44) for a solid paint; for a gradient paint
        # Paint::is_opaque (core/paint/paint.cpp:9-19): every stop opaque, at least one stop.
        op = ((self.path_rgba >> 24) == 255).astype(np.uint8)
        if self.path_gradient is not None:
            for p in np.nonzero(self.path_gradient >= 0)[0]:
                gi = int(self.path_gradient[p])
                g = self.gradients[gi]
                a, n = int(g["stop_start"]), int(g["stop_count"])
                if n < 0:  # a given ramp: opaque when every texel is
                    op[p] = 1 if ((self.gradient_ramps[gi] >> 24) == 255).all() else 0
                else:
                    op[p] = 1 if n > 0 and ((self.gradient_stop_rgba[a:a + n] >> 24) == 255).all() else 0
        return op

    def subset(self, paths) -> "Scene":
        """The scene restricted to the given draw paths (in the given order); clip paths are dropped."""
        assert self.path_clip is None or (self.path_clip < 0).all(), "subset of a scene with clipped paths"
        paths = np.asarray(paths, np.int64)
        pts, flg, cst, ccl, pcs = [], [], [0], [], [0]
        for p in paths:
            for c in range(self.path_contour_start[p], self.path_contour_start[p + 1]):
                a, b = self.contour_start[c], self.contour_start[c + 1]
                pts.append(self.points[a:b])
                flg.append(self.point_flags[a:b])
                cst.append(cst[-1] + (b - a))
                ccl.append(self.contour_closed[c])
            pcs.append(len(ccl))
        return Scene(self.width, self.height,
                     np.concatenate(pts) if pts else np.zeros((0, 2), np.float32),
                     np.concatenate(flg) if flg else np.zeros(0, np.uint8),
                     np.array(cst), np.array(ccl, np.uint8), np.array(pcs),
                     self.path_fill_rule[paths], self.path_rgba[paths], self.view_box.copy())


# pfh_gradient (include/pfh.h): kind (0 linear, 1 radial), wrap (0 clamp, 1 repeat), line (from, to),
# radii, radial transform (m11, m21, m12, m22, m13, m23), stop range
GRADIENT = np.dtype([("kind", np.int32), ("wrap", np.int32), ("line", np.float32, 4), ("radii", np.float32, 2),
                     ("transform", np.float32, 6), ("stop_start", np.int32), ("stop_count", np.int32)])


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed & ((1 << 64) - 1)))


def _pack_rgba(r, g, b, a) -> np.ndarray:
    return (r.astype(np.uint32) | (g.astype(np.uint32) << 8) | (b.astype(np.uint32) << 16)
            | (a.astype(np.uint32) << 24))


def cubic_scene(seed: int = BASE_SEED ^ 0, size: int = 512, n_paths: int = 64, n_cubics: int = 8) -> Scene:
    """move_to + n_cubics x cubic_to with every point uniform in the canvas, then close_path."""
    rng = _rng(seed)
    npp = 1 + 3 * n_cubics
    pts = rng.uniform(0.0, size, size=(n_paths, npp, 2)).astype(np.float32)
    flags = np.zeros(npp, np.uint8)
    flags[1::3] = 1
    flags[2::3] = 2
    rgb = rng.integers(0, 256, size=(n_paths, 3))
    rgba = _pack_rgba(rgb[:, 0], rgb[:, 1], rgb[:, 2], np.full(n_paths, 255))
    return Scene(size, size, pts.reshape(-1, 2), np.tile(flags, n_paths),
                 np.arange(n_paths + 1, dtype=np.int32) * npp, np.ones(n_paths, np.uint8),
                 np.arange(n_paths + 1, dtype=np.int32), np.zeros(n_paths, np.uint8), rgba)


def blob_scene(seed: int, width: int = 4096, height: int = 4096, n_paths: int = 10000,
               r_min: float = 4.0, r_max: float = 512.0, seg_min: int = 4, seg_max: int = 64,
               alpha_min: float = 0.2) -> Scene:
    """Closed star-shaped blobs: n vertices at sorted random angles, radius r*U(0.6,1);
    each edge a line, quadratic or cubic with control points at radius r*U(0.5,1.2)
    between the edge's end angles."""
    rng = _rng(seed)
    P = int(n_paths)
    cx = rng.uniform(0, width, P)
    cy = rng.uniform(0, height, P)
    r = np.exp(rng.uniform(math.log(r_min), math.log(max(r_max, r_min)), P))
    n = rng.integers(seg_min, seg_max + 1, P)
    alpha = np.rint(255.0 * rng.uniform(alpha_min, 1.0, P)).astype(np.int64)
    rgb = rng.integers(0, 256, size=(P, 3))
    rule = rng.integers(0, 2, P).astype(np.uint8)

    V = int(n.sum())
    pid = np.repeat(np.arange(P), n)
    vstart = np.concatenate([[0], np.cumsum(n)[:-1]])
    ang = rng.uniform(0, 2 * math.pi, V)
    order = np.lexsort((ang, pid))
    ang = ang[order]
    vrad = r[pid] * rng.uniform(0.6, 1.0, V)
    vx = cx[pid] + vrad * np.cos(ang)
    vy = cy[pid] + vrad * np.sin(ang)
    kind = rng.integers(0, 3, V)  # 0 line, 1 quad, 2 cubic: edge from vertex i to i+1
    local = np.arange(V) - vstart[pid]
    nxt = np.where(local + 1 == n[pid], vstart[pid], np.arange(V) + 1)
    ang_next = np.where(local + 1 == n[pid], ang[nxt] + 2 * math.pi, ang[nxt])
    f = np.sort(rng.uniform(0, 1, (V, 2)), axis=1)
    crad = r[pid][:, None] * rng.uniform(0.5, 1.2, (V, 2))
    ca = ang[:, None] + f * (ang_next - ang)[:, None]
    cpx = cx[pid][:, None] + crad * np.cos(ca)
    cpy = cy[pid][:, None] + crad * np.sin(ca)
    quad_single = ang + 0.5 * (ang_next - ang)
    qx = cx[pid] + crad[:, 0] * np.cos(quad_single)
    qy = cy[pid] + crad[:, 0] * np.sin(quad_single)

    is_last = local + 1 == n[pid]
    # The final edge back to vertex 0: a line is left to the implicit close
    # (contour.cpp:155-170); a curve is written out and ends on a copy of vertex 0,
    # so the closing line is skipped by the approx_eq test.
    last_curve = is_last & (kind != 0)
    npts_v = 1 + kind + last_curve.astype(np.int64)
    off = np.concatenate([[0], np.cumsum(npts_v)])
    N = int(off[-1])
    X = np.empty(N)
    Y = np.empty(N)
    F = np.zeros(N, np.uint8)
    o = off[:-1]
    X[o] = vx
    Y[o] = vy
    q = kind == 1
    X[o[q] + 1] = qx[q]
    Y[o[q] + 1] = qy[q]
    F[o[q] + 1] = 1
    c = kind == 2
    X[o[c] + 1] = cpx[c, 0]
    Y[o[c] + 1] = cpy[c, 0]
    F[o[c] + 1] = 1
    X[o[c] + 2] = cpx[c, 1]
    Y[o[c] + 2] = cpy[c, 1]
    F[o[c] + 2] = 2
    lc = last_curve
    endi = o[lc] + 1 + kind[lc]
    X[endi] = vx[vstart[pid[lc]]]
    Y[endi] = vy[vstart[pid[lc]]]
    pts = np.stack([X, Y], axis=1).astype(np.float32)
    # Path point ranges: vertices of path p are contiguous.
    path_pt_start = off[np.concatenate([vstart, [V]])]
    rgba = _pack_rgba(rgb[:, 0], rgb[:, 1], rgb[:, 2], alpha)
    return Scene(width, height, pts, F, path_pt_start.astype(np.int32), np.ones(P, np.uint8),
                 np.arange(P + 1, dtype=np.int32), rule, rgba)


def single_scene_4k(seed: int = BASE_SEED ^ 1, n_paths: int = 10000, size: int = 4096) -> Scene:
    """The 4096x4096, 10 000-path blob workload."""
    return blob_scene(seed, size, size, n_paths)


def batch_scenes(n_scenes: int = 256, seed: int = BASE_SEED ^ 3, side_min: int = 256, side_max: int = 4096,
                 paths_min: int = 10, paths_max: int = 20000):
    """Scenes with side log-uniform in [side_min, side_max] and path count log-uniform in
    [paths_min, paths_max]; each scene k is blob_scene(seed + k + 1) with radii
    log-uniform in [4, min(512, side/2)]."""
    rng = _rng(seed)
    sides = np.rint(np.exp(rng.uniform(math.log(side_min), math.log(side_max), n_scenes))).astype(int)
    counts = np.rint(np.exp(rng.uniform(math.log(paths_min), math.log(paths_max), n_scenes))).astype(int)
    return [blob_scene(seed + k + 1, int(sides[k]), int(sides[k]), int(counts[k]),
                       r_max=min(512.0, sides[k] / 2.0)) for k in range(n_scenes)]


def clip_scene(seed: int, size: int = 1024, n_draw: int = 200, n_clip: int = 8, clipped: float = 0.6,
               r_max: float = 256.0) -> Scene:
    """Blob draw paths, the last n_clip blobs of the generator serving as clip paths, each draw path
    clipped by a random one of them with probability `clipped`."""
    base = blob_scene(seed, size, size, n_draw + n_clip, r_max=r_max)
    rng = _rng(seed ^ 0x5EED)
    pc = rng.integers(0, n_clip, n_draw).astype(np.int32)
    pc[rng.random(n_draw) >= clipped] = -1
    base.n_clip_paths = n_clip
    base.path_clip = np.concatenate([pc, np.full(n_clip, -1, np.int32)])
    return base


def concat_paths(a: Scene, b: Scene) -> Scene:
    """The paths of a followed by those of b, on a's canvas (stroke styles carried, clip information not)."""
    strokes, dashes = None, None
    if a.path_stroke is not None or b.path_stroke is not None:
        def cols7(sc):
            ps = sc.path_stroke if sc.path_stroke is not None else np.zeros((sc.n_paths, 4), np.float32)
            return np.concatenate([ps, np.zeros((ps.shape[0], 7 - ps.shape[1]), np.float32)], axis=1)

        sa, sb = cols7(a), cols7(b)
        da = a.stroke_dashes if a.stroke_dashes is not None else np.zeros(0, np.float32)
        db = b.stroke_dashes if b.stroke_dashes is not None else np.zeros(0, np.float32)
        sb[:, 4] += np.where(sb[:, 5] > 0, da.shape[0], 0)
        strokes = np.concatenate([sa, sb])
        dashes = np.concatenate([da, db]) if da.shape[0] + db.shape[0] else None
    return Scene(a.width, a.height, np.concatenate([a.points, b.points]),
                 np.concatenate([a.point_flags, b.point_flags]),
                 np.concatenate([a.contour_start, b.contour_start[1:] + a.contour_start[-1]]),
                 np.concatenate([a.contour_closed, b.contour_closed]),
                 np.concatenate([a.path_contour_start, b.path_contour_start[1:] + a.path_contour_start[-1]]),
                 np.concatenate([a.path_fill_rule, b.path_fill_rule]), np.concatenate([a.path_rgba, b.path_rgba]),
                 a.view_box.copy(), path_stroke=strokes, stroke_dashes=dashes)


def stress_clip_scene(seed: int = BASE_SEED ^ 4, size: int = 4096, n_stars: int = 2000, n_max: int = 2000,
                      n_clip: int = 200, clipped: float = 0.5) -> Scene:
    """The stress workload's star polygons (heavy overlap, windings far above 1) under clip paths:
    n_clip blob outlines as clip paths, each star clipped by a random one with probability
    `clipped` (single-level clips; strokes reach the rasterizer as outlines like any path)."""
    stars = star_scene(seed=seed, size=size, n_stars=n_stars, n_max=n_max)
    clips = blob_scene(seed ^ 0xC11, size, size, n_clip, r_max=size / 4)
    sc = concat_paths(stars, clips)
    rng = _rng(seed ^ 0xC12)
    pc = rng.integers(0, n_clip, n_stars).astype(np.int32)
    pc[rng.random(n_stars) >= clipped] = -1
    sc.n_clip_paths = n_clip
    sc.path_clip = np.concatenate([pc, np.full(n_clip, -1, np.int32)])
    return sc


def star_clip_scene(seed: int, size: int = 2048, n_stars: int = 300, n_clip: int = 40, n_max: int = 400,
                    clipped: float = 0.6, clip_rule: int = 1) -> Scene:
    """Star polygons clipped by star polygons: windings far above 1 under both fill rules at every
    level, where the clip paths' own rules (clip_rule 1) and the reference's |winding| differ most.
    The clip stars form nesting chains (each, by a coin, clipped by the one before it, to depth 4);
    each draw star is clipped by a random clip star with probability `clipped`."""
    stars = star_scene(seed=seed, size=size, n_stars=n_stars, n_max=n_max)
    clips = star_scene(seed=seed ^ 0xC11, size=size, n_stars=n_clip, n_min=20, n_max=200, m_max=12,
                       r_min=size / 16, r_max=size / 3)
    sc = concat_paths(stars, clips)
    rng = _rng(seed ^ 0xC12)
    pc = rng.integers(0, n_clip, n_stars).astype(np.int32)
    pc[rng.random(n_stars) >= clipped] = -1
    parent, depth = np.full(n_clip, -1, np.int32), np.ones(n_clip, np.int64)
    for q in range(1, n_clip):
        if depth[q - 1] < 4 and rng.integers(0, 2):
            parent[q], depth[q] = q - 1, depth[q - 1] + 1
    sc.n_clip_paths = n_clip
    sc.path_clip = np.concatenate([pc, parent])
    sc.clip_rule = clip_rule
    return sc


def star_scene(seed: int = BASE_SEED ^ 2, size: int = 4096, n_stars: int = 2000, n_min: int = 100,
               n_max: int = 2000, m_max: int = 50, r_min: float = 16.0, r_max: float = 512.0) -> Scene:
    """Star polygons {n/m} with gcd(n, m) = 1: centre winding +-m (sign by orientation)."""
    rng = _rng(seed)
    pts, flags, cst, rules, rgba = [], [], [0], [], []
    for _ in range(n_stars):
        while True:
            n = int(rng.integers(n_min, n_max + 1))
            m = int(rng.integers(1, m_max + 1))
            if math.gcd(n, m) == 1 and 2 * m < n:
                break
        cx, cy = rng.uniform(0, size, 2)
        r = math.exp(rng.uniform(math.log(r_min), math.log(r_max)))
        orient = 1 if rng.integers(0, 2) else -1
        k = np.arange(n)
        a = orient * 2 * math.pi * ((k * m) % n) / n
        p = np.stack([cx + r * np.cos(a), cy + r * np.sin(a)], 1)
        pts.append(p)
        flags.append(np.zeros(n, np.uint8))
        cst.append(cst[-1] + n)
        rules.append(int(rng.integers(0, 2)))
        c = rng.integers(0, 256, 3)
        rgba.append(int(c[0]) | int(c[1]) << 8 | int(c[2]) << 16 | int(round(255 * rng.uniform(0.2, 1.0))) << 24)
    P = n_stars
    return Scene(size, size, np.concatenate(pts).astype(np.float32), np.concatenate(flags), np.array(cst),
                 np.ones(P, np.uint8), np.arange(P + 1), np.array(rules, np.uint8), np.array(rgba, np.uint32))


def stats(scene: Scene) -> dict:
    f = scene.point_flags
    return {"paths": scene.n_paths, "points": int(f.shape[0]), "contours": int(scene.contour_closed.shape[0]),
            "canvas": [scene.width, scene.height]}


def cubic_polylines(seed: int, size: int = 4096, n: int = 500, seg_min: int = 4, seg_max: int = 16,
                    step_min: float = 20.0, step_max: float = 200.0):
    """Open polylines of seg_min..seg_max cubic Béziers each: a random walk from a uniform start point,
    each cubic's end a step of length U(step_min, step_max) in a direction that turns by up to 90
    degrees, its controls at 1/3 and 2/3 of the step pushed off the chord by up to half its length.
    Returns outlines (points [N, 2], flags, contour_start [2], contour_closed [1])."""
    rng = _rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.integers(seg_min, seg_max + 1))
        p = rng.uniform(0, size, 2)
        ang = rng.uniform(0, 2 * math.pi)
        pts, flg = [p], [0]
        for _ in range(k):
            ang += rng.uniform(-math.pi / 2, math.pi / 2)
            step = rng.uniform(step_min, step_max)
            d = np.array([math.cos(ang), math.sin(ang)])
            nrm = np.array([-d[1], d[0]])
            q = p + step * d
            c0 = p + d * step / 3 + nrm * rng.uniform(-0.5, 0.5) * step
            c1 = p + d * 2 * step / 3 + nrm * rng.uniform(-0.5, 0.5) * step
            pts += [c0, c1, q]
            flg += [1, 2, 0]
            p = q
        out.append((np.array(pts, np.float32), np.array(flg, np.uint8), np.array([0, len(flg)], np.int32),
                    np.zeros(1, np.uint8)))
    return out


def stroked_polylines(seed: int, size: int = 4096, n: int = 500, width_min: float = 1.0, width_max: float = 20.0):
    """cubic_polylines with a stroke style each: width U(width_min, width_max), cap and join drawn
    uniformly from butt / square / round and miter / bevel / round, miter limit 10 (StrokeStyle's
    default, core/stroke.h:12-26).  Returns [(outline, (width, cap, join, miter_limit))]."""
    lines = cubic_polylines(seed, size, n)
    rng = _rng(seed ^ 0x57)
    return [(o, (float(rng.uniform(width_min, width_max)), int(rng.integers(0, 3)), int(rng.integers(0, 3)), 10.0))
            for o in lines]


def outlines_scene(width, height, outlines, rules, rgba) -> Scene:
    """One path per outline (points, flags, contour_start, contour_closed)."""
    pts, flg, cst, ccl, pcs = [], [], [0], [], [0]
    for p, f, cs, cc in outlines:
        pts.append(np.asarray(p, np.float32).reshape(-1, 2))
        flg.append(np.asarray(f, np.uint8))
        cst.extend((cst[-1] + np.asarray(cs[1:], np.int64)).tolist())
        ccl.extend(np.asarray(cc, np.uint8).tolist())
        pcs.append(len(ccl))
    return Scene(width, height, np.concatenate(pts) if pts else np.zeros((0, 2), np.float32),
                 np.concatenate(flg) if flg else np.zeros(0, np.uint8), np.array(cst), np.array(ccl, np.uint8),
                 np.array(pcs), np.asarray(rules, np.uint8), np.asarray(rgba, np.uint32))


def stroke_scene(seed: int, size: int = 4096, n: int = 500) -> Scene:
    """n stroked cubic polylines (stroked_polylines): the open polylines as the paths' outlines with
    their stroke styles (path_stroke: stroked at upload, the stroke filled with the nonzero rule, as
    Canvas::stroke_path pushes it, core/canvas.cpp:272-301), colours with alpha U(0.2, 1)."""
    lines = stroked_polylines(seed, size, n)
    rng = _rng(seed ^ 0xC0)
    rgb = rng.integers(0, 256, size=(n, 3))
    alpha = np.rint(255.0 * rng.uniform(0.2, 1.0, n)).astype(np.int64)
    sc = outlines_scene(size, size, [o for o, _ in lines], np.zeros(n, np.uint8),
                        _pack_rgba(rgb[:, 0], rgb[:, 1], rgb[:, 2], alpha))
    sc.path_stroke = np.array([st for _, st in lines], np.float32)
    return sc


def blend_scene(seed: int, size: int = 1024, n: int = 200, painted: float = 0.3, src_over: float = 0.3) -> Scene:
    """gradient_scene (a fraction `painted` of the paths with gradient paints) whose paths take blend
    modes: src-over with probability `src_over`, else one of the 27 BlendMode values uniformly;
    a third of the solid paths made opaque so that the occlusion rules matter."""
    sc = gradient_scene(seed, size, n, painted)
    rng = _rng(seed ^ 0xB1E)
    pb = rng.integers(0, 27, n).astype(np.uint8)
    pb[rng.random(n) < src_over] = 4
    sc.path_blend = pb
    opaque = rng.random(n) < 0.33
    sc.path_rgba = np.where(opaque, sc.path_rgba | np.uint32(0xff000000), sc.path_rgba).astype(np.uint32)
    return sc


def gradient_scene(seed: int, size: int = 1024, n: int = 200, painted: float = 0.6) -> Scene:
    """blob_scene with a fraction `painted` of the paths given gradient paints: half linear (a line
    of length U(size/16, size) through the blob), half radial (focal circle inside the outer one,
    start radius 0 or U(0, r/2), outer radius U(r/2, 2r), a rotation-and-scale transform on a third
    of them), clamp or repeat, 1-5 stops at offsets in [0, 1] (sometimes equal, sometimes out of
    order), opaque stops for half of the gradients."""
    sc = blob_scene(seed, size, size, n, r_max=size / 6)
    rng = _rng(seed ^ 0x6EAD)
    pg = np.full(n, -1, np.int32)
    grads, offs, cols = [], [], []
    for p in range(n):
        if rng.random() >= painted:
            continue
        c0, c1 = int(sc.path_contour_start[p]), int(sc.path_contour_start[p + 1])
        pts = sc.points[int(sc.contour_start[c0]):int(sc.contour_start[c1])]
        cx, cy = pts.mean(axis=0)
        r = float(np.abs(pts - [cx, cy]).max()) + 1.0
        g = np.zeros((), GRADIENT)
        g["kind"] = int(rng.integers(0, 2))
        g["wrap"] = int(rng.integers(0, 2))
        g["transform"] = (1, 0, 0, 1, 0, 0)
        if g["kind"] == 0:
            ang, ln = rng.uniform(0, 2 * math.pi), rng.uniform(size / 16, size)
            x0, y0 = cx - 0.5 * ln * math.cos(ang), cy - 0.5 * ln * math.sin(ang)
            g["line"] = (x0, y0, x0 + ln * math.cos(ang), y0 + ln * math.sin(ang))
        else:
            r1 = rng.uniform(0.5 * r, 2.0 * r)
            r0 = 0.0 if rng.random() < 0.5 else rng.uniform(0.0, 0.5 * r1)
            fx, fy = rng.uniform(-0.3, 0.3, 2) * r1
            if rng.random() < 1 / 3:
                th, sx = rng.uniform(0, 2 * math.pi), rng.uniform(0.5, 2.0)
                ct, st = math.cos(th) * sx, math.sin(th) * sx
                g["transform"] = (ct, st, -st, ct, cx, cy)  # rotate and scale about the blob's centre
                g["line"] = (0.0, 0.0, fx, fy)
            else:
                g["line"] = (cx, cy, cx + fx, cy + fy)
            g["radii"] = (r0, r1)
        k = int(rng.integers(1, 6))
        o = rng.uniform(0, 1, k)
        if k > 2 and rng.random() < 0.3:
            o[1] = o[2]  # equal offsets
        if rng.random() < 0.5:
            o = np.sort(o)
        opaque = rng.random() < 0.5
        a = np.full(k, 255) if opaque else rng.integers(0, 256, k)
        rgb = rng.integers(0, 256, (k, 3))
        g["stop_start"], g["stop_count"] = len(offs), k
        offs.extend(o.astype(np.float32).tolist())
        cols.extend(_pack_rgba(rgb[:, 0], rgb[:, 1], rgb[:, 2], a).tolist())
        pg[p] = len(grads)
        grads.append(g)
    sc.path_gradient = pg
    sc.gradients = np.array(grads, GRADIENT) if grads else np.zeros(0, GRADIENT)
    sc.gradient_stop_offsets = np.array(offs, np.float32)
    sc.gradient_stop_rgba = np.array(cols, np.uint32)
    return sc


def dashed_stroke_scene(seed: int, size: int = 4096, n: int = 500, dashed: float = 0.5) -> Scene:
    """stroke_scene with a fraction `dashed` of the paths dashed before stroking (Canvas::stroke_path,
    core/canvas.cpp:284-289): each such path gets a pattern of 2 or 4 lengths U(2, 40) px and an
    offset U(0, 30) px (OutlineDash's offset argument)."""
    sc = stroke_scene(seed, size, n)
    rng = _rng(seed ^ 0xDA5)
    styles = np.concatenate([sc.path_stroke, np.zeros((n, 3), np.float32)], axis=1)
    dashes = []
    for p in range(n):
        if rng.random() < dashed:
            k = int(rng.choice([2, 4]))
            styles[p, 4], styles[p, 5], styles[p, 6] = len(dashes), k, rng.uniform(0.0, 30.0)
            dashes.extend(rng.uniform(2.0, 40.0, k).tolist())
    sc.path_stroke = styles
    sc.stroke_dashes = np.array(dashes, np.float32)
    return sc


def _circle(cx, cy, r):
    """A closed circle of 4 cubic arcs (kappa 0.5523), clockwise in y-down coordinates."""
    k = 0.5522847498 * r
    p = [(cx + r, cy), (cx + r, cy + k), (cx + k, cy + r), (cx, cy + r), (cx - k, cy + r), (cx - r, cy + k),
         (cx - r, cy), (cx - r, cy - k), (cx - k, cy - r), (cx, cy - r), (cx + k, cy - r), (cx + r, cy - k),
         (cx + r, cy)]
    return (np.array(p, np.float32), np.array([0, 1, 2] * 4 + [0], np.uint8), np.array([0, 13], np.int32),
            np.ones(1, np.uint8))


def _rect(x0, y0, x1, y1):
    return (np.array([(x0, y0), (x1, y0), (x1, y1), (x0, y1)], np.float32), np.zeros(4, np.uint8),
            np.array([0, 4], np.int32), np.ones(1, np.uint8))


def nested_clips(seed: int, size: int = 4096, n_clip: int = 200, depth_max: int = 4):
    """Clip paths in nesting chains: chains of depth U{1..depth_max} until n_clip paths exist; a
    chain's first clip is a circle or a rect (a coin) of radius log-U[64, size/4] around a uniform
    centre, each further one a circle or rect around a point inside its parent, 0.5-0.95 of its
    size.  Returns (outlines, parent index among the clips or -1, chain of every clip)."""
    rng = _rng(seed)
    outlines, parent, chain = [], [], []
    while len(outlines) < n_clip:
        d = min(int(rng.integers(1, depth_max + 1)), n_clip - len(outlines))
        cx, cy = rng.uniform(0, size, 2)
        r = math.exp(rng.uniform(math.log(64.0), math.log(size / 4)))
        prev = -1
        cid = len(chain) and (max(chain) + 1)
        for _ in range(d):
            if rng.integers(0, 2):
                outlines.append(_circle(cx, cy, r))
            else:
                a = r * rng.uniform(0.6, 1.0)
                outlines.append(_rect(cx - r, cy - a, cx + r, cy + a))
            parent.append(prev)
            chain.append(cid)
            prev = len(outlines) - 1
            cx += rng.uniform(-0.4, 0.4) * r
            cy += rng.uniform(-0.4, 0.4) * r
            r *= rng.uniform(0.5, 0.95)
    return outlines, np.array(parent, np.int32), np.array(chain, np.int32)


def stress_scene(seed: int = BASE_SEED ^ 2, size: int = 4096, n_stars: int = 2000, n_strokes: int = 500,
                 n_clip: int = 200, clipped: float = 0.5, n_max: int = 2000) -> Scene:
    """The stress workload: n_stars star polygons {n/m} of 100..n_max vertices (star_scene), then n_strokes
    stroked cubic polylines (stroke_scene), then n_clip clip paths in nesting chains of depth 1..4
    (nested_clips).  Each draw path is clipped with probability `clipped` by a clip drawn uniformly
    from all clip paths (so by chains of every depth)."""
    stars = star_scene(seed=seed, size=size, n_stars=n_stars, n_max=n_max)
    strokes = stroke_scene(seed ^ 0x5EED5, size=size, n=n_strokes)
    clips, parent, _ = nested_clips(seed ^ 0xC11, size=size, n_clip=n_clip)
    rng = _rng(seed ^ 0xC12)
    crgb = rng.integers(0, 256, size=(n_clip, 3))
    clip_sc = outlines_scene(size, size, clips, rng.integers(0, 2, n_clip),
                             _pack_rgba(crgb[:, 0], crgb[:, 1], crgb[:, 2], np.full(n_clip, 255)))
    sc = concat_paths(concat_paths(stars, strokes), clip_sc)
    n_draw = n_stars + n_strokes
    pc = rng.integers(0, n_clip, n_draw).astype(np.int32)
    pc[rng.random(n_draw) >= clipped] = -1
    sc.n_clip_paths = n_clip
    sc.path_clip = np.concatenate([pc, parent])
    return sc
