"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg).  Python bindings for

  * oracle/liboracle.so   — the C++ restatement of the reference's hybrid CPU
                            tiler plus the FP64 coverage / composite model;
  * oracle/_ref/libpfref.so — the reference's own Tiler compiled from its
                            unchanged sources (present only where it was built).

Both return the same canonical form (see ``canonical`` below).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libpfref.so")

_P = ctypes.c_void_p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


_oracle = None
_ref = None


def build():
    subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True, capture_output=True)


def load_oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            build()
        lib = ctypes.CDLL(ORACLE_SO)
        lib.pfo_run.restype = _P
        lib.pfo_run.argtypes = [ctypes.c_int] + [_P] * 8 + [ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, _P]
        lib.pfo_fetch_clip.argtypes = [_P] * 4
        lib.pfo_apply_rule_array.argtypes = [ctypes.c_int64, _P, _P, _P]
        lib.pfo_counts.argtypes = [_P, _P]
        lib.pfo_fetch.argtypes = [_P] * 17
        lib.pfo_free.argtypes = [_P]
        lib.pfo_path_lines.argtypes = [_P, _P]
        lib.pfo_tiling_seconds.restype = ctypes.c_double
        lib.pfo_tiling_seconds.argtypes = [_P]
        lib.pfo_fill_area.restype = ctypes.c_double
        lib.pfo_fill_area.argtypes = [_P, ctypes.c_int, ctypes.c_int]
        lib.pfo_coverage.argtypes = [ctypes.c_int64, _P, _P, _P, _P, _P]
        lib.pfo_half_value.restype = ctypes.c_double
        lib.pfo_half_value.argtypes = [ctypes.c_float]
        lib.pfo_gradient_ramp.argtypes = [_P, _P, ctypes.c_int, _P]
        lib.pfo_gradient_record.restype = ctypes.c_int
        lib.pfo_gradient_record.argtypes = [ctypes.c_int, ctypes.c_int, _P, _P, _P, ctypes.c_int, _P]
        lib.pfo_paint_colour.restype = ctypes.c_int
        lib.pfo_paint_colour.argtypes = [_P, _P, ctypes.c_float, ctypes.c_float, _P]
        lib.pfo_composite_paint.argtypes = [ctypes.c_int64] + [_P] * 10 + [ctypes.c_int] * 3 + [_P, _P] + [ctypes.c_int] * 2
        lib.pfo_composite_op.argtypes = [_P, _P, ctypes.c_int, _P]
        lib.pfo_coverage_raw.argtypes = [ctypes.c_int64, _P, _P, _P, _P]
        lib.pfo_apply_rule.restype = ctypes.c_double
        lib.pfo_apply_rule.argtypes = [ctypes.c_double, ctypes.c_int]
        lib.pfo_composite.argtypes = [ctypes.c_int64] + [_P] * 7 + [ctypes.c_int] * 4 + [_P] + [ctypes.c_int] * 2
        lib.pfo_stroke_outline.argtypes = [_P, _P, _P, _P, ctypes.c_int64, _P, _P, ctypes.c_int64, _P, _P]
        lib.pfo_dash_outline.argtypes = [_P, _P, ctypes.c_int32, ctypes.c_float, _P, _P, ctypes.c_int64, _P, _P,
                                         ctypes.c_int64, _P, _P]
        lib.pfo_raster_cpu.restype = ctypes.c_double
        lib.pfo_raster_cpu.argtypes = [_P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def load_ref():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_SO)
        lib.pfref_run.restype = _P
        lib.pfref_run.argtypes = [ctypes.c_int] + [_P] * 8 + [ctypes.c_int, _P, ctypes.c_int, _P, _P]
        lib.pfref_fetch_clip.argtypes = [_P] * 3
        lib.pfref_counts.argtypes = [_P, _P]
        lib.pfref_fetch.argtypes = [_P] * 14
        lib.pfref_free.argtypes = [_P]
        lib.pfref_time.restype = ctypes.c_double
        lib.pfref_time.argtypes = [ctypes.c_int] + [_P] * 7 + [ctypes.c_int, ctypes.c_int]
        lib.pfref_stroke.restype = _P
        lib.pfref_stroke.argtypes = [_P] * 4 + [ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_float]
        lib.pfref_stroke_counts.argtypes = [_P, _P]
        lib.pfref_stroke_fetch.argtypes = [_P] * 5
        lib.pfref_stroke_free.argtypes = [_P]
        lib.pfref_dash.restype = _P
        lib.pfref_dash.argtypes = [_P] * 4 + [ctypes.c_int, _P, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_float]
        lib.pfref_gradient_ramp.argtypes = [_P, _P, ctypes.c_int, _P]
        lib.pfref_half_bits.argtypes = [_P, ctypes.c_int64, _P]
        lib.pfref_transform_inverse.argtypes = [_P, _P]
        lib.pfref_svg.restype = _P
        lib.pfref_svg.argtypes = [ctypes.c_char_p, ctypes.c_int64]
        lib.pfref_svg_counts.argtypes = [_P, _P]
        lib.pfref_svg_fetch.argtypes = [_P] * 15
        lib.pfref_svg_free.argtypes = [_P]
        lib.pfref_svg_gradient_counts.argtypes = [_P, _P]
        lib.pfref_svg_gradient_fetch.argtypes = [_P] * 6
        _ref = lib
    return _ref


class _Outline(ctypes.Structure):
    _fields_ = [("n_contours", ctypes.c_int32), ("n_points", ctypes.c_int32), ("points", _P), ("point_flags", _P),
                ("contour_start", _P), ("contour_closed", _P)]


class _Style(ctypes.Structure):
    _fields_ = [("line_width", ctypes.c_float), ("line_cap", ctypes.c_int32), ("line_join", ctypes.c_int32),
                ("miter_limit", ctypes.c_float)]


def stroke(points, point_flags, contour_start, contour_closed, width, cap=0, join=0, miter_limit=10.0):
    """The oracle's stroke-to-fill (oracle/stroke_oracle.cpp, a restatement of OutlineStrokeToFill,
    core/stroke.cpp:124-195) on one outline: (points [N, 2] f32, point_flags u8, contour_start i32,
    contour_closed u8)."""
    lib = load_oracle()
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 2)
    flg = np.ascontiguousarray(point_flags, np.uint8)
    cst = np.ascontiguousarray(contour_start, np.int32)
    ccl = np.ascontiguousarray(contour_closed, np.uint8)
    o = _Outline(int(ccl.shape[0]), int(flg.shape[0]), _ptr(pts), _ptr(flg), _ptr(cst), _ptr(ccl))
    st = _Style(float(width), int(cap), int(join), float(miter_limit))
    n_pts, n_ct = np.zeros(1, np.int64), np.zeros(1, np.int64)
    cs0 = np.zeros(1, np.int32)
    rc = lib.pfo_stroke_outline(ctypes.byref(o), ctypes.byref(st), None, None, 0, _ptr(cs0), None, 0, _ptr(n_pts),
                                _ptr(n_ct))
    if rc not in (0, -5):
        raise ValueError(f"pfo_stroke_outline failed ({rc})")
    out = (np.zeros((n_pts[0], 2), np.float32), np.zeros(n_pts[0], np.uint8), np.zeros(n_ct[0] + 1, np.int32),
           np.zeros(n_ct[0], np.uint8))
    rc = lib.pfo_stroke_outline(ctypes.byref(o), ctypes.byref(st), _ptr(out[0]), _ptr(out[1]), int(n_pts[0]),
                                _ptr(out[2]), _ptr(out[3]), int(n_ct[0]), _ptr(n_pts), _ptr(n_ct))
    if rc != 0:
        raise ValueError(f"pfo_stroke_outline failed ({rc})")
    return out


def dash(points, point_flags, contour_start, contour_closed, dashes, offset=0.0):
    """The oracle's dasher (oracle/stroke_oracle.cpp, a restatement of OutlineDash, core/dash.cpp:9-126)
    on one outline: its dashes as open contours (points [N, 2] f32, point_flags u8, contour_start i32,
    contour_closed u8)."""
    lib = load_oracle()
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 2)
    flg = np.ascontiguousarray(point_flags, np.uint8)
    cst = np.ascontiguousarray(contour_start, np.int32)
    ccl = np.ascontiguousarray(contour_closed, np.uint8)
    dsh = np.ascontiguousarray(dashes, np.float32)
    o = _Outline(int(ccl.shape[0]), int(flg.shape[0]), _ptr(pts), _ptr(flg), _ptr(cst), _ptr(ccl))
    n_pts, n_ct = np.zeros(1, np.int64), np.zeros(1, np.int64)
    cs0 = np.zeros(1, np.int32)
    rc = lib.pfo_dash_outline(ctypes.byref(o), _ptr(dsh), int(dsh.shape[0]), float(offset), None, None, 0, _ptr(cs0),
                              None, 0, _ptr(n_pts), _ptr(n_ct))
    if rc not in (0, -5):
        raise ValueError(f"pfo_dash_outline failed ({rc})")
    out = (np.zeros((n_pts[0], 2), np.float32), np.zeros(n_pts[0], np.uint8), np.zeros(n_ct[0] + 1, np.int32),
           np.zeros(n_ct[0], np.uint8))
    rc = lib.pfo_dash_outline(ctypes.byref(o), _ptr(dsh), int(dsh.shape[0]), float(offset), _ptr(out[0]), _ptr(out[1]),
                              int(n_pts[0]), _ptr(out[2]), _ptr(out[3]), int(n_ct[0]), _ptr(n_pts), _ptr(n_ct))
    if rc != 0:
        raise ValueError(f"pfo_dash_outline failed ({rc})")
    return out


def stroke_styles_array(ps):
    """Rows (line_width, cap, join, miter_limit, dash_start, dash_count, dash_offset) of a scene's
    path_stroke (a structured array or rows of 4 or 7 numbers)."""
    ps = np.asarray(ps)
    if ps.dtype.names:
        ps = np.stack([ps[k].astype(np.float64) for k in ps.dtype.names], axis=1)
    ps = np.asarray(ps, np.float64)
    ps = ps.reshape(ps.shape[0], -1)
    if ps.shape[1] == 4:
        ps = np.concatenate([ps, np.zeros((ps.shape[0], 3))], axis=1)
    return ps


def apply_strokes(scene):
    """The scene with every stroked path (path_stroke width > 0) replaced by its stroked outline, filled
    with the nonzero rule (core/canvas.cpp:272-301); the scene itself when nothing is stroked."""
    ps = getattr(scene, "path_stroke", None)
    if ps is None:
        return scene
    import scenegen  # the Scene container only

    ps = stroke_styles_array(ps)
    dashes = np.asarray(getattr(scene, "stroke_dashes", None) if getattr(scene, "stroke_dashes", None) is not None
                        else np.zeros(0), np.float32)
    outlines, rules = [], scene.path_fill_rule.copy()
    for p in range(scene.n_paths):
        c0, c1 = int(scene.path_contour_start[p]), int(scene.path_contour_start[p + 1])
        a, b = int(scene.contour_start[c0]), int(scene.contour_start[c1])
        o = (scene.points[a:b], scene.point_flags[a:b], scene.contour_start[c0:c1 + 1] - a,
             scene.contour_closed[c0:c1])
        if ps[p, 0] > 0:
            if ps[p, 5] > 0:  # dashed before stroking (core/canvas.cpp:284-289)
                d0 = int(ps[p, 4])
                o = dash(*o, dashes[d0:d0 + int(ps[p, 5])], ps[p, 6])
            o = stroke(*o, ps[p, 0], int(ps[p, 1]), int(ps[p, 2]), ps[p, 3])
            rules[p] = 0
        outlines.append(o)
    out = scenegen.outlines_scene(scene.width, scene.height, outlines, rules, scene.path_rgba)
    out.view_box = scene.view_box.copy()
    out.n_clip_paths = scene.n_clip_paths
    out.path_clip = scene.path_clip
    out.clip_rule = int(getattr(scene, "clip_rule", 0))
    out.path_blend = getattr(scene, "path_blend", None)
    for k in ("path_gradient", "gradients", "gradient_stop_offsets", "gradient_stop_rgba", "gradient_ramps"):
        setattr(out, k, getattr(scene, k, None))
    return out


def _scene_args(scene):
    return [ctypes.c_int(scene.n_paths), _ptr(scene.path_contour_start), _ptr(scene.path_fill_rule),
            _ptr(scene.path_opaque), _ptr(scene.contour_start), _ptr(scene.contour_closed),
            _ptr(scene.points), _ptr(scene.point_flags), _ptr(scene.view_box)]


# BlendMode (core/paint/effects.h:56-123) -> the tile shader's composite op
# (blend_mode_to_composite_ctrl, effects.h:184-232): the Porter-Duff modes 0..11 draw over (0).
COMPOSITE_OP = np.zeros(27, np.uint8)
COMPOSITE_OP[12:] = [4, 5, 1, 2, 8, 3, 6, 7, 9, 10, 11, 12, 13, 14, 15]


def _blend_arg(scene):
    pb = getattr(scene, "path_blend", None)
    if pb is None:
        return None
    scene._blend_keep = np.ascontiguousarray(pb, np.uint8)
    return _ptr(scene._blend_keep)


def path_ops(scene):
    """Per path the composite op of its blend mode (clip paths: none), or None without blend modes."""
    pb = getattr(scene, "path_blend", None)
    if pb is None:
        return None
    ops = COMPOSITE_OP[np.asarray(pb, np.int64)].astype(np.uint8)
    ops[scene.n_paths - int(scene.n_clip_paths):] = 0
    return np.ascontiguousarray(ops)


def composite_op(dest, src, op: int) -> np.ndarray:
    """B(dest, src) of one composite op on single RGB colours in [0, 1] (the oracle's restatement)."""
    d = np.ascontiguousarray(dest, np.float64)
    s = np.ascontiguousarray(src, np.float64)
    out = np.zeros(3, np.float64)
    load_oracle().pfo_composite_op(_ptr(d), _ptr(s), ctypes.c_int(op), _ptr(out))
    return out


def _clip_args(scene):
    return [ctypes.c_int(scene.n_clip_paths), _ptr(scene.path_clip)]


def _alloc(counts):
    T, W, F, E, NFB, P = [int(x) for x in counts[:6]]
    return dict(
        path_rect=np.zeros((P, 4), np.int32), tile_off=np.zeros(P + 1, np.int64),
        col_off=np.zeros(P + 1, np.int64), fill_off=np.zeros(P + 1, np.int64),
        tile_backdrop=np.zeros(T, np.int8), tile_alpha=np.zeros(T, np.uint8), col_sum=np.zeros(W, np.int32),
        fill_seg=np.zeros((F, 4), np.uint16), fill_tile=np.zeros(F, np.uint32),
        entry_fbtile=np.zeros(E, np.uint32), entry_path=np.zeros(E, np.uint32), zbuf=np.zeros(NFB, np.uint32),
        fb_rect=np.zeros(4, np.int32))


CANONICAL_KEYS = ("path_rect", "tile_off", "col_off", "fill_off", "tile_backdrop", "tile_alpha", "col_sum",
                  "fill_seg", "fill_tile", "entry_fbtile", "entry_path", "zbuf", "fb_rect")


def run_oracle(scene, zmode: int = 1) -> dict:
    """Canonical tiling of ``scene`` by the oracle restatement.

    zmode 0: the reference's z-buffer rule; 1: even-odd tiles with an even
    backdrop never occlude (the rule the HIP path implements)."""
    lib = load_oracle()
    scene = apply_strokes(scene)
    h = lib.pfo_run(*_scene_args(scene), ctypes.c_int(zmode), *_clip_args(scene),
                    ctypes.c_int(int(getattr(scene, "clip_rule", 0))), _blend_arg(scene))
    if not h:
        raise ValueError("clip paths nest deeper than 4 or in a cycle")
    counts = np.zeros(8, np.int64)
    lib.pfo_counts(h, _ptr(counts))
    out = _alloc(counts)
    T, W, E = int(counts[0]), int(counts[1]), int(counts[3])
    out["tile_delta"] = np.zeros(T, np.int8)
    out["col_above"] = np.zeros(W, np.int32)
    out["entry_tile"] = np.zeros(E, np.int64)
    lib.pfo_fetch(h, *[_ptr(out[k]) for k in CANONICAL_KEYS], _ptr(out["tile_delta"]), _ptr(out["col_above"]),
                  _ptr(out["entry_tile"]))
    out["clip_combine"] = np.zeros(T, np.uint8)
    out["clip_comb_bd"] = np.zeros(T, np.int8)
    out["clip_src"] = np.zeros(T, np.int64)
    lib.pfo_fetch_clip(h, _ptr(out["clip_combine"]), _ptr(out["clip_comb_bd"]), _ptr(out["clip_src"]))
    out["n_alpha"] = int(counts[6])
    out["guard_hits"] = int(counts[7])
    out["path_lines"] = np.zeros(scene.n_paths, np.int32)
    lib.pfo_path_lines(h, _ptr(out["path_lines"]))
    out["tiling_seconds"] = lib.pfo_tiling_seconds(h)
    lib.pfo_free(h)
    return out


def run_ref(scene, threads: int = 1) -> dict:
    """Canonical tiling of ``scene`` by the reference's own compiled Tiler."""
    lib = load_ref()
    el = np.zeros(1, np.float64)
    h = lib.pfref_run(*_scene_args(scene), ctypes.c_int(threads), _ptr(el), *_clip_args(scene), _blend_arg(scene))
    counts = np.zeros(6, np.int64)
    lib.pfref_counts(h, _ptr(counts))
    out = _alloc(counts)
    lib.pfref_fetch(h, *[_ptr(out[k]) for k in CANONICAL_KEYS])
    T = out["tile_alpha"].shape[0]
    out["clip_combine"] = np.zeros(T, np.uint8)
    out["clip_comb_bd"] = np.zeros(T, np.int8)
    lib.pfref_fetch_clip(h, _ptr(out["clip_combine"]), _ptr(out["clip_comb_bd"]))
    lib.pfref_free(h)
    out["tiling_seconds"] = float(el[0])
    return out


def raster_cpu(scene, threads: int = 1):
    """The CPU raster half on the oracle's tiling (identical to the reference's): float fill and
    per-layer src-over composite on `threads` host threads.  Returns (RGBA8 [H, W, 4], seconds of
    fill + composite).  Scenes without clip paths."""
    assert not scene.n_clip_paths, "the CPU raster baseline covers scenes without clip paths"
    lib = load_oracle()
    h = lib.pfo_run(*_scene_args(scene), ctypes.c_int(1), ctypes.c_int(0), None, ctypes.c_int(0), None)
    out = np.zeros((scene.height, scene.width, 4), np.uint8)
    secs = lib.pfo_raster_cpu(h, _ptr(scene.path_fill_rule), _ptr(scene.path_rgba), scene.width, scene.height,
                              int(threads), _ptr(out))
    lib.pfo_free(h)
    return out, float(secs)


def time_ref(scene, threads: int = 4, repeats: int = 1) -> float:
    """Wall seconds for the reference Tiler to tile every path of ``scene`` ``repeats`` times."""
    lib = load_ref()
    return lib.pfref_time(ctypes.c_int(scene.n_paths), _ptr(scene.path_contour_start), _ptr(scene.path_fill_rule),
                          _ptr(scene.contour_start), _ptr(scene.contour_closed), _ptr(scene.points),
                          _ptr(scene.point_flags), _ptr(scene.view_box), ctypes.c_int(threads),
                          ctypes.c_int(repeats))


def ref_dash(points, point_flags, contour_start, contour_closed, dashes, offset=0.0, width=0.0, cap=0, join=0,
             miter_limit=10.0):
    """The reference's OutlineDash (core/dash.cpp:48-126), compiled from its sources, on one outline, then
    (width > 0) its OutlineStrokeToFill of the dashed outline: (points [N, 2] f32, point_flags u8,
    contour_start i32, contour_closed u8)."""
    lib = load_ref()
    pts = np.ascontiguousarray(points, np.float32)
    flg = np.ascontiguousarray(point_flags, np.uint8)
    cst = np.ascontiguousarray(contour_start, np.int32)
    ccl = np.ascontiguousarray(contour_closed, np.uint8)
    dsh = np.ascontiguousarray(dashes, np.float32)
    h = lib.pfref_dash(_ptr(pts), _ptr(flg), _ptr(cst), _ptr(ccl), int(ccl.shape[0]), _ptr(dsh), int(dsh.shape[0]),
                       float(offset), float(width), int(cap), int(join), float(miter_limit))
    counts = np.zeros(2, np.int64)
    lib.pfref_stroke_counts(h, _ptr(counts))
    out = (np.zeros((counts[0], 2), np.float32), np.zeros(counts[0], np.uint8), np.zeros(counts[1] + 1, np.int32),
           np.zeros(counts[1], np.uint8))
    lib.pfref_stroke_fetch(h, _ptr(out[0]), _ptr(out[1]), _ptr(out[2]), _ptr(out[3]))
    lib.pfref_stroke_free(h)
    return out


def ref_svg(text: bytes):
    """The reference's SVG ingest (nanosvg + SvgScene's steps, core/svg.cpp:128-206), compiled from its
    sources, on one SVG document.  Returns a dict: size (w, h); per draw path `path_i32` rows (kind 0
    fill / 1 stroke, fill rule, nanosvg paint type, rgba, cap, join, dash start, dash count) and
    `path_f32` rows (stroke width, miter limit); `dashes`; the outlines before stroking (`src_*`,
    paths delimited by src_pcs) and the outlines the scene receives (`out_*`, out_pcs)."""
    lib = load_ref()
    h = lib.pfref_svg(text, len(text))
    if not h:
        raise ValueError("nanosvg could not parse the document")
    c = np.zeros(6, np.int64)
    lib.pfref_svg_counts(h, _ptr(c))
    n, sp, sc, op, oc, nd = [int(v) for v in c]
    r = dict(size=np.zeros(2, np.float32), path_i32=np.zeros((n, 8), np.int32), path_f32=np.zeros((n, 2), np.float32),
             dashes=np.zeros(nd, np.float32), src_pcs=np.zeros(n + 1, np.int32), src_points=np.zeros((sp, 2), np.float32),
             src_flags=np.zeros(sp, np.uint8), src_cs=np.zeros(sc + 1, np.int32), src_closed=np.zeros(sc, np.uint8),
             out_pcs=np.zeros(n + 1, np.int32), out_points=np.zeros((op, 2), np.float32),
             out_flags=np.zeros(op, np.uint8), out_cs=np.zeros(oc + 1, np.int32), out_closed=np.zeros(oc, np.uint8))
    lib.pfref_svg_fetch(h, *[_ptr(r[k]) for k in ("size", "path_i32", "path_f32", "dashes", "src_pcs", "src_points",
                                                   "src_flags", "src_cs", "src_closed", "out_pcs", "out_points",
                                                   "out_flags", "out_cs", "out_closed")])
    g = np.zeros(2, np.int64)
    lib.pfref_svg_gradient_counts(h, _ptr(g))
    ng, ns = int(g[0]), int(g[1])
    r.update(path_grad=np.zeros(n, np.int32), grad_i32=np.zeros((ng, 4), np.int32),
             grad_f32=np.zeros((ng, 12), np.float32), stop_off=np.zeros(ns, np.float32),
             stop_rgba=np.zeros(ns, np.uint32))
    lib.pfref_svg_gradient_fetch(h, *[_ptr(r[k]) for k in ("path_grad", "grad_i32", "grad_f32", "stop_off",
                                                            "stop_rgba")])
    lib.pfref_svg_free(h)
    return r


def ref_gradient_ramp(offsets, rgba) -> np.ndarray:
    """The reference's ramp (Gradient::add_color_stop + GradientTileBuilder::allocate), 256 u32 texels."""
    off = np.ascontiguousarray(offsets, np.float32)
    col = np.ascontiguousarray(rgba, np.uint32)
    out = np.zeros(256, np.uint32)
    load_ref().pfref_gradient_ramp(_ptr(off) if off.size else None, _ptr(col) if col.size else None,
                                   int(off.shape[0]), _ptr(out))
    return out


def ref_half_bits(values) -> np.ndarray:
    """umHalf's float-to-half conversion (the reference's metadata texture): the halves' bits."""
    v = np.ascontiguousarray(values, np.float32)
    out = np.zeros(v.shape[0], np.uint16)
    load_ref().pfref_half_bits(_ptr(v), int(v.shape[0]), _ptr(out))
    return out


def ref_transform_inverse(t) -> np.ndarray:
    """Transform2::inverse of (m11, m21, m12, m22, m13, m23)."""
    v = np.ascontiguousarray(t, np.float32)
    out = np.zeros(6, np.float32)
    load_ref().pfref_transform_inverse(_ptr(v), _ptr(out))
    return out


def ref_stroke(points, point_flags, contour_start, contour_closed, width, cap=0, join=0, miter_limit=10.0):
    """The reference's OutlineStrokeToFill (core/stroke.cpp:124-195), compiled from its sources, on one
    outline: (points [N, 2] f32, point_flags u8, contour_start i32, contour_closed u8)."""
    lib = load_ref()
    pts = np.ascontiguousarray(points, np.float32)
    flg = np.ascontiguousarray(point_flags, np.uint8)
    cst = np.ascontiguousarray(contour_start, np.int32)
    ccl = np.ascontiguousarray(contour_closed, np.uint8)
    h = lib.pfref_stroke(_ptr(pts), _ptr(flg), _ptr(cst), _ptr(ccl), int(ccl.shape[0]), float(width), int(cap),
                         int(join), float(miter_limit))
    counts = np.zeros(2, np.int64)
    lib.pfref_stroke_counts(h, _ptr(counts))
    out = (np.zeros((counts[0], 2), np.float32), np.zeros(counts[0], np.uint8), np.zeros(counts[1] + 1, np.int32),
           np.zeros(counts[1], np.uint8))
    lib.pfref_stroke_fetch(h, *[_ptr(a) for a in out])
    lib.pfref_stroke_free(h)
    return out


def alpha_tiles(canon: dict, scene) -> dict:
    """Alpha tiles in canonical order (path, row, column) with their fills grouped,
    their backdrop and fill rule."""
    T = canon["tile_alpha"].shape[0]
    tile_path = np.repeat(np.arange(scene.n_paths), np.diff(canon["tile_off"]))
    alpha_tiles = np.nonzero(canon["tile_alpha"])[0]
    alpha_index = np.full(T, -1, np.int64)
    alpha_index[alpha_tiles] = np.arange(alpha_tiles.shape[0])
    fill_path = np.repeat(np.arange(scene.n_paths), np.diff(canon["fill_off"]))
    fill_global_tile = canon["tile_off"][fill_path] + canon["fill_tile"].astype(np.int64)
    fa = alpha_index[fill_global_tile]
    # fills of tiles that clip resolution culled draw nothing (core/d3d9/tiler.cpp:425-428)
    keep = fa >= 0
    fa = fa[keep]
    order = np.nonzero(keep)[0][np.argsort(fa, kind="stable")]
    counts = np.bincount(fa, minlength=alpha_tiles.shape[0]) if fa.size else np.zeros(alpha_tiles.shape[0], np.int64)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return dict(tile=alpha_tiles, alpha_index=alpha_index, fill_off=off,
                fill_seg=np.ascontiguousarray(canon["fill_seg"][order]),
                backdrop=np.ascontiguousarray(canon["tile_backdrop"][alpha_tiles]),
                even_odd=np.ascontiguousarray(scene.path_fill_rule[tile_path[alpha_tiles]]),
                path=tile_path[alpha_tiles])


def coverage_fp64(at: dict, raw: bool = False) -> np.ndarray:
    lib = load_oracle()
    A = at["tile"].shape[0]
    out = np.zeros((A, 16, 16), np.float64)
    if raw:
        lib.pfo_coverage_raw(ctypes.c_int64(A), _ptr(at["fill_off"]), _ptr(at["fill_seg"]), _ptr(at["backdrop"]),
                             _ptr(out))
    else:
        lib.pfo_coverage(ctypes.c_int64(A), _ptr(at["fill_off"]), _ptr(at["fill_seg"]), _ptr(at["backdrop"]),
                         _ptr(at["even_odd"]), _ptr(out))
    return out


def gradient_ramp(offsets, rgba) -> np.ndarray:
    """The oracle's ramp of a gradient (256 RGBA8 texels as u32), stops in the order given."""
    off = np.ascontiguousarray(offsets, np.float32)
    col = np.ascontiguousarray(rgba, np.uint32)
    out = np.zeros(256, np.uint32)
    load_oracle().pfo_gradient_ramp(_ptr(off) if off.size else None, _ptr(col) if col.size else None,
                                    int(off.shape[0]), _ptr(out))
    return out


def gradient_record(g, row: int = 0) -> np.ndarray:
    """The oracle's 16-float record of one pfh_gradient-shaped record (a numpy structured scalar with
    kind, wrap, line, radii, transform), as the half metadata texture holds it."""
    line = np.ascontiguousarray(g["line"], np.float32)
    radii = np.ascontiguousarray(g["radii"], np.float32)
    xf = np.ascontiguousarray(g["transform"], np.float32)
    rec = np.zeros(16, np.float32)
    rc = load_oracle().pfo_gradient_record(int(g["kind"]), int(g["wrap"]), _ptr(line), _ptr(radii), _ptr(xf),
                                           int(row), _ptr(rec))
    if rc != 0:
        raise ValueError("radial gradient transform without inverse")
    return rec


def half_value(v: float) -> float:
    return load_oracle().pfo_half_value(float(v))


def paint_colour(rec, ramp, px: float, py: float):
    """The oracle's un-premultiplied colour of a gradient paint at pixel centre (px, py), or None."""
    out = np.zeros(4, np.float64)
    r = np.ascontiguousarray(rec, np.float32)
    rp = np.ascontiguousarray(ramp, np.uint32)
    ok = load_oracle().pfo_paint_colour(_ptr(r), _ptr(rp), float(px), float(py), _ptr(out))
    return out if ok else None


def scene_paints(scene):
    """(path_paint i32 [n_paths] (-1 solid), records [G, 16] f32, ramps [G, 256] u32) of a scene's
    gradient paints, each used gradient compiled once in order of first use (its ramp row: its
    index modulo 256), or None without gradients."""
    pg = getattr(scene, "path_gradient", None)
    if pg is None:
        return None
    slot, recs, ramps = {}, [], []
    paint = np.full(scene.n_paths, -1, np.int32)
    for p in range(scene.n_paths):
        gi = int(pg[p])
        if gi < 0:
            continue
        if gi not in slot:
            g = scene.gradients[gi]
            a, n = int(g["stop_start"]), int(g["stop_count"])
            slot[gi] = len(recs)
            recs.append(gradient_record(g, len(recs) % 256))
            ramps.append(np.asarray(scene.gradient_ramps[gi], np.uint32) if n < 0 else
                         gradient_ramp(scene.gradient_stop_offsets[a:a + n], scene.gradient_stop_rgba[a:a + n]))
        paint[p] = slot[gi]
    if not recs:
        return None
    return paint, np.ascontiguousarray(np.stack(recs)), np.ascontiguousarray(np.stack(ramps))


def composite_fp64(canon: dict, at: dict, coverage: np.ndarray, scene) -> np.ndarray:
    """FP64 premultiplied RGBA [H, W, 4] of the scene from the canonical entries."""
    lib = load_oracle()
    paints = scene_paints(scene)
    ops = path_ops(scene)
    if paints is not None or ops is not None:
        if paints is None:
            paints = (None, None, None)
        ent_tile = canon["entry_tile"]
        ent_alpha = np.ascontiguousarray(at["alpha_index"][ent_tile])
        ent_bd = np.ascontiguousarray(canon["tile_backdrop"][ent_tile])
        fbw = int(canon["fb_rect"][2] - canon["fb_rect"][0])
        out = np.zeros((scene.height, scene.width, 4), np.float64)
        cov = np.ascontiguousarray(coverage, np.float64)
        lib.pfo_composite_paint(ctypes.c_int64(ent_tile.shape[0]), _ptr(canon["entry_fbtile"]),
                                _ptr(canon["entry_path"]), _ptr(ent_alpha), _ptr(ent_bd), _ptr(scene.path_fill_rule),
                                _ptr(scene.path_rgba), _ptr(paints[0]), _ptr(paints[1]), _ptr(paints[2]), _ptr(cov),
                                fbw, scene.width, scene.height, _ptr(out), _ptr(ops), int(canon["fb_rect"][0]),
                                int(canon["fb_rect"][1]))
        return out
    ent_tile = canon["entry_tile"]
    ent_alpha = np.ascontiguousarray(at["alpha_index"][ent_tile])
    ent_bd = np.ascontiguousarray(canon["tile_backdrop"][ent_tile])
    fb = canon["fb_rect"]
    fbw, fbh = int(fb[2] - fb[0]), int(fb[3] - fb[1])
    out = np.zeros((scene.height, scene.width, 4), np.float64)
    cov = np.ascontiguousarray(coverage, np.float64)
    lib.pfo_composite(ctypes.c_int64(ent_tile.shape[0]), _ptr(canon["entry_fbtile"]), _ptr(canon["entry_path"]),
                      _ptr(ent_alpha), _ptr(ent_bd), _ptr(scene.path_fill_rule), _ptr(scene.path_rgba), _ptr(cov),
                      fbw, fbh, scene.width, scene.height, _ptr(out), int(fb[0]), int(fb[1]))
    return out


def apply_rule(c: np.ndarray, even_odd: np.ndarray) -> np.ndarray:
    """The fill rule (nonzero 0, even-odd 1) applied elementwise to FP64 coverage sums."""
    c = np.ascontiguousarray(c, np.float64)
    eo = np.ascontiguousarray(np.broadcast_to(np.asarray(even_odd, np.uint8).reshape(
        (-1,) + (1,) * (c.ndim - 1)), c.shape))
    out = np.empty_like(c)
    load_oracle().pfo_apply_rule_array(ctypes.c_int64(c.size), _ptr(c), _ptr(eo), _ptr(out))
    return out


def clip_levels(scene) -> np.ndarray:
    """Per path: 0 for a draw path, 1 + the nesting level for a clip path (1: no clip of its own,
    2: clipped by a level-1 clip path, ... up to 4)."""
    n_clip = int(scene.n_clip_paths)
    lev = np.zeros(scene.n_paths, np.int64)
    if n_clip == 0:
        return lev
    n_draw = scene.n_paths - n_clip
    pc = scene.path_clip if scene.path_clip is not None else np.full(scene.n_paths, -1, np.int32)
    for q in range(n_clip):
        k, c = 1, int(pc[n_draw + q])
        while c >= 0:
            k += 1
            c = int(pc[n_draw + c])
        lev[n_draw + q] = k
    return lev


def clipped_coverage(canon: dict, at: dict, scene) -> np.ndarray:
    """FP64 coverage of every alpha tile after clip resolution (core/d3d9/tiler.cpp:385-431) and the
    clip combine pass (shaders/d3d9/tile_clip_combine.frag:25-27, then the draw path's fill rule in
    shaders/d3d9/tile.frag).  A clip path's coverage value K, per pixel, is |own sum + backdrop| for
    a clip path without a clip of its own (its fill rule is not used), and follows the same cases
    against its own clip for a nested one; a draw path then takes:
      * a tile with a combine job: rule(min(|own sum + own backdrop|, K of the clip tile));
      * a solid tile that took the clip tile's mask: rule(K of the clip tile);
      * any other alpha tile: rule(own sum + backdrop).
    A clip path's own masks (its dump rows) hold the nonzero rule of its K: min(K, 1).
    With scene.clip_rule 1 (ClipPath::fill_rule honoured) every path's coverage is taken under its
    own rule first: K = rule_clip(own sum + backdrop), a combine job gives min(rule(own sum + own
    backdrop), K of the clip tile) and a tile that took the clip tile's mask K of it; a clip path's
    dump rows hold its K."""
    raw = coverage_fp64(at, raw=True)
    src = canon["clip_src"][at["tile"]]
    comb = canon["clip_combine"][at["tile"]].astype(bool)
    comb_bd = canon["clip_comb_bd"][at["tile"]].astype(np.float64)[:, None, None]
    level = clip_levels(scene)[at["path"]]
    if int(getattr(scene, "clip_rule", 0)) == 1:
        K = apply_rule(raw, at["even_odd"])
        for lev in sorted(set(int(v) for v in level[level > 0])) + [0]:
            sel = np.nonzero((level == lev) & (src >= 0))[0]
            if sel.size == 0:
                continue
            k_src = K[at["alpha_index"][src[sel]]]
            own = apply_rule(raw[sel] + comb_bd[sel], at["even_odd"][sel])
            K[sel] = np.where(comb[sel][:, None, None], np.minimum(own, k_src), k_src)
        return K
    eff = raw.copy()
    K = np.abs(raw)
    # outermost clip level first: a tile's source clip tile lies one level further out, so its K
    # is final by the time the tile reads it; draw paths (level 0) last
    for lev in sorted(set(int(v) for v in level[level > 0])) + [0]:
        sel = np.nonzero((level == lev) & (src >= 0))[0]
        if sel.size == 0:
            continue
        k_src = K[at["alpha_index"][src[sel]]]
        own = np.abs(raw[sel] + comb_bd[sel])
        val = np.where(comb[sel][:, None, None], np.minimum(own, k_src), k_src)
        K[sel] = val
        eff[sel] = val
    eo = np.where(level > 0, 0, at["even_odd"]).astype(np.uint8)
    return apply_rule(eff, eo)


def full_model(scene, zmode: int = 1) -> dict:
    """Canonical tiling, FP64 coverage and FP64 composite of a scene (its stroked paths stroked first)."""
    scene = apply_strokes(scene)
    canon = run_oracle(scene, zmode)
    canon["n_clip_paths"] = int(scene.n_clip_paths)
    at = alpha_tiles(canon, scene)
    cov = clipped_coverage(canon, at, scene) if scene.n_clip_paths else coverage_fp64(at)
    img = composite_fp64(canon, at, cov, scene)
    return dict(canon=canon, alpha=at, coverage=cov, image=img)


def path_coverage_image(scene, model: dict, path: int) -> np.ndarray:
    """FP64 coverage (after the fill rule) of one draw path over the canvas [H, W]."""
    canon, at, cov = model["canon"], model["alpha"], model["coverage"]
    img = np.zeros((scene.height + 16, scene.width + 16), np.float64)
    l, t, r, b = [int(v) for v in canon["path_rect"][path]]
    w = max(r - l, 0)
    t0, t1 = int(canon["tile_off"][path]), int(canon["tile_off"][path + 1])
    for i in range(t1 - t0):
        tx, ty = l + i % w, t + i // w
        x, y = tx * 16, ty * 16
        a = at["alpha_index"][t0 + i]
        if a >= 0:
            img[y:y + 16, x:x + 16] = cov[a]
        else:
            img[y:y + 16, x:x + 16] = load_oracle().pfo_apply_rule(float(canon["tile_backdrop"][t0 + i]),
                                                                  int(scene.path_fill_rule[path]))
    return img[:scene.height, :scene.width]
