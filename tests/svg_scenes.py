"""Scenes from tests/golden/svg_fixtures.npz, what the reference's SVG ingest (nanosvg + SvgScene,
core/svg.cpp:128-206) makes of its assets tiger.svg and features.svg.

* scene_src(name): the draw paths as the scene's caller hands them to the library -- outlines
  before stroking, each stroke with its style and dash pattern (path_stroke / stroke_dashes), so
  the device dashes and strokes them at upload.
* scene_out(name): the same draw paths as the reference's scene holds them -- strokes already
  dashed and stroked by the reference (filled nonzero).
With solid_only (the default) the paths with gradient paints are left out; otherwise they carry
their gradients (path_gradient / gradients / stops, as the reference's convert_nsvg_paint builds
them, core/svg.cpp:64-124).  Canvas = the document's size, rounded up."""
import math
import os

import numpy as np

import scenegen

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "svg_fixtures.npz")
PAINT_COLOR = 1  # NSVGpaintType: 0 none, 1 colour, 2 linear gradient, 3 radial gradient


def load(name):
    z = np.load(FIX, allow_pickle=False)
    return {k[len(name) + 1:]: z[k] for k in z.files if k.startswith(name + "/")}


def _paths(d, which, keep):
    pts, flg, cs, cl, pcs = (d[f"{which}_points"], d[f"{which}_flags"], d[f"{which}_cs"], d[f"{which}_closed"],
                             d[f"{which}_pcs"])
    out = []
    for p in keep:
        c0, c1 = int(pcs[p]), int(pcs[p + 1])
        a, b = int(cs[c0]), int(cs[c1])
        out.append((pts[a:b], flg[a:b], cs[c0:c1 + 1] - a, cl[c0:c1]))
    return out


def _kept(d, solid_only):
    pi = d["path_i32"]
    return [p for p in range(pi.shape[0]) if not solid_only or pi[p, 2] == PAINT_COLOR]


def _scene(d, outlines, rules, rgba, keep):
    w, h = (int(math.ceil(float(v))) for v in d["size"])
    sc = scenegen.outlines_scene(w, h, outlines, rules, rgba)
    pg = d["path_grad"][keep]
    if (pg >= 0).any():
        gi, gf = d["grad_i32"], d["grad_f32"]
        g = np.zeros(gi.shape[0], scenegen.GRADIENT)
        g["kind"], g["wrap"], g["stop_start"], g["stop_count"] = gi[:, 0], gi[:, 1], gi[:, 2], gi[:, 3]
        g["line"], g["radii"], g["transform"] = gf[:, 0:4], gf[:, 4:6], gf[:, 6:12]
        sc.path_gradient = pg.astype(np.int32)
        sc.gradients = g
        sc.gradient_stop_offsets = d["stop_off"].astype(np.float32)
        sc.gradient_stop_rgba = d["stop_rgba"].astype(np.uint32)
    return sc


def scene_src(name, solid_only=True):
    d = load(name)
    keep = _kept(d, solid_only)
    pi, pf = d["path_i32"], d["path_f32"]
    sc = _scene(d, _paths(d, "src", keep), pi[keep, 1].astype(np.uint8), pi[keep, 3].astype(np.uint32), keep)
    st = np.zeros((len(keep), 7), np.float32)
    for k, p in enumerate(keep):
        if pi[p, 0] == 1:
            st[k] = (pf[p, 0], pi[p, 4], pi[p, 5], pf[p, 1], pi[p, 6], pi[p, 7], 0.0)
    sc.path_stroke = st
    sc.stroke_dashes = d["dashes"].astype(np.float32) if d["dashes"].size else None
    return sc


def scene_out(name, solid_only=True):
    d = load(name)
    keep = _kept(d, solid_only)
    pi = d["path_i32"]
    rules = np.where(pi[keep, 0] == 1, 0, pi[keep, 1]).astype(np.uint8)
    return _scene(d, _paths(d, "out", keep), rules, pi[keep, 3].astype(np.uint32), keep)


def strokes(name):
    """The stroke paths: [(source outline, (w, cap, join, miter), dash pattern or None, reference outline)]."""
    d = load(name)
    pi, pf = d["path_i32"], d["path_f32"]
    idx = [p for p in range(pi.shape[0]) if pi[p, 0] == 1]
    src, out = _paths(d, "src", idx), _paths(d, "out", idx)
    res = []
    for k, p in enumerate(idx):
        pat = d["dashes"][pi[p, 6]:pi[p, 6] + pi[p, 7]] if pi[p, 7] > 0 else None
        res.append((src[k], (float(pf[p, 0]), int(pi[p, 4]), int(pi[p, 5]), float(pf[p, 1])), pat, out[k]))
    return res
