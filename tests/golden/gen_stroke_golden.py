"""Writes tests/golden/stroke_fixtures.npz: outlines, stroke styles and the stroked outlines the
reference's own OutlineStrokeToFill (core/stroke.cpp, compiled from its unchanged sources into
oracle/_ref/libpfref.so by oracle/build_ref.sh) produces for them.

Run in the build container, where the reference tree is present:
    python tests/golden/gen_stroke_golden.py
tests/test_stroke.py checks the library's stroker (pfh_stroke_outline) against these bit for bit.
"""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "pathfinder-cpp_amd"))

import oracle  # noqa: E402
import scenegen  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stroke_fixtures.npz")

BUTT, SQUARE, ROUND = 0, 1, 2
MITER, BEVEL, RJOIN = 0, 1, 2


def outline(*contours):
    """contours: (points, flags or None, closed)"""
    pts, flg, cst, ccl = [], [], [0], []
    for p, f, closed in contours:
        p = np.asarray(p, np.float32).reshape(-1, 2)
        pts.append(p)
        flg.append(np.zeros(len(p), np.uint8) if f is None else np.asarray(f, np.uint8))
        cst.append(cst[-1] + len(p))
        ccl.append(1 if closed else 0)
    return (np.concatenate(pts) if pts else np.zeros((0, 2), np.float32),
            np.concatenate(flg) if flg else np.zeros(0, np.uint8), np.array(cst, np.int32), np.array(ccl, np.uint8))


def cases():
    out = []
    line = outline(([(10, 20), (110, 20)], None, False))
    for cap in (BUTT, SQUARE, ROUND):
        out.append((f"line_cap{cap}", line, (8.0, cap, MITER, 10.0)))
    diag = outline(([(5.5, 7.25), (63.125, 91.5)], None, False))
    out.append(("diag_round", diag, (3.0, ROUND, RJOIN, 10.0)))
    zig = outline(([(0, 0), (40, 30), (80, 0), (120, 30), (125, -20), (60, -40)], None, False))
    for join in (MITER, BEVEL, RJOIN):
        for cap in (BUTT, ROUND):
            out.append((f"zigzag_join{join}_cap{cap}", zig, (6.0, cap, join, 4.0)))
    out.append(("zigzag_miter_limit_hit", zig, (6.0, BUTT, MITER, 1.05)))
    sharp = outline(([(0, 0), (100, 5), (0, 10)], None, False))  # a near-reversal
    for join in (MITER, BEVEL, RJOIN):
        out.append((f"sharp_turn_join{join}", sharp, (5.0, SQUARE, join, 10.0)))
    reverse = outline(([(0, 0), (50, 0), (20, 0)], None, False))  # exact 180-degree turn
    out.append(("reversal", reverse, (4.0, ROUND, RJOIN, 10.0)))
    square = outline(([(10, 10), (90, 10), (90, 90), (10, 90)], None, True))
    for join in (MITER, BEVEL, RJOIN):
        out.append((f"closed_square_join{join}", square, (10.0, BUTT, join, 10.0)))
    closed_dup = outline(([(10, 10), (90, 10), (90, 90), (10, 10)], None, True))  # last == first
    out.append(("closed_repeat_first", closed_dup, (4.0, BUTT, MITER, 10.0)))
    cub = outline(([(0, 0), (30, 80), (70, -40), (100, 50), (120, 90), (150, 90), (160, 40)], [0, 1, 2, 0, 1, 2, 0],
                   False))
    for cap, join in ((BUTT, MITER), (SQUARE, BEVEL), (ROUND, RJOIN)):
        for w in (1.0, 7.5, 20.0):
            out.append((f"cubics_cap{cap}_join{join}_w{w}", cub, (w, cap, join, 10.0)))
    loop = outline(([(50, 50), (150, 0), (-50, 0), (50, 50)], [0, 1, 2, 0], False))  # self-intersecting cubic
    out.append(("cubic_loop", loop, (6.0, ROUND, RJOIN, 10.0)))
    cusp = outline(([(0, 0), (100, 60), (0, 60), (100, 0)], [0, 1, 2, 0], False))
    out.append(("cubic_cusp", cusp, (5.0, BUTT, MITER, 10.0)))
    coincident_ctrl = outline(([(0, 0), (0, 0), (60, 40), (90, 10)], [0, 1, 2, 0], False),
                              ([(0, 50), (40, 90), (90, 60), (90, 60)], [0, 1, 2, 0], False))
    out.append(("cubic_coincident_controls", coincident_ctrl, (4.0, BUTT, MITER, 10.0)))
    quads = outline(([(0, 0), (50, 80), (100, 0), (130, -40), (160, 20)], [0, 1, 0, 1, 0], False))
    for join in (MITER, RJOIN):
        out.append((f"quads_join{join}", quads, (6.0, ROUND, join, 10.0)))
    closed_mix = outline(([(20, 20), (80, -10), (120, 40), (100, 100), (60, 120), (40, 90), (30, 60)],
                          [0, 1, 0, 1, 2, 0, 0], True))
    out.append(("closed_mixed", closed_mix, (5.0, BUTT, RJOIN, 10.0)))
    degenerate = outline(([(5, 5), (5, 5), (5, 5), (40, 5), (40, 5), (40, 30)], None, False))
    out.append(("zero_length_segments", degenerate, (3.0, ROUND, RJOIN, 10.0)))
    tiny = outline(([(0, 0), (0.05, 0.02), (0.09, 0.0), (30, 10)], None, False))  # below STROKE_TOL
    out.append(("tiny_segments", tiny, (2.0, SQUARE, MITER, 10.0)))
    single = outline(([(3, 4)], None, False))
    out.append(("single_point", single, (4.0, ROUND, RJOIN, 10.0)))
    two_same = outline(([(3, 4), (3, 4)], None, False))
    out.append(("two_coincident_points", two_same, (4.0, SQUARE, MITER, 10.0)))
    multi = outline(([(0, 0), (50, 20), (80, 0)], None, False), ([(10, 40), (70, 40), (40, 90)], None, True),
                    ([(100, 100), (130, 150), (160, 80), (190, 120)], [0, 1, 2, 0], False))
    out.append(("multi_contour", multi, (5.0, ROUND, MITER, 10.0)))
    wide = outline(([(0, 0), (10, 30), (20, -30), (30, 0)], [0, 1, 2, 0], False))  # radius above curvature radius
    out.append(("wide_stroke_tight_curve", wide, (40.0, BUTT, RJOIN, 10.0)))
    big = outline(([(1000.5, 2000.25), (1800, 2600), (2600, 1400), (3400, 2000)], [0, 1, 2, 0], False))
    out.append(("large_coordinates", big, (15.0, ROUND, RJOIN, 10.0)))
    # the stress workload's stroked polylines: a few of them as generated
    pl = scenegen.stroked_polylines(seed=0x5170, size=1024, n=6)
    for i, (o, style) in enumerate(pl):
        out.append((f"polyline_{i}", o, style))
    return out


def main():
    arrays = {}
    names = []
    for name, (pts, flg, cst, ccl), (w, cap, join, ml) in cases():
        ref = oracle.ref_stroke(pts, flg, cst, ccl, w, cap, join, ml)
        names.append(name)
        for k, v in zip(("points", "point_flags", "contour_start", "contour_closed"), (pts, flg, cst, ccl)):
            arrays[f"{name}/in/{k}"] = v
        arrays[f"{name}/style"] = np.array([w, cap, join, ml], np.float64)
        for k, v in zip(("points", "point_flags", "contour_start", "contour_closed"), ref):
            arrays[f"{name}/ref/{k}"] = v
    arrays["names"] = np.array(names)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {len(names)} stroke fixtures to {OUT}")


if __name__ == "__main__":
    main()
