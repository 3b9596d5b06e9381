"""Writes tests/golden/dash_fixtures.npz: outlines, dash patterns and offsets, and what the
reference's own OutlineDash (core/dash.cpp) makes of them -- the dashed outline alone, and the
dashed outline stroked by its OutlineStrokeToFill (core/stroke.cpp) as Canvas::stroke_path does
(core/canvas.cpp:284-296) -- both compiled from the unchanged sources into
oracle/_ref/libpfref.so by oracle/build_ref.sh.

Run in the build container, where the reference tree is present:
    python tests/golden/gen_dash_golden.py
tests/test_dash.py checks the oracle's dasher against these bit for bit, tests/test_gpu_stroke.py
the device dasher (pfh_stroke_outlines with dash patterns).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "pathfinder-cpp_amd"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import oracle  # noqa: E402
import scenegen  # noqa: E402
from gen_stroke_golden import outline  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dash_fixtures.npz")

BUTT, SQUARE, ROUND = 0, 1, 2
MITER, BEVEL, RJOIN = 0, 1, 2


def cases():
    """(name, outline, dashes, offset, (width, cap, join, miter_limit))"""
    out = []
    line = outline(([(0, 0), (100, 0)], None, False))
    out.append(("line_10_5", line, [10, 5], 0.0, (4.0, BUTT, MITER, 10.0)))
    out.append(("line_offset", line, [10, 5], 7.5, (4.0, SQUARE, MITER, 10.0)))
    out.append(("line_offset_past_total", line, [10, 5], 37.0, (4.0, ROUND, RJOIN, 10.0)))
    out.append(("line_offset_huge", line, [10, 5], 1.0e6 + 3.25, (4.0, BUTT, MITER, 10.0)))
    out.append(("line_single_length", line, [6], 0.0, (3.0, BUTT, MITER, 10.0)))
    out.append(("line_odd_pattern", line, [5, 3, 2], 0.0, (3.0, ROUND, RJOIN, 10.0)))
    out.append(("line_dots", line, [0, 8], 0.0, (5.0, ROUND, RJOIN, 10.0)))
    out.append(("line_dots_square", line, [0, 8], 2.0, (5.0, SQUARE, MITER, 10.0)))
    out.append(("line_fine", line, [0.5, 0.5], 0.0, (0.4, BUTT, MITER, 10.0)))
    out.append(("line_dash_longer_than_path", line, [500, 10], 0.0, (4.0, BUTT, MITER, 10.0)))
    out.append(("line_off_first", line, [10, 5], 12.0, (4.0, BUTT, MITER, 10.0)))
    out.append(("line_exact_multiple", line, [25, 25], 0.0, (4.0, BUTT, MITER, 10.0)))
    diag = outline(([(5.5, 7.25), (63.125, 91.5)], None, False))
    out.append(("diag", diag, [7.3, 2.9], 1.1, (3.0, ROUND, RJOIN, 10.0)))
    zig = outline(([(0, 0), (40, 30), (80, 0), (120, 30), (125, -20), (60, -40)], None, False))
    for join in (MITER, BEVEL, RJOIN):
        out.append((f"zigzag_join{join}", zig, [13, 7], 0.0, (6.0, BUTT, join, 4.0)))
    out.append(("zigzag_long_dashes", zig, [70, 4], 3.0, (6.0, ROUND, MITER, 10.0)))
    square = outline(([(10, 10), (90, 10), (90, 90), (10, 90)], None, True))
    out.append(("closed_square", square, [20, 10, 5, 10], 3.0, (5.0, BUTT, MITER, 10.0)))
    out.append(("closed_square_round", square, [33, 11], 0.0, (5.0, ROUND, RJOIN, 10.0)))
    closed_dup = outline(([(10, 10), (90, 10), (90, 90), (10, 10)], None, True))  # last == first
    out.append(("closed_repeat_first", closed_dup, [15, 5], 0.0, (4.0, BUTT, MITER, 10.0)))
    cub = outline(([(0, 0), (30, 80), (70, -40), (100, 50), (120, 90), (150, 90), (160, 40)], [0, 1, 2, 0, 1, 2, 0],
                   False))
    out.append(("cubics_9_4", cub, [9, 4], 0.0, (3.0, BUTT, MITER, 10.0)))
    out.append(("cubics_30_12", cub, [30, 12], 5.0, (7.5, ROUND, RJOIN, 10.0)))
    out.append(("cubics_bevel", cub, [17, 6, 2, 6], 0.0, (4.0, SQUARE, BEVEL, 10.0)))
    quads = outline(([(0, 0), (50, 80), (100, 0), (130, -40), (160, 20)], [0, 1, 0, 1, 0], False))
    out.append(("quads", quads, [6, 6], 0.0, (3.0, ROUND, RJOIN, 10.0)))
    out.append(("quads_long", quads, [40, 3], 11.0, (6.0, BUTT, MITER, 10.0)))
    loop = outline(([(50, 50), (150, 0), (-50, 0), (50, 50)], [0, 1, 2, 0], False))
    out.append(("cubic_loop", loop, [12, 5], 0.0, (4.0, ROUND, RJOIN, 10.0)))
    closed_mix = outline(([(20, 20), (80, -10), (120, 40), (100, 100), (60, 120), (40, 90), (30, 60)],
                          [0, 1, 0, 1, 2, 0, 0], True))
    out.append(("closed_mixed", closed_mix, [25, 8], 4.0, (5.0, BUTT, RJOIN, 10.0)))
    multi = outline(([(0, 0), (50, 20), (80, 0)], None, False), ([(10, 40), (70, 40), (40, 90)], None, True),
                    ([(100, 100), (130, 150), (160, 80), (190, 120)], [0, 1, 2, 0], False))
    out.append(("multi_contour_carry", multi, [17, 9, 3, 9], 0.0, (5.0, ROUND, MITER, 10.0)))
    out.append(("multi_contour_long_dash", multi, [120, 30], 0.0, (5.0, BUTT, MITER, 10.0)))
    degenerate = outline(([(5, 5), (5, 5), (5, 5), (40, 5), (40, 5), (40, 30)], None, False))
    out.append(("zero_length_segments", degenerate, [7, 3], 0.0, (3.0, ROUND, RJOIN, 10.0)))
    single = outline(([(3, 4)], None, False))
    out.append(("single_point", single, [5, 5], 0.0, (4.0, ROUND, RJOIN, 10.0)))
    two_same = outline(([(3, 4), (3, 4)], None, False))
    out.append(("two_coincident_points", two_same, [5, 5], 0.0, (4.0, SQUARE, MITER, 10.0)))
    big = outline(([(1000.5, 2000.25), (1800, 2600), (2600, 1400), (3400, 2000)], [0, 1, 2, 0], False))
    out.append(("large_coordinates", big, [40, 25], 10.0, (15.0, ROUND, RJOIN, 10.0)))
    # stroked polylines as the dashed workload generates them
    sc = scenegen.dashed_stroke_scene(seed=0xDA51, size=1024, n=8, dashed=1.0)
    for p in range(sc.n_paths):
        c0, c1 = int(sc.path_contour_start[p]), int(sc.path_contour_start[p + 1])
        a, b = int(sc.contour_start[c0]), int(sc.contour_start[c1])
        o = (sc.points[a:b], sc.point_flags[a:b], sc.contour_start[c0:c1 + 1] - a, sc.contour_closed[c0:c1])
        st = sc.path_stroke[p]
        d0, dn = int(st[4]), int(st[5])
        out.append((f"polyline_{p}", o, sc.stroke_dashes[d0:d0 + dn].tolist(), float(st[6]),
                    (float(st[0]), int(st[1]), int(st[2]), float(st[3]))))
    return out


def main():
    arrays = {}
    names = []
    keys = ("points", "point_flags", "contour_start", "contour_closed")
    for name, (pts, flg, cst, ccl), dashes, offset, (w, cap, join, ml) in cases():
        dashes = np.asarray(dashes, np.float32)
        offset = np.float32(offset)
        dashed = oracle.ref_dash(pts, flg, cst, ccl, dashes, offset)
        stroked = oracle.ref_dash(pts, flg, cst, ccl, dashes, offset, w, cap, join, ml)
        names.append(name)
        for k, v in zip(keys, (pts, flg, cst, ccl)):
            arrays[f"{name}/in/{k}"] = v
        arrays[f"{name}/dashes"] = dashes
        arrays[f"{name}/offset"] = np.array([offset], np.float32)
        arrays[f"{name}/style"] = np.array([w, cap, join, ml], np.float64)
        for k, v in zip(keys, dashed):
            arrays[f"{name}/dash/{k}"] = v
        for k, v in zip(keys, stroked):
            arrays[f"{name}/ref/{k}"] = v
    arrays["names"] = np.array(names)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {len(names)} dash fixtures to {OUT}")


if __name__ == "__main__":
    main()
