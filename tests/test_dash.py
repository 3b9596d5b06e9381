"""The oracle's dasher (oracle/stroke_oracle.cpp, OutlineDash of core/dash.cpp), the checker of the
GPU dasher, against the reference.

* Bit for bit against tests/golden/dash_fixtures.npz, written by the reference's own compiled
  OutlineDash: the dashed outline alone and the dashed outline stroked (Canvas::stroke_path's order,
  core/canvas.cpp:284-296).  The fixtures cover offsets inside, past and far past the pattern
  total, patterns of one, two, three and four lengths, zero-length dashes (dots), dashes longer
  than the path, the dash state carried across the vertices of polylines, the closing line of
  closed contours (skipped where the first point repeats), quadratics, cubics and a looped cubic,
  several contours of one path (the state carries across them), degenerate contours and large
  coordinates.
* Closed forms the geometry fixes (independent of the reference's code): along a straight line
  the dashes lie at the "on" pieces of the pattern's walk (its state rules written out in
  pattern_pieces: the walk starts in the dash the offset lands in with the offset's remainder left,
  so an offset of 0 begins with a zero-length dash, and the index keeps its parity across repeats
  of an odd pattern); a one-length pattern cuts the line into consecutive pieces that cover it end
  to end; the on-length over a closed square's perimeter is the measure of the "on" pieces; shifting
  the offset by the pattern total changes nothing.
* Input checks: negative, non-finite and all-zero patterns and negative offsets are refused.
Runs without a GPU; tests/test_gpu_dash.py holds the GPU dasher to the same fixtures."""
import math
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "dash_fixtures.npz")
KEYS = ("points", "point_flags", "contour_start", "contour_closed")


def load_cases():
    z = np.load(FIX, allow_pickle=False)
    out = {}
    for name in [str(n) for n in z["names"]]:
        w, cap, join, ml = [float(v) for v in z[f"{name}/style"]]
        out[name] = dict(inp=[z[f"{name}/in/{k}"] for k in KEYS], dashes=z[f"{name}/dashes"],
                         offset=float(z[f"{name}/offset"][0]), style=(w, int(cap), int(join), ml),
                         dash=[z[f"{name}/dash/{k}"] for k in KEYS], ref=[z[f"{name}/ref/{k}"] for k in KEYS])
    return out


CASES = load_cases()


def assert_same(got, want, what):
    for k, g, r in zip(KEYS, got, want):
        assert g.dtype == r.dtype and g.shape == r.shape, f"{what}: {k} {g.shape} vs {r.shape}"
        np.testing.assert_array_equal(g.view(np.uint8), r.view(np.uint8), err_msg=f"{what}: {k}")


@pytest.mark.parametrize("name", list(CASES))
def test_dash_matches_reference_fixture(name):
    c = CASES[name]
    assert_same(oracle.dash(*c["inp"], c["dashes"], c["offset"]), c["dash"], name)


@pytest.mark.parametrize("name", list(CASES))
def test_dash_then_stroke_matches_reference_fixture(name):
    c = CASES[name]
    assert_same(oracle.stroke(*oracle.dash(*c["inp"], c["dashes"], c["offset"]), *c["style"]), c["ref"], name)


def line(x0, y0, x1, y1):
    return (np.array([[x0, y0], [x1, y1]], np.float32), np.zeros(2, np.uint8), np.array([0, 2], np.int32),
            np.zeros(1, np.uint8))


def dashes_of(out):
    """The contours of a dashed outline as arrays of points."""
    pts, _, cs, _ = out
    return [pts[cs[i]:cs[i + 1]].astype(np.float64) for i in range(len(cs) - 1)]


def pattern_pieces(pattern, offset, length):
    """The (start, end, on) pieces of [0, length) as DashState walks them (core/dash.cpp:9-29,
    :74-124): the offset is reduced modulo the pattern total, whole dashes are skipped, and the walk
    starts in the dash it lands in with distance_left set to what remains of the offset (so an
    offset of 0 begins with a zero-length dash); the dash index then advances through the pattern,
    even indices on, wrapping to 0 (an odd pattern does not swap on and off between repeats)."""
    total = float(sum(pattern))
    rem = math.fmod(offset, total)
    idx = 0
    while idx < len(pattern) and rem >= pattern[idx]:
        rem -= pattern[idx]
        idx += 1
    x, run, out = 0.0, rem, []
    while x < length:
        out.append((x, min(x + run, length), idx % 2 == 0))
        x += run
        idx = (idx + 1) % len(pattern)
        run = pattern[idx]
    return out


# lengths chosen so that no dash boundary falls within 1e-3 of the line's end
@pytest.mark.parametrize("a,b,offset", [(10, 5, 0), (10, 5, 7.5), (7, 3, 2), (2.5, 0.5, 0), (30, 20, 45)])
def test_dashes_of_a_line_follow_the_pattern(a, b, offset):
    L = 201.3
    got = dashes_of(oracle.dash(*line(0, 0, L, 0), [a, b], offset))
    want = [(lo, hi) for lo, hi, on in pattern_pieces([a, b], offset, L) if on]
    assert len(got) == len(want), (got, want)
    for d, (lo, hi) in zip(got, want):
        assert np.all(d[:, 1] == 0)
        assert abs(d[0, 0] - lo) < 1e-3 and abs(d[-1, 0] - hi) < 1e-3, (d, lo, hi)


def test_zero_offset_begins_with_a_zero_length_dash():
    # distance_left starts at the offset's remainder (core/dash.cpp:27): 0 makes the first dash a dot
    got = dashes_of(oracle.dash(*line(0, 0, 50, 0), [10, 5], 0.0))
    np.testing.assert_array_equal(got[0], [[0, 0], [0, 0]])
    assert abs(got[1][0, 0] - 5.0) < 1e-4 and abs(got[1][-1, 0] - 15.0) < 1e-4


def test_single_length_pattern_cuts_the_line_into_consecutive_pieces():
    got = dashes_of(oracle.dash(*line(3, 4, 83, 64), [7.0], 0.0))
    assert len(got) == 1 + math.ceil(100.0 / 7.0)  # the zero-length first dash, then 7-long pieces
    assert np.allclose(got[0], [[3, 4], [3, 4]]) and np.allclose(got[-1][-1], [83, 64], atol=1e-4)
    for d0, d1 in zip(got, got[1:]):
        np.testing.assert_array_equal(d0[-1], d1[0])  # each piece starts where the last ended
    for d in got[1:-1]:
        assert abs(np.linalg.norm(d[-1] - d[0]) - 7.0) < 1e-3


@pytest.mark.parametrize("pattern,offset", [([20, 10, 5, 10], 3.0), ([33, 11], 0.0), ([6, 6], 1.0), ([5, 3, 2], 4.0)])
def test_on_length_over_a_closed_square(pattern, offset):
    sq = (np.array([[10, 10], [90, 10], [90, 90], [10, 90]], np.float32), np.zeros(4, np.uint8),
          np.array([0, 4], np.int32), np.ones(1, np.uint8))
    got = dashes_of(oracle.dash(*sq, pattern, offset))
    on = sum(np.sum(np.linalg.norm(np.diff(d, axis=0), axis=1)) for d in got)
    want = sum(hi - lo for lo, hi, o in pattern_pieces(pattern, offset, 320.0) if o)
    assert abs(on - want) < 1e-2, (on, want)


def test_offset_shift_by_the_total_changes_nothing():
    cub = (np.array([[0, 0], [30, 80], [70, -40], [100, 50]], np.float32), np.array([0, 1, 2, 0], np.uint8),
           np.array([0, 4], np.int32), np.zeros(1, np.uint8))
    a = oracle.dash(*cub, [9, 4], 3.0)
    b = oracle.dash(*cub, [9, 4], 3.0 + 13.0 * 5)
    assert_same(a, b, "offset + 5 x total")


def test_odd_pattern_keeps_the_parity_of_the_index():
    # (5, 3, 2) from an offset of 0: a dot, off 3, on 2, then index 0 again: on 5 -- two "on" dashes
    # in a row, as separate contours that touch
    got = dashes_of(oracle.dash(*line(0, 0, 40, 0), [5, 3, 2], 0.0))
    lengths = [d[-1, 0] - d[0, 0] for d in got]
    np.testing.assert_allclose(lengths[:5], [0, 2, 5, 2, 5], atol=1e-4)
    np.testing.assert_array_equal(got[1][-1], got[2][0])


@pytest.mark.parametrize("pattern,offset", [([-1, 5], 0.0), ([0, 0], 0.0), ([np.nan, 3], 0.0), ([np.inf, 3], 0.0),
                                            ([4, 3], -1.0), ([4, 3], np.inf)])
def test_bad_patterns_are_refused(pattern, offset):
    with pytest.raises(ValueError, match="pfo_dash_outline"):
        oracle.dash(*line(0, 0, 10, 0), np.array(pattern, np.float32), offset)


def test_too_fine_pattern_is_refused():
    # 2^24 walk steps and more: the oracle and the device report it instead of walking on
    with pytest.raises(ValueError, match=r"\(-4\)"):
        oracle.dash(*line(0, 0, 1.0e5, 0), [1.0e-3, 1.0e-3], 0.0)
