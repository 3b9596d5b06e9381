"""The GPU stroke-to-fill (pfh_stroke_outlines, pfh_stroke.hip) through the C ABI, bit for bit:

* against the reference's own OutlineStrokeToFill on every fixture of tests/golden/stroke_fixtures.npz,
  one path per call and all of them in one call;
* against the oracle's stroker (itself pinned to those fixtures) on 1 500 random outlines of lines,
  quadratics and cubics, open and closed, with every cap and join, widths 0.5-30, in one call;
* a scene whose paths carry stroke styles is stroked at upload: tiled bit-exact like the oracle's
  stroking of the same scene (tests/test_gpu_parity.py covers that at the scene level too).
"""
import numpy as np
import pytest

import oracle
import pfh
from test_stroke import CASES, KEYS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = pfh.Context(0)
    yield c
    c.close()


def as_bytes(a):
    return a.view(np.uint8) if a.dtype == np.float32 else a


def batch(outlines):
    """Concatenate (points, flags, contour_start, closed) outlines as paths of one call."""
    pts, flg, cst, ccl, pcs = [], [], [0], [], [0]
    for p, f, cs, cc in outlines:
        pts.append(np.asarray(p, np.float32).reshape(-1, 2))
        flg.append(np.asarray(f, np.uint8))
        cst.extend((cst[-1] + np.asarray(cs[1:], np.int64)).tolist())
        ccl.extend(np.asarray(cc, np.uint8).tolist())
        pcs.append(len(ccl))
    return (np.concatenate(pts), np.concatenate(flg), np.array(cst, np.int32), np.array(ccl, np.uint8),
            np.array(pcs, np.int32))


def split(out, k):
    """Path k of a stroked batch as (points, flags, contour_start, closed)."""
    pts, flg, cst, ccl, pcs = out
    c0, c1 = pcs[k], pcs[k + 1]
    a, b = cst[c0], cst[c1]
    return pts[a:b], flg[a:b], (cst[c0:c1 + 1] - a).astype(np.int32), ccl[c0:c1]


@pytest.mark.parametrize("name", list(CASES))
def test_gpu_stroke_matches_reference_fixture(ctx, name):
    inp, style, ref = CASES[name]
    pts, flg, cst, ccl = inp
    out = ctx.stroke_outlines(pts, flg, cst, ccl, np.array([0, len(ccl)], np.int32), [style])
    for k, g, r in zip(KEYS, split(out, 0), ref):
        assert g.shape == r.shape, f"{name}: {k} {g.shape} vs {r.shape}"
        np.testing.assert_array_equal(as_bytes(g), as_bytes(r), err_msg=f"{name}: {k}")


def test_gpu_stroke_all_fixtures_in_one_call(ctx):
    names = list(CASES)
    b = batch([CASES[n][0] for n in names])
    out = ctx.stroke_outlines(*b, [CASES[n][1] for n in names])
    for k, n in enumerate(names):
        for key, g, r in zip(KEYS, split(out, k), CASES[n][2]):
            np.testing.assert_array_equal(as_bytes(g), as_bytes(r), err_msg=f"{n}: {key}")


def random_outline(rng):
    cont = []
    for _ in range(int(rng.integers(1, 3))):
        kinds = rng.integers(0, 3, int(rng.integers(1, 8)))
        pts, flg = [rng.uniform(0, 300, 2)], [0]
        for kd in kinds:
            if kd == 1:
                pts.append(rng.uniform(0, 300, 2))
                flg.append(1)
            elif kd == 2:
                pts += [rng.uniform(0, 300, 2), rng.uniform(0, 300, 2)]
                flg += [1, 2]
            pts.append(rng.uniform(0, 300, 2))
            flg.append(0)
        if rng.random() < 0.15:
            pts[-1] = pts[0].copy()
        cont.append((np.array(pts, np.float32), np.array(flg, np.uint8), int(rng.integers(0, 2))))
    p = np.concatenate([c[0] for c in cont])
    f = np.concatenate([c[1] for c in cont])
    cs = np.concatenate([[0], np.cumsum([len(c[1]) for c in cont])]).astype(np.int32)
    return p, f, cs, np.array([c[2] for c in cont], np.uint8)


def test_gpu_stroke_random_outlines_match_oracle(ctx):
    rng = np.random.default_rng(2024)
    outlines = [random_outline(rng) for _ in range(1500)]
    styles = [(float(rng.uniform(0.5, 30)), int(rng.integers(0, 3)), int(rng.integers(0, 3)), float(rng.uniform(1, 10)))
              for _ in outlines]
    out = ctx.stroke_outlines(*batch(outlines), styles)
    for k, (o, st) in enumerate(zip(outlines, styles)):
        want = oracle.stroke(*o, *st)
        for key, g, r in zip(KEYS, split(out, k), want):
            np.testing.assert_array_equal(as_bytes(g), as_bytes(r), err_msg=f"outline {k}: {key}")


def test_gpu_stroke_rejects_bad_style(ctx):
    p, f, cs, cc = random_outline(np.random.default_rng(1))
    with pytest.raises(pfh.PfhError, match="stroke style"):
        ctx.stroke_outlines(p, f, cs, cc, np.array([0, len(cc)], np.int32), [(2.0, 5, 0, 10.0)])
