"""The GPU dasher (k_dash in pfh_stroke.hip, reached through pfh_stroke_outlines with dash patterns)
through the C ABI, bit for bit:

* dashed then stroked, against the reference's own OutlineDash + OutlineStrokeToFill on every
  fixture of tests/golden/dash_fixtures.npz, one path per call and all of them in one call (each
  path with its own pattern at its own place in the dash array);
* against the oracle's dasher and stroker (pinned to those fixtures) on 800 random outlines of
  lines, quadratics and cubics, open and closed, one to three contours, half of them dashed with
  patterns of one to four lengths and offsets up to twice the pattern, in one call;
* the input checks: patterns outside the dash array, negative or non-finite lengths, a zero total,
  negative offsets (PFH_E_ARG) and a pattern too fine for its path (PFH_E_LIMIT), after which the
  context keeps working.
A scene whose stroked paths carry dash patterns is covered at the scene level in
tests/test_gpu_parity.py."""
import numpy as np
import pytest

import oracle
import pfh
from test_dash import CASES, KEYS
from test_gpu_stroke import as_bytes, batch, random_outline, split

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = pfh.Context(0)
    yield c
    c.close()


def style7(style, start, count, offset):
    return tuple(style) + (start, count, offset)


@pytest.mark.parametrize("name", list(CASES))
def test_gpu_dash_matches_reference_fixture(ctx, name):
    c = CASES[name]
    pts, flg, cst, ccl = c["inp"]
    d = c["dashes"]
    out = ctx.stroke_outlines(pts, flg, cst, ccl, np.array([0, len(ccl)], np.int32),
                              [style7(c["style"], 0, len(d), c["offset"])], dashes=d)
    for k, g, r in zip(KEYS, split(out, 0), c["ref"]):
        assert g.shape == r.shape, f"{name}: {k} {g.shape} vs {r.shape}"
        np.testing.assert_array_equal(as_bytes(g), as_bytes(r), err_msg=f"{name}: {k}")


def test_gpu_dash_all_fixtures_in_one_call(ctx):
    names = list(CASES)
    dashes, styles = [], []
    for n in names:
        c = CASES[n]
        styles.append(style7(c["style"], len(dashes), len(c["dashes"]), c["offset"]))
        dashes.extend(c["dashes"].tolist())
    out = ctx.stroke_outlines(*batch([CASES[n]["inp"] for n in names]), styles, dashes=np.array(dashes, np.float32))
    for k, n in enumerate(names):
        for key, g, r in zip(KEYS, split(out, k), CASES[n]["ref"]):
            np.testing.assert_array_equal(as_bytes(g), as_bytes(r), err_msg=f"{n}: {key}")


def test_gpu_dash_random_outlines_match_oracle(ctx):
    rng = np.random.default_rng(77)
    outlines = [random_outline(rng) for _ in range(800)]
    dashes, styles, pats = [], [], []
    for _ in outlines:
        st = (float(rng.uniform(0.5, 12)), int(rng.integers(0, 3)), int(rng.integers(0, 3)), float(rng.uniform(1, 10)))
        if rng.random() < 0.5:
            pat = rng.uniform(0.0, 40.0, int(rng.integers(1, 5))).astype(np.float32)
            pat[0] = max(pat[0], np.float32(0.5))
            off = float(np.float32(rng.uniform(0.0, 2.0 * float(pat.sum()))))
            styles.append(st + (len(dashes), len(pat), off))
            dashes.extend(pat.tolist())
            pats.append((pat, off))
        else:
            styles.append(st + (0, 0, 0.0))
            pats.append(None)
    out = ctx.stroke_outlines(*batch(outlines), styles, dashes=np.array(dashes, np.float32))
    for k, (o, st, pat) in enumerate(zip(outlines, styles, pats)):
        src = o if pat is None else oracle.dash(*o, pat[0], pat[1])
        want = oracle.stroke(*src, *st[:4])
        for key, g, r in zip(KEYS, split(out, k), want):
            np.testing.assert_array_equal(as_bytes(g), as_bytes(r), err_msg=f"outline {k} ({pat}): {key}")


@pytest.mark.parametrize("style,dashes,match", [
    ((2.0, 0, 0, 10.0, 1, 2, 0.0), [4.0, 2.0], "outside the dash array"),
    ((2.0, 0, 0, 10.0, -1, 2, 0.0), [4.0, 2.0], "outside the dash array"),
    ((2.0, 0, 0, 10.0, 0, -1, 0.0), [4.0, 2.0], "outside the dash array"),
    ((2.0, 0, 0, 10.0, 0, 2, 0.0), [4.0, -2.0], "negative"),
    ((2.0, 0, 0, 10.0, 0, 2, 0.0), [np.nan, 2.0], "negative"),
    ((2.0, 0, 0, 10.0, 0, 2, 0.0), [0.0, 0.0], "zero or infinite"),
    ((2.0, 0, 0, 10.0, 0, 2, -3.0), [4.0, 2.0], "offset"),
    ((2.0, 0, 0, 10.0, 0, 2, np.inf), [4.0, 2.0], "offset"),
])
def test_gpu_dash_rejects_bad_patterns(ctx, style, dashes, match):
    p, f, cs, cc = random_outline(np.random.default_rng(3))
    with pytest.raises(pfh.PfhError, match=match):
        ctx.stroke_outlines(p, f, cs, cc, np.array([0, len(cc)], np.int32), [style],
                            dashes=np.array(dashes, np.float32))


def test_gpu_dash_too_fine_pattern_is_refused_then_context_works(ctx):
    line = (np.array([[0, 0], [1.0e5, 0]], np.float32), np.zeros(2, np.uint8), np.array([0, 2], np.int32),
            np.zeros(1, np.uint8))
    with pytest.raises(pfh.PfhError, match="2\\^24"):
        ctx.stroke_outlines(*line, np.array([0, 1], np.int32), [(1.0, 0, 0, 10.0, 0, 2, 0.0)],
                            dashes=np.array([1.0e-3, 1.0e-3], np.float32))
    c = CASES["closed_square"]
    out = ctx.stroke_outlines(*c["inp"], np.array([0, 1], np.int32),
                              [style7(c["style"], 0, len(c["dashes"]), c["offset"])], dashes=c["dashes"])
    for key, g, r in zip(KEYS, split(out, 0), c["ref"]):
        np.testing.assert_array_equal(as_bytes(g), as_bytes(r), err_msg=key)
