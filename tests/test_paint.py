"""Gradient paints on the CPU: the oracle's restatement (oracle/paint_oracle.cpp) against the
reference, and the library's host step (pfh_gradient_compile) against the oracle.

* Ramps bit for bit against the reference's own Gradient + GradientTileBuilder
  (oracle/_ref: core/paint/gradient.cpp compiled from its sources) on 400 random stop sets: none,
  one, equal offsets, out of order, offsets outside [0, 1], translucent colours.
* The metadata texture's float-to-half conversion bit for bit against third_party/umHalf (as
  core/renderer.cpp uses it) over normals, denormals, the underflow and overflow edges, zeros,
  infinities and NaN; the radial transform's inverse against Transform2::inverse.
* Closed forms the geometry fixes: a black-to-white linear gradient along x gives grey level
  (x - x0) / (x1 - x0) at pixel centres, a single-circle radial one |p - c| / r, repeat wraps with
  period 1, clamp holds the end colours -- each within the ramp's byte steps and the half texture's
  truncation (2/255).
* pfh_gradient_compile (libpfh, host only) equals the oracle's ramps and records bit for bit.
Runs without a GPU; tests/test_gpu_paint.py rasterizes gradient scenes on the device."""
import math

import numpy as np
import pytest

import oracle
import pfh
import scenegen

ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built")


def random_stops(rng):
    k = int(rng.integers(0, 6))
    o = rng.uniform(-0.2, 1.2, k).astype(np.float32)
    if k > 2 and rng.random() < 0.3:
        o[1] = o[2]
    if k > 1 and rng.random() < 0.2:
        o[0] = 0.0
    cols = rng.integers(0, 2 ** 32, k, dtype=np.uint64).astype(np.uint32)
    return o, cols


@ref
def test_ramp_matches_reference():
    rng = np.random.default_rng(5)
    for i in range(400):
        o, c = random_stops(rng)
        np.testing.assert_array_equal(oracle.gradient_ramp(o, c), oracle.ref_gradient_ramp(o, c), err_msg=f"set {i}")


@ref
def test_half_conversion_matches_reference():
    rng = np.random.default_rng(6)
    v = np.concatenate([
        rng.standard_normal(2000).astype(np.float32) * np.float32(2.0) ** rng.integers(-30, 20, 2000),
        np.float32([0.0, -0.0, 1.0, -1.0, 65504.0, 65520.0, 70000.0, 2.0 ** -14, 2.0 ** -15, 2.0 ** -24, 2.0 ** -25,
                    1e-8, 3.14159, np.inf, -np.inf, np.nan]),
    ]).astype(np.float32)
    bits = oracle.ref_half_bits(v)
    want = bits.view(np.float16).astype(np.float64)
    got = np.array([oracle.half_value(x) for x in v])
    both_nan = np.isnan(want) & np.isnan(got)
    assert (both_nan | (want == got)).all(), v[~(both_nan | (want == got))][:10]
    # umHalf truncates: never above the float in magnitude
    fin = np.isfinite(v) & np.isfinite(got)
    assert (np.abs(got[fin]) <= np.abs(v[fin].astype(np.float64))).all()


@ref
def test_radial_inverse_matches_reference():
    rng = np.random.default_rng(7)
    for _ in range(200):
        t = rng.uniform(-3, 3, 6).astype(np.float32)
        g = np.zeros((), scenegen.GRADIENT)
        g["kind"], g["transform"], g["radii"] = 1, t, (0, 1)
        rec = oracle.gradient_record(g)
        inv = oracle.ref_transform_inverse(t)
        want = [oracle.half_value(x) for x in inv]
        np.testing.assert_array_equal(rec[:6], np.float32(want))


def grey_linear(x0, x1, wrap):
    g = np.zeros((), scenegen.GRADIENT)
    g["kind"], g["wrap"], g["line"] = 0, wrap, (x0, 10.0, x1, 10.0)
    ramp = oracle.gradient_ramp([0.0, 1.0], [0xff000000, 0xffffffff])
    return oracle.gradient_record(g), ramp


@pytest.mark.parametrize("x0,x1", [(20.0, 220.0), (300.0, 40.0), (0.0, 1000.0)])
def test_linear_grey_ramp_closed_form(x0, x1):
    rec, ramp = grey_linear(x0, x1, 0)
    for px in np.linspace(min(x0, x1), max(x0, x1), 97)[1:-1]:
        t = (px - x0) / (x1 - x0)
        c = oracle.paint_colour(rec, ramp, px, 37.5)
        assert c is not None and abs(c[0] - t) <= 2 / 255 and c[3] == 1.0, (px, t, c)
    # clamp: the end texels beyond the line (the last holds sample(255.5 / 256), truncated to 254)
    lo = oracle.paint_colour(rec, ramp, x0 - 50 * np.sign(x1 - x0), 3.5)
    hi = oracle.paint_colour(rec, ramp, x1 + 50 * np.sign(x1 - x0), 3.5)
    assert lo[0] == 0.0 and hi[0] == 254 / 255


def test_linear_repeat_has_period_one():
    rec, ramp = grey_linear(0.0, 128.0, 1)
    for px in (10.5, 40.5, 77.5, 120.5):
        a = oracle.paint_colour(rec, ramp, px, 0.5)
        b = oracle.paint_colour(rec, ramp, px + 128.0, 0.5)
        assert np.abs(a - b).max() <= 1 / 255 and abs(a[0] - px / 128.0) <= 2 / 255


@pytest.mark.parametrize("cx,cy,r", [(100.0, 80.0, 60.0), (512.0, 300.0, 200.0)])
def test_radial_single_circle_closed_form(cx, cy, r):
    g = np.zeros((), scenegen.GRADIENT)
    g["kind"], g["line"], g["radii"], g["transform"] = 1, (cx, cy, cx, cy), (0.0, r), (1, 0, 0, 1, 0, 0)
    rec = oracle.gradient_record(g)
    ramp = oracle.gradient_ramp([0.0, 1.0], [0xff000000, 0xffffffff])
    rng = np.random.default_rng(8)
    for _ in range(200):
        px, py = cx + rng.uniform(-0.95, 0.95) * r, cy + rng.uniform(-0.95, 0.95) * r
        d = math.hypot(px - cx, py - cy) / r
        if d >= 0.99 or d < 0.01:
            continue
        c = oracle.paint_colour(rec, ramp, px, py)
        assert c is not None and abs(c[0] - d) <= 2 / 255, (px, py, d, c)
    # the centre's discriminant is 0: the reference's filter leaves the pixel transparent
    g["line"] = (100.5, 80.5, 100.5, 80.5)
    assert oracle.paint_colour(oracle.gradient_record(g), ramp, 100.5, 80.5) is None


def test_library_compile_matches_oracle():
    rng = np.random.default_rng(9)
    for i in range(300):
        o, c = random_stops(rng)
        g = np.zeros((), pfh.GRADIENT)
        g["kind"], g["wrap"] = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        g["line"] = rng.uniform(-500, 1500, 4)
        g["radii"] = rng.uniform(0, 300, 2)
        th, s = rng.uniform(0, 6.3), rng.uniform(0.3, 3)
        g["transform"] = (math.cos(th) * s, math.sin(th) * s, -math.sin(th) * s, math.cos(th) * s,
                          rng.uniform(-100, 100), rng.uniform(-100, 100))
        g["stop_start"], g["stop_count"] = 0, len(o)
        row = int(rng.integers(0, 256))
        ramp, rec = pfh.gradient_compile(g, o, c, row)
        np.testing.assert_array_equal(ramp, oracle.gradient_ramp(o, c), err_msg=f"ramp {i}")
        np.testing.assert_array_equal(rec.view(np.uint32), oracle.gradient_record(g, row).view(np.uint32),
                                      err_msg=f"record {i}")


def test_library_compile_rejects_bad_gradients():
    g = np.zeros((), pfh.GRADIENT)
    g["kind"] = 3
    with pytest.raises(pfh.PfhError):
        pfh.gradient_compile(g, [], [])
    g["kind"], g["transform"] = 1, (1, 2, 2, 4, 0, 0)  # singular
    with pytest.raises(pfh.PfhError):
        pfh.gradient_compile(g, [], [])
    g["kind"], g["line"] = 0, (0, np.nan, 1, 1)
    with pytest.raises(pfh.PfhError):
        pfh.gradient_compile(g, [], [])
