"""The CPU raster half that bench.py's cpu_baseline leg times (oracle.raster_cpu: float fill and
per-layer src-over composite on host threads) draws the scene: within 2 LSB of the FP64 model's
RGBA8 (per-layer float blending and float areas), whatever the thread count."""
import numpy as np

import oracle
import scenegen


def test_cpu_raster_matches_fp64_model():
    sc = scenegen.blob_scene(901, 300, 220, 120, r_max=90.0)
    want = np.rint(np.clip(oracle.full_model(sc)["image"], 0, 1) * 255.0).astype(np.int64)
    one, _ = oracle.raster_cpu(sc, threads=1)
    assert np.abs(one.astype(np.int64) - want).max() <= 2
    four, _ = oracle.raster_cpu(sc, threads=4)
    np.testing.assert_array_equal(one, four)
