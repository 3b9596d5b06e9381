"""The oracle restatement of the hybrid tiler against fixtures written by the
reference's own compiled Tiler (tests/golden/gen_golden.py): bit-exact on every
canonical array, including the reference z-buffer rule."""
import numpy as np
import pytest

import oracle

from conftest import load_golden

NAMES = list(load_golden().keys())


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference_fixture(golden, name):
    scene, ref = golden[name]
    got = oracle.run_oracle(scene, zmode=0)
    assert got["guard_hits"] == 0
    for k in oracle.CANONICAL_KEYS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=f"{name}: {k}")
    if scene.n_clip_paths:  # the reference's clip combine jobs (core/d3d9/tiler.cpp:393-403)
        for k in ("clip_combine", "clip_comb_bd"):
            np.testing.assert_array_equal(got[k], ref[k], err_msg=f"{name}: {k}")


def test_fixed_zrule_only_relaxes_evenodd_even(golden):
    """zmode 1 differs from the reference only where an opaque even-odd solid tile has an even backdrop."""
    scene, ref = golden["zcull_evenodd_even"]
    fixed = oracle.run_oracle(scene, zmode=1)
    assert (fixed["zbuf"] <= ref["zbuf"]).all()
    assert (fixed["zbuf"] < ref["zbuf"]).any()
    for name in ("zcull_opaque", "cubic_512_64x8", "blob_small_3"):
        scene, ref = golden[name]
        np.testing.assert_array_equal(oracle.run_oracle(scene, zmode=1)["zbuf"], ref["zbuf"])
