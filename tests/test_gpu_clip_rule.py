"""Clip paths under their own fill rules (pfh_scene_desc::clip_rule = PFH_CLIP_RULE_OWN:
ClipPath::fill_rule honoured, which neither of the reference's renderers reads) on the device,
against the oracle: tiles, fills, backdrops and order bit-exact, masks within 1/255, RGBA8 within
1 LSB.

* The closed-form rect scenes of tests/test_oracle_model.py (every pair of draw and clip rules, a
  chain of two clips, even windings over whole tiles) -- where the oracle itself is pinned to
  min(rule_draw, rule_clip, ...) per pixel.
* Star polygons clipped by star polygons in nesting chains (windings far above 1 under both rules
  at every level), plain and batched.
* The stress workload at full size (strokes, 200 nested clip paths) under the own rules.
* Scenes of both clip modes in one batched launch: each image equals its plain run, and the two
  modes of one scene differ (the mode reaches the kernels per scene).
"""
import numpy as np
import pytest

import pfh
import scenegen
from test_gpu_parity import check_scene
from test_oracle_model import OWN_C1, OWN_C2, OWN_D, _rect_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = pfh.Context(0)
    yield c
    c.close()


def _rect_scenes():
    out = []
    for rd in (0, 1):
        for rc in (0, 1):
            out.append((f"rects d{rd} c{rc}", _rect_scene(96, 96, [OWN_D, OWN_C1], [rd, rc], 1, [0, -1])))
    for rules in ((1, 1, 1), (0, 1, 0), (1, 0, 1)):
        out.append((f"chain {rules}", _rect_scene(96, 96, [OWN_D, OWN_C1, OWN_C2], list(rules), 2, [1, -1, 0])))
    out.append(("even clip tiles", _rect_scene(96, 96, [[(0, 0, 96, 96)], [(8, 8, 88, 88), (32, 32, 72, 72)]],
                                               [0, 1], 1, [0, -1])))
    out.append(("even draw tiles", _rect_scene(96, 96, [[(0, 0, 96, 96), (16, 16, 80, 80)], [(4.5, 4.5, 91.5, 50.5)]],
                                               [1, 0], 1, [0, -1])))
    for _, sc in out:
        sc.clip_rule = pfh.CLIP_RULE_OWN
    return out


@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH])
def test_own_rules_closed_form_scenes(ctx, flags):
    for name, sc in _rect_scenes():
        check_scene(ctx, sc, flags=flags, name=f"own rules: {name}")


@pytest.mark.parametrize("seed", [61, 62])
@pytest.mark.parametrize("flags", [pfh.PFH_PLAIN, pfh.PFH_BATCHED])
def test_star_clips_own_rules(ctx, seed, flags):
    check_scene(ctx, scenegen.star_clip_scene(seed), flags=flags, name=f"star clips own rules {seed}")


def test_stress_scene_full_size_own_rules(ctx):
    scene = scenegen.stress_scene()
    scene.clip_rule = pfh.CLIP_RULE_OWN
    check_scene(ctx, scene, flags=pfh.PFH_BATCHED, name="stress full, own clip rules")


def test_clip_modes_in_one_batch(ctx):
    own = scenegen.star_clip_scene(63, size=1024, n_stars=150, n_clip=20)
    ref = scenegen.star_clip_scene(63, size=1024, n_stars=150, n_clip=20, clip_rule=pfh.CLIP_RULE_ABS)
    plain = scenegen.blob_scene(64, 512, 384, 200, r_max=128.0)
    ids = [ctx.upload(s) for s in (own, ref, plain)]
    alone = []
    for sid in ids:
        ctx.rasterize([sid], pfh.PFH_PLAIN)
        alone.append(ctx.readback(sid))
    # the oracle has the two modes differ by up to 46 LSB on 1.4 % of the pixels of this scene
    diff = np.abs(alone[0].astype(np.int64) - alone[1]).max(-1)
    assert diff.max() > 16 and (diff > 4).sum() > 1000
    for flags in (pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH, pfh.PFH_BATCHED, pfh.PFH_BATCHED):
        ctx.rasterize(ids, flags)
        for sid, img in zip(ids, alone):
            np.testing.assert_array_equal(ctx.readback(sid), img)
    for sid in ids:
        ctx.release(sid)


def test_bad_clip_rule_is_refused(ctx):
    sc = scenegen.star_clip_scene(65, size=256, n_stars=10, n_clip=3)
    sc.clip_rule = 2
    with pytest.raises(pfh.PfhError):
        ctx.upload(sc)
