"""Many small seeded scenes on the device against the oracle, each drawn from one feature family with
random sizes and counts: fill rules and curve kinds (blobs), windings far above 1 (stars), nested
clip paths under either clip rule, dashed and stroked polylines, gradient paints, well-conditioned
blend modes (the ops that do not divide: see tests/test_gpu_blend.py for the others), odd canvas
sizes and view boxes (inside the canvas: the target outside its tiles cleared; past it on every side:
tiles at negative positions).  Tiles, fills, backdrops and order bit-exact, masks within 1/255, RGBA8 within
1 LSB (2 with gradients or blend ops), plain or batched by the seed."""
import numpy as np
import pytest

import pfh
import scenegen
from test_gpu_parity import check_scene

pytestmark = pytest.mark.gpu
FAMILIES = ("blobs", "stars", "clips", "strokes", "gradients", "blends")
# blend modes whose ops do not divide by a colour: Porter-Duff, darken .. overlay, soft light ..
# exclusion (dodge and burn, and the HSL ops, amplify the masks' quantisation: tests/test_gpu_blend.py)
WELL = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 20, 21, 22], np.uint8)


@pytest.fixture(scope="module")
def ctx():
    c = pfh.Context(0)
    yield c
    c.close()


def make(seed):
    rng = np.random.default_rng(seed)
    fam = FAMILIES[seed % len(FAMILIES)]
    w, h = int(rng.integers(17, 300)), int(rng.integers(17, 300))
    s = max(w, h)
    tol = 1
    if fam == "blobs":
        sc = scenegen.blob_scene(seed, w, h, int(rng.integers(1, 60)), r_min=2.0, r_max=float(rng.uniform(8, 160)))
    elif fam == "stars":
        sc = scenegen.star_scene(seed=seed, size=s, n_stars=int(rng.integers(1, 8)), n_min=5, n_max=200, m_max=20,
                                 r_min=4.0, r_max=float(rng.uniform(20, s)))
    elif fam == "clips":
        sc = scenegen.clip_scene(seed, size=s, n_draw=int(rng.integers(1, 40)), n_clip=int(rng.integers(1, 6)),
                                 r_max=float(rng.uniform(20, s / 2)))
        n_clip = sc.n_clip_paths
        pc = sc.path_clip.copy()
        depth = [1] * n_clip
        for q in range(1, n_clip):  # nest some clip paths in the previous one (chains of at most 4)
            if depth[q - 1] < 4 and rng.random() < 0.5:
                pc[sc.n_paths - n_clip + q] = q - 1
                depth[q] = depth[q - 1] + 1
        sc.path_clip = pc
        sc.clip_rule = int(rng.integers(0, 2))
    elif fam == "strokes":
        sc = scenegen.dashed_stroke_scene(seed, size=s, n=int(rng.integers(1, 12)))
    elif fam == "gradients":
        sc = scenegen.gradient_scene(seed, size=s, n=int(rng.integers(1, 40)))
        tol = 2
    else:
        sc = scenegen.blend_scene(seed, size=s, n=int(rng.integers(2, 40)))
        sc.path_blend = WELL[rng.integers(0, len(WELL), sc.n_paths)]
        tol = 2
    u = rng.random()
    if u < 0.2:  # a view box inside the canvas
        vb = np.sort(rng.uniform(0, [sc.width, sc.height, sc.width, sc.height]).reshape(2, 2), axis=0)
        sc.view_box = np.array([vb[0, 0], vb[0, 1], vb[1, 0] + 1, vb[1, 1] + 1], np.float32)
    elif u < 0.3:  # one reaching past the canvas on every side (negative origin)
        sc.view_box = np.array([-rng.uniform(1, 40), -rng.uniform(1, 40), sc.width + rng.uniform(1, 40),
                                sc.height + rng.uniform(1, 40)], np.float32)
    return fam, sc, tol


@pytest.mark.parametrize("seed", range(120))
def test_random_scene(ctx, seed):
    fam, sc, tol = make(seed)
    flags = pfh.PFH_PLAIN if seed % 2 == 0 else pfh.PFH_BATCHED
    check_scene(ctx, sc, flags=flags, name=f"random {fam} seed {seed}", rgba_tol=tol)
