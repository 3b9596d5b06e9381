import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "pathfinder-cpp_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden", "hybrid_fixtures.npz")
SCENE_KEYS = ("points", "point_flags", "contour_start", "contour_closed", "path_contour_start", "path_fill_rule",
              "path_rgba", "view_box")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_golden():
    import scenegen

    z = np.load(GOLDEN, allow_pickle=False)
    out = {}
    for name in [str(n) for n in z["names"]]:
        w, h = [int(v) for v in z[f"{name}/size"]]
        kw = {k: z[f"{name}/scene/{k}"] for k in SCENE_KEYS}
        sc = scenegen.Scene(w, h, kw["points"], kw["point_flags"], kw["contour_start"], kw["contour_closed"],
                            kw["path_contour_start"], kw["path_fill_rule"], kw["path_rgba"], kw["view_box"])
        if f"{name}/scene/path_blend" in z.files:
            sc.path_blend = np.ascontiguousarray(z[f"{name}/scene/path_blend"], np.uint8)
        if f"{name}/scene/path_clip" in z.files:
            sc.n_clip_paths = int(z[f"{name}/scene/n_clip_paths"])
            sc.path_clip = np.ascontiguousarray(z[f"{name}/scene/path_clip"], np.int32)
        ref = {k[len(name) + 5:]: z[k] for k in z.files if k.startswith(f"{name}/ref/")}
        out[name] = (sc, ref)
    return out


@pytest.fixture(scope="session")
def golden():
    return load_golden()
