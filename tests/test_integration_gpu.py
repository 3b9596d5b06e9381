"""RendererHip (the reference-side binding, integration/) draws a Scene built through the reference's
API on a DeviceHip into a TextureHip, read back through CommandEncoder::read_texture and a QueueHip
submit: within 2 LSB of the oracle's FP64 composite of the packed scene on the 300x220 canvas (the
bound for gradient paints; see test_integration.py for the packing).  The driver binary is built beforehand by integration/build.sh."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import scenegen
from integration_desc import read_desc
from test_integration import CHECK

pytestmark = pytest.mark.gpu


def test_renderer_hip_draws_the_reference_scene(tmp_path):
    if not os.path.exists(CHECK):
        pytest.skip("integration/_build/check not built (integration/build.sh needs the reference tree)")
    desc, img = str(tmp_path / "desc.bin"), str(tmp_path / "img.bin")
    print("[integration] check draw", flush=True)
    r = subprocess.run([CHECK, "draw", desc, img], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = read_desc(desc)
    w, h = 300, 220  # the destination texture's size
    got = np.fromfile(img, np.uint8).reshape(h, w, 4)
    sc = scenegen.Scene(w, h, d["points"], d["point_flags"], d["contour_start"], d["contour_closed"],
                        d["path_contour_start"], d["path_fill_rule"], d["path_rgba"], d["view_box"])
    sc.n_clip_paths = d["n_clip_paths"]
    sc.path_clip = d["path_clip"]
    for k in ("path_gradient", "gradients", "gradient_stop_offsets", "gradient_stop_rgba", "gradient_ramps",
              "path_blend"):
        setattr(sc, k, d.get(k))
    want = np.rint(np.clip(oracle.full_model(sc, zmode=1)["image"], 0.0, 1.0) * 255.0).astype(np.int64)
    # 2 LSB: the scene has gradient paints (the stated bound of tests/test_gpu_paint.py)
    assert np.abs(got.astype(np.int64) - want).max() <= 2
    assert (got[..., 3] > 0).mean() > 0.2  # it drew something
