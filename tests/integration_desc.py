"""Reads the packed pfh_scene_desc that pathfinder-cpp_amd/integration/_build/check writes."""
import numpy as np


def read_desc(path):
    raw = open(path, "rb").read()
    head = np.frombuffer(raw, np.int32, 7)
    w, h, n_paths, n_contours, n_points, n_clip = [int(v) for v in head[:6]]
    off = 28
    out = {"width": w, "height": h, "n_clip_paths": n_clip}

    def take(name, dtype, count, shape=None):
        nonlocal off
        a = np.frombuffer(raw, dtype, count, off).copy()
        off += a.nbytes
        out[name] = a if shape is None else a.reshape(shape)

    take("view_box", np.float32, 4)
    take("points", np.float32, 2 * n_points, (n_points, 2))
    take("point_flags", np.uint8, n_points)
    take("contour_start", np.int32, n_contours + 1)
    take("contour_closed", np.uint8, n_contours)
    take("path_contour_start", np.int32, n_paths + 1)
    take("path_fill_rule", np.uint8, n_paths)
    take("path_rgba", np.uint32, n_paths)
    take("path_clip", np.int32, n_paths)
    n_grad, n_stops = [int(v) for v in np.frombuffer(raw, np.int32, 2, off)]
    off += 8
    if n_grad:
        import scenegen
        take("path_gradient", np.int32, n_paths)
        take("gradients", scenegen.GRADIENT, n_grad)
        take("gradient_stop_offsets", np.float32, n_stops)
        take("gradient_stop_rgba", np.uint32, n_stops)
        has_ramps = int(np.frombuffer(raw, np.int32, 1, off)[0])
        off += 4
        if has_ramps:
            take("gradient_ramps", np.uint32, n_grad * 256, (n_grad, 256))
    has_blend = int(np.frombuffer(raw, np.int32, 1, off)[0])
    off += 4
    if has_blend:
        take("path_blend", np.uint8, n_paths)
    assert off == len(raw), (off, len(raw))
    return out
