"""ctypes binding of the C ABI in include/pfh.h (libpfh.so, built in-tree).

The product path is the HIP library; this module only marshals numpy arrays
across the ABI.  It fails loudly when libpfh.so is missing: there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfh.so")

PFH_PLAIN = 0x0
PFH_BATCHED = 0x1
PFH_PREDICT = 0x2
PFH_DUMP = 0x4
PFH_LOAD_DST = 0x8
PFH_NO_GRAPH = 0x10
PFH_NOALLOC_SCRATCH_BYTES = 128 << 20  # the static work array of pfh_rasterize_noalloc, per GPU

N_FEATURES = 12
MLP_OUTPUTS = 3
MLP_PARAMS = 12 * 64 + 64 + 64 * 64 + 64 + 64 * MLP_OUTPUTS + MLP_OUTPUTS
PREDICTOR_FLOATS = 24 + MLP_PARAMS

_P = ctypes.c_void_p


class SceneDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("view_box", ctypes.c_float * 4),
                ("n_paths", ctypes.c_int32), ("n_contours", ctypes.c_int32), ("n_points", ctypes.c_int32),
                ("points", _P), ("point_flags", _P), ("contour_start", _P), ("contour_closed", _P),
                ("path_contour_start", _P), ("path_fill_rule", _P), ("path_rgba", _P),
                ("n_clip_paths", ctypes.c_int32), ("path_clip", _P), ("path_stroke", _P),
                ("stroke_dashes", _P), ("n_stroke_dashes", ctypes.c_int32),
                ("path_gradient", _P), ("gradients", _P), ("n_gradients", ctypes.c_int32),
                ("gradient_stop_offsets", _P), ("gradient_stop_rgba", _P), ("n_gradient_stops", ctypes.c_int32),
                ("gradient_ramps", _P), ("clip_rule", ctypes.c_int32), ("path_blend", _P)]


class Outline(ctypes.Structure):
    _fields_ = [("n_contours", ctypes.c_int32), ("n_points", ctypes.c_int32), ("points", _P), ("point_flags", _P),
                ("contour_start", _P), ("contour_closed", _P)]


class StrokeStyle(ctypes.Structure):
    _fields_ = [("line_width", ctypes.c_float), ("line_cap", ctypes.c_int32), ("line_join", ctypes.c_int32),
                ("miter_limit", ctypes.c_float)]


class OutlineBuf(ctypes.Structure):
    _fields_ = [("points", _P), ("point_flags", _P), ("cap_points", ctypes.c_int64), ("contour_start", _P),
                ("contour_closed", _P), ("cap_contours", ctypes.c_int64), ("path_contour_start", _P),
                ("n_points", ctypes.c_int64), ("n_contours", ctypes.c_int64)]


PFH_CAP_BUTT, PFH_CAP_SQUARE, PFH_CAP_ROUND = 0, 1, 2
CLIP_RULE_ABS, CLIP_RULE_OWN = 0, 1  # pfh_scene_desc::clip_rule (PFH_CLIP_RULE_*)
# pfh_scene_desc::path_blend (PFH_BLEND_*, the order of BlendMode, core/paint/effects.h:56-123)
BLEND_MODES = ("clear", "copy", "src_in", "src_out", "src_over", "src_atop", "dest_in", "dest_out", "dest_over",
               "dest_atop", "xor", "lighter", "darken", "lighten", "multiply", "screen", "hard_light", "overlay",
               "color_dodge", "color_burn", "soft_light", "difference", "exclusion", "hue", "saturation", "color",
               "luminosity")
BLEND_SRC_OVER = 4
PFH_JOIN_MITER, PFH_JOIN_BEVEL, PFH_JOIN_ROUND = 0, 1, 2


class Stats(ctypes.Structure):
    _fields_ = [("paths", ctypes.c_int64), ("segments", ctypes.c_int64), ("tiles", ctypes.c_int64),
                ("columns", ctypes.c_int64), ("fills", ctypes.c_int64), ("alpha_tiles", ctypes.c_int64),
                ("entries", ctypes.c_int64), ("fb_tiles", ctypes.c_int64), ("retries", ctypes.c_int64),
                ("guard_hits", ctypes.c_int64), ("ms_total", ctypes.c_double), ("ms_stage", ctypes.c_double * 8),
                ("capacity_fills", ctypes.c_int64), ("capacity_alpha", ctypes.c_int64),
                ("capacity_entries", ctypes.c_int64), ("alpha_entries", ctypes.c_int64), ("lines", ctypes.c_int64),
                ("capacity_lines", ctypes.c_int64), ("host_syncs", ctypes.c_int64),
                ("fills_masked", ctypes.c_int64), ("alpha_masked", ctypes.c_int64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "ms_stage"}
        d["ms_stage"] = list(self.ms_stage)
        return d


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("queued", ctypes.c_int64),
                ("in_flight", ctypes.c_int64), ("launches", ctypes.c_int64), ("retries", ctypes.c_int64),
                ("max_concurrent_launches", ctypes.c_int64), ("mean_batch", ctypes.c_double),
                ("latency_ms_p50", ctypes.c_double), ("latency_ms_p99", ctypes.c_double),
                ("latency_ms_mean", ctypes.c_double), ("latency_ms_max", ctypes.c_double),
                ("reserved_bytes", ctypes.c_int64), ("peak_reserved_bytes", ctypes.c_int64),
                ("budget_bytes", ctypes.c_int64), ("device_bytes", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Schedule(ctypes.Structure):
    _fields_ = [("fill_grid", ctypes.c_int32), ("composite_block", ctypes.c_int32),
                ("column_block", ctypes.c_int32), ("xcd_order", ctypes.c_int32), ("flatten_segs", ctypes.c_int32),
                ("flatten_take", ctypes.c_int32), ("column_rows", ctypes.c_int32), ("flatten_deep", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


EXPORTS = ("pfh_create", "pfh_destroy", "pfh_last_error", "pfh_scene_upload", "pfh_scene_release",
           "pfh_scene_features", "pfh_rasterize", "pfh_scene_set_target", "pfh_scene_target", "pfh_readback_rgba8",
           "pfh_dump_counts", "pfh_dump", "pfh_stats_get", "pfh_set_predictor", "pfh_scene_prediction", "pfh_path_counts",
           "pfh_mlp_eval", "pfh_mlp_grad", "pfh_mlp_train", "pfh_server_start", "pfh_server_submit",
           "pfh_server_poll", "pfh_server_stats_get", "pfh_server_stop", "pfh_get_schedule",
           "pfh_set_schedule", "pfh_autotune", "pfh_stroke_outlines", "pfh_rasterize_noalloc", "pfh_gradient_compile")

_lib = None


class PfhError(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PfhError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    i32, i64, f32 = ctypes.c_int32, ctypes.c_int64, ctypes.c_float
    sig = {
        "pfh_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
        "pfh_destroy": (None, [_P]),
        "pfh_last_error": (ctypes.c_char_p, [_P]),
        "pfh_scene_upload": (ctypes.c_int, [_P, ctypes.POINTER(SceneDesc), ctypes.POINTER(i32)]),
        "pfh_scene_release": (ctypes.c_int, [_P, i32]),
        "pfh_scene_features": (ctypes.c_int, [_P, i32, _P]),
        "pfh_rasterize": (ctypes.c_int, [_P, _P, i32, ctypes.c_uint32]),
        "pfh_rasterize_noalloc": (ctypes.c_int, [_P, _P, i32, ctypes.c_uint32]),
        "pfh_gradient_compile": (ctypes.c_int, [_P, _P, _P, i32, _P, _P]),
        "pfh_stroke_outlines": (ctypes.c_int, [_P, ctypes.POINTER(Outline), _P, i32, _P, _P, i32,
                                                ctypes.POINTER(OutlineBuf)]),
        "pfh_scene_set_target": (ctypes.c_int, [_P, i32, _P, i64]),
        "pfh_scene_target": (ctypes.c_int, [_P, i32, ctypes.POINTER(_P), ctypes.POINTER(i64)]),
        "pfh_readback_rgba8": (ctypes.c_int, [_P, i32, _P, i64]),
        "pfh_dump_counts": (ctypes.c_int, [_P, i32, _P]),
        "pfh_dump": (ctypes.c_int, [_P, i32] + [_P] * 14),
        "pfh_stats_get": (ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
        "pfh_set_predictor": (ctypes.c_int, [_P, _P, i64, f32]),
        "pfh_scene_prediction": (ctypes.c_int, [_P, i32, _P]),
        "pfh_path_counts": (ctypes.c_int, [_P, i32, _P, _P, _P]),
        "pfh_mlp_eval": (ctypes.c_int, [_P, _P, _P, i64, _P]),
        "pfh_mlp_grad": (ctypes.c_int, [_P, _P, _P, _P, i64, _P, _P]),
        "pfh_mlp_train": (ctypes.c_int, [_P, _P, _P, i64, i32, i32, f32, ctypes.c_uint64, _P, _P]),
        "pfh_server_start": (ctypes.c_int, [_P, i32, i64, i32, ctypes.c_uint32]),
        "pfh_server_submit": (ctypes.c_int, [_P, i32, ctypes.POINTER(i64)]),
        "pfh_server_poll": (ctypes.c_int, [_P, ctypes.c_double, _P, i32, ctypes.POINTER(i32)]),
        "pfh_server_stats_get": (ctypes.c_int, [_P, ctypes.POINTER(ServerStats)]),
        "pfh_server_stop": (ctypes.c_int, [_P]),
        "pfh_get_schedule": (ctypes.c_int, [_P, ctypes.POINTER(Schedule)]),
        "pfh_set_schedule": (ctypes.c_int, [_P, ctypes.POINTER(Schedule)]),
        "pfh_autotune": (ctypes.c_int, [_P, _P, i32, ctypes.c_uint32, i32, ctypes.POINTER(Schedule),
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


STROKE_STYLE = np.dtype([("line_width", np.float32), ("line_cap", np.int32), ("line_join", np.int32),
                         ("miter_limit", np.float32), ("dash_start", np.int32), ("dash_count", np.int32),
                         ("dash_offset", np.float32)])


# pfh_gradient (include/pfh.h): kind, wrap, line, radii, radial transform, stop range
GRADIENT = np.dtype([("kind", np.int32), ("wrap", np.int32), ("line", np.float32, 4), ("radii", np.float32, 2),
                     ("transform", np.float32, 6), ("stop_start", np.int32), ("stop_count", np.int32)])
GRADIENT_LINEAR, GRADIENT_RADIAL = 0, 1
WRAP_CLAMP, WRAP_REPEAT = 0, 1


def gradient_compile(g, stop_offsets, stop_rgba, ramp_row=0):
    """pfh_gradient_compile on one GRADIENT record (host only): (ramp [256] u32, record [16] f32)."""
    lib = load()
    rec = np.ascontiguousarray(np.asarray(g, GRADIENT).reshape(1))
    off = np.ascontiguousarray(stop_offsets, np.float32)
    rgba = np.ascontiguousarray(stop_rgba, np.uint32)
    ramp = np.zeros(256, np.uint32)
    out = np.zeros(16, np.float32)
    rc = lib.pfh_gradient_compile(_ptr(rec), _ptr(off) if off.size else None, _ptr(rgba) if rgba.size else None,
                                  int(ramp_row), _ptr(ramp), _ptr(out))
    if rc != 0:
        raise PfhError(f"pfh_gradient_compile failed ({rc})")
    return ramp, out


def stroke_styles(styles) -> np.ndarray:
    """Rows of (line_width, line_cap, line_join, miter_limit) or of those and (dash_start,
    dash_count, dash_offset) as pfh_stroke_style records (4 columns: undashed)."""
    a = np.asarray(styles)
    if a.dtype == STROKE_STYLE:
        return a
    a = np.asarray(styles, np.float64)
    cols = a.shape[-1] if a.ndim == 2 else 4
    if cols not in (4, 7):
        raise ValueError("stroke styles have 4 or 7 columns")
    a = a.reshape(-1, cols)
    out = np.zeros(a.shape[0], STROKE_STYLE)
    out["line_width"], out["line_cap"], out["line_join"], out["miter_limit"] = a[:, 0], a[:, 1], a[:, 2], a[:, 3]
    if cols == 7:
        out["dash_start"], out["dash_count"], out["dash_offset"] = a[:, 4], a[:, 5], a[:, 6]
    return out


class Context:
    """One HIP context of the rasterizer (pfh_create)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = _P()
        rc = self.lib.pfh_create(device, ctypes.byref(h))
        if rc != 0:
            raise PfhError(f"pfh_create({device}) failed with {rc}")
        self.h = h
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.pfh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.pfh_last_error(self.h)
            raise PfhError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def upload(self, scene) -> int:
        d = SceneDesc()
        d.width, d.height = int(scene.width), int(scene.height)
        for i in range(4):
            d.view_box[i] = float(scene.view_box[i])
        d.n_paths = scene.n_paths
        d.n_contours = int(scene.contour_closed.shape[0])
        d.n_points = int(scene.point_flags.shape[0])
        d.points = _ptr(scene.points)
        d.point_flags = _ptr(scene.point_flags)
        d.contour_start = _ptr(scene.contour_start)
        d.contour_closed = _ptr(scene.contour_closed)
        d.path_contour_start = _ptr(scene.path_contour_start)
        d.path_fill_rule = _ptr(scene.path_fill_rule)
        d.path_rgba = _ptr(scene.path_rgba)
        d.n_clip_paths = int(getattr(scene, "n_clip_paths", 0))
        path_clip = getattr(scene, "path_clip", None)
        d.path_clip = None if path_clip is None else _ptr(path_clip)
        d.clip_rule = int(getattr(scene, "clip_rule", CLIP_RULE_ABS))
        path_blend = getattr(scene, "path_blend", None)
        if path_blend is not None:
            path_blend = np.ascontiguousarray(path_blend, np.uint8)
            d.path_blend = _ptr(path_blend)
        path_stroke = getattr(scene, "path_stroke", None)
        if path_stroke is not None:
            path_stroke = np.ascontiguousarray(stroke_styles(path_stroke))
        d.path_stroke = None if path_stroke is None else _ptr(path_stroke)
        dashes = getattr(scene, "stroke_dashes", None)
        if dashes is not None:
            dashes = np.ascontiguousarray(dashes, np.float32)
            d.stroke_dashes, d.n_stroke_dashes = _ptr(dashes), int(dashes.shape[0])
        pg = getattr(scene, "path_gradient", None)
        keep = []
        if pg is not None:
            pg = np.ascontiguousarray(pg, np.int32)
            grads = np.ascontiguousarray(np.asarray(scene.gradients, GRADIENT).reshape(-1))
            offs = np.ascontiguousarray(scene.gradient_stop_offsets, np.float32)
            rgba = np.ascontiguousarray(scene.gradient_stop_rgba, np.uint32)
            ramps = getattr(scene, "gradient_ramps", None)
            if ramps is not None:
                ramps = np.ascontiguousarray(ramps, np.uint32)
                d.gradient_ramps = _ptr(ramps)
            keep = [pg, grads, offs, rgba, ramps]
            d.path_gradient, d.gradients, d.n_gradients = _ptr(pg), _ptr(grads), int(grads.shape[0])
            d.gradient_stop_offsets, d.gradient_stop_rgba = _ptr(offs), _ptr(rgba)
            d.n_gradient_stops = int(offs.shape[0])
        sid = ctypes.c_int32()
        self._check(self.lib.pfh_scene_upload(self.h, ctypes.byref(d), ctypes.byref(sid)), "pfh_scene_upload")
        self._keep[sid.value] = (int(scene.width), int(scene.height), scene.n_paths)
        return sid.value

    def stroke_outlines(self, points, point_flags, contour_start, contour_closed, path_contour_start, styles,
                        dashes=None):
        """pfh_stroke_outlines: every path stroked on the device with its style (rows of
        (line_width, line_cap, line_join, miter_limit[, dash_start, dash_count, dash_offset])),
        dashed first where dash_count > 0 with the lengths `dashes`.  Returns (points [N, 2] f32,
        point_flags u8, contour_start i32, contour_closed u8, path_contour_start i32)."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 2)
        flg = np.ascontiguousarray(point_flags, np.uint8)
        cst = np.ascontiguousarray(contour_start, np.int32)
        ccl = np.ascontiguousarray(contour_closed, np.uint8)
        pcs = np.ascontiguousarray(path_contour_start, np.int32)
        st = np.ascontiguousarray(stroke_styles(styles))
        dsh = None if dashes is None else np.ascontiguousarray(dashes, np.float32)
        nd = 0 if dsh is None else int(dsh.shape[0])
        n = int(pcs.shape[0] - 1)
        o = Outline(int(ccl.shape[0]), int(flg.shape[0]), _ptr(pts), _ptr(flg), _ptr(cst), _ptr(ccl))
        out_pcs = np.zeros(n + 1, np.int32)
        cs0 = np.zeros(1, np.int32)
        buf = OutlineBuf(None, None, 0, _ptr(cs0), None, 0, _ptr(out_pcs), 0, 0)
        rc = self.lib.pfh_stroke_outlines(self.h, ctypes.byref(o), _ptr(pcs), n, _ptr(st), _ptr(dsh), nd,
                                          ctypes.byref(buf))
        if rc not in (0, -5):
            self._check(rc, "pfh_stroke_outlines")
        out = (np.zeros((buf.n_points, 2), np.float32), np.zeros(buf.n_points, np.uint8),
               np.zeros(buf.n_contours + 1, np.int32), np.zeros(buf.n_contours, np.uint8))
        buf = OutlineBuf(_ptr(out[0]), _ptr(out[1]), buf.n_points, _ptr(out[2]), _ptr(out[3]), buf.n_contours,
                         _ptr(out_pcs), 0, 0)
        self._check(self.lib.pfh_stroke_outlines(self.h, ctypes.byref(o), _ptr(pcs), n, _ptr(st), _ptr(dsh), nd,
                                                 ctypes.byref(buf)), "pfh_stroke_outlines")
        return out + (out_pcs,)

    def release(self, sid: int):
        self._check(self.lib.pfh_scene_release(self.h, sid), "pfh_scene_release")
        self._keep.pop(sid, None)

    def rasterize(self, ids, flags: int = PFH_PLAIN):
        arr = np.ascontiguousarray(np.atleast_1d(np.asarray(ids, np.int32)))
        self._check(self.lib.pfh_rasterize(self.h, _ptr(arr), int(arr.shape[0]), flags), "pfh_rasterize")

    def rasterize_noalloc(self, ids, flags: int = PFH_PLAIN):
        """pfh_rasterize_noalloc: the plain pipeline with every work buffer in the library's static
        scratch array (no device allocation during the call)."""
        arr = np.ascontiguousarray(np.atleast_1d(np.asarray(ids, np.int32)))
        self._check(self.lib.pfh_rasterize_noalloc(self.h, _ptr(arr), int(arr.shape[0]), flags),
                    "pfh_rasterize_noalloc")

    def readback(self, sid: int) -> np.ndarray:
        w, h, _ = self._keep[sid]
        out = np.zeros((h, w, 4), np.uint8)
        self._check(self.lib.pfh_readback_rgba8(self.h, sid, _ptr(out), w * 4), "pfh_readback_rgba8")
        return out

    def target(self, sid: int):
        p, pitch = _P(), ctypes.c_int64()
        self._check(self.lib.pfh_scene_target(self.h, sid, ctypes.byref(p), ctypes.byref(pitch)), "pfh_scene_target")
        return p.value, pitch.value

    def set_target(self, sid: int, device_ptr, pitch: int):
        self._check(self.lib.pfh_scene_set_target(self.h, sid, _P(device_ptr), pitch), "pfh_scene_set_target")

    def features(self, sid: int) -> np.ndarray:
        n = self._keep[sid][2]
        out = np.zeros((n, N_FEATURES), np.float32)
        self._check(self.lib.pfh_scene_features(self.h, sid, _ptr(out)), "pfh_scene_features")
        return out

    def stats(self) -> dict:
        s = Stats()
        self._check(self.lib.pfh_stats_get(self.h, ctypes.byref(s)), "pfh_stats_get")
        return s.as_dict()

    def dump(self, sid: int) -> dict:
        """Canonical intermediates (the oracle's layout) of the last PFH_DUMP rasterize."""
        counts = np.zeros(8, np.int64)
        self._check(self.lib.pfh_dump_counts(self.h, sid, _ptr(counts)), "pfh_dump_counts")
        T, W, F, E, NFB, P, A = [int(x) for x in counts[:7]]
        out = dict(
            path_rect=np.zeros((P, 4), np.int32), tile_off=np.zeros(P + 1, np.int64),
            col_off=np.zeros(P + 1, np.int64), fill_off=np.zeros(P + 1, np.int64),
            tile_backdrop=np.zeros(T, np.int8), tile_alpha=np.zeros(T, np.uint8), col_sum=np.zeros(W, np.int32),
            fill_seg=np.zeros((F, 4), np.uint16), fill_tile=np.zeros(F, np.uint32),
            entry_fbtile=np.zeros(E, np.uint32), entry_path=np.zeros(E, np.uint32), zbuf=np.zeros(NFB, np.uint32),
            fb_rect=np.zeros(4, np.int32), masks=np.zeros((A, 16, 16), np.uint8))
        keys = ("path_rect", "tile_off", "col_off", "fill_off", "tile_backdrop", "tile_alpha", "col_sum", "fill_seg",
                "fill_tile", "entry_fbtile", "entry_path", "zbuf", "fb_rect", "masks")
        self._check(self.lib.pfh_dump(self.h, sid, *[_ptr(out[k]) for k in keys]), "pfh_dump")
        out["guard_hits"] = int(counts[7])
        return out

    # launch schedule
    def schedule(self) -> dict:
        sc = Schedule()
        self._check(self.lib.pfh_get_schedule(self.h, ctypes.byref(sc)), "pfh_get_schedule")
        return sc.as_dict()

    def set_schedule(self, **kw):
        sc = Schedule(**{**self.schedule(), **kw})
        self._check(self.lib.pfh_set_schedule(self.h, ctypes.byref(sc)), "pfh_set_schedule")

    def autotune(self, ids, flags: int = PFH_PREDICT, reps: int = 5):
        arr = np.ascontiguousarray(ids, np.int32)
        sc, a, b = Schedule(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.pfh_autotune(self.h, _ptr(arr), len(arr), flags, reps, ctypes.byref(sc), ctypes.byref(a),
                                          ctypes.byref(b)), "pfh_autotune")
        return sc.as_dict(), a.value, b.value

    # request loop (continuous batching)
    def server_start(self, slots: int = 2, memory_budget: int = 8 << 30, max_batch: int = 64,
                     flags: int = PFH_PREDICT | PFH_NO_GRAPH):
        self._check(self.lib.pfh_server_start(self.h, slots, int(memory_budget), max_batch, flags), "pfh_server_start")

    def submit(self, sid: int) -> int:
        rid = ctypes.c_int64()
        self._check(self.lib.pfh_server_submit(self.h, sid, ctypes.byref(rid)), "pfh_server_submit")
        return rid.value

    def poll(self, wait_ms: float = 0.0, capacity: int = 1024) -> list:
        ids = np.zeros(capacity, np.int64)
        n = ctypes.c_int32()
        self._check(self.lib.pfh_server_poll(self.h, float(wait_ms), _ptr(ids), capacity, ctypes.byref(n)),
                    "pfh_server_poll")
        return [int(x) for x in ids[:n.value]]

    def server_stats(self) -> dict:
        st = ServerStats()
        self._check(self.lib.pfh_server_stats_get(self.h, ctypes.byref(st)), "pfh_server_stats_get")
        return st.as_dict()

    def server_stop(self):
        self._check(self.lib.pfh_server_stop(self.h), "pfh_server_stop")

    def set_predictor(self, weights, safety: float = 1.25):
        if weights is None:
            self._check(self.lib.pfh_set_predictor(self.h, None, 0, safety), "pfh_set_predictor")
            return
        w = np.ascontiguousarray(weights, np.float32)
        self._check(self.lib.pfh_set_predictor(self.h, _ptr(w), int(w.shape[0]), safety), "pfh_set_predictor")

    def prediction(self, sid: int) -> np.ndarray:
        """Predicted (chords, fills, alpha tiles) of an uploaded scene (-1 without a predictor)."""
        out = np.zeros(3, np.float32)
        self._check(self.lib.pfh_scene_prediction(self.h, sid, _ptr(out)), "pfh_scene_prediction")
        return out

    def path_counts(self, sid: int):
        """Exact per-path (chords, fills, alpha tiles) of the last PFH_DUMP rasterize."""
        n = self._keep[sid][2]
        out = [np.zeros(n, np.int32) for _ in range(3)]
        self._check(self.lib.pfh_path_counts(self.h, sid, *[_ptr(a) for a in out]), "pfh_path_counts")
        return np.stack(out, axis=1)

    # the sizing MLP's HIP kernels on host arrays
    def mlp_eval(self, params, x):
        params = np.ascontiguousarray(params, np.float32)
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_FEATURES)
        out = np.zeros((x.shape[0], MLP_OUTPUTS), np.float32)
        self._check(self.lib.pfh_mlp_eval(self.h, _ptr(params), _ptr(x), x.shape[0], _ptr(out)), "pfh_mlp_eval")
        return out

    def mlp_grad(self, params, x, target):
        params = np.ascontiguousarray(params, np.float32)
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_FEATURES)
        target = np.ascontiguousarray(target, np.float32).reshape(-1, MLP_OUTPUTS)
        grad = np.zeros(MLP_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self.lib.pfh_mlp_grad(self.h, _ptr(params), _ptr(x), _ptr(target), x.shape[0], _ptr(grad),
                                          _ptr(loss)), "pfh_mlp_grad")
        return grad, float(loss[0])

    def mlp_train(self, params, x, target, epochs=10, batch=4096, lr=1e-3, seed=1):
        params = np.ascontiguousarray(params, np.float32).copy()
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_FEATURES)
        target = np.ascontiguousarray(target, np.float32).reshape(-1, MLP_OUTPUTS)
        hist = np.zeros(max(epochs, 1), np.float32)
        self._check(self.lib.pfh_mlp_train(self.h, _ptr(x), _ptr(target), x.shape[0], int(epochs), int(batch),
                                           float(lr), int(seed), _ptr(params), _ptr(hist)), "pfh_mlp_train")
        return params, hist[:epochs]
