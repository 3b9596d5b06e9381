// Host runtime behind include/pfh.h: scene packing and residency, batch
// planning, the plain (stage-by-stage) and batched (graph-replayed) pipelines,
// capacity control, and canonical dumps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pfh.h"
#include "pfh_internal.h"

namespace pfh {
hipError_t launch_fill_u32(uint32_t *p, uint32_t value, uint64_t n, hipStream_t st);
size_t scan_temp_elems(uint64_t n);
hipError_t scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *tmp, hipStream_t s);
hipError_t scan_entries(uint32_t *flags_inout, uint32_t n, uint32_t *counters, uint32_t *tmp, hipStream_t s);
hipError_t scan_tiles(const BatchDev &b, void *tmp, hipStream_t s);
hipError_t launch_flatten(const BatchDev &b, bool emit, hipStream_t s);
hipError_t launch_bin(const BatchDev &b, bool emit, uint32_t grid_lines, hipStream_t s);
hipError_t scan_counted(uint32_t *inout, const uint32_t *n_ptr, uint32_t cap, uint32_t *counters, int counter,
                        uint32_t *tmp, hipStream_t s);
hipError_t scan_total(uint32_t *inout, uint32_t n, uint32_t *counters, int counter, uint32_t *tmp, hipStream_t s);
hipError_t launch_propagate(const BatchDev &b, hipStream_t s);
hipError_t launch_group_fills(const BatchDev &b, uint32_t grid_fills, hipStream_t s);
hipError_t launch_entries(const BatchDev &b, bool write, hipStream_t s);
size_t radix_temp_elems(uint64_t cap);
hipError_t radix_sort(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt, const uint32_t *n_ptr,
                      uint32_t cap, int bits, uint32_t *tmp, bool *result_alt, hipStream_t s);
hipError_t launch_fill_masks(const BatchDev &b, uint32_t grid_alpha, hipStream_t s);
hipError_t launch_composite(const BatchDev &b, const uint32_t *sorted_vals, hipStream_t s);
hipError_t mlp_forward(const float *params, const float *x, int64_t n, float *hidden, float *out, hipStream_t s);
hipError_t mlp_backward(const float *params, const float *x, int64_t n, const float *hidden, const float *out,
                        const float *target, float *grad, float *loss, float *scratch, hipStream_t s);
hipError_t mlp_adam(float *params, float *m, float *v, const float *grad, int64_t n, float lr, int step,
                    hipStream_t s);
int64_t mlp_param_count();
hipError_t mlp_predict_sum(const float *params, const float *x, int64_t n, float *sums, hipStream_t s);
}  // namespace pfh

using namespace pfh;

namespace {

constexpr int N_FEATURES = 12;

struct PfhError {
    int code;
    std::string msg;
};

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) throw PfhError{PFH_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)}; \
    } while (0)

// Bumped whenever a device buffer is (re)allocated: captured graphs hold raw pointers.
static uint64_t g_alloc_generation = 0;

// Grow-only device buffer (the reference grows buffers to 2^k, core/d3d11/renderer.cpp:719-724).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    void ensure(size_t need) {
        if (need <= bytes && ptr) return;
        size_t cap = 256;
        while (cap < need) cap <<= 1;
        if (ptr) HIP_TRY(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
        g_alloc_generation++;
        hipError_t e = hipMalloc(&ptr, cap);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            ptr = nullptr;
            throw PfhError{PFH_E_NOMEM, "hipMalloc(" + std::to_string(cap) + ") failed"};
        }
        bytes = cap;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const {
        return reinterpret_cast<T *>(ptr);
    }
};

template <typename T>
static void upload(DevBuf &buf, const std::vector<T> &v) {
    buf.ensure(std::max<size_t>(v.size() * sizeof(T), 4));
    if (!v.empty()) HIP_TRY(hipMemcpy(buf.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

struct SceneHost {
    bool live = false;
    int32_t width = 0, height = 0;
    float vb[4] = {0, 0, 0, 0};
    int32_t fb_x0 = 0, fb_y0 = 0, fb_w = 0, fb_h = 0;
    uint32_t n_paths = 0, n_segs = 0, n_tiles = 0, n_cols = 0;
    std::vector<int4> path_rect;
    std::vector<uint32_t> path_tile_off, path_col_off, path_seg_off, path_rgba;
    std::vector<uint8_t> path_rule;
    std::vector<float> features;
    DevBuf d_features;
    DevBuf d_pts, d_segs, d_seg_path, d_path_rect, d_path_tile_off, d_path_col_off, d_col_path, d_path_rgba,
        d_path_rule, d_out_owned;
    uint8_t *d_out = nullptr;
    int64_t out_pitch = 0;
    // placement in the last rasterized batch (for dumps)
    bool dumped = false;
    uint32_t seg_base = 0, tile_base = 0, col_base = 0, fb_base = 0;
    // exact counts of the last run (capacity planning)
    bool counted = false;
    uint32_t last_fills = 0, last_alpha = 0, last_entries = 0, last_lines = 0;
    float pred_fills = -1, pred_alpha = -1, pred_lines = -1;
};

struct GraphKey {
    std::vector<int32_t> ids;
    uint32_t flags;
    uint32_t cap_f, cap_a, cap_e, cap_l;
    bool operator<(const GraphKey &o) const {
        if (ids != o.ids) return ids < o.ids;
        if (flags != o.flags) return flags < o.flags;
        if (cap_f != o.cap_f) return cap_f < o.cap_f;
        if (cap_a != o.cap_a) return cap_a < o.cap_a;
        if (cap_l != o.cap_l) return cap_l < o.cap_l;
        return cap_e < o.cap_e;
    }
};

}  // namespace

struct pfh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    std::vector<std::unique_ptr<SceneHost>> scenes;
    // batch work buffers
    DevBuf scene_table, bases, tile_delta, tile_cnt, tile_fill_off, tile_alpha_idx, tile_bd, tile_entry, col_bd,
        seg_cnt, lines, line_seg, line_cnt, fills, dump_fills, alpha_rec, masks, zbuf, fb_cnt, keys, vals, keys_alt, vals_alt, entry_rec,
        counters, scan_tmp, sort_tmp;
    uint32_t *h_counters = nullptr;  // pinned
    SceneDev *h_table = nullptr;     // pinned scene table
    size_t h_table_cap = 0;
    uint32_t *h_bases = nullptr;
    size_t h_bases_cap = 0;
    bool last_sorted_alt = false;
    std::map<GraphKey, hipGraphExec_t> graphs;
    uint64_t graphs_generation = 0;  // g_alloc_generation the cached graphs were captured under
    // predictor
    std::vector<float> predictor;
    float predictor_safety = 1.25f;
    DevBuf mlp_params, mlp_x, mlp_h, mlp_out;
    pfh_stats stats{};
};

namespace {

// core/scene.cpp:170-179 + common/math/basic.h:93-108: right edge rounded up to a multiple of 16.
static void scene_view_box(const float *in, float *out) {
    out[0] = in[0];
    out[1] = in[1];
    out[3] = in[3];
    int number = (int)std::ceil(in[2]);
    int rem = std::abs(number) % 16;
    int rounded = number;
    if (rem != 0) rounded = number < 0 ? -(std::abs(number) - rem) : number + 16 - rem;
    out[2] = (float)rounded;
}

struct RectF {
    float l = 1, t = 1, r = 0, b = 0;
    bool valid() const { return l <= r && t <= b; }
};

// common/math/rect.h:279-291
static RectF intersect(const RectF &a, const RectF &o) {
    if (a.l > o.r || a.r < o.l || a.t > o.b || a.b < o.t) return RectF{};
    return RectF{std::max(a.l, o.l), std::max(a.t, o.t), std::min(a.r, o.r), std::min(a.b, o.b)};
}

static void validate_desc(const pfh_scene_desc *d) {
    if (!d) throw PfhError{PFH_E_ARG, "null scene desc"};
    if (d->width <= 0 || d->height <= 0) throw PfhError{PFH_E_ARG, "canvas size must be positive"};
    if (d->n_paths < 0 || d->n_contours < 0 || d->n_points < 0) throw PfhError{PFH_E_ARG, "negative counts"};
    if ((d->n_points && (!d->points || !d->point_flags)) || !d->contour_start || !d->path_contour_start ||
        (d->n_contours && !d->contour_closed) || (d->n_paths && (!d->path_fill_rule || !d->path_rgba)))
        throw PfhError{PFH_E_ARG, "null array"};
    if (d->contour_start[0] != 0 || d->contour_start[d->n_contours] != d->n_points)
        throw PfhError{PFH_E_ARG, "contour_start must run from 0 to n_points"};
    if (d->path_contour_start[0] != 0 || d->path_contour_start[d->n_paths] != d->n_contours)
        throw PfhError{PFH_E_ARG, "path_contour_start must run from 0 to n_contours"};
    for (int c = 0; c < d->n_contours; c++)
        if (d->contour_start[c + 1] < d->contour_start[c]) throw PfhError{PFH_E_ARG, "contour_start not monotone"};
    for (int p = 0; p < d->n_paths; p++)
        if (d->path_contour_start[p + 1] < d->path_contour_start[p])
            throw PfhError{PFH_E_ARG, "path_contour_start not monotone"};
    if ((int64_t)d->width * d->height > (int64_t)1 << 30) throw PfhError{PFH_E_LIMIT, "canvas larger than 2^30 pixels"};
    if (d->n_paths >= (1 << 24)) throw PfhError{PFH_E_LIMIT, "more than 2^24 paths in a scene"};
}

// Features for the sizing predictor, per path (see pfh_scene_features).
static void path_features(const SceneHost &s, const std::vector<float2> &pts, const std::vector<uint2> &segs,
                          uint32_t p, float *f) {
    const int4 &r = s.path_rect[p];
    float side = (float)std::max(s.width, s.height);
    float nl = 0, nq = 0, nc = 0, poly = 0, turn = 0;
    float bx0 = 1e30f, by0 = 1e30f, bx1 = -1e30f, by1 = -1e30f;
    float2 prev_dir = make_float2(0, 0);
    bool have_prev = false;
    for (uint32_t sg = s.path_seg_off[p]; sg < s.path_seg_off[p + 1]; sg++) {
        uint32_t k = segs[sg].y, fp = segs[sg].x;
        int np = k == SEG_LINE ? 2 : (k == SEG_QUAD ? 3 : 4);
        if (k == SEG_LINE) nl += 1;
        else if (k == SEG_QUAD) nq += 1;
        else nc += 1;
        for (int i = 0; i < np; i++) {
            float2 q = pts[fp + i];
            bx0 = std::min(bx0, q.x);
            by0 = std::min(by0, q.y);
            bx1 = std::max(bx1, q.x);
            by1 = std::max(by1, q.y);
            if (i > 0) {
                float2 a = pts[fp + i - 1];
                float dx = q.x - a.x, dy = q.y - a.y;
                float len = std::sqrt(dx * dx + dy * dy);
                poly += len;
                if (len > 0) {
                    float2 dir = make_float2(dx / len, dy / len);
                    if (have_prev) {
                        float c = std::max(-1.0f, std::min(1.0f, dir.x * prev_dir.x + dir.y * prev_dir.y));
                        turn += std::acos(c);
                    }
                    prev_dir = dir;
                    have_prev = true;
                }
            }
        }
    }
    if (bx0 > bx1) bx0 = by0 = bx1 = by1 = 0;
    f[0] = bx0 / side;
    f[1] = by0 / side;
    f[2] = bx1 / side;
    f[3] = by1 / side;
    f[4] = std::log1p(poly);
    f[5] = std::log1p(turn);
    f[6] = std::log1p(nl);
    f[7] = std::log1p(nq);
    f[8] = std::log1p(nc);
    f[9] = std::log1p((float)std::max(0, r.z - r.x) * (float)std::max(0, r.w - r.y));
    f[10] = std::log(side);
    f[11] = (float)s.path_rule[p];
}

static void predict_scene(pfh_ctx *ctx, SceneHost &s);

static int scene_upload(pfh_ctx *ctx, const pfh_scene_desc *d) {
    validate_desc(d);
    auto s = std::make_unique<SceneHost>();
    s->width = d->width;
    s->height = d->height;
    scene_view_box(d->view_box, s->vb);
    // framebuffer tile rect: round_out(view_box / 16)
    s->fb_x0 = (int32_t)std::floor(s->vb[0] * (1.0f / 16.0f));
    s->fb_y0 = (int32_t)std::floor(s->vb[1] * (1.0f / 16.0f));
    s->fb_w = (int32_t)std::ceil(s->vb[2] * (1.0f / 16.0f)) - s->fb_x0;
    s->fb_h = (int32_t)std::ceil(s->vb[3] * (1.0f / 16.0f)) - s->fb_y0;
    if (s->fb_w <= 0 || s->fb_h <= 0) throw PfhError{PFH_E_ARG, "empty view box"};
    RectF vb{s->vb[0], s->vb[1], s->vb[2], s->vb[3]};

    const float2 *P = reinterpret_cast<const float2 *>(d->points);
    std::vector<float2> pts(P, P + d->n_points);
    std::vector<uint2> segs;
    std::vector<uint32_t> seg_path;
    s->n_paths = (uint32_t)d->n_paths;
    s->path_rect.resize(d->n_paths);
    s->path_tile_off.assign(d->n_paths + 1, 0);
    s->path_col_off.assign(d->n_paths + 1, 0);
    s->path_seg_off.assign(d->n_paths + 1, 0);
    s->path_rgba.assign(d->path_rgba, d->path_rgba + d->n_paths);
    s->path_rule.assign(d->path_fill_rule, d->path_fill_rule + d->n_paths);
    std::vector<uint32_t> col_path;
    uint64_t tiles = 0, cols = 0;
    for (int p = 0; p < d->n_paths; p++) {
        RectF bounds;  // Outline::bounds: min/max of every point (core/data/contour.cpp:73-81)
        for (int c = d->path_contour_start[p]; c < d->path_contour_start[p + 1]; c++) {
            int a = d->contour_start[c], e = d->contour_start[c + 1];
            for (int i = a; i < e; i++) {
                float x = P[i].x, y = P[i].y;
                if (!bounds.valid()) bounds = RectF{x, y, x, y};
                else {
                    bounds.l = std::min(bounds.l, x);
                    bounds.t = std::min(bounds.t, y);
                    bounds.r = std::max(bounds.r, x);
                    bounds.b = std::max(bounds.b, y);
                }
            }
            // SegmentsIter::get_next(force_closed = true), core/data/contour.cpp:110-173
            int n = e - a;
            if (n == 0) continue;
            const uint8_t *fl = d->point_flags + a;
            int head = 0;
            while (true) {
                if (head >= n || fl[head] != 0) break;
                uint32_t kind = 0, first = (uint32_t)(a + head);
                if (head + 1 < n) {
                    if (fl[head + 1] == 0) {
                        kind = SEG_LINE;
                        head += 1;
                    } else if (head + 2 < n) {
                        if (fl[head + 1] == 1 && fl[head + 2] == 0) {
                            kind = SEG_QUAD;
                            head += 2;
                        } else if (head + 3 < n && fl[head + 1] == 1 && fl[head + 2] == 2 && fl[head + 3] == 0) {
                            kind = SEG_CUBIC;
                            head += 3;
                        }
                    }
                    if (kind == 0) break;
                    segs.push_back(make_uint2(first, kind));
                    seg_path.push_back((uint32_t)p);
                } else {
                    if (n > 1) {
                        float2 f0 = P[a], fl1 = P[e - 1];
                        float dx = f0.x - fl1.x, dy = f0.y - fl1.y;
                        float len = std::sqrt(dx * dx + dy * dy);
                        if (!(len <= 0.0001f)) {
                            // closing line (last -> first): its two points are appended
                            uint32_t fp = (uint32_t)pts.size();
                            pts.push_back(fl1);
                            pts.push_back(f0);
                            segs.push_back(make_uint2(fp, SEG_LINE));
                            seg_path.push_back((uint32_t)p);
                        }
                    }
                    break;
                }
            }
        }
        s->path_seg_off[p + 1] = (uint32_t)segs.size();
        // core/data/built_path.cpp:8-34 + data.h:53-55: tile rect of bounds ∩ view box
        RectF bi = intersect(bounds, vb);
        int4 r;
        r.x = (int32_t)std::floor(bi.l * (1.0f / 16.0f));
        r.y = (int32_t)std::floor(bi.t * (1.0f / 16.0f));
        r.z = (int32_t)std::ceil(bi.r * (1.0f / 16.0f));
        r.w = (int32_t)std::ceil(bi.b * (1.0f / 16.0f));
        // DenseTileMap w x h tiles and BuiltPathData::backdrops of width w (built_path.cpp:25-27)
        int w = std::max(0, r.z - r.x), h = std::max(0, r.w - r.y);
        s->path_rect[p] = r;
        tiles += (uint64_t)w * h;
        for (int x = 0; x < w; x++) col_path.push_back((uint32_t)p);
        cols += (uint64_t)w;
        if (tiles >= 0xF0000000ull) throw PfhError{PFH_E_LIMIT, "scene has more than ~2^32 tiles"};
        s->path_tile_off[p + 1] = (uint32_t)tiles;
        s->path_col_off[p + 1] = (uint32_t)cols;
    }
    if (segs.size() >= 0xF0000000ull || pts.size() >= 0xF0000000ull) throw PfhError{PFH_E_LIMIT, "too many segments"};
    s->n_segs = (uint32_t)segs.size();
    s->n_tiles = (uint32_t)tiles;
    s->n_cols = (uint32_t)cols;
    s->features.resize((size_t)d->n_paths * N_FEATURES);
    for (int p = 0; p < d->n_paths; p++) path_features(*s, pts, segs, (uint32_t)p, &s->features[(size_t)p * N_FEATURES]);

    HIP_TRY(hipSetDevice(ctx->device));
    upload(s->d_features, s->features);
    upload(s->d_pts, pts);
    upload(s->d_segs, segs);
    upload(s->d_seg_path, seg_path);
    upload(s->d_path_rect, s->path_rect);
    upload(s->d_path_tile_off, s->path_tile_off);
    upload(s->d_path_col_off, s->path_col_off);
    upload(s->d_col_path, col_path);
    upload(s->d_path_rgba, s->path_rgba);
    upload(s->d_path_rule, s->path_rule);
    s->out_pitch = (int64_t)s->width * 4;
    s->d_out_owned.ensure((size_t)s->out_pitch * s->height);
    s->d_out = s->d_out_owned.as<uint8_t>();
    s->live = true;
    predict_scene(ctx, *s);
    ctx->scenes.push_back(std::move(s));
    return (int)ctx->scenes.size() - 1;
}

static SceneHost &get_scene(pfh_ctx *ctx, int32_t id) {
    if (id < 0 || id >= (int32_t)ctx->scenes.size() || !ctx->scenes[id] || !ctx->scenes[id]->live)
        throw PfhError{PFH_E_ARG, "unknown scene id " + std::to_string(id)};
    return *ctx->scenes[id];
}

// ---------------------------------------------------------------------------
// Predictor: device inference at upload / set-predictor time, never inside a rasterize
// ---------------------------------------------------------------------------

static void predict_scene(pfh_ctx *ctx, SceneHost &s) {
    if (ctx->predictor.empty()) {
        s.pred_fills = s.pred_alpha = s.pred_lines = -1;
        return;
    }
    float sums[3] = {0, 0, 0};
    if (s.n_paths > 0) {
        ctx->mlp_out.ensure(16);
        HIP_TRY(mlp_predict_sum(ctx->mlp_params.as<float>(), s.d_features.as<float>(), s.n_paths,
                                ctx->mlp_out.as<float>(), ctx->stream));
        HIP_TRY(hipMemcpyAsync(sums, ctx->mlp_out.ptr, 12, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    s.pred_lines = sums[0];
    s.pred_fills = sums[1];
    s.pred_alpha = sums[2];
}

// ---------------------------------------------------------------------------
// Pipeline
// ---------------------------------------------------------------------------

struct Plan {
    std::vector<int32_t> ids;
    uint64_t S = 0, T = 0, W = 0, FB = 0;
    uint32_t cap_f = 0, cap_a = 0, cap_e = 0, cap_l = 0;
    int key_bits = 1;
};

static uint32_t clamp_cap(uint64_t v) { return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(v, 64), 0xF0000000ull); }

static void ensure_sizes(pfh_ctx *ctx, const Plan &pl, bool dump) {
    ctx->tile_delta.ensure(4 * pl.T + 4);
    ctx->tile_cnt.ensure(4 * pl.T + 4);
    ctx->tile_fill_off.ensure(4 * pl.T + 8);
    ctx->tile_alpha_idx.ensure(4 * pl.T + 8);
    ctx->tile_bd.ensure(pl.T + 4);
    ctx->tile_entry.ensure(4 * pl.T + 8);
    ctx->col_bd.ensure(4 * pl.W + 4);
    ctx->seg_cnt.ensure(4 * pl.S + 8);
    ctx->lines.ensure(16ull * pl.cap_l);
    ctx->line_seg.ensure(4ull * pl.cap_l);
    ctx->line_cnt.ensure(4ull * pl.cap_l + 8);
    ctx->zbuf.ensure(4 * pl.FB + 4);
    ctx->fb_cnt.ensure(4 * pl.FB + 8);
    ctx->counters.ensure(4 * C_COUNT);
    uint64_t scan_n = std::max<uint64_t>({pl.T + 1, pl.S + 1, pl.FB + 1, (uint64_t)pl.cap_l + 1, 1});
    ctx->scan_tmp.ensure(16 * scan_temp_elems(scan_n) + 64);
    ctx->fills.ensure(8ull * pl.cap_f);
    if (dump) ctx->dump_fills.ensure(16ull * pl.cap_f);
    ctx->alpha_rec.ensure(16ull * pl.cap_a);
    ctx->masks.ensure(256ull * pl.cap_a);
    ctx->keys.ensure(4ull * pl.cap_e);
    ctx->vals.ensure(4ull * pl.cap_e);
    ctx->keys_alt.ensure(4ull * pl.cap_e);
    ctx->vals_alt.ensure(4ull * pl.cap_e);
    ctx->entry_rec.ensure(16ull * pl.cap_e);
    ctx->sort_tmp.ensure(4 * radix_temp_elems(pl.cap_e) + 64);
}

// Capacities the current allocations hold, by the same per-buffer sizes ensure_sizes asks
// for, so that ensure_sizes on these capacities never reallocates (captured graphs stay valid).
struct Caps {
    uint32_t l, f, a;
};
static Caps held_caps(const pfh_ctx *ctx, bool dump) {
    auto fit = [](size_t bytes, size_t per, size_t extra) -> uint64_t {
        return bytes > extra ? (bytes - extra) / per : 0;
    };
    uint64_t l = std::min({fit(ctx->lines.bytes, 16, 0), fit(ctx->line_seg.bytes, 4, 0), fit(ctx->line_cnt.bytes, 4, 8)});
    uint64_t f = fit(ctx->fills.bytes, 8, 0);
    if (dump) f = std::min(f, fit(ctx->dump_fills.bytes, 16, 0));
    uint64_t a = std::min(fit(ctx->alpha_rec.bytes, 16, 0), fit(ctx->masks.bytes, 256, 0));
    return Caps{clamp_cap(l), clamp_cap(f), clamp_cap(a)};
}

static void write_tables(pfh_ctx *ctx, Plan &pl, bool load_dst) {
    size_t n = pl.ids.size();
    if (ctx->h_table_cap < n) {
        if (ctx->h_table) (void)hipHostFree(ctx->h_table);
        HIP_TRY(hipHostMalloc((void **)&ctx->h_table, sizeof(SceneDev) * n));
        ctx->h_table_cap = n;
    }
    if (ctx->h_bases_cap < 4 * (n + 1)) {
        if (ctx->h_bases) (void)hipHostFree(ctx->h_bases);
        HIP_TRY(hipHostMalloc((void **)&ctx->h_bases, sizeof(uint32_t) * 4 * (n + 1)));
        ctx->h_bases_cap = 4 * (n + 1);
    }
    uint64_t S = 0, T = 0, W = 0, FB = 0;
    for (size_t k = 0; k < n; k++) {
        SceneHost &s = get_scene(ctx, pl.ids[k]);
        SceneDev &d = ctx->h_table[k];
        d.pts = s.d_pts.as<float2>();
        d.segs = s.d_segs.as<uint2>();
        d.seg_path = s.d_seg_path.as<uint32_t>();
        d.path_rect = s.d_path_rect.as<int4>();
        d.path_tile_off = s.d_path_tile_off.as<uint32_t>();
        d.path_col_off = s.d_path_col_off.as<uint32_t>();
        d.col_path = s.d_col_path.as<uint32_t>();
        d.path_rgba = s.d_path_rgba.as<uint32_t>();
        d.path_rule = s.d_path_rule.as<uint8_t>();
        d.vb_l = s.vb[0];
        d.vb_t = s.vb[1];
        d.vb_r = s.vb[2];
        d.vb_b = s.vb[3];
        d.fb_x0 = s.fb_x0;
        d.fb_y0 = s.fb_y0;
        d.fb_w = s.fb_w;
        d.fb_h = s.fb_h;
        d.width = s.width;
        d.height = s.height;
        d.n_paths = s.n_paths;
        d.n_segs = s.n_segs;
        d.n_tiles = s.n_tiles;
        d.n_cols = s.n_cols;
        d.n_fb = (uint32_t)(s.fb_w * s.fb_h);
        d.seg_base = (uint32_t)S;
        d.path_base = 0;
        d.tile_base = (uint32_t)T;
        d.col_base = (uint32_t)W;
        d.fb_base = (uint32_t)FB;
        d.out = s.d_out;
        d.out_pitch = s.out_pitch;
        d.load_dst = load_dst ? 1u : 0u;
        d.pad = 0;
        s.seg_base = d.seg_base;
        s.tile_base = d.tile_base;
        s.col_base = d.col_base;
        s.fb_base = d.fb_base;
        ctx->h_bases[0 * (n + 1) + k] = (uint32_t)S;
        ctx->h_bases[1 * (n + 1) + k] = (uint32_t)T;
        ctx->h_bases[2 * (n + 1) + k] = (uint32_t)W;
        ctx->h_bases[3 * (n + 1) + k] = (uint32_t)FB;
        S += s.n_segs;
        T += s.n_tiles;
        W += s.n_cols;
        FB += d.n_fb;
        if (S >= 0xF0000000ull || T >= 0xF0000000ull || W >= 0xF0000000ull || FB >= 0xF0000000ull)
            throw PfhError{PFH_E_LIMIT, "batch exceeds 2^32 segments/tiles"};
    }
    ctx->h_bases[0 * (n + 1) + n] = (uint32_t)S;
    ctx->h_bases[1 * (n + 1) + n] = (uint32_t)T;
    ctx->h_bases[2 * (n + 1) + n] = (uint32_t)W;
    ctx->h_bases[3 * (n + 1) + n] = (uint32_t)FB;
    pl.S = S;
    pl.T = T;
    pl.W = W;
    pl.FB = FB;
    pl.key_bits = 1;
    while ((1ull << pl.key_bits) < FB) pl.key_bits++;
}

static void upload_tables(pfh_ctx *ctx, const Plan &pl) {
    size_t n = pl.ids.size();
    ctx->scene_table.ensure(sizeof(SceneDev) * n);
    ctx->bases.ensure(sizeof(uint32_t) * 4 * (n + 1));
    HIP_TRY(hipMemcpyAsync(ctx->scene_table.ptr, ctx->h_table, sizeof(SceneDev) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->bases.ptr, ctx->h_bases, sizeof(uint32_t) * 4 * (n + 1), hipMemcpyHostToDevice,
                           ctx->stream));
}

static BatchDev batch_dev(pfh_ctx *ctx, const Plan &pl, bool dump) {
    BatchDev b{};
    size_t n = pl.ids.size();
    b.scenes = ctx->scene_table.as<SceneDev>();
    b.n_scenes = (int32_t)n;
    const uint32_t *bases = ctx->bases.as<uint32_t>();
    b.seg_bases = bases;
    b.tile_bases = bases + (n + 1);
    b.col_bases = bases + 2 * (n + 1);
    b.fb_bases = bases + 3 * (n + 1);
    b.n_segs = (uint32_t)pl.S;
    b.n_tiles = (uint32_t)pl.T;
    b.n_cols = (uint32_t)pl.W;
    b.n_fb = (uint32_t)pl.FB;
    b.cap_fills = pl.cap_f;
    b.cap_alpha = pl.cap_a;
    b.cap_entries = pl.cap_e;
    b.cap_lines = pl.cap_l;
    b.tile_delta = ctx->tile_delta.as<int32_t>();
    b.tile_cnt = ctx->tile_cnt.as<uint32_t>();
    b.tile_fill_off = ctx->tile_fill_off.as<uint32_t>();
    b.tile_alpha_idx = ctx->tile_alpha_idx.as<uint32_t>();
    b.tile_bd = ctx->tile_bd.as<int8_t>();
    b.tile_entry = ctx->tile_entry.as<uint32_t>();
    b.col_bd = ctx->col_bd.as<int32_t>();
    b.seg_cnt = ctx->seg_cnt.as<uint32_t>();
    b.lines = ctx->lines.as<float4>();
    b.line_seg = ctx->line_seg.as<uint32_t>();
    b.line_cnt = ctx->line_cnt.as<uint32_t>();
    b.fills = ctx->fills.as<uint2>();
    b.dump_fills = dump ? ctx->dump_fills.as<uint4>() : nullptr;
    b.alpha_rec = ctx->alpha_rec.as<uint4>();
    b.masks = ctx->masks.as<uint8_t>();
    b.zbuf = ctx->zbuf.as<uint32_t>();
    b.fb_cnt = ctx->fb_cnt.as<uint32_t>();
    b.keys = ctx->keys.as<uint32_t>();
    b.vals = ctx->vals.as<uint32_t>();
    b.keys_alt = ctx->keys_alt.as<uint32_t>();
    b.vals_alt = ctx->vals_alt.as<uint32_t>();
    b.entry_rec = ctx->entry_rec.as<uint4>();
    b.counters = ctx->counters.as<uint32_t>();
    b.dump = dump ? 1 : 0;
    return b;
}

// Stage groups; every launch goes to ctx->stream so the batched sequence can be captured.
static void stage_clear(pfh_ctx *ctx, const BatchDev &b) {
    hipStream_t s = ctx->stream;
    HIP_TRY(launch_fill_u32((uint32_t *)b.tile_delta, 0, b.n_tiles, s));
    HIP_TRY(launch_fill_u32(b.tile_cnt, 0, b.n_tiles, s));
    HIP_TRY(launch_fill_u32((uint32_t *)b.col_bd, 0, b.n_cols, s));
    HIP_TRY(launch_fill_u32(b.zbuf, 0, b.n_fb, s));
    HIP_TRY(launch_fill_u32(b.fb_cnt, 0, (uint64_t)b.n_fb + 1, s));
    HIP_TRY(launch_fill_u32(b.counters, 0, C_COUNT, s));
}

// flatten (count, scan, emit) and bin count: chords and exact fill / alpha-tile counts
static void stage_count(pfh_ctx *ctx, const BatchDev &b) {
    hipStream_t s = ctx->stream;
    uint32_t *tmp = ctx->scan_tmp.as<uint32_t>();
    HIP_TRY(launch_flatten(b, false, s));
    HIP_TRY(scan_total(b.seg_cnt, b.n_segs, b.counters, C_LINES, tmp, s));
    HIP_TRY(launch_flatten(b, true, s));
    HIP_TRY(launch_bin(b, false, b.cap_lines, s));
    HIP_TRY(scan_tiles(b, ctx->scan_tmp.ptr, s));
}

static void stage_emit(pfh_ctx *ctx, const BatchDev &b) {
    hipStream_t s = ctx->stream;
    uint32_t *tmp = ctx->scan_tmp.as<uint32_t>();
    if (b.dump) HIP_TRY(scan_counted(b.line_cnt, b.counters + C_LINES, b.cap_lines, b.counters, -1, tmp, s));
    HIP_TRY(launch_bin(b, true, b.cap_lines, s));
    HIP_TRY(launch_propagate(b, s));
    HIP_TRY(launch_entries(b, false, s));
    HIP_TRY(scan_entries(b.tile_entry, b.n_tiles, b.counters, tmp, s));
}

static void stage_finish(pfh_ctx *ctx, const BatchDev &b, const Plan &pl) {
    HIP_TRY(launch_entries(b, true, ctx->stream));
    HIP_TRY(scan_u32(b.fb_cnt, b.fb_cnt, b.n_fb, ctx->scan_tmp.as<uint32_t>(), ctx->stream));
    bool alt = false;
    HIP_TRY(radix_sort(b.keys, b.vals, b.keys_alt, b.vals_alt, b.counters + C_ENTRIES, b.cap_entries, pl.key_bits,
                       ctx->sort_tmp.as<uint32_t>(), &alt, ctx->stream));
    ctx->last_sorted_alt = alt;
    HIP_TRY(launch_fill_masks(b, b.cap_alpha, ctx->stream));
    HIP_TRY(launch_composite(b, alt ? b.vals_alt : b.vals, ctx->stream));
}

static void read_counters(pfh_ctx *ctx) {
    ctx->stats.host_syncs++;
    HIP_TRY(hipMemcpyAsync(ctx->h_counters, ctx->counters.ptr, 4 * C_COUNT, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
}

static double elapsed_ms(pfh_ctx *ctx, hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, a, b));
    return ms;
}

static void record_counts(pfh_ctx *ctx, const Plan &pl) {
    ctx->stats.fills = ctx->h_counters[C_FILLS];
    ctx->stats.alpha_tiles = ctx->h_counters[C_ALPHA];
    ctx->stats.entries = ctx->h_counters[C_ENTRIES];
    ctx->stats.guard_hits = ctx->h_counters[C_GUARD];
    ctx->stats.alpha_entries = ctx->h_counters[C_ALPHA_ENTRIES];
    ctx->stats.lines = ctx->h_counters[C_LINES];
    ctx->stats.capacity_lines = pl.cap_l;
    ctx->stats.tiles = (int64_t)pl.T;
    ctx->stats.columns = (int64_t)pl.W;
    ctx->stats.segments = (int64_t)pl.S;
    ctx->stats.fb_tiles = (int64_t)pl.FB;
    ctx->stats.capacity_fills = pl.cap_f;
    ctx->stats.capacity_alpha = pl.cap_a;
    ctx->stats.capacity_entries = pl.cap_e;
}

// Plain: one scene, stage by stage, exact sizes read back between stages
// (RendererD3D11::prepare_tiles, core/d3d11/renderer.cpp:510-616).
static void rasterize_plain(pfh_ctx *ctx, int32_t id, uint32_t flags) {
    bool dump = (flags & PFH_DUMP) != 0;
    Plan pl;
    pl.ids = {id};
    write_tables(ctx, pl, (flags & PFH_LOAD_DST) != 0);
    Caps held = held_caps(ctx, dump);
    pl.cap_f = held.f;
    pl.cap_a = held.a;
    pl.cap_l = held.l;
    pl.cap_e = clamp_cap(std::max<uint64_t>(pl.T, 1));
    ensure_sizes(ctx, pl, dump);
    upload_tables(ctx, pl);
    BatchDev b = batch_dev(ctx, pl, dump);
    hipStream_t st = ctx->stream;
    // scratch is re-read at every use: ensure_sizes may reallocate it when a count grows
    auto tmp = [ctx]() { return ctx->scan_tmp.as<uint32_t>(); };
    hipEvent_t ev[9];
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(ev[0], st));
    stage_clear(ctx, b);
    HIP_TRY(launch_flatten(b, false, st));
    HIP_TRY(scan_total(b.seg_cnt, b.n_segs, b.counters, C_LINES, tmp(), st));
    read_counters(ctx);  // chord count (renderer.cpp:707-724)
    uint32_t nl = ctx->h_counters[C_LINES];
    if (nl > pl.cap_l) {
        pl.cap_l = clamp_cap(nl);
        ensure_sizes(ctx, pl, dump);
        b = batch_dev(ctx, pl, dump);
    }
    HIP_TRY(launch_flatten(b, true, st));
    HIP_TRY(hipEventRecord(ev[1], st));
    HIP_TRY(launch_bin(b, false, nl, st));
    HIP_TRY(scan_tiles(b, ctx->scan_tmp.ptr, st));
    read_counters(ctx);  // fill / alpha counts (renderer.cpp:834-845)
    HIP_TRY(hipEventRecord(ev[2], st));
    uint32_t nf = ctx->h_counters[C_FILLS], na = ctx->h_counters[C_ALPHA];
    if (nf > pl.cap_f || na > pl.cap_a) {
        pl.cap_f = clamp_cap(std::max<uint64_t>(nf, pl.cap_f));
        pl.cap_a = clamp_cap(std::max<uint64_t>(na, pl.cap_a));
        ensure_sizes(ctx, pl, dump);
        b = batch_dev(ctx, pl, dump);
    }
    if (dump) HIP_TRY(scan_counted(b.line_cnt, b.counters + C_LINES, nl, b.counters, -1, tmp(), st));
    HIP_TRY(launch_bin(b, true, nl, st));
    HIP_TRY(hipEventRecord(ev[3], st));
    HIP_TRY(launch_propagate(b, st));
    HIP_TRY(hipEventRecord(ev[4], st));
    HIP_TRY(launch_entries(b, false, st));
    HIP_TRY(scan_entries(b.tile_entry, b.n_tiles, b.counters, tmp(), st));
    HIP_TRY(launch_entries(b, true, st));
    HIP_TRY(hipEventRecord(ev[5], st));
    HIP_TRY(scan_u32(b.fb_cnt, b.fb_cnt, b.n_fb, tmp(), st));
    bool alt = false;
    HIP_TRY(radix_sort(b.keys, b.vals, b.keys_alt, b.vals_alt, b.counters + C_ENTRIES, b.cap_entries, pl.key_bits,
                       ctx->sort_tmp.as<uint32_t>(), &alt, st));
    ctx->last_sorted_alt = alt;
    HIP_TRY(hipEventRecord(ev[6], st));
    HIP_TRY(launch_fill_masks(b, na, st));
    HIP_TRY(hipEventRecord(ev[7], st));
    HIP_TRY(launch_composite(b, alt ? b.vals_alt : b.vals, st));
    HIP_TRY(hipEventRecord(ev[8], st));
    read_counters(ctx);
    for (int i = 0; i < 8; i++) ctx->stats.ms_stage[i] += elapsed_ms(ctx, ev[i], ev[i + 1]);
    for (auto &e : ev) (void)hipEventDestroy(e);
    record_counts(ctx, pl);
    SceneHost &s = get_scene(ctx, id);
    s.counted = true;
    s.last_lines = nl;
    s.last_fills = ctx->h_counters[C_FILLS];
    s.last_alpha = ctx->h_counters[C_ALPHA];
    s.last_entries = ctx->h_counters[C_ENTRIES];
    s.dumped = dump;
}

// Batched: every scene in one pass, no host round trip inside; the launch sequence is
// captured once per (scene set, capacities, flags) and replayed from a HIP graph.
static void rasterize_batched(pfh_ctx *ctx, const int32_t *ids, int32_t n, uint32_t flags) {
    bool dump = (flags & PFH_DUMP) != 0;
    bool use_graph = (flags & PFH_NO_GRAPH) == 0;
    Plan pl;
    pl.ids.assign(ids, ids + n);
    write_tables(ctx, pl, (flags & PFH_LOAD_DST) != 0);
    // Capacities: entries are bounded by tiles exactly; fills and alpha tiles come from
    // the predictor, else from the last exact counts, else from a count pass read back.
    uint64_t pf = 0, pa = 0, plc = 0;
    bool known = true;
    for (int32_t id : pl.ids) {
        SceneHost &s = get_scene(ctx, id);
        if ((flags & PFH_PREDICT) && !ctx->predictor.empty()) {
            plc += (uint64_t)std::ceil(s.pred_lines * ctx->predictor_safety) + 64;
            pf += (uint64_t)std::ceil(s.pred_fills * ctx->predictor_safety) + 64;
            pa += (uint64_t)std::ceil(s.pred_alpha * ctx->predictor_safety) + 16;
        } else if (s.counted) {
            plc += s.last_lines;
            pf += s.last_fills;
            pa += s.last_alpha;
        } else {
            known = false;
        }
    }
    pl.cap_e = clamp_cap(std::max<uint64_t>(pl.T, 1));
    for (int attempt = 0; attempt < 4; attempt++) {
        if (!known) {
            // size from an eager count pass (read back), as the plain pipeline does
            Caps held = held_caps(ctx, dump);
            pl.cap_l = held.l;
            pl.cap_f = held.f;
            pl.cap_a = held.a;
            ensure_sizes(ctx, pl, dump);
            upload_tables(ctx, pl);
            BatchDev b = batch_dev(ctx, pl, dump);
            stage_clear(ctx, b);
            HIP_TRY(launch_flatten(b, false, ctx->stream));
            HIP_TRY(scan_total(b.seg_cnt, b.n_segs, b.counters, C_LINES, ctx->scan_tmp.as<uint32_t>(), ctx->stream));
            read_counters(ctx);
            plc = ctx->h_counters[C_LINES];
            pl.cap_l = clamp_cap(std::max<uint64_t>(plc, pl.cap_l));
            ensure_sizes(ctx, pl, dump);
            b = batch_dev(ctx, pl, dump);
            stage_clear(ctx, b);
            stage_count(ctx, b);
            read_counters(ctx);
            pf = ctx->h_counters[C_FILLS];
            pa = ctx->h_counters[C_ALPHA];
            known = true;
        }
        Caps held = held_caps(ctx, dump);
        pl.cap_l = clamp_cap(std::max<uint64_t>(plc, held.l));
        pl.cap_f = clamp_cap(std::max<uint64_t>(pf, held.f));
        pl.cap_a = clamp_cap(std::max<uint64_t>(pa, held.a));
        ensure_sizes(ctx, pl, dump);
        // ensure() rounds to 2^k: use the full allocation, then size the buffers that
        // depend on it (scan scratch) for exactly those capacities, so that a repeat call
        // asks for nothing new and replays its graph
        for (int fix = 0; fix < 4; fix++) {
            held = held_caps(ctx, dump);
            if (fix > 0 && held.l == pl.cap_l && held.f == pl.cap_f && held.a == pl.cap_a) break;
            pl.cap_l = held.l;
            pl.cap_f = held.f;
            pl.cap_a = held.a;
            ensure_sizes(ctx, pl, dump);
        }
        upload_tables(ctx, pl);
        BatchDev b = batch_dev(ctx, pl, dump);
        HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
        if (use_graph) {
            GraphKey key{pl.ids, flags & ~(uint32_t)PFH_PREDICT, pl.cap_f, pl.cap_a, pl.cap_e, pl.cap_l};
            if (ctx->graphs_generation != g_alloc_generation) {
                // some buffer moved since these graphs were captured: never replay them
                for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
                ctx->graphs.clear();
                ctx->graphs_generation = g_alloc_generation;
            }
            auto it = ctx->graphs.find(key);
            if (it == ctx->graphs.end()) {
                if (ctx->graphs.size() > 32) {
                    for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
                    ctx->graphs.clear();
                }
                hipGraph_t graph;
                HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                try {
                    stage_clear(ctx, b);
                    stage_count(ctx, b);
                    stage_emit(ctx, b);
                    stage_finish(ctx, b, pl);
                } catch (...) {
                    hipGraph_t dummy;
                    (void)hipStreamEndCapture(ctx->stream, &dummy);
                    throw;
                }
                HIP_TRY(hipStreamEndCapture(ctx->stream, &graph));
                hipGraphExec_t exec;
                HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                (void)hipGraphDestroy(graph);
                it = ctx->graphs.emplace(key, exec).first;
            } else {
                // the sort's result buffer parity depends only on the key width
                ctx->last_sorted_alt = ((pl.key_bits + 7) / 8) % 2 == 1;
            }
            HIP_TRY(hipGraphLaunch(it->second, ctx->stream));
        } else {
            stage_clear(ctx, b);
            stage_count(ctx, b);
            stage_emit(ctx, b);
            stage_finish(ctx, b, pl);
        }
        HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
        read_counters(ctx);
        uint32_t nl = ctx->h_counters[C_LINES], nf = ctx->h_counters[C_FILLS], na = ctx->h_counters[C_ALPHA];
        if (nl <= pl.cap_l && nf <= pl.cap_f && na <= pl.cap_a) {
            ctx->stats.ms_total = elapsed_ms(ctx, ctx->ev0, ctx->ev1);
            record_counts(ctx, pl);
            for (int32_t id : pl.ids) {
                SceneHost &s = get_scene(ctx, id);
                s.dumped = dump;
            }
            return;
        }
        // under-prediction: a retry with the exact counts, never a wrong pixel.  When the
        // chord buffer overflowed, the fill and alpha counts are partial: keep growing.
        ctx->stats.retries++;
        plc = std::max<uint64_t>(plc, nl);
        pf = std::max<uint64_t>(pf, nl > pl.cap_l ? 2ull * nf : nf);
        pa = std::max<uint64_t>(pa, nl > pl.cap_l ? 2ull * na : na);
    }
    throw PfhError{PFH_E_OVERFLOW, "capacity still exceeded after retries"};
}


static int run_guarded(pfh_ctx *ctx, const std::function<void()> &f) {
    try {
        f();
        return PFH_OK;
    } catch (const PfhError &e) {
        if (ctx) ctx->err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        return PFH_E_ARG;
    }
}
}  // namespace

extern "C" {

int pfh_create(int device, pfh_ctx **out) {
    if (!out) return PFH_E_ARG;
    *out = nullptr;
    auto *ctx = new pfh_ctx();
    ctx->device = device;
    int rc = run_guarded(ctx, [&] {
        int n = 0;
        HIP_TRY(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw PfhError{PFH_E_ARG, "no HIP device " + std::to_string(device)};
        HIP_TRY(hipSetDevice(device));
        HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&ctx->ev0));
        HIP_TRY(hipEventCreate(&ctx->ev1));
        HIP_TRY(hipHostMalloc((void **)&ctx->h_counters, 4 * C_COUNT));
    });
    if (rc != PFH_OK) {
        fprintf(stderr, "pfh_create: %s\n", ctx->err.c_str());
        pfh_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return PFH_OK;
}

void pfh_destroy(pfh_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
    for (auto &s : ctx->scenes) {
        if (!s) continue;
        for (DevBuf *b : {&s->d_features, &s->d_pts, &s->d_segs, &s->d_seg_path, &s->d_path_rect, &s->d_path_tile_off,
                          &s->d_path_col_off, &s->d_col_path, &s->d_path_rgba, &s->d_path_rule, &s->d_out_owned})
            b->release();
    }
    for (DevBuf *b : {&ctx->scene_table, &ctx->bases, &ctx->tile_delta, &ctx->tile_cnt, &ctx->tile_fill_off,
                      &ctx->tile_alpha_idx, &ctx->tile_bd, &ctx->tile_entry, &ctx->col_bd, &ctx->seg_cnt, &ctx->lines,
                      &ctx->line_seg, &ctx->line_cnt, &ctx->fills,
                      &ctx->dump_fills, &ctx->alpha_rec, &ctx->masks, &ctx->zbuf, &ctx->fb_cnt, &ctx->keys, &ctx->vals,
                      &ctx->keys_alt, &ctx->vals_alt, &ctx->entry_rec, &ctx->counters, &ctx->scan_tmp, &ctx->sort_tmp,
                      &ctx->mlp_params, &ctx->mlp_x, &ctx->mlp_h, &ctx->mlp_out})
        b->release();
    if (ctx->h_counters) (void)hipHostFree(ctx->h_counters);
    if (ctx->h_table) (void)hipHostFree(ctx->h_table);
    if (ctx->h_bases) (void)hipHostFree(ctx->h_bases);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *pfh_last_error(const pfh_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pfh_scene_upload(pfh_ctx *ctx, const pfh_scene_desc *desc, int32_t *scene_id) {
    if (!ctx || !scene_id) return PFH_E_ARG;
    return run_guarded(ctx, [&] { *scene_id = scene_upload(ctx, desc); });
}

int pfh_scene_release(pfh_ctx *ctx, int32_t scene_id) {
    if (!ctx) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        SceneHost &s = get_scene(ctx, scene_id);
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
        ctx->graphs.clear();
        for (DevBuf *b : {&s.d_features, &s.d_pts, &s.d_segs, &s.d_seg_path, &s.d_path_rect, &s.d_path_tile_off, &s.d_path_col_off,
                          &s.d_col_path, &s.d_path_rgba, &s.d_path_rule, &s.d_out_owned})
            b->release();
        ctx->scenes[scene_id].reset();
    });
}

int pfh_scene_features(pfh_ctx *ctx, int32_t scene_id, float *out) {
    if (!ctx || !out) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        SceneHost &s = get_scene(ctx, scene_id);
        std::memcpy(out, s.features.data(), s.features.size() * sizeof(float));
    });
}

int pfh_rasterize(pfh_ctx *ctx, const int32_t *scene_ids, int32_t n, uint32_t flags) {
    if (!ctx || (n > 0 && !scene_ids) || n < 0) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        HIP_TRY(hipSetDevice(ctx->device));
        for (int32_t i = 0; i < n; i++) get_scene(ctx, scene_ids[i]);
        if (n == 0) return;
        ctx->stats.host_syncs = 0;
        ctx->stats.paths = 0;
        for (int32_t i = 0; i < n; i++) ctx->stats.paths += get_scene(ctx, scene_ids[i]).n_paths;
        if (flags & PFH_BATCHED) {
            rasterize_batched(ctx, scene_ids, n, flags);
        } else {
            for (double &m : ctx->stats.ms_stage) m = 0;
            HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
            for (int32_t i = 0; i < n; i++) rasterize_plain(ctx, scene_ids[i], flags);
            HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
            HIP_TRY(hipEventSynchronize(ctx->ev1));
            ctx->stats.ms_total = elapsed_ms(ctx, ctx->ev0, ctx->ev1);
        }
    });
}

int pfh_scene_set_target(pfh_ctx *ctx, int32_t scene_id, void *device_rgba8, int64_t pitch_bytes) {
    if (!ctx) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        SceneHost &s = get_scene(ctx, scene_id);
        if (device_rgba8 == nullptr) {
            s.d_out = s.d_out_owned.as<uint8_t>();
            s.out_pitch = (int64_t)s.width * 4;
            return;
        }
        if (pitch_bytes < (int64_t)s.width * 4 || (pitch_bytes & 3)) throw PfhError{PFH_E_ARG, "bad pitch"};
        s.d_out = reinterpret_cast<uint8_t *>(device_rgba8);
        s.out_pitch = pitch_bytes;
        for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
        ctx->graphs.clear();
    });
}

int pfh_scene_target(pfh_ctx *ctx, int32_t scene_id, void **device_rgba8, int64_t *pitch_bytes) {
    if (!ctx || !device_rgba8 || !pitch_bytes) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        SceneHost &s = get_scene(ctx, scene_id);
        *device_rgba8 = s.d_out;
        *pitch_bytes = s.out_pitch;
    });
}

int pfh_readback_rgba8(pfh_ctx *ctx, int32_t scene_id, void *dst, int64_t pitch_bytes) {
    if (!ctx || !dst) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        SceneHost &s = get_scene(ctx, scene_id);
        if (pitch_bytes < (int64_t)s.width * 4) throw PfhError{PFH_E_ARG, "bad pitch"};
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipMemcpy2D(dst, (size_t)pitch_bytes, s.d_out, (size_t)s.out_pitch, (size_t)s.width * 4,
                            (size_t)s.height, hipMemcpyDeviceToHost));
    });
}

int pfh_dump_counts(pfh_ctx *ctx, int32_t scene_id, int64_t *counts) {
    if (!ctx || !counts) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        SceneHost &s = get_scene(ctx, scene_id);
        if (!s.dumped) throw PfhError{PFH_E_STATE, "scene was not rasterized with PFH_DUMP"};
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        uint32_t f0 = 0, f1 = 0, a0 = 0, a1 = 0, e0 = 0, e1 = 0, l0 = 0, l1 = 0;
        uint32_t sb = s.seg_base, tb = s.tile_base, fb = s.fb_base, nfb = (uint32_t)(s.fb_w * s.fb_h);
        HIP_TRY(hipMemcpy(&l0, ctx->seg_cnt.as<uint32_t>() + sb, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&l1, ctx->seg_cnt.as<uint32_t>() + sb + s.n_segs, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&f0, ctx->line_cnt.as<uint32_t>() + l0, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&f1, ctx->line_cnt.as<uint32_t>() + l1, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&a0, ctx->tile_alpha_idx.as<uint32_t>() + tb, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&a1, ctx->tile_alpha_idx.as<uint32_t>() + tb + s.n_tiles, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&e0, ctx->fb_cnt.as<uint32_t>() + fb, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&e1, ctx->fb_cnt.as<uint32_t>() + fb + nfb, 4, hipMemcpyDeviceToHost));
        counts[0] = s.n_tiles;
        counts[1] = s.n_cols;
        counts[2] = f1 - f0;
        counts[3] = e1 - e0;
        counts[4] = nfb;
        counts[5] = s.n_paths;
        counts[6] = a1 - a0;
        counts[7] = ctx->h_counters[C_GUARD];
    });
}

int pfh_dump(pfh_ctx *ctx, int32_t scene_id, int32_t *path_rect, int64_t *tile_off, int64_t *col_off,
             int64_t *fill_off, int8_t *tile_backdrop, uint8_t *tile_alpha, int32_t *col_sum, uint16_t *fill_seg,
             uint32_t *fill_tile, uint32_t *entry_fbtile, uint32_t *entry_path, uint32_t *zbuf, int32_t *fb_rect,
             uint8_t *masks) {
    if (!ctx) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        SceneHost &s = get_scene(ctx, scene_id);
        if (!s.dumped) throw PfhError{PFH_E_STATE, "scene was not rasterized with PFH_DUMP"};
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        uint32_t sb = s.seg_base, tb = s.tile_base, cb = s.col_base, fbb = s.fb_base;
        uint32_t nfb = (uint32_t)(s.fb_w * s.fb_h);
        auto d2h = [](void *dst, const void *src, size_t bytes) {
            if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        };
        // chords of a path start at its first segment's chord offset; fills of a chord at
        // that chord's scanned fill offset
        std::vector<uint32_t> seg_line(s.n_segs + 1);
        d2h(seg_line.data(), ctx->seg_cnt.as<uint32_t>() + sb, 4 * (s.n_segs + 1));
        uint32_t l0 = seg_line[0], l1 = seg_line[s.n_segs];
        std::vector<uint32_t> line_fill(l1 - l0 + 1);
        d2h(line_fill.data(), ctx->line_cnt.as<uint32_t>() + l0, 4 * (size_t)(l1 - l0 + 1));
        for (uint32_t p = 0; p < s.n_paths; p++) std::memcpy(path_rect + 4 * p, &s.path_rect[p], 16);
        for (uint32_t p = 0; p <= s.n_paths; p++) {
            tile_off[p] = s.path_tile_off[p];
            col_off[p] = s.path_col_off[p];
            uint32_t first_line = seg_line[s.path_seg_off[p]] - l0;
            fill_off[p] = (int64_t)line_fill[first_line] - line_fill[0];
        }
        d2h(tile_backdrop, ctx->tile_bd.as<int8_t>() + tb, s.n_tiles);
        std::vector<uint32_t> foff(s.n_tiles + 1);
        d2h(foff.data(), ctx->tile_fill_off.as<uint32_t>() + tb, 4 * (s.n_tiles + 1));
        for (uint32_t t = 0; t < s.n_tiles; t++) tile_alpha[t] = foff[t + 1] > foff[t] ? 1 : 0;
        d2h(col_sum, ctx->col_bd.as<int32_t>() + cb, 4 * s.n_cols);
        uint32_t nf = line_fill[l1 - l0] - line_fill[0];
        std::vector<uint4> df(nf);
        d2h(df.data(), ctx->dump_fills.as<uint4>() + line_fill[0], 16 * (size_t)nf);
        uint32_t p = 0;
        for (uint32_t i = 0; i < nf; i++) {
            while (p + 1 <= s.n_paths && (int64_t)i >= fill_off[p + 1]) p++;
            fill_seg[4 * i + 0] = (uint16_t)(df[i].x & 0xffff);
            fill_seg[4 * i + 1] = (uint16_t)(df[i].x >> 16);
            fill_seg[4 * i + 2] = (uint16_t)(df[i].y & 0xffff);
            fill_seg[4 * i + 3] = (uint16_t)(df[i].y >> 16);
            fill_tile[i] = df[i].z - tb - s.path_tile_off[p];
        }
        std::vector<uint32_t> fboff(nfb + 1);
        d2h(fboff.data(), ctx->fb_cnt.as<uint32_t>() + fbb, 4 * (nfb + 1));
        uint32_t ne = fboff[nfb] - fboff[0];
        std::vector<uint32_t> sv(ne);
        const uint32_t *sorted = ctx->last_sorted_alt ? ctx->vals_alt.as<uint32_t>() : ctx->vals.as<uint32_t>();
        d2h(sv.data(), sorted + fboff[0], 4 * (size_t)ne);
        uint32_t elo = 0xffffffffu, ehi = 0;
        for (uint32_t v : sv) {
            elo = std::min(elo, v);
            ehi = std::max(ehi, v + 1);
        }
        if (ne) {
            std::vector<uint4> recs(ehi - elo);
            d2h(recs.data(), ctx->entry_rec.as<uint4>() + elo, 16 * (size_t)(ehi - elo));
            for (uint32_t i = 0; i < ne; i++) {
                entry_fbtile[i] = recs[sv[i] - elo].w;
                entry_path[i] = recs[sv[i] - elo].z;
            }
        }
        d2h(zbuf, ctx->zbuf.as<uint32_t>() + fbb, 4 * (size_t)nfb);
        fb_rect[0] = s.fb_x0;
        fb_rect[1] = s.fb_y0;
        fb_rect[2] = s.fb_x0 + s.fb_w;
        fb_rect[3] = s.fb_y0 + s.fb_h;
        if (masks) {
            uint32_t a0 = 0, a1 = 0;
            d2h(&a0, ctx->tile_alpha_idx.as<uint32_t>() + tb, 4);
            d2h(&a1, ctx->tile_alpha_idx.as<uint32_t>() + tb + s.n_tiles, 4);
            d2h(masks, ctx->masks.as<uint8_t>() + 256ull * a0, 256ull * (a1 - a0));
        }
    });
}

int pfh_stats_get(pfh_ctx *ctx, pfh_stats *out) {
    if (!ctx || !out) return PFH_E_ARG;
    *out = ctx->stats;
    return PFH_OK;
}

int pfh_set_predictor(pfh_ctx *ctx, const float *weights, int64_t n_floats, float safety) {
    if (!ctx) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        const int64_t np = mlp_param_count();
        if (!(safety >= 1.0f)) throw PfhError{PFH_E_ARG, "safety factor must be >= 1"};
        if (weights == nullptr) {
            ctx->predictor.clear();
        } else {
            if (n_floats != 24 + np) throw PfhError{PFH_E_ARG, "predictor needs " + std::to_string(24 + np) + " floats"};
            ctx->predictor.assign(weights, weights + n_floats);
            // fold the feature normalisation (x - mean) * inv_std into the first layer
            std::vector<float> p(weights + 24, weights + 24 + np);
            const float *mean = weights, *istd = weights + 12;
            float *W1 = p.data(), *b1 = p.data() + 12 * 64;
            for (int j = 0; j < 64; j++) {
                double shift = 0;
                for (int i = 0; i < 12; i++) {
                    shift += (double)mean[i] * istd[i] * W1[i * 64 + j];
                    W1[i * 64 + j] *= istd[i];
                }
                b1[j] = (float)(b1[j] - shift);
            }
            ctx->mlp_params.ensure(4 * np);
            HIP_TRY(hipMemcpy(ctx->mlp_params.ptr, p.data(), 4 * np, hipMemcpyHostToDevice));
        }
        ctx->predictor_safety = safety;
        for (auto &s : ctx->scenes)
            if (s && s->live) predict_scene(ctx, *s);
    });
}

/* Predicted (chords, fills, alpha tiles) of an uploaded scene, or -1 without a predictor. */
int pfh_scene_prediction(pfh_ctx *ctx, int32_t scene_id, float *out3) {
    if (!ctx || !out3) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        SceneHost &s = get_scene(ctx, scene_id);
        out3[0] = s.pred_lines;
        out3[1] = s.pred_fills;
        out3[2] = s.pred_alpha;
    });
}

int pfh_path_counts(pfh_ctx *ctx, int32_t scene_id, int32_t *lines, int32_t *fills, int32_t *alpha_tiles) {
    if (!ctx || !lines || !fills || !alpha_tiles) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        SceneHost &s = get_scene(ctx, scene_id);
        if (!s.dumped) throw PfhError{PFH_E_STATE, "scene was not rasterized with PFH_DUMP"};
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::vector<uint32_t> seg_line(s.n_segs + 1);
        HIP_TRY(hipMemcpy(seg_line.data(), ctx->seg_cnt.as<uint32_t>() + s.seg_base, 4 * (s.n_segs + 1),
                          hipMemcpyDeviceToHost));
        uint32_t l0 = seg_line[0], l1 = seg_line[s.n_segs];
        std::vector<uint32_t> line_fill(l1 - l0 + 1);
        HIP_TRY(hipMemcpy(line_fill.data(), ctx->line_cnt.as<uint32_t>() + l0, 4 * (size_t)(l1 - l0 + 1),
                          hipMemcpyDeviceToHost));
        std::vector<uint32_t> fill_off(s.n_tiles + 1);
        HIP_TRY(hipMemcpy(fill_off.data(), ctx->tile_fill_off.as<uint32_t>() + s.tile_base, 4 * (s.n_tiles + 1),
                          hipMemcpyDeviceToHost));
        for (uint32_t p = 0; p < s.n_paths; p++) {
            uint32_t a = seg_line[s.path_seg_off[p]] - l0, e = seg_line[s.path_seg_off[p + 1]] - l0;
            lines[p] = (int32_t)(e - a);
            fills[p] = (int32_t)(line_fill[e] - line_fill[a]);
            int32_t na = 0;
            for (uint32_t t = s.path_tile_off[p]; t < s.path_tile_off[p + 1]; t++) na += fill_off[t + 1] > fill_off[t];
            alpha_tiles[p] = na;
        }
    });
}

namespace {
struct MlpDevEval {
    DevBuf params, x, h, out;
};
}  // namespace

int pfh_mlp_eval(pfh_ctx *ctx, const float *params, const float *x, int64_t n, float *out) {
    if (!ctx || !params || (n > 0 && (!x || !out)) || n < 0) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        const int64_t np = mlp_param_count();
        if (n == 0) return;
        MlpDevEval d;
        d.params.ensure(4 * np);
        d.x.ensure(4 * n * 12);
        d.h.ensure(4 * n * 128);
        d.out.ensure(4 * n * 3);
        HIP_TRY(hipMemcpy(d.params.ptr, params, 4 * np, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.x.ptr, x, 4 * n * 12, hipMemcpyHostToDevice));
        HIP_TRY(mlp_forward(d.params.as<float>(), d.x.as<float>(), n, d.h.as<float>(), d.out.as<float>(),
                            ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipMemcpy(out, d.out.ptr, 4 * n * 3, hipMemcpyDeviceToHost));
        for (DevBuf *bf : {&d.params, &d.x, &d.h, &d.out}) bf->release();
    });
}

namespace {
struct MlpDev {
    DevBuf params, m, v, grad, loss, x, t, h, out, scratch;
};
static void mlp_grad_dev(pfh_ctx *ctx, MlpDev &d, const float *x, const float *t, int64_t n) {
    HIP_TRY(mlp_forward(d.params.as<float>(), x, n, d.h.as<float>(), d.out.as<float>(), ctx->stream));
    HIP_TRY(mlp_backward(d.params.as<float>(), x, n, d.h.as<float>(), d.out.as<float>(), t, d.grad.as<float>(),
                         d.loss.as<float>(), d.scratch.as<float>(), ctx->stream));
}
}  // namespace

int pfh_mlp_grad(pfh_ctx *ctx, const float *params, const float *x, const float *target, int64_t n, float *grad,
                 float *loss) {
    if (!ctx || !params || !x || !target || !grad || !loss || n <= 0) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        const int64_t np = mlp_param_count();
        MlpDev d;
        d.params.ensure(4 * np);
        d.grad.ensure(4 * np);
        d.loss.ensure(4);
        d.x.ensure(4 * n * 12);
        d.t.ensure(4 * n * 3);
        d.h.ensure(4 * n * 128);
        d.out.ensure(4 * n * 3);
        d.scratch.ensure(4 * n * 131);
        HIP_TRY(hipMemcpy(d.params.ptr, params, 4 * np, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.x.ptr, x, 4 * n * 12, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.t.ptr, target, 4 * n * 3, hipMemcpyHostToDevice));
        mlp_grad_dev(ctx, d, d.x.as<float>(), d.t.as<float>(), n);
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipMemcpy(grad, d.grad.ptr, 4 * np, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(loss, d.loss.ptr, 4, hipMemcpyDeviceToHost));
        for (DevBuf *bf : {&d.params, &d.grad, &d.loss, &d.x, &d.t, &d.h, &d.out, &d.scratch}) bf->release();
    });
}

int pfh_mlp_train(pfh_ctx *ctx, const float *x, const float *target, int64_t n, int32_t epochs, int32_t batch,
                  float lr, uint64_t seed, float *params, float *loss_history) {
    if (!ctx || !x || !target || !params || n <= 0 || epochs < 0 || batch <= 0) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        const int64_t np = mlp_param_count();
        const int64_t bs = std::min<int64_t>(batch, n);
        MlpDev d;
        d.params.ensure(4 * np);
        d.m.ensure(4 * np);
        d.v.ensure(4 * np);
        d.grad.ensure(4 * np);
        d.loss.ensure(4);
        d.x.ensure(4 * n * 12);
        d.t.ensure(4 * n * 3);
        d.h.ensure(4 * n * 128);
        d.out.ensure(4 * n * 3);
        d.scratch.ensure(4 * n * 131);
        HIP_TRY(hipMemcpy(d.params.ptr, params, 4 * np, hipMemcpyHostToDevice));
        std::vector<float> zeros(np, 0.0f);
        HIP_TRY(hipMemcpy(d.m.ptr, zeros.data(), 4 * np, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.v.ptr, zeros.data(), 4 * np, hipMemcpyHostToDevice));
        // shuffled copies of the data per epoch (host permutation, one upload per epoch)
        std::vector<float> xs((size_t)n * 12), ts((size_t)n * 3);
        std::vector<int64_t> perm(n);
        for (int64_t i = 0; i < n; i++) perm[i] = i;
        uint64_t st = seed ? seed : 0x9E3779B97F4A7C15ull;
        auto rnd = [&]() {  // splitmix64
            st += 0x9E3779B97F4A7C15ull;
            uint64_t z = st;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        int step = 0;
        for (int32_t ep = 0; ep < epochs; ep++) {
            for (int64_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
            for (int64_t i = 0; i < n; i++) {
                std::memcpy(&xs[(size_t)i * 12], x + perm[i] * 12, 48);
                std::memcpy(&ts[(size_t)i * 3], target + perm[i] * 3, 12);
            }
            HIP_TRY(hipMemcpy(d.x.ptr, xs.data(), 4 * n * 12, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(d.t.ptr, ts.data(), 4 * n * 3, hipMemcpyHostToDevice));
            for (int64_t b0 = 0; b0 + bs <= n; b0 += bs) {
                mlp_grad_dev(ctx, d, d.x.as<float>() + b0 * 12, d.t.as<float>() + b0 * 3, bs);
                HIP_TRY(mlp_adam(d.params.as<float>(), d.m.as<float>(), d.v.as<float>(), d.grad.as<float>(), np, lr,
                                 ++step, ctx->stream));
            }
            if (loss_history) {
                HIP_TRY(mlp_forward(d.params.as<float>(), d.x.as<float>(), n, d.h.as<float>(), d.out.as<float>(),
                                    ctx->stream));
                HIP_TRY(mlp_backward(d.params.as<float>(), d.x.as<float>(), n, d.h.as<float>(), d.out.as<float>(),
                                     d.t.as<float>(), d.grad.as<float>(), d.loss.as<float>(), d.scratch.as<float>(),
                                     ctx->stream));
                HIP_TRY(hipStreamSynchronize(ctx->stream));
                HIP_TRY(hipMemcpy(&loss_history[ep], d.loss.ptr, 4, hipMemcpyDeviceToHost));
            }
        }
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipMemcpy(params, d.params.ptr, 4 * np, hipMemcpyDeviceToHost));
        for (DevBuf *bf : {&d.params, &d.m, &d.v, &d.grad, &d.loss, &d.x, &d.t, &d.h, &d.out, &d.scratch})
            bf->release();
    });
}

}  // extern "C"
