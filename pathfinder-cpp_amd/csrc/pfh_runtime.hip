// Host runtime behind include/pfh.h: scene packing and residency, batch
// planning, the plain (stage-by-stage) and batched (graph-replayed) pipelines,
// capacity control, and canonical dumps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pfh.h"
#include "pfh_internal.h"

namespace pfh {
hipError_t launch_zero_spans(const ZeroSpans &z, hipStream_t st);
size_t scan_temp_elems(uint64_t n);
hipError_t scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *tmp, hipStream_t s);
hipError_t scan_entries(uint2 *inout, uint32_t n, uint32_t *counters, void *tmp, hipStream_t s);
hipError_t scan_paths(const BatchDev &b, void *tmp, hipStream_t s);
hipError_t launch_flatten(const BatchDev &b, bool emit, uint32_t segs_per_wave, uint32_t take, hipStream_t s);
hipError_t launch_bin(const BatchDev &b, bool emit, hipStream_t s);
hipError_t scan_counted(uint32_t *inout, const uint32_t *n_ptr, uint32_t cap, uint32_t *counters, int counter,
                        uint32_t *tmp, hipStream_t s);
hipError_t scan_total(uint32_t *inout, uint32_t n, uint32_t *counters, int counter, uint32_t *tmp, hipStream_t s);
hipError_t launch_propagate(const BatchDev &b, uint32_t block, uint32_t rows, hipStream_t s);
hipError_t launch_entries(const BatchDev &b, bool write, hipStream_t s);
hipError_t launch_bin_items(const BatchDev &b, hipStream_t s);
hipError_t launch_fb_bounds(const uint32_t *sorted_keys, const uint32_t *n_ptr, uint32_t cap, uint32_t n_fb,
                            uint2 *range, hipStream_t s);
size_t radix_temp_elems(uint64_t cap, uint32_t n_scenes);
hipError_t radix_sort(const BatchDev &b, uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                      int bits, uint32_t *tmp, bool *result_alt, hipStream_t s);
hipError_t launch_fill_masks(const BatchDev &b, uint32_t grid_alpha, uint32_t max_blocks, hipStream_t s);
hipError_t launch_fill_clipped(const BatchDev &b, hipStream_t s);
hipError_t launch_composite(const BatchDev &b, const uint32_t *sorted_vals, uint32_t block, hipStream_t s);
hipError_t launch_clear_outside(const BatchDev &b, uint32_t max_rows, hipStream_t s);
hipError_t mlp_forward(const float *params, const float *x, int64_t n, float *hidden, float *out, hipStream_t s);
hipError_t mlp_backward(const float *params, const float *x, int64_t n, const float *hidden, const float *out,
                        const float *target, float *grad, float *loss, float *scratch, hipStream_t s);
hipError_t mlp_adam(float *params, float *m, float *v, const float *grad, int64_t n, float lr, int step,
                    hipStream_t s);
int64_t mlp_param_count();
hipError_t mlp_predict_sum(const float *params, const float *x, int64_t n, float *sums, hipStream_t s);
struct StrokeSegment {
    float2 p0, c0, c1, p1;
    uint32_t kind;
    int32_t join;
    float distance;
    uint32_t pad;
};
struct StrokeLeaf {
    float2 p0, c0, c1, p1;
    float2 join_point;
    uint32_t kind;
    uint32_t pad;
};
struct StrokeContour {
    uint32_t seg0, seg1, seg2, seg3;
    uint32_t closed;
    int32_t cap, join;
    float width, miter;
    float2 input_first;
    uint32_t pad;
};
hipError_t launch_offset_pieces(const StrokeSegment *segs, uint32_t n, uint32_t *leaf_cnt, const uint32_t *leaf_off,
                                StrokeLeaf *leaves, bool emit, hipStream_t s);
hipError_t launch_assemble(const StrokeContour *contours, uint32_t n, const uint32_t *leaf_off, const StrokeLeaf *leaves,
                           const StrokeSegment *segs, const uint32_t *region_off, float2 *region_pts,
                           uint8_t *region_flags, uint32_t *point_cnt, hipStream_t s);
hipError_t launch_compact(const uint32_t *region_off, const uint32_t *out_off, uint32_t n, const float2 *region_pts,
                          const uint8_t *region_flags, float2 *out_pts, uint8_t *out_flags, hipStream_t s);
struct DashPath {
    uint32_t c0, c1;
    uint32_t d0, d1;
    float offset;
    uint32_t pad;
};
hipError_t launch_dash(const float2 *pts, const uint8_t *flags, const int32_t *contour_start, const uint8_t *closed,
                       const DashPath *paths, uint32_t n, const float *dashes, uint32_t *cnt, const uint32_t *off,
                       float2 *out_pts, uint8_t *out_flags, uint32_t *out_cstart, bool emit, hipStream_t s);
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

// A fixed region of device memory handed out front to back (the allocation-free entry point's
// work buffers live in the library's static scratch array); reset at the start of every call.
struct Arena {
    uint8_t *base = nullptr;
    size_t cap = 0, top = 0;
};

// Grow-only device buffer (the reference grows buffers to 2^k, core/d3d11/renderer.cpp:719-724).
// With an arena the buffer is carved from it instead (no hipMalloc / hipFree): a regrowth takes
// fresh space above the arena's top, and running past the arena's end is an error.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    Arena *arena = nullptr;
    void ensure(size_t need) {
        if (need <= bytes && ptr) return;
        if (arena) {
            const size_t cap = (std::max<size_t>(need, 4) + 255) & ~(size_t)255;
            if (cap > arena->cap - arena->top)
                throw PfhError{PFH_E_LIMIT, "the scene needs more work memory than the static scratch holds (" +
                                                std::to_string(arena->cap >> 20) + " MiB)"};
            ptr = arena->base + arena->top;
            arena->top += cap;
            bytes = cap;
            return;
        }
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
        if (ptr && !arena) (void)hipFree(ptr);
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
    uint32_t n_clip = 0;                // trailing clip paths
    bool clip_own_rules = false;        // PFH_CLIP_RULE_OWN
    std::vector<uint32_t> path_clip;    // clip reference per path (CLIP_*)
    std::vector<uint8_t> path_level;    // clip level per path (0 draw path, 1..4 clip path)
    uint32_t clip_levels = 0;           // deepest clip level
    std::vector<uint32_t> path_breaks;  // chord chains per path (two extra stream points each)
    uint64_t n_breaks = 0;
    std::vector<float> features;
    DevBuf d_features;
    DevBuf d_pts, d_segs, d_seg_path, d_path_rect, d_path_tile_off, d_path_col_off, d_col_path, d_path_rgba,
        d_path_rule, d_path_seg_off, d_bin_paths, d_out_owned, d_path_clip, d_path_level;
    // paint records (none: path_paint empty)
    std::vector<uint32_t> path_paint;
    bool has_ops = false;  // some record composites through an op
    DevBuf d_path_paint, d_paints, d_ramps;
    uint32_t n_bin_wave = 0, n_bin_group = 0;  // paths binned by one wave / by one workgroup
    uint8_t *d_out = nullptr;
    int64_t out_pitch = 0;
    // placement in the last rasterized batch (for dumps)
    bool dumped = false;
    uint32_t seg_base = 0, tile_base = 0, col_base = 0, fb_base = 0, path_base = 0;
    // exact counts of the last run (capacity planning)
    bool counted = false;
    uint32_t last_fills = 0, last_alpha = 0, last_entries = 0, last_lines = 0;
    float pred_fills = -1, pred_alpha = -1, pred_lines = -1;
};

struct GraphKey {
    std::vector<int32_t> ids;
    uint32_t flags;
    uint32_t cap_f, cap_a, cap_e, cap_l;
    uint32_t sched[8];  // the launch schedule the graph was captured with
    bool operator<(const GraphKey &o) const {
        if (ids != o.ids) return ids < o.ids;
        for (int i = 0; i < 8; i++)
            if (sched[i] != o.sched[i]) return sched[i] < o.sched[i];
        if (flags != o.flags) return flags < o.flags;
        if (cap_f != o.cap_f) return cap_f < o.cap_f;
        if (cap_a != o.cap_a) return cap_a < o.cap_a;
        if (cap_l != o.cap_l) return cap_l < o.cap_l;
        return cap_e < o.cap_e;
    }
};

struct Server;

// A captured batched launch.  Its event-record nodes bracket the fill and composite kernels with
// the workspace's events (Workspace::kev), which outlive every executable graph.
struct CapturedGraph {
    hipGraphExec_t exec = nullptr;
};

// Device working set of one launch: work buffers, stream, events and captured graphs.
// The context's own workspace serves pfh_rasterize; the request loop owns one per slot so
// that launches in different slots overlap on the device.
struct Workspace {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // around the fill and composite kernels of a batched launch; recorded as event nodes
    // inside captured graphs, so replays time those kernels too
    hipEvent_t kev[3] = {nullptr, nullptr, nullptr};
    DevBuf scene_table, bases, tile_delta, tile_cnt, tile_fill_off, tile_alpha_idx, tile_bd, tile_st, path_ent, ent_quarter, path_tot, bin_items, col_item, col_bd, tall_cols, curve_mask,
        deep_list, seg_cnt, lines, line_cnt, fills, dump_fills, alpha_rec, alpha_vis, masks, zbuf, fb_range, keys, vals, keys_alt,
        vals_alt, entry_rec, entry_meta, counters, scan_tmp, sort_tmp, tile_clip_ai, clip_jobs, clip_rec, clip_seen;
    uint32_t *h_counters = nullptr;  // pinned
    SceneDev *h_table = nullptr;     // pinned scene table
    size_t h_table_cap = 0;
    uint32_t *h_bases = nullptr;
    size_t h_bases_cap = 0;
    bool last_sorted_alt = false;
    std::map<GraphKey, CapturedGraph> graphs;
    uint64_t graphs_generation = 0;  // g_alloc_generation the cached graphs were captured under
    hipEvent_t stage_ev[9] = {};     // plain mode: between its stages
    Arena arena;                     // work buffers carved from a fixed region (arena.base set)

    // Stream, events and pinned host buffers for `table_scenes` scenes per launch, made once so
    // that a launch of at most that many scenes creates nothing.
    void create(size_t table_scenes = 0) {
        HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&ev0));
        HIP_TRY(hipEventCreate(&ev1));
        for (auto &e : kev) HIP_TRY(hipEventCreate(&e));
        // An event whose first record is an external record node inside a stream capture is
        // refused (invalid argument): record each once, eagerly, before any capture.
        for (auto &e : kev) HIP_TRY(hipEventRecord(e, stream));
        for (auto &e : stage_ev) HIP_TRY(hipEventCreate(&e));
        HIP_TRY(hipHostMalloc((void **)&h_counters, 4 * C_COUNT));
        if (table_scenes) {
            HIP_TRY(hipHostMalloc((void **)&h_table, sizeof(SceneDev) * table_scenes));
            h_table_cap = table_scenes;
            HIP_TRY(hipHostMalloc((void **)&h_bases, sizeof(uint32_t) * 6 * (table_scenes + 1)));
            h_bases_cap = 6 * (table_scenes + 1);
        }
    }
    // Every work buffer is carved from [base, base + cap) from now on (no device allocation).
    void use_arena(void *base, size_t cap) {
        arena.base = reinterpret_cast<uint8_t *>(base);
        arena.cap = cap;
        arena.top = 0;
        for (DevBuf *b : all_buffers()) {
            b->release();
            b->arena = &arena;
        }
    }
    // Forgets every carved buffer: the next launch lays them out afresh from the arena's start.
    void reset_arena() {
        for (DevBuf *b : all_buffers()) b->release();
        arena.top = 0;
    }
    static void destroy_graph(CapturedGraph &g) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = CapturedGraph{};
    }
    // Executable graphs launch on this workspace's stream only: once it is idle none of them
    // runs, and they can be destroyed (an executable graph is never destroyed while it may run).
    void drop_graphs() {
        if (graphs.empty()) return;
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &g : graphs) destroy_graph(g.second);
        graphs.clear();
    }
    // Drops the graphs whose scene set contains scene `id` (after the stream is idle).
    void drop_graphs_of(int32_t id) {
        bool any = false;
        for (auto &g : graphs) any |= std::find(g.first.ids.begin(), g.first.ids.end(), id) != g.first.ids.end();
        if (!any) return;
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto it = graphs.begin(); it != graphs.end();) {
            if (std::find(it->first.ids.begin(), it->first.ids.end(), id) != it->first.ids.end()) {
                destroy_graph(it->second);
                it = graphs.erase(it);
            } else {
                ++it;
            }
        }
    }
    std::vector<DevBuf *> all_buffers() {
        return {&scene_table, &bases, &tile_delta, &tile_cnt, &tile_fill_off, &tile_alpha_idx, &tile_bd, &tile_st,
                &path_ent, &ent_quarter, &path_tot, &bin_items, &col_item, &col_bd, &tall_cols, &curve_mask, &deep_list, &seg_cnt, &lines, &line_cnt, &fills, &dump_fills, &alpha_rec, &alpha_vis,
                &masks, &zbuf, &fb_range, &keys, &vals, &keys_alt, &vals_alt, &entry_rec, &entry_meta, &counters, &scan_tmp,
                &sort_tmp, &tile_clip_ai, &clip_jobs, &clip_rec, &clip_seen};
    }
    size_t device_bytes() {
        size_t n = 0;
        for (DevBuf *b : all_buffers()) n += b->bytes;
        return n;
    }
    void release_buffers() {
        drop_graphs();
        for (DevBuf *b : all_buffers()) b->release();
    }
    void destroy() {
        if (stream) (void)hipStreamSynchronize(stream);
        release_buffers();
        if (h_counters) (void)hipHostFree(h_counters);
        if (h_table) (void)hipHostFree(h_table);
        if (h_bases) (void)hipHostFree(h_bases);
        h_counters = nullptr;
        h_table = nullptr;
        h_bases = nullptr;
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (auto &e : kev)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : stage_ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

}  // namespace

// Work memory of the allocation-free entry point (pfh_rasterize_noalloc): one array per GPU,
// held until the process ends and shared by every context on that GPU.  Host code never uses
// this name as an address (it is a host-side stand-in): the device address comes from
// hipGetSymbolAddress, once per device, in scratch_on_device.
#define PFH_SCRATCH_BYTES ((size_t)PFH_NOALLOC_SCRATCH_BYTES)
static __device__ __align__(256) uint8_t scratch[PFH_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on `device` (the current device): looked up on the first call for
// that device, then kept.
void *scratch_on_device(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw PfhError{PFH_E_LIMIT, "device index above 63"};
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw PfhError{PFH_E_HIP, "hipGetSymbolAddress(scratch) returned no address"};
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}
}  // namespace

struct pfh_ctx {
    int device = 0;
    std::string err;
    std::vector<std::unique_ptr<SceneHost>> scenes;
    Workspace main;
    Workspace fixed;       // pfh_rasterize_noalloc: buffers carved from `scratch`, events made at create
    Workspace *w = &main;  // workspace the pipeline functions enqueue into
    // predictor
    std::vector<float> predictor;
    float predictor_safety = 1.25f;
    DevBuf mlp_params, mlp_x, mlp_h, mlp_out;
    pfh_stats stats{};
    Schedule sched;                  // launch schedule (pfh_set_schedule / pfh_autotune)
    std::unique_ptr<Server> server;  // request loop, when started
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

// ---------------------------------------------------------------------------
// Blend modes as the reference's hybrid renderer treats them (core/paint/effects.h): the modes
// whose opaque paths occlude (blend_mode_occludes_backdrop, :127-160), the destructive ones, which
// tile the whole view box (is_blend_mode_destructive, :234-243, core/data/built_path.cpp:21), and
// the tile shader's composite op of each (blend_mode_to_composite_ctrl, :184-232; 0: over).
static bool blend_occludes(uint32_t m) { return m == PFH_BLEND_SRC_OVER || m == PFH_BLEND_CLEAR; }
static bool blend_destructive(uint32_t m) {
    return m == PFH_BLEND_CLEAR || m == PFH_BLEND_COPY || m == PFH_BLEND_SRC_IN || m == PFH_BLEND_DEST_IN ||
           m == PFH_BLEND_SRC_OUT || m == PFH_BLEND_DEST_ATOP;
}
static uint32_t blend_composite_op(uint32_t m) {
    switch (m) {
        case PFH_BLEND_MULTIPLY: return 0x1;
        case PFH_BLEND_SCREEN: return 0x2;
        case PFH_BLEND_OVERLAY: return 0x3;
        case PFH_BLEND_DARKEN: return 0x4;
        case PFH_BLEND_LIGHTEN: return 0x5;
        case PFH_BLEND_COLOR_DODGE: return 0x6;
        case PFH_BLEND_COLOR_BURN: return 0x7;
        case PFH_BLEND_HARD_LIGHT: return 0x8;
        case PFH_BLEND_SOFT_LIGHT: return 0x9;
        case PFH_BLEND_DIFFERENCE: return 0xa;
        case PFH_BLEND_EXCLUSION: return 0xb;
        case PFH_BLEND_HUE: return 0xc;
        case PFH_BLEND_SATURATION: return 0xd;
        case PFH_BLEND_COLOR: return 0xe;
        case PFH_BLEND_LUMINOSITY: return 0xf;
        default: return 0;  // the Porter-Duff modes: the tile pass blends them over
    }
}

// Gradient paints: ramps and colour-texture records, built at upload
// ---------------------------------------------------------------------------

// third_party/umHalf.inl:49-111, the float-to-half conversion the metadata texture goes through:
// the mantissa truncated (not rounded), magnitudes below 2^-24 to zero, 2^-24 .. 2^-14 to
// denormals, above 2^16 to infinity; returned as the float value the half holds.
static float half_trunc(float v) {
    uint32_t f;
    std::memcpy(&f, &v, 4);
    const uint32_t sign = f >> 31, fe = (f >> 23) & 0xffu, ff = f & 0x7fffffu;
    uint32_t he = 0, hf = 0;
    if (fe == 0xffu) {
        he = 31;
        hf = ff != 0 ? 1u : 0u;
    } else if (fe != 0) {
        const int e = (int)fe - 127;
        if (e < -24) {
        } else if (e < -14) {
            const int k = -14 - e;  // 1 .. 10
            hf = (1024u >> k) + (ff >> (13 + k));
        } else if (e > 15) {
            he = 31;
        } else {
            he = (uint32_t)(e + 15);
            hf = ff >> 13;
        }
    }
    float mag;
    if (he == 31) mag = hf ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
    else if (he == 0) mag = std::ldexp((float)hf, -24);
    else mag = std::ldexp(1.0f + (float)hf / 1024.0f, (int)he - 15);
    return sign ? -mag : mag;
}

struct ColorStopH {
    float offset;
    uint32_t rgba;
};

// Gradient::sample (core/paint/gradient.cpp:23-56) over stops in Gradient::add order: the two stops
// around t, their colours as ColorF (channel * (1 / 255)), ColorF::lerp (common/color.cpp:61-72),
// and ColorU(ColorF)'s truncation to bytes (common/color.cpp:15-21).
static uint32_t gradient_sample(const std::vector<ColorStopH> &stops, float t) {
    if (stops.empty()) return 0u;
    t = std::min(std::max(t, 0.0f), 1.0f);
    const size_t last = stops.size() - 1;
    size_t it = 0;
    while (it < stops.size() && (stops[it].offset < t || stops[it].offset == 0)) it++;
    const size_t upper = std::min(it, last), lower = upper > 0 ? upper - 1 : upper;
    const ColorStopH &lo = stops[lower], &hi = stops[upper];
    const float denom = hi.offset - lo.offset;
    if (denom == 0) return lo.rgba;
    float ratio = std::min((t - lo.offset) / denom, 1.0f);
    ratio = std::min(std::max(ratio, 0.0f), 1.0f);
    const float k = 1.f / 255.f;
    uint32_t out = 0;
    for (int c = 0; c < 4; c++) {
        const float a = (float)((lo.rgba >> (8 * c)) & 255u) * k, b = (float)((hi.rgba >> (8 * c)) & 255u) * k;
        const float v = a + (b - a) * ratio;
        out |= (uint32_t)(uint8_t)(v * 255.f) << (8 * c);
    }
    return out;
}

// One gradient: its ramp (GradientTileBuilder::allocate, core/paint/gradient.cpp:86-123: texel x
// holds sample((x + 0.5) / 256)) and its record (Palette::calculate_texture_transforms,
// core/paint/palette.cpp:398-440; PaintMetadata::filter, core/paint/paint.cpp:45-65), every float
// of the record as the RGBA16F metadata texture holds it (core/renderer.cpp:167-250).
static void compile_gradient(const pfh_gradient &g, const float *stop_offsets, const uint32_t *stop_rgba,
                             const uint32_t *given_ramp, int32_t row, uint32_t *ramp, PaintRec &rec) {
    if ((g.kind != PFH_GRADIENT_LINEAR && g.kind != PFH_GRADIENT_RADIAL) ||
        (g.wrap != PFH_WRAP_CLAMP && g.wrap != PFH_WRAP_REPEAT))
        throw PfhError{PFH_E_ARG, "gradient kind or wrap out of range"};
    for (float v : g.line)
        if (!std::isfinite(v)) throw PfhError{PFH_E_ARG, "gradient line not finite"};
    std::vector<ColorStopH> stops;
    if (g.stop_count < 0 && !given_ramp) throw PfhError{PFH_E_ARG, "gradient without stops needs its ramp"};
    for (int32_t i = 0; i < g.stop_count; i++) {
        const ColorStopH st{stop_offsets[g.stop_start + i], stop_rgba[g.stop_start + i]};
        if (!std::isfinite(st.offset)) throw PfhError{PFH_E_ARG, "gradient stop offset not finite"};
        size_t at = 0;  // Gradient::add: before the first stop whose offset is not below the new one's
        while (at < stops.size() && stops[at].offset < st.offset) at++;
        stops.insert(stops.begin() + (ptrdiff_t)at, st);
    }
    if (g.stop_count < 0) std::memcpy(ramp, given_ramp, 4 * RAMP_TEXELS);
    else
        for (int x = 0; x < RAMP_TEXELS; x++) ramp[x] = gradient_sample(stops, ((float)x + 0.5f) / (float)RAMP_TEXELS);
    const float scale = 1.f / (float)RAMP_TEXELS;
    const float v0 = ((float)row + (float)(row + 1)) * 0.5f * scale;  // the row's centre
    float m11, m21, m12, m22, m13, m23;
    rec.from_x = rec.from_y = rec.vec_x = rec.vec_y = rec.r0 = rec.r1 = rec.uv_x = rec.uv_y = 0.0f;
    if (g.kind == PFH_GRADIENT_LINEAR) {
        const float fx = g.line[0], fy = g.line[1], dx = g.line[2] - fx, dy = g.line[3] - fy;
        const float sq = dx * dx + dy * dy;
        const float m0x = dx / sq, m0y = dy / sq;
        const float ax = m0x * -fx, ay = m0y * -fy;
        m11 = m0x, m21 = 0.0f, m12 = m0y, m22 = 0.0f, m13 = ax + ay, m23 = v0;
        rec.kind = 1;
    } else {
        for (float v : g.radii)
            if (!std::isfinite(v)) throw PfhError{PFH_E_ARG, "gradient radii not finite"};
        for (float v : g.transform)
            if (!std::isfinite(v)) throw PfhError{PFH_E_ARG, "gradient transform not finite"};
        // Transform2::inverse (common/math/transform2.cpp:32-36, Mat2::inverse: adjugate * (1 / det))
        const float a11 = g.transform[0], a21 = g.transform[1], a12 = g.transform[2], a22 = g.transform[3];
        const float det = a11 * a22 - a21 * a12;
        if (det == 0.0f || !std::isfinite(1.0f / det)) throw PfhError{PFH_E_ARG, "radial gradient transform not invertible"};
        const float id = 1.0f / det;
        m11 = (a22 * 1.0f) * id, m21 = (a21 * -1.0f) * id, m12 = (a12 * -1.0f) * id, m22 = (a11 * 1.0f) * id;
        const float vx = g.transform[4], vy = g.transform[5];
        m13 = -(m11 * vx + m12 * vy), m23 = -(m21 * vx + m22 * vy);
        rec.from_x = half_trunc(g.line[0]);
        rec.from_y = half_trunc(g.line[1]);
        rec.vec_x = half_trunc(g.line[2] - g.line[0]);
        rec.vec_y = half_trunc(g.line[3] - g.line[1]);
        rec.r0 = half_trunc(g.radii[0]);
        rec.r1 = half_trunc(g.radii[1]);
        rec.uv_x = half_trunc(0.0f * scale);
        rec.uv_y = half_trunc((float)row * scale + scale * 0.5f);  // the row's top, moved half a texel in
        rec.kind = 2;
    }
    rec.m11 = half_trunc(m11);
    rec.m21 = half_trunc(m21);
    rec.m12 = half_trunc(m12);
    rec.m22 = half_trunc(m22);
    rec.m13 = half_trunc(m13);
    rec.m23 = half_trunc(m23);
    rec.wrap = (uint32_t)g.wrap;
}

// Paint::is_opaque for a gradient paint (core/paint/paint.cpp:9-19, Gradient::is_opaque,
// core/paint/gradient.cpp:58-70): white base colour, every stop opaque, at least one stop.
static bool gradient_opaque(const pfh_gradient &g, const uint32_t *stop_rgba, const uint32_t *given_ramp) {
    if (g.stop_count < 0) {  // a given ramp: opaque when every texel is
        for (int x = 0; x < RAMP_TEXELS; x++)
            if ((given_ramp[x] >> 24) != 255u) return false;
        return true;
    }
    if (g.stop_count == 0) return false;
    for (int32_t i = 0; i < g.stop_count; i++)
        if ((stop_rgba[g.stop_start + i] >> 24) != 255u) return false;
    return true;
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
    for (int p = 0; p < d->n_paths; p++)
        if (d->path_fill_rule[p] > 1) throw PfhError{PFH_E_ARG, "path_fill_rule must be 0 (nonzero) or 1 (even-odd)"};
    if (d->n_clip_paths < 0 || d->n_clip_paths > d->n_paths)
        throw PfhError{PFH_E_ARG, "n_clip_paths must lie in [0, n_paths]"};
    if (d->clip_rule != PFH_CLIP_RULE_ABS && d->clip_rule != PFH_CLIP_RULE_OWN)
        throw PfhError{PFH_E_ARG, "clip_rule must be PFH_CLIP_RULE_ABS (0) or PFH_CLIP_RULE_OWN (1)"};
    if (d->path_blend)
        for (int p = 0; p < d->n_paths; p++)
            if (d->path_blend[p] > PFH_BLEND_LUMINOSITY) throw PfhError{PFH_E_ARG, "path_blend names no blend mode"};
    if (d->path_gradient) {
        if (d->n_gradients < 0 || d->n_gradient_stops < 0 || (d->n_gradients && !d->gradients) ||
            (d->n_gradient_stops && (!d->gradient_stop_offsets || !d->gradient_stop_rgba)))
            throw PfhError{PFH_E_ARG, "null gradient array"};
        for (int p = 0; p < d->n_paths; p++)
            if (d->path_gradient[p] < -1 || d->path_gradient[p] >= d->n_gradients)
                throw PfhError{PFH_E_ARG, "path_gradient names no gradient"};
        for (int i = 0; i < d->n_gradients; i++) {
            const pfh_gradient &g = d->gradients[i];
            if (g.stop_count < 0 && !d->gradient_ramps)
                throw PfhError{PFH_E_ARG, "gradient without stops needs gradient_ramps"};
            if (g.stop_count >= 0 &&
                (g.stop_start < 0 || (int64_t)g.stop_start + g.stop_count > d->n_gradient_stops))
                throw PfhError{PFH_E_ARG, "gradient stops outside the stop arrays"};
        }
    }
    if (d->path_clip) {
        const int n_draw = d->n_paths - d->n_clip_paths;
        for (int p = 0; p < d->n_paths; p++) {
            const int32_t c = d->path_clip[p];
            if (c < -1 || c >= d->n_clip_paths) throw PfhError{PFH_E_ARG, "path_clip names no clip path"};
        }
        // A clip path may be clipped by another (nesting; undefined in the reference's hybrid tiler,
        // core/d3d9/scene_builder.cpp:209-216, given the level structure of the GPU-driven renderer,
        // core/d3d11/scene_builder.cpp:92-133): chains of at most MAX_CLIP_LEVELS clip paths.
        for (int q = 0; q < d->n_clip_paths; q++) {
            uint32_t k = 1;
            for (int32_t c = d->path_clip[n_draw + q]; c >= 0; c = d->path_clip[n_draw + c])
                if (++k > MAX_CLIP_LEVELS)
                    throw PfhError{PFH_E_ARG, "clip paths nest deeper than 4 (or in a cycle)"};
        }
    }
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

// ---------------------------------------------------------------------------
// Stroke-to-fill on the device (kernels in pfh_stroke.hip)
// ---------------------------------------------------------------------------

// Outlines as flat arrays (points, flags, contours, paths).
struct HostOutlines {
    std::vector<float2> pts;
    std::vector<uint8_t> flags;
    std::vector<int32_t> contour_start{0};
    std::vector<uint8_t> closed;
    std::vector<int32_t> path_contour_start{0};
};

static void validate_style(const pfh_stroke_style &st) {
    if (st.line_cap < PFH_CAP_BUTT || st.line_cap > PFH_CAP_ROUND || st.line_join < PFH_JOIN_MITER ||
        st.line_join > PFH_JOIN_ROUND || !std::isfinite(st.line_width) || !std::isfinite(st.miter_limit))
        throw PfhError{PFH_E_ARG, "invalid stroke style"};
}

// Strokes every path of `in` with its style (OutlineStrokeToFill::offset + into_outline,
// core/stroke.cpp:124-195).  The host only lays out each contour's strands of segments in the
// reference's offsetting order (forward: SegmentsIter order with the closing line unless the
// contour ends where it starts, core/data/contour.cpp:110-173; backward: that closing line
// reversed, then the segments from the tail reversed, core/stroke.cpp:54-122); offsetting,
// joins and caps run on the device.
static HostOutlines stroke_on_device(pfh_ctx *ctx, const float2 *P, const uint8_t *F, const int32_t *contour_start,
                                     const uint8_t *contour_closed, const int32_t *path_contour_start, int32_t n_paths,
                                     const pfh_stroke_style *styles) {
    std::vector<StrokeSegment> segs;
    std::vector<StrokeContour> outs;
    std::vector<int32_t> path_out(n_paths + 1, 0);
    auto approx_eq = [](float2 a, float2 b) {
        const float dx = a.x - b.x, dy = a.y - b.y;
        return std::sqrt(dx * dx + dy * dy) <= 0.0001f;
    };
    for (int32_t p = 0; p < n_paths; p++) {
        const pfh_stroke_style &st = styles[p];
        validate_style(st);
        const float radius = st.line_width * 0.5f;
        for (int32_t c = path_contour_start[p]; c < path_contour_start[p + 1]; c++) {
            const int32_t a = contour_start[c], n = contour_start[c + 1] - a;
            const float2 *pt = P + a;
            const uint8_t *fl = F + a;
            const bool closed = contour_closed[c] != 0;
            auto add = [&](float2 p0, float2 c0, float2 c1, float2 p1, uint32_t kind, uint32_t index) {
                segs.push_back(StrokeSegment{p0, c0, c1, p1, kind, index == 0 ? (int32_t)PFH_JOIN_BEVEL : st.line_join,
                                             -radius, 0});
            };
            const float2 z = make_float2(0.0f, 0.0f);
            // forward strand (core/stroke.cpp:28-52)
            const uint32_t f0 = (uint32_t)segs.size();
            uint32_t index = 0;
            for (int32_t head = 0;;) {
                if (head >= n || fl[head] != 0) break;
                if (head + 1 < n) {
                    if (fl[head + 1] == 0) {
                        add(pt[head], z, z, pt[head + 1], 1, index++);
                        head += 1;
                    } else if (head + 2 < n && fl[head + 1] == 1 && fl[head + 2] == 0) {
                        add(pt[head], pt[head + 1], z, pt[head + 2], 2, index++);
                        head += 2;
                    } else if (head + 3 < n && fl[head + 1] == 1 && fl[head + 2] == 2 && fl[head + 3] == 0) {
                        add(pt[head], pt[head + 1], pt[head + 2], pt[head + 3], 3, index++);
                        head += 3;
                    } else {
                        break;
                    }
                } else {
                    if (closed && n > 1 && !approx_eq(pt[0], pt[n - 1])) add(pt[n - 1], z, z, pt[0], 1, index++);
                    break;
                }
            }
            const uint32_t f1 = (uint32_t)segs.size();
            // backward strand (core/stroke.cpp:54-122)
            index = 0;
            if (closed && n > 1 && !approx_eq(pt[0], pt[n - 1])) add(pt[0], z, z, pt[n - 1], 1, index++);
            for (int32_t tail = n - 1; tail >= 0;) {
                if (fl[tail] != 0) break;
                if (tail >= 3 && fl[tail - 1] == 2 && fl[tail - 2] == 1 && fl[tail - 3] == 0) {
                    add(pt[tail], pt[tail - 1], pt[tail - 2], pt[tail - 3], 3, index++);
                    tail -= 3;
                } else if (tail >= 2 && fl[tail - 1] == 1 && fl[tail - 2] == 0) {
                    add(pt[tail], pt[tail - 1], z, pt[tail - 2], 2, index++);
                    tail -= 2;
                } else if (tail >= 1 && fl[tail - 1] == 0) {
                    add(pt[tail], z, z, pt[tail - 1], 1, index++);
                    tail -= 1;
                } else {
                    break;
                }
            }
            const uint32_t b1 = (uint32_t)segs.size();
            const float2 first = n > 0 ? pt[0] : z;
            auto contour = [&](uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3) {
                outs.push_back(StrokeContour{s0, s1, s2, s3, closed ? 1u : 0u, st.line_cap, st.line_join, st.line_width,
                                             st.miter_limit, first, 0});
            };
            if (closed) {
                contour(f0, f1, f1, f1);
                contour(f1, b1, b1, b1);
            } else {
                contour(f0, f1, f1, b1);
            }
        }
        path_out[p + 1] = (int32_t)outs.size();
    }
    HostOutlines res;
    res.path_contour_start = path_out;
    const uint32_t ns = (uint32_t)segs.size(), nc = (uint32_t)outs.size();
    if (nc == 0) return res;
    hipStream_t st = ctx->w->stream;
    DevBuf d_segs, d_cnt, d_off, d_leaves, d_outs, d_region_off, d_region_pts, d_region_flags, d_pcnt, d_out_off,
        d_pts, d_flags;
    struct Release {
        std::vector<DevBuf *> bufs;
        ~Release() {
            for (DevBuf *b : bufs) b->release();
        }
    } release{{&d_segs, &d_cnt, &d_off, &d_leaves, &d_outs, &d_region_off, &d_region_pts, &d_region_flags, &d_pcnt,
               &d_out_off, &d_pts, &d_flags}};
    upload(d_segs, segs);
    d_cnt.ensure(4ull * ns + 4);
    HIP_TRY(launch_offset_pieces(d_segs.as<StrokeSegment>(), ns, d_cnt.as<uint32_t>(), nullptr, nullptr, false, st));
    std::vector<uint32_t> cnt(ns);
    if (ns) HIP_TRY(hipMemcpyAsync(cnt.data(), d_cnt.ptr, 4ull * ns, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<uint32_t> leaf_off(ns + 1, 0);
    for (uint32_t i = 0; i < ns; i++) {
        if ((uint64_t)leaf_off[i] + cnt[i] >= 0xF0000000ull) throw PfhError{PFH_E_LIMIT, "too many stroke pieces"};
        leaf_off[i + 1] = leaf_off[i] + cnt[i];
    }
    upload(d_off, leaf_off);
    d_leaves.ensure(sizeof(StrokeLeaf) * ((size_t)leaf_off[ns] + 1));
    HIP_TRY(launch_offset_pieces(d_segs.as<StrokeSegment>(), ns, nullptr, d_off.as<uint32_t>(),
                                 d_leaves.as<StrokeLeaf>(), true, st));
    // regions: an upper bound of every output contour's points
    std::vector<uint32_t> region_off(nc + 1, 0);
    for (uint32_t c = 0; c < nc; c++) {
        const StrokeContour &k = outs[c];
        const uint64_t leaves = (uint64_t)leaf_off[k.seg1] - leaf_off[k.seg0] + leaf_off[k.seg3] - leaf_off[k.seg2];
        const uint64_t bound = 20 * leaves + 3 * 16;
        if (region_off[c] + bound >= 0xF0000000ull) throw PfhError{PFH_E_LIMIT, "stroked outline too large"};
        region_off[c + 1] = region_off[c] + (uint32_t)bound;
    }
    upload(d_outs, outs);
    upload(d_region_off, region_off);
    d_region_pts.ensure(8ull * region_off[nc] + 8);
    d_region_flags.ensure((size_t)region_off[nc] + 4);
    d_pcnt.ensure(4ull * nc);
    HIP_TRY(launch_assemble(d_outs.as<StrokeContour>(), nc, d_off.as<uint32_t>(), d_leaves.as<StrokeLeaf>(),
                            d_segs.as<StrokeSegment>(), d_region_off.as<uint32_t>(), d_region_pts.as<float2>(),
                            d_region_flags.as<uint8_t>(), d_pcnt.as<uint32_t>(), st));
    std::vector<uint32_t> pcnt(nc);
    HIP_TRY(hipMemcpyAsync(pcnt.data(), d_pcnt.ptr, 4ull * nc, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<uint32_t> out_off(nc + 1, 0);
    for (uint32_t c = 0; c < nc; c++) {
        if (pcnt[c] > region_off[c + 1] - region_off[c]) throw PfhError{PFH_E_STATE, "stroke region overflow"};
        out_off[c + 1] = out_off[c] + pcnt[c];
    }
    upload(d_out_off, out_off);
    d_pts.ensure(8ull * out_off[nc] + 8);
    d_flags.ensure((size_t)out_off[nc] + 4);
    HIP_TRY(launch_compact(d_region_off.as<uint32_t>(), d_out_off.as<uint32_t>(), nc, d_region_pts.as<float2>(),
                           d_region_flags.as<uint8_t>(), d_pts.as<float2>(), d_flags.as<uint8_t>(), st));
    res.pts.resize(out_off[nc]);
    res.flags.resize(out_off[nc]);
    if (out_off[nc]) {
        HIP_TRY(hipMemcpyAsync(res.pts.data(), d_pts.ptr, 8ull * out_off[nc], hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(res.flags.data(), d_flags.ptr, out_off[nc], hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    res.contour_start.assign(out_off.begin(), out_off.end());
    res.closed.assign(nc, 1);
    return res;
}

// Whether a path's style carries a dash pattern, after checking it: the pattern lies inside the
// dash array, its lengths are finite and non-negative with a positive finite total, and the
// offset is finite and non-negative.  OutlineDash (core/dash.cpp:9-29) takes the pattern
// unchecked; with a zero total its fmod is NaN and its walk never advances.
static bool dashed(const pfh_stroke_style &st, const float *dashes, int32_t n_dashes) {
    if (st.dash_count == 0) return false;
    if (st.dash_count < 0 || st.dash_start < 0 || (int64_t)st.dash_start + st.dash_count > n_dashes || !dashes)
        throw PfhError{PFH_E_ARG, "dash pattern outside the dash array"};
    float total = 0.0f;
    for (int32_t i = st.dash_start; i < st.dash_start + st.dash_count; i++) {
        if (!std::isfinite(dashes[i]) || dashes[i] < 0.0f) throw PfhError{PFH_E_ARG, "dash length not finite or negative"};
        total += dashes[i];
    }
    if (!(total > 0.0f) || !std::isfinite(total)) throw PfhError{PFH_E_ARG, "dash pattern of zero or infinite length"};
    if (!std::isfinite(st.dash_offset) || st.dash_offset < 0.0f) throw PfhError{PFH_E_ARG, "dash offset not finite or negative"};
    return true;
}

// Dashes the paths whose style carries a pattern (OutlineDash::dash + into_outline,
// core/dash.cpp:48-126), as Canvas::stroke_path does before it strokes (core/canvas.cpp:284-289):
// a thread per dashed path walks its contours with the dash state carried across them (k_dash),
// a count pass sizing each path's dashes and an emit pass writing them at the scanned offsets.
// The dashes are open contours; a path without a pattern passes through unchanged.
static HostOutlines dash_on_device(pfh_ctx *ctx, const float2 *P, const uint8_t *F, const int32_t *contour_start,
                                   const uint8_t *contour_closed, const int32_t *path_contour_start, int32_t n_paths,
                                   const pfh_stroke_style *styles, const float *dashes, int32_t n_dashes) {
    std::vector<DashPath> dp;
    std::vector<int32_t> which;
    for (int32_t p = 0; p < n_paths; p++)
        if (dashed(styles[p], dashes, n_dashes)) {
            const pfh_stroke_style &st = styles[p];
            dp.push_back(DashPath{(uint32_t)path_contour_start[p], (uint32_t)path_contour_start[p + 1],
                                  (uint32_t)st.dash_start, (uint32_t)(st.dash_start + st.dash_count), st.dash_offset, 0});
            which.push_back(p);
        }
    const uint32_t n = (uint32_t)dp.size();
    const int32_t n_contours = path_contour_start[n_paths];
    const int32_t n_points = contour_start[n_contours];
    std::vector<uint32_t> off(2 * (size_t)n + 2, 0);
    std::vector<float2> dpts;
    std::vector<uint8_t> dflags;
    std::vector<uint32_t> dcs;
    if (n) {
        hipStream_t st = ctx->w->stream;
        DevBuf d_pts, d_flags, d_cs, d_closed, d_paths, d_dashes, d_cnt, d_off, d_opts, d_oflags, d_ocs;
        struct Release {
            std::vector<DevBuf *> bufs;
            ~Release() {
                for (DevBuf *b : bufs) b->release();
            }
        } release{{&d_pts, &d_flags, &d_cs, &d_closed, &d_paths, &d_dashes, &d_cnt, &d_off, &d_opts, &d_oflags, &d_ocs}};
        upload(d_pts, std::vector<float2>(P, P + n_points));
        upload(d_flags, std::vector<uint8_t>(F, F + n_points));
        upload(d_cs, std::vector<int32_t>(contour_start, contour_start + n_contours + 1));
        upload(d_closed, std::vector<uint8_t>(contour_closed, contour_closed + n_contours));
        upload(d_paths, dp);
        upload(d_dashes, std::vector<float>(dashes, dashes + n_dashes));
        d_cnt.ensure(8ull * n);
        HIP_TRY(launch_dash(d_pts.as<float2>(), d_flags.as<uint8_t>(), d_cs.as<int32_t>(), d_closed.as<uint8_t>(),
                            d_paths.as<DashPath>(), n, d_dashes.as<float>(), d_cnt.as<uint32_t>(), nullptr, nullptr,
                            nullptr, nullptr, false, st));
        std::vector<uint32_t> cnt(2 * (size_t)n);
        HIP_TRY(hipMemcpyAsync(cnt.data(), d_cnt.ptr, 8ull * n, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (uint32_t k = 0; k < n; k++) {
            if (cnt[2 * k] == 0xFFFFFFFFu)
                throw PfhError{PFH_E_LIMIT, "dash pattern too fine for its path (more than 2^24 dash steps)"};
            const uint64_t np = (uint64_t)off[2 * k] + cnt[2 * k], nc = (uint64_t)off[2 * k + 1] + cnt[2 * k + 1];
            if (np >= 0x7FFFFFFFull || nc >= 0x7FFFFFFFull) throw PfhError{PFH_E_LIMIT, "dashed outlines too large"};
            off[2 * k + 2] = (uint32_t)np;
            off[2 * k + 3] = (uint32_t)nc;
        }
        upload(d_off, off);
        d_opts.ensure(8ull * off[2 * n] + 8);
        d_oflags.ensure((size_t)off[2 * n] + 4);
        d_ocs.ensure(4ull * off[2 * n + 1] + 4);
        HIP_TRY(launch_dash(d_pts.as<float2>(), d_flags.as<uint8_t>(), d_cs.as<int32_t>(), d_closed.as<uint8_t>(),
                            d_paths.as<DashPath>(), n, d_dashes.as<float>(), nullptr, d_off.as<uint32_t>(),
                            d_opts.as<float2>(), d_oflags.as<uint8_t>(), d_ocs.as<uint32_t>(), true, st));
        dpts.resize(off[2 * n]);
        dflags.resize(off[2 * n]);
        dcs.resize(off[2 * n + 1]);
        if (!dpts.empty()) {
            HIP_TRY(hipMemcpyAsync(dpts.data(), d_opts.ptr, 8 * dpts.size(), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(dflags.data(), d_oflags.ptr, dflags.size(), hipMemcpyDeviceToHost, st));
        }
        if (!dcs.empty()) HIP_TRY(hipMemcpyAsync(dcs.data(), d_ocs.ptr, 4 * dcs.size(), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    HostOutlines res;
    size_t k = 0;
    for (int32_t p = 0; p < n_paths; p++) {
        if (k < which.size() && which[k] == p) {
            const uint32_t c0 = off[2 * k + 1], c1 = off[2 * k + 3], pend = off[2 * k + 2];
            for (uint32_t c = c0; c < c1; c++) {
                const uint32_t a = dcs[c], b = c + 1 < c1 ? dcs[c + 1] : pend;
                res.pts.insert(res.pts.end(), dpts.begin() + a, dpts.begin() + b);
                res.flags.insert(res.flags.end(), dflags.begin() + a, dflags.begin() + b);
                res.contour_start.push_back((int32_t)res.pts.size());
                res.closed.push_back(0);
            }
            k++;
        } else {
            for (int32_t c = path_contour_start[p]; c < path_contour_start[p + 1]; c++) {
                res.pts.insert(res.pts.end(), P + contour_start[c], P + contour_start[c + 1]);
                res.flags.insert(res.flags.end(), F + contour_start[c], F + contour_start[c + 1]);
                res.contour_start.push_back((int32_t)res.pts.size());
                res.closed.push_back(contour_closed[c]);
            }
        }
        if (res.pts.size() >= 0x7FFFFFFFull) throw PfhError{PFH_E_LIMIT, "dashed outlines too large"};
        res.path_contour_start.push_back((int32_t)res.closed.size());
    }
    return res;
}

// Strokes the paths, dashing first those whose style carries a pattern.
static HostOutlines dash_and_stroke(pfh_ctx *ctx, const float2 *P, const uint8_t *F, const int32_t *contour_start,
                                    const uint8_t *contour_closed, const int32_t *path_contour_start, int32_t n_paths,
                                    const pfh_stroke_style *styles, const float *dashes, int32_t n_dashes) {
    bool any = false;
    for (int32_t p = 0; p < n_paths; p++) any |= dashed(styles[p], dashes, n_dashes);
    if (!any)
        return stroke_on_device(ctx, P, F, contour_start, contour_closed, path_contour_start, n_paths, styles);
    HostOutlines d = dash_on_device(ctx, P, F, contour_start, contour_closed, path_contour_start, n_paths, styles,
                                    dashes, n_dashes);
    return stroke_on_device(ctx, d.pts.data(), d.flags.data(), d.contour_start.data(), d.closed.data(),
                            d.path_contour_start.data(), n_paths, styles);
}

static int scene_upload_desc(pfh_ctx *ctx, const pfh_scene_desc *d);

// A scene whose paths carry stroke styles: those paths are stroked on the device first and
// their stroked outlines (filled with the nonzero rule, core/canvas.cpp:297-299) take their place.
static int scene_upload(pfh_ctx *ctx, const pfh_scene_desc *d) {
    if (!d || !d->path_stroke) return scene_upload_desc(ctx, d);
    if (d->n_paths < 0 || d->n_contours < 0 || d->n_points < 0 || !d->contour_start || !d->path_contour_start ||
        (d->n_points && (!d->points || !d->point_flags)) || (d->n_contours && !d->contour_closed) ||
        (d->n_paths && !d->path_fill_rule))
        throw PfhError{PFH_E_ARG, "null array"};
    if (d->contour_start[0] != 0 || d->contour_start[d->n_contours] != d->n_points ||
        d->path_contour_start[0] != 0 || d->path_contour_start[d->n_paths] != d->n_contours)
        throw PfhError{PFH_E_ARG, "contour / path ranges do not cover the arrays"};
    for (int c = 0; c < d->n_contours; c++)
        if (d->contour_start[c + 1] < d->contour_start[c]) throw PfhError{PFH_E_ARG, "contour_start not monotone"};
    for (int p = 0; p < d->n_paths; p++)
        if (d->path_contour_start[p + 1] < d->path_contour_start[p])
            throw PfhError{PFH_E_ARG, "path_contour_start not monotone"};
    const float2 *P = reinterpret_cast<const float2 *>(d->points);
    // the stroked paths as a batch of their own
    std::vector<int32_t> stroked;
    for (int32_t p = 0; p < d->n_paths; p++)
        if (d->path_stroke[p].line_width > 0.0f) stroked.push_back(p);
    std::vector<int32_t> spcs{0};
    std::vector<pfh_stroke_style> sst;
    for (int32_t p : stroked) {
        spcs.push_back(spcs.back() + d->path_contour_start[p + 1] - d->path_contour_start[p]);
        sst.push_back(d->path_stroke[p]);
    }
    // a stroked path's contours are contiguous: index them through a remapped contour table
    std::vector<int32_t> scs{0};
    std::vector<uint8_t> scl;
    std::vector<float2> spts;
    std::vector<uint8_t> sfl;
    for (int32_t p : stroked)
        for (int32_t c = d->path_contour_start[p]; c < d->path_contour_start[p + 1]; c++) {
            spts.insert(spts.end(), P + d->contour_start[c], P + d->contour_start[c + 1]);
            sfl.insert(sfl.end(), d->point_flags + d->contour_start[c], d->point_flags + d->contour_start[c + 1]);
            scs.push_back((int32_t)spts.size());
            scl.push_back(d->contour_closed[c]);
        }
    if (d->n_stroke_dashes < 0) throw PfhError{PFH_E_ARG, "negative n_stroke_dashes"};
    HostOutlines so = dash_and_stroke(ctx, spts.data(), sfl.data(), scs.data(), scl.data(), spcs.data(),
                                      (int32_t)stroked.size(), sst.data(), d->stroke_dashes, d->n_stroke_dashes);
    // the scene with the stroked outlines in place of the stroked paths' own
    std::vector<float2> pts;
    std::vector<uint8_t> fl, closed, rule(d->path_fill_rule, d->path_fill_rule + d->n_paths);
    std::vector<int32_t> cs{0}, pcs{0};
    size_t k = 0;
    for (int32_t p = 0; p < d->n_paths; p++) {
        if (k < stroked.size() && stroked[k] == p) {
            for (int32_t c = so.path_contour_start[k]; c < so.path_contour_start[k + 1]; c++) {
                pts.insert(pts.end(), so.pts.begin() + so.contour_start[c], so.pts.begin() + so.contour_start[c + 1]);
                fl.insert(fl.end(), so.flags.begin() + so.contour_start[c], so.flags.begin() + so.contour_start[c + 1]);
                cs.push_back((int32_t)pts.size());
                closed.push_back(so.closed[c]);
            }
            rule[p] = PFH_FILL_NONZERO;
            k++;
        } else {
            for (int32_t c = d->path_contour_start[p]; c < d->path_contour_start[p + 1]; c++) {
                pts.insert(pts.end(), P + d->contour_start[c], P + d->contour_start[c + 1]);
                fl.insert(fl.end(), d->point_flags + d->contour_start[c], d->point_flags + d->contour_start[c + 1]);
                cs.push_back((int32_t)pts.size());
                closed.push_back(d->contour_closed[c]);
            }
        }
        pcs.push_back((int32_t)closed.size());
    }
    if (pts.size() >= 0x7FFFFFFFull) throw PfhError{PFH_E_LIMIT, "stroked scene has more than 2^31 points"};
    pfh_scene_desc e = *d;
    e.n_contours = (int32_t)closed.size();
    e.n_points = (int32_t)pts.size();
    e.points = reinterpret_cast<const float *>(pts.data());
    e.point_flags = fl.data();
    e.contour_start = cs.data();
    e.contour_closed = closed.data();
    e.path_contour_start = pcs.data();
    e.path_fill_rule = rule.data();
    e.path_stroke = nullptr;
    return scene_upload_desc(ctx, &e);
}

static int scene_upload_desc(pfh_ctx *ctx, const pfh_scene_desc *d) {
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
    std::vector<PaintRec> paints;
    std::vector<uint32_t> ramps;
    const int n_draw_paths = d->n_paths - d->n_clip_paths;
    auto blend_of = [&](int p) -> uint32_t {
        return d->path_blend && p < n_draw_paths ? d->path_blend[p] : (uint32_t)PFH_BLEND_SRC_OVER;
    };
    bool records = d->path_gradient != nullptr;
    for (int p = 0; p < n_draw_paths && !records; p++)
        records = !blend_occludes(blend_of(p)) || blend_composite_op(blend_of(p)) != 0;
    if (records) {
        // Paint records: each gradient compiled once per composite op (its ramp row the gradient's
        // ordinal in order of first use), and one per solid path whose blend mode does not occlude
        // or composites through an op.  A path with a record keeps in its colour word only whether
        // it occludes (alpha 255 or 0), which the z-buffer reads.
        std::map<std::pair<int32_t, uint32_t>, int32_t> slot;  // (gradient, op) -> record
        std::vector<int32_t> row(d->path_gradient ? d->n_gradients : 0, -1);
        int32_t n_rows = 0;
        s->path_paint.assign(d->n_paths, PAINT_NONE);
        for (int p = 0; p < d->n_paths; p++) {
            const int32_t gi = d->path_gradient ? d->path_gradient[p] : -1;
            const uint32_t bm = blend_of(p), op = blend_composite_op(bm);
            const bool occludes = blend_occludes(bm);
            if (gi >= 0) {
                if (row[gi] < 0) row[gi] = n_rows++;
                auto it = slot.find({gi, op});
                if (it == slot.end()) {
                    it = slot.emplace(std::make_pair(gi, op), (int32_t)paints.size()).first;
                    paints.emplace_back();
                    ramps.resize(ramps.size() + RAMP_TEXELS);
                    compile_gradient(d->gradients[gi], d->gradient_stop_offsets, d->gradient_stop_rgba,
                                     d->gradient_ramps ? d->gradient_ramps + (size_t)gi * RAMP_TEXELS : nullptr,
                                     row[gi] % RAMP_TEXELS, ramps.data() + ramps.size() - RAMP_TEXELS, paints.back());
                    paints.back().op = op;
                    s->has_ops = s->has_ops || op != 0;
                }
                s->path_paint[p] = (uint32_t)it->second;
                const uint32_t *given = d->gradient_ramps ? d->gradient_ramps + (size_t)gi * RAMP_TEXELS : nullptr;
                s->path_rgba[p] =
                    occludes && gradient_opaque(d->gradients[gi], d->gradient_stop_rgba, given) ? 0xff000000u : 0u;
            } else if (!occludes || op != 0) {
                PaintRec rec{};
                rec.kind = 0;
                rec.colour = d->path_rgba[p];
                rec.op = op;
                s->has_ops = s->has_ops || op != 0;
                s->path_paint[p] = (uint32_t)paints.size();
                paints.push_back(rec);
                ramps.resize(ramps.size() + RAMP_TEXELS);  // unused: records and ramp rows share an index
                s->path_rgba[p] = 0u;
            }
        }
        if (paints.empty()) s->path_paint.clear();
    }
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
        // core/data/built_path.cpp:8-34 + data.h:53-55: tile rect of bounds ∩ view box, of the
        // view box for a draw path of a destructive blend mode (built_path.cpp:21)
        RectF bi = blend_destructive(blend_of(p)) ? vb : intersect(bounds, vb);
        int4 r;
        r.x = (int32_t)std::floor(bi.l * (1.0f / 16.0f));
        r.y = (int32_t)std::floor(bi.t * (1.0f / 16.0f));
        r.z = (int32_t)std::ceil(bi.r * (1.0f / 16.0f));
        r.w = (int32_t)std::ceil(bi.b * (1.0f / 16.0f));
        // DenseTileMap w x h tiles and BuiltPathData::backdrops of width w (built_path.cpp:25-27)
        int w = std::max(0, r.z - r.x), h = std::max(0, r.w - r.y);
        // the tile walk indexes a rect with 24-bit multiplies
        if (w >= (1 << 24) || h >= (1 << 24)) throw PfhError{PFH_E_LIMIT, "a path's tile rect has a side of 2^24 tiles or more"};
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
    // binning classes by what the path owns (its rect's tiles and its columns) and by its walk: a
    // path of more segments than a wave has lanes goes to the group class's 256 lanes
    std::vector<uint32_t> bin_wave, bin_group;
    for (int p = 0; p < d->n_paths; p++) {
        const int4 &r = s->path_rect[p];
        const uint64_t w = (uint64_t)std::max(0, r.z - r.x), h = (uint64_t)std::max(0, r.w - r.y);
        if (w == 0) continue;  // no columns: nothing is binned (k_flatten emits no chords)
        const uint64_t nseg = s->path_seg_off[p + 1] - s->path_seg_off[p];
        if (w * h <= (uint64_t)BIN_WAVE_TILES && w <= (uint64_t)BIN_WAVE_COLS && nseg <= (uint64_t)BIN_WAVE_SEGS)
            bin_wave.push_back((uint32_t)p);
        else bin_group.push_back((uint32_t)p);
    }
    s->n_bin_wave = (uint32_t)bin_wave.size();
    s->n_bin_group = (uint32_t)bin_group.size();
    // clip references: a clip path's binning position (or CLIP_EMPTY without tile columns)
    s->n_clip = (uint32_t)d->n_clip_paths;
    s->clip_own_rules = d->clip_rule == PFH_CLIP_RULE_OWN;
    s->path_clip.assign(d->n_paths, CLIP_NONE);
    if (s->n_clip > 0) {
        const int n_draw = d->n_paths - d->n_clip_paths;
        std::vector<uint32_t> pos(d->n_paths, CLIP_EMPTY);
        for (size_t i = 0; i < bin_wave.size(); i++) pos[bin_wave[i]] = (uint32_t)i;
        for (size_t i = 0; i < bin_group.size(); i++) pos[bin_group[i]] = CLIP_GROUP | (uint32_t)i;
        s->path_level.assign(d->n_paths, 0);
        for (int p = 0; p < d->n_paths; p++) {
            if (d->path_clip && d->path_clip[p] >= 0) s->path_clip[p] = pos[n_draw + d->path_clip[p]];
            if (p >= n_draw) {
                uint32_t k = 1;
                for (int32_t c = d->path_clip ? d->path_clip[p] : -1; c >= 0; c = d->path_clip[n_draw + c]) k++;
                s->path_level[p] = (uint8_t)k;
                s->clip_levels = std::max(s->clip_levels, k);
            }
        }
    }
    bin_wave.insert(bin_wave.end(), bin_group.begin(), bin_group.end());
    s->features.resize((size_t)d->n_paths * N_FEATURES);
    for (int p = 0; p < d->n_paths; p++) path_features(*s, pts, segs, (uint32_t)p, &s->features[(size_t)p * N_FEATURES]);

    // On the device, segments of paths without tile columns carry kind 0: nothing can land
    // from them (add_fill culls, adjust_alpha_tile_backdrop ignores x outside [0, width)), so
    // the flattener skips them without looking up their path.  So do segments with a
    // non-finite point (process_line_segment drops every piece of them, tiler.cpp:133-136).
    // A segment that does not continue the previous emitting segment of its path bit for bit
    // is marked SEG_CHAIN_BREAK: its chord points start with a separator and its start point.
    std::vector<uint2> segs_dev = segs;
    s->path_breaks.assign(d->n_paths, 0);
    s->n_breaks = 0;
    auto seg_np = [](uint32_t k) { return k == SEG_LINE ? 2u : (k == SEG_QUAD ? 3u : 4u); };
    auto bits = [](float2 v) { uint64_t u; std::memcpy(&u, &v, 8); return u; };
    for (size_t i = 0; i < segs_dev.size(); i++) {
        const uint32_t p = seg_path[i];
        const int4 &r = s->path_rect[p];
        const uint32_t np = seg_np(segs[i].y);
        bool finite = true;
        for (uint32_t q = 0; q < np; q++) finite = finite && std::isfinite(pts[segs[i].x + q].x) && std::isfinite(pts[segs[i].x + q].y);
        if (r.z - r.x <= 0 || !finite) {
            segs_dev[i].y = 0;
            continue;
        }
        bool chained = false;
        for (size_t j = i; j-- > 0 && seg_path[j] == p;) {  // the previous emitting segment of the path
            if (segs_dev[j].y == 0) break;
            chained = bits(pts[segs[j].x + seg_np(segs[j].y) - 1]) == bits(pts[segs[i].x]);
            break;
        }
        if (!chained) {
            segs_dev[i].y |= SEG_CHAIN_BREAK;
            s->path_breaks[p]++;
            s->n_breaks++;
        }
    }
    HIP_TRY(hipSetDevice(ctx->device));
    upload(s->d_features, s->features);
    upload(s->d_pts, pts);
    upload(s->d_segs, segs_dev);
    upload(s->d_seg_path, seg_path);
    upload(s->d_path_rect, s->path_rect);
    upload(s->d_path_tile_off, s->path_tile_off);
    upload(s->d_path_col_off, s->path_col_off);
    upload(s->d_col_path, col_path);
    upload(s->d_path_rgba, s->path_rgba);
    upload(s->d_path_rule, s->path_rule);
    if (!s->path_paint.empty()) {
        upload(s->d_path_paint, s->path_paint);
        upload(s->d_paints, paints);
        upload(s->d_ramps, ramps);
    }
    upload(s->d_path_seg_off, s->path_seg_off);
    upload(s->d_bin_paths, bin_wave);
    if (s->n_clip > 0) {
        upload(s->d_path_clip, s->path_clip);
        upload(s->d_path_level, s->path_level);
    }
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
                                ctx->mlp_out.as<float>(), ctx->w->stream));
        HIP_TRY(hipMemcpyAsync(sums, ctx->mlp_out.ptr, 12, hipMemcpyDeviceToHost, ctx->w->stream));
        HIP_TRY(hipStreamSynchronize(ctx->w->stream));
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
    uint64_t S = 0, T = 0, W = 0, FB = 0, BW = 0, BG = 0, P = 0;
    uint64_t breaks = 0;  // chord chains of the batch (two stream points each besides the chords)
    bool clips = false;   // some scene has clip paths
    bool paints = false;  // some scene has gradient paints
    bool clear_outside = false;  // some scene's target reaches past its view box's tiles (and is cleared)
    uint32_t max_rows = 0;       // tile rows of the tallest target
    uint32_t clip_levels = 0;  // deepest clip nesting level of the batch
    uint32_t cap_f = 0, cap_a = 0, cap_e = 0, cap_l = 0;
    int key_bits = 1;
    Schedule sched;
};

// Segments per flatten wave: the schedule's choice, or enough for ~4 waves per SIMD (1024 SIMDs),
// 64 to 1024 (a wave balances its lanes over its whole range; past 1024 too few waves share the
// curves of a large batch).
static uint32_t flatten_segs(const Plan &pl) {
    if (pl.sched.flatten_segs) return pl.sched.flatten_segs;
    const uint64_t per = (pl.S + 4096 * 64 - 1) / (4096 * 64);
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(64 * per, 64), 1024);
}

// Entries per deep-list region: 64 for every wave of k_flatten_lines (64 segments each) that maps
// to the region (waves w, w + DEEP_SLOTS, ...).
static uint32_t deep_slot_cap(uint64_t segs) {
    const uint64_t waves = (segs + 63) / 64;
    return (uint32_t)(64 * ((waves + DEEP_SLOTS - 1) / DEEP_SLOTS));
}

static uint32_t clamp_cap(uint64_t v) { return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(v, 64), 0xF0000000ull); }

static void ensure_sizes(Workspace &w, const Plan &pl, bool dump) {
    w.tile_delta.ensure(4 * pl.T + 4);
    w.tile_cnt.ensure(4 * pl.T + 4);
    w.tile_fill_off.ensure(4 * pl.T + 8);
    w.tile_alpha_idx.ensure(4 * pl.T + 8);
    if (dump || pl.clips) w.tile_bd.ensure(pl.T + 4);
    if (pl.clips) {
        w.tile_clip_ai.ensure(4 * pl.T + 4);
        w.clip_jobs.ensure(8ull * pl.cap_a + 8);
        w.clip_rec.ensure(16ull * pl.cap_a + 16);
        w.clip_seen.ensure(4ull * (pl.cap_a / 32 + 1));
    }
    w.tile_st.ensure(pl.T + 4);
    w.path_ent.ensure(8 * pl.P + 8);
    w.ent_quarter.ensure(32 * pl.BG + 32);
    w.path_tot.ensure(8 * pl.P + 8);
    w.bin_items.ensure(sizeof(BinItem) * (pl.BW + pl.BG) + 64);
    w.col_item.ensure(4 * pl.W + 4);
    w.col_bd.ensure(4 * pl.W + 4);
    w.tall_cols.ensure(4 * pl.W + 4);
    w.seg_cnt.ensure(4 * pl.S + 8);
    w.curve_mask.ensure(8 * (pl.S / 64) + 16);
    if (pl.sched.flatten_deep) w.deep_list.ensure(sizeof(DeepCurve) * DEEP_SLOTS * (uint64_t)deep_slot_cap(pl.S) + 16);
    w.lines.ensure(8ull * pl.cap_l);
    w.line_cnt.ensure(4ull * pl.cap_l + 8);
    w.zbuf.ensure(4 * pl.FB + 4);
    w.fb_range.ensure(8 * pl.FB + 8);
    w.counters.ensure(4 * C_COUNT);
    uint64_t scan_n = std::max<uint64_t>({pl.T + 1, pl.S + 1, pl.FB + 1, pl.P + 1, pl.W + 1, (uint64_t)pl.cap_l + 1, 1});
    w.scan_tmp.ensure(16 * scan_temp_elems(scan_n) + 64);
    w.fills.ensure(8ull * pl.cap_f);
    if (dump) w.dump_fills.ensure(16ull * pl.cap_f);
    w.alpha_rec.ensure(16ull * pl.cap_a);
    w.alpha_vis.ensure(4ull * (pl.clips ? std::max(pl.cap_a, pl.cap_e) : pl.cap_a));
    w.masks.ensure(256ull * pl.cap_a + 256);  // + the all-ones row
    w.keys.ensure(4ull * pl.cap_e);
    w.vals.ensure(4ull * pl.cap_e);
    w.keys_alt.ensure(4ull * pl.cap_e);
    w.vals_alt.ensure(4ull * pl.cap_e);
    w.entry_rec.ensure(8ull * pl.cap_e);
    if (dump) w.entry_meta.ensure(8ull * pl.cap_e);
    w.sort_tmp.ensure(4 * radix_temp_elems(pl.cap_e, (uint32_t)pl.ids.size()) + 64);
}

// Capacities the current allocations hold, by the same per-buffer sizes ensure_sizes asks
// for, so that ensure_sizes on these capacities never reallocates (captured graphs stay valid).
struct Caps {
    uint32_t l, f, a;
};
static Caps held_caps(Workspace &w, bool dump) {
    auto fit = [](size_t bytes, size_t per, size_t extra) -> uint64_t {
        return bytes > extra ? (bytes - extra) / per : 0;
    };
    uint64_t l = std::min({fit(w.lines.bytes, 8, 0), fit(w.line_cnt.bytes, 4, 8)});
    uint64_t f = fit(w.fills.bytes, 8, 0);
    if (dump) f = std::min(f, fit(w.dump_fills.bytes, 16, 0));
    uint64_t a = std::min({fit(w.alpha_rec.bytes, 16, 0), fit(w.alpha_vis.bytes, 4, 0), fit(w.masks.bytes, 256, 256)});
    return Caps{clamp_cap(l), clamp_cap(f), clamp_cap(a)};
}

// base table: segment, tile, column, fb-tile, wave-class and group-class item offsets per scene
constexpr size_t N_BASES = 6;

static void write_tables(pfh_ctx *ctx, Plan &pl, bool load_dst) {
    pl.sched = ctx->sched;
    size_t n = pl.ids.size();
    if (ctx->w->h_table_cap < n) {
        if (ctx->w->h_table) (void)hipHostFree(ctx->w->h_table);
        HIP_TRY(hipHostMalloc((void **)&ctx->w->h_table, sizeof(SceneDev) * n));
        ctx->w->h_table_cap = n;
    }
    if (ctx->w->h_bases_cap < N_BASES * (n + 1)) {
        if (ctx->w->h_bases) (void)hipHostFree(ctx->w->h_bases);
        HIP_TRY(hipHostMalloc((void **)&ctx->w->h_bases, sizeof(uint32_t) * N_BASES * (n + 1)));
        ctx->w->h_bases_cap = N_BASES * (n + 1);
    }
    uint64_t S = 0, T = 0, W = 0, FB = 0, BW = 0, BG = 0, P = 0;
    for (size_t k = 0; k < n; k++) {
        SceneHost &s = get_scene(ctx, pl.ids[k]);
        SceneDev &d = ctx->w->h_table[k];
        d.pts = s.d_pts.as<float2>();
        d.segs = s.d_segs.as<uint2>();
        d.seg_path = s.d_seg_path.as<uint32_t>();
        d.path_rect = s.d_path_rect.as<int4>();
        d.path_tile_off = s.d_path_tile_off.as<uint32_t>();
        d.path_col_off = s.d_path_col_off.as<uint32_t>();
        d.col_path = s.d_col_path.as<uint32_t>();
        d.path_rgba = s.d_path_rgba.as<uint32_t>();
        d.path_seg_off = s.d_path_seg_off.as<uint32_t>();
        d.bin_paths = s.d_bin_paths.as<uint32_t>();
        d.n_bin_wave = s.n_bin_wave;
        d.n_bin_group = s.n_bin_group;
        d.bin_wave_base = (uint32_t)BW;
        d.bin_group_base = (uint32_t)BG;
        d.path_rule = s.d_path_rule.as<uint8_t>();
        d.clip_own_rules = s.clip_own_rules ? 1u : 0u;
        d.path_clip = s.n_clip > 0 ? s.d_path_clip.as<uint32_t>() : nullptr;
        d.path_level = s.n_clip > 0 ? s.d_path_level.as<uint8_t>() : nullptr;
        const bool painted = !s.path_paint.empty();
        d.path_paint = painted ? s.d_path_paint.as<uint32_t>() : nullptr;
        d.paints = painted ? s.d_paints.as<PaintRec>() : nullptr;
        d.has_ops = s.has_ops ? 1u : 0u;
        d.ramps = painted ? s.d_ramps.as<uint32_t>() : nullptr;
        d.vb_l = s.vb[0];
        d.vb_t = s.vb[1];
        d.vb_r = s.vb[2];
        d.vb_b = s.vb[3];
        d.fb_x0 = s.fb_x0;
        d.fb_y0 = s.fb_y0;
        d.fb_w = s.fb_w;
        d.fb_inv_w = s.fb_w > 0 ? 1.0f / (float)s.fb_w : 0.0f;
        d.fb_h = s.fb_h;
        d.width = s.width;
        d.height = s.height;
        d.n_paths = s.n_paths;
        d.n_segs = s.n_segs;
        d.n_tiles = s.n_tiles;
        d.n_cols = s.n_cols;
        d.n_fb = (uint32_t)(s.fb_w * s.fb_h);
        d.seg_base = (uint32_t)S;
        d.path_base = (uint32_t)P;
        d.tile_base = (uint32_t)T;
        d.col_base = (uint32_t)W;
        d.fb_base = (uint32_t)FB;
        d.out = s.d_out;
        d.out_pitch = s.out_pitch;
        d.load_dst = load_dst ? 1u : 0u;
        // the pixels of the target outside the view box's tiles are cleared by k_clear_outside
        // (the reference clears the whole destination texture, clear_dst)
        const bool covered = s.fb_x0 <= 0 && s.fb_y0 <= 0 && (int64_t)(s.fb_x0 + s.fb_w) * 16 >= s.width &&
                             (int64_t)(s.fb_y0 + s.fb_h) * 16 >= s.height;
        pl.clear_outside = pl.clear_outside || (!load_dst && !covered);
        pl.max_rows = std::max<uint32_t>(pl.max_rows, (uint32_t)((s.height + 15) / 16));
        s.seg_base = d.seg_base;
        s.path_base = d.path_base;
        s.tile_base = d.tile_base;
        s.col_base = d.col_base;
        s.fb_base = d.fb_base;
        ctx->w->h_bases[0 * (n + 1) + k] = (uint32_t)S;
        ctx->w->h_bases[1 * (n + 1) + k] = (uint32_t)T;
        ctx->w->h_bases[2 * (n + 1) + k] = (uint32_t)W;
        ctx->w->h_bases[3 * (n + 1) + k] = (uint32_t)FB;
        ctx->w->h_bases[4 * (n + 1) + k] = (uint32_t)BW;
        ctx->w->h_bases[5 * (n + 1) + k] = (uint32_t)BG;
        S += s.n_segs;
        pl.breaks += s.n_breaks;
        pl.clips = pl.clips || s.n_clip > 0;
        pl.paints = pl.paints || !s.path_paint.empty();
        pl.clip_levels = std::max(pl.clip_levels, s.clip_levels);
        T += s.n_tiles;
        W += s.n_cols;
        FB += d.n_fb;
        BW += s.n_bin_wave;
        BG += s.n_bin_group;
        P += s.n_paths;
        if (S >= 0xF0000000ull || T >= 0xF0000000ull || W >= 0xF0000000ull || FB >= 0xF0000000ull)
            throw PfhError{PFH_E_LIMIT, "batch exceeds 2^32 segments/tiles"};
    }
    ctx->w->h_bases[0 * (n + 1) + n] = (uint32_t)S;
    ctx->w->h_bases[1 * (n + 1) + n] = (uint32_t)T;
    ctx->w->h_bases[2 * (n + 1) + n] = (uint32_t)W;
    ctx->w->h_bases[3 * (n + 1) + n] = (uint32_t)FB;
    ctx->w->h_bases[4 * (n + 1) + n] = (uint32_t)BW;
    ctx->w->h_bases[5 * (n + 1) + n] = (uint32_t)BG;
    pl.BW = BW;
    pl.BG = BG;
    pl.S = S;
    pl.T = T;
    pl.W = W;
    pl.FB = FB;
    pl.P = P;
    // the sort runs per scene on scene-local framebuffer tiles
    uint64_t max_fb = 1;
    for (size_t k = 0; k < n; k++) max_fb = std::max<uint64_t>(max_fb, ctx->w->h_table[k].n_fb);
    pl.key_bits = 1;
    while ((1ull << pl.key_bits) < max_fb) pl.key_bits++;
}

static void upload_tables(Workspace &w, const Plan &pl) {
    size_t n = pl.ids.size();
    w.scene_table.ensure(sizeof(SceneDev) * n);
    w.bases.ensure(sizeof(uint32_t) * N_BASES * (n + 1));
    HIP_TRY(hipMemcpyAsync(w.scene_table.ptr, w.h_table, sizeof(SceneDev) * n, hipMemcpyHostToDevice, w.stream));
    HIP_TRY(hipMemcpyAsync(w.bases.ptr, w.h_bases, sizeof(uint32_t) * N_BASES * (n + 1), hipMemcpyHostToDevice,
                           w.stream));
}

static BatchDev batch_dev(Workspace &w, const Plan &pl, bool dump) {
    BatchDev b{};
    size_t n = pl.ids.size();
    b.scenes = w.scene_table.as<SceneDev>();
    b.n_scenes = (int32_t)n;
    const uint32_t *bases = w.bases.as<uint32_t>();
    b.seg_bases = bases;
    b.tile_bases = bases + (n + 1);
    b.col_bases = bases + 2 * (n + 1);
    b.fb_bases = bases + 3 * (n + 1);
    b.bin_wave_bases = bases + 4 * (n + 1);
    b.bin_group_bases = bases + 5 * (n + 1);
    b.n_bin_wave = (uint32_t)pl.BW;
    b.n_bin_group = (uint32_t)pl.BG;
    b.n_segs = (uint32_t)pl.S;
    b.n_tiles = (uint32_t)pl.T;
    b.n_cols = (uint32_t)pl.W;
    b.n_fb = (uint32_t)pl.FB;
    b.cap_fills = pl.cap_f;
    // a gradient entry keeps its mask row in 30 bits (ENTRY_ROW_MASK)
    if (pl.cap_a >= ENTRY_ROW_EMPTY) throw PfhError{PFH_E_LIMIT, "more than 2^30 alpha tiles in one launch"};
    b.cap_alpha = pl.cap_a;
    b.cap_entries = pl.cap_e;
    b.cap_lines = pl.cap_l;
    b.cap_vis = pl.clips ? std::max(pl.cap_a, pl.cap_e) : pl.cap_a;
    b.paints = pl.paints ? 1u : 0u;
    b.tile_delta = w.tile_delta.as<int32_t>();
    b.tile_cnt = w.tile_cnt.as<uint32_t>();
    b.tile_fill_off = w.tile_fill_off.as<uint32_t>();
    b.tile_alpha_idx = w.tile_alpha_idx.as<uint32_t>();
    b.tile_bd = (dump || pl.clips) ? w.tile_bd.as<int8_t>() : nullptr;
    b.tile_clip_ai = pl.clips ? w.tile_clip_ai.as<uint32_t>() : nullptr;
    b.clip_jobs = pl.clips ? w.clip_jobs.as<uint2>() : nullptr;
    b.clip_rec = pl.clips ? w.clip_rec.as<uint4>() : nullptr;
    b.clip_seen = pl.clips ? w.clip_seen.as<uint32_t>() : nullptr;
    b.clip_levels = pl.clip_levels;
    b.has_clips = pl.clips ? 1 : 0;
    b.tile_st = w.tile_st.as<uint8_t>();
    b.path_ent = w.path_ent.as<uint2>();
    b.ent_quarter = w.ent_quarter.as<uint2>();
    b.path_tot = w.path_tot.as<uint2>();
    b.bin_items = w.bin_items.as<BinItem>();
    b.col_item = w.col_item.as<uint32_t>();
    b.n_paths = (uint32_t)pl.P;
    b.col_bd = w.col_bd.as<int32_t>();
    b.tall_cols = w.tall_cols.as<uint32_t>();
    b.seg_cnt = w.seg_cnt.as<uint32_t>();
    b.curve_mask = w.curve_mask.as<uint64_t>();
    b.deep_pred = pl.sched.flatten_deep;
    b.deep_list = pl.sched.flatten_deep ? w.deep_list.as<DeepCurve>() : nullptr;
    b.deep_slot_cap = deep_slot_cap(pl.S);
    b.lines = w.lines.as<float2>();
    b.line_cnt = w.line_cnt.as<uint32_t>();
    b.fills = w.fills.as<uint2>();
    b.dump_fills = dump ? w.dump_fills.as<uint4>() : nullptr;
    b.alpha_rec = w.alpha_rec.as<uint4>();
    b.alpha_vis = w.alpha_vis.as<uint32_t>();
    b.masks = w.masks.as<uint8_t>();
    b.zbuf = w.zbuf.as<uint32_t>();
    b.fb_range = w.fb_range.as<uint2>();
    b.keys = w.keys.as<uint32_t>();
    b.vals = w.vals.as<uint32_t>();
    b.keys_alt = w.keys_alt.as<uint32_t>();
    b.vals_alt = w.vals_alt.as<uint32_t>();
    b.entry_rec = w.entry_rec.as<uint2>();
    b.entry_meta = dump ? w.entry_meta.as<uint2>() : nullptr;
    b.counters = w.counters.as<uint32_t>();
    b.dump = dump ? 1 : 0;
    b.xcd_order = pl.sched.xcd_order;
    return b;
}

// Stage groups; every launch goes to ctx->w->stream so the batched sequence can be captured.
static void stage_clear(Workspace &w, const BatchDev &b) {
    hipStream_t s = w.stream;
    // tile_delta / tile_cnt / col_bd need no clearing: the bin count pass stores every
    // tile and column of every binned path (paths with no columns own none)
    // one launch for every array (no memsets: the clears are captured into graphs)
    ZeroSpans z{};
    z.add(b.zbuf, b.n_fb);
    // paths with no columns are not binned: their totals and entry counts stay zero
    z.add(reinterpret_cast<uint32_t *>(b.path_tot), 2 * (uint64_t)b.n_paths);
    z.add(reinterpret_cast<uint32_t *>(b.path_ent), 2 * (uint64_t)b.n_paths);
    z.add(reinterpret_cast<uint32_t *>(b.fb_range), 2 * (uint64_t)b.n_fb);
    z.add(b.counters, C_COUNT);
    if (b.clip_seen) z.add(b.clip_seen, (uint64_t)b.cap_alpha / 32 + 1);
    HIP_TRY(launch_zero_spans(z, s));
}

// flatten (count, scan, emit) and bin count: chords and exact fill / alpha-tile counts
static void stage_count(Workspace &w, const BatchDev &b, const Plan &pl) {
    hipStream_t s = w.stream;
    uint32_t *tmp = w.scan_tmp.as<uint32_t>();
    HIP_TRY(launch_flatten(b, false, flatten_segs(pl), pl.sched.flatten_take, s));
    HIP_TRY(scan_total(b.seg_cnt, b.n_segs, b.counters, C_LINES, tmp, s));
    HIP_TRY(launch_flatten(b, true, flatten_segs(pl), pl.sched.flatten_take, s));
    HIP_TRY(launch_bin_items(b, s));
    HIP_TRY(launch_bin(b, false, s));
    HIP_TRY(scan_paths(b, w.scan_tmp.ptr, s));
}

static void stage_emit(Workspace &w, const BatchDev &b, const Plan &pl) {
    hipStream_t s = w.stream;
    uint32_t *tmp = w.scan_tmp.as<uint32_t>();
    if (b.dump) HIP_TRY(scan_counted(b.line_cnt, b.counters + C_LINES, b.cap_lines, b.counters, -1, tmp, s));
    HIP_TRY(launch_bin(b, true, s));
    HIP_TRY(launch_propagate(b, pl.sched.column_block, pl.sched.column_rows, s));
    HIP_TRY(launch_entries(b, false, s));
    HIP_TRY(scan_entries(b.path_ent, b.n_paths, b.counters, w.scan_tmp.ptr, s));
}

static void stage_finish(Workspace &w, const BatchDev &b, const Plan &pl) {
    HIP_TRY(launch_entries(b, true, w.stream));
    bool alt = false;
    HIP_TRY(radix_sort(b, b.keys, b.vals, b.keys_alt, b.vals_alt, pl.key_bits, w.sort_tmp.as<uint32_t>(), &alt,
                       w.stream));
    w.last_sorted_alt = alt;
    HIP_TRY(launch_fb_bounds(alt ? b.keys_alt : b.keys, b.counters + C_ENTRIES, b.cap_entries, b.n_fb, b.fb_range,
                             w.stream));
    HIP_TRY(hipEventRecordWithFlags(w.kev[0], w.stream, hipEventRecordExternal));
    HIP_TRY(launch_fill_masks(b, b.cap_alpha, pl.sched.fill_grid, w.stream));
    HIP_TRY(launch_fill_clipped(b, w.stream));
    HIP_TRY(hipEventRecordWithFlags(w.kev[1], w.stream, hipEventRecordExternal));
    if (pl.clear_outside) HIP_TRY(launch_clear_outside(b, pl.max_rows, w.stream));
    HIP_TRY(launch_composite(b, alt ? b.vals_alt : b.vals, pl.sched.composite_block, w.stream));
    HIP_TRY(hipEventRecordWithFlags(w.kev[2], w.stream, hipEventRecordExternal));
}

static void read_counters(pfh_ctx *ctx) {
    ctx->stats.host_syncs++;
    HIP_TRY(hipMemcpyAsync(ctx->w->h_counters, ctx->w->counters.ptr, 4 * C_COUNT, hipMemcpyDeviceToHost, ctx->w->stream));
    HIP_TRY(hipStreamSynchronize(ctx->w->stream));
}

static double elapsed_ms(pfh_ctx *ctx, hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, a, b));
    return ms;
}

static void record_counts(pfh_ctx *ctx, const Plan &pl, bool dump) {
    ctx->stats.fills = ctx->w->h_counters[C_FILLS];
    int64_t drawn = 0;
    for (int i = 0; i < DRAWN_SLOTS; i++) drawn += ctx->w->h_counters[C_DRAWN + i];
    ctx->stats.fills_masked = drawn;
    // the fill kernel rasterizes the alpha tiles of draw entries (every alpha tile with dumps)
    // (listed mode: one row per alpha entry, less the clip mask rows another entry listed first)
    ctx->stats.alpha_masked = std::min<int64_t>(
        dump ? ctx->w->h_counters[C_ALPHA]
             : (int64_t)ctx->w->h_counters[C_ALPHA_ENTRIES] - (int64_t)ctx->w->h_counters[C_CLIP_DUPS],
        pl.cap_a);
    ctx->stats.alpha_tiles = ctx->w->h_counters[C_ALPHA];
    ctx->stats.entries = ctx->w->h_counters[C_ENTRIES];
    ctx->stats.guard_hits = ctx->w->h_counters[C_GUARD];
    ctx->stats.alpha_entries = ctx->w->h_counters[C_ALPHA_ENTRIES];
    // C_LINES counts stream points: two per chord chain besides one per chord
    ctx->stats.lines = (int64_t)ctx->w->h_counters[C_LINES] - 2 * (int64_t)pl.breaks;
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
    Caps held = held_caps(*ctx->w, dump);
    pl.cap_f = held.f;
    pl.cap_a = held.a;
    pl.cap_l = held.l;
    pl.cap_e = clamp_cap(std::max<uint64_t>(pl.T, 1));
    ensure_sizes(*ctx->w, pl, dump);
    upload_tables(*ctx->w, pl);
    BatchDev b = batch_dev(*ctx->w, pl, dump);
    hipStream_t st = ctx->w->stream;
    // scratch is re-read at every use: ensure_sizes may reallocate it when a count grows
    auto tmp = [ctx]() { return ctx->w->scan_tmp.as<uint32_t>(); };
    hipEvent_t *ev = ctx->w->stage_ev;
    HIP_TRY(hipEventRecord(ev[0], st));
    stage_clear(*ctx->w, b);
    HIP_TRY(launch_flatten(b, false, flatten_segs(pl), pl.sched.flatten_take, st));
    HIP_TRY(scan_total(b.seg_cnt, b.n_segs, b.counters, C_LINES, tmp(), st));
    read_counters(ctx);  // chord count (renderer.cpp:707-724)
    uint32_t nl = ctx->w->h_counters[C_LINES];
    if (nl > pl.cap_l) {
        pl.cap_l = clamp_cap(nl);
        ensure_sizes(*ctx->w, pl, dump);
        b = batch_dev(*ctx->w, pl, dump);
    }
    HIP_TRY(launch_flatten(b, true, flatten_segs(pl), pl.sched.flatten_take, st));
    HIP_TRY(hipEventRecord(ev[1], st));
    HIP_TRY(launch_bin_items(b, st));
    HIP_TRY(launch_bin(b, false, st));
    HIP_TRY(scan_paths(b, ctx->w->scan_tmp.ptr, st));
    read_counters(ctx);  // fill / alpha counts (renderer.cpp:834-845)
    HIP_TRY(hipEventRecord(ev[2], st));
    uint32_t nf = ctx->w->h_counters[C_FILLS], na = ctx->w->h_counters[C_ALPHA];
    if (nf > pl.cap_f || na > pl.cap_a) {
        pl.cap_f = clamp_cap(std::max<uint64_t>(nf, pl.cap_f));
        pl.cap_a = clamp_cap(std::max<uint64_t>(na, pl.cap_a));
        ensure_sizes(*ctx->w, pl, dump);
        b = batch_dev(*ctx->w, pl, dump);
    }
    if (dump) HIP_TRY(scan_counted(b.line_cnt, b.counters + C_LINES, nl, b.counters, -1, tmp(), st));
    HIP_TRY(launch_bin(b, true, st));
    HIP_TRY(hipEventRecord(ev[3], st));
    HIP_TRY(launch_propagate(b, pl.sched.column_block, pl.sched.column_rows, st));
    HIP_TRY(hipEventRecord(ev[4], st));
    HIP_TRY(launch_entries(b, false, st));
    HIP_TRY(scan_entries(b.path_ent, b.n_paths, b.counters, ctx->w->scan_tmp.ptr, st));
    HIP_TRY(launch_entries(b, true, st));
    HIP_TRY(hipEventRecord(ev[5], st));
    bool alt = false;
    HIP_TRY(radix_sort(b, b.keys, b.vals, b.keys_alt, b.vals_alt, pl.key_bits, ctx->w->sort_tmp.as<uint32_t>(), &alt,
                       st));
    ctx->w->last_sorted_alt = alt;
    HIP_TRY(launch_fb_bounds(alt ? b.keys_alt : b.keys, b.counters + C_ENTRIES, b.cap_entries, b.n_fb, b.fb_range, st));
    HIP_TRY(hipEventRecord(ev[6], st));
    HIP_TRY(launch_fill_masks(b, na, pl.sched.fill_grid, st));
    HIP_TRY(launch_fill_clipped(b, st));
    HIP_TRY(hipEventRecord(ev[7], st));
    if (pl.clear_outside) HIP_TRY(launch_clear_outside(b, pl.max_rows, st));
    HIP_TRY(launch_composite(b, alt ? b.vals_alt : b.vals, pl.sched.composite_block, st));
    HIP_TRY(hipEventRecord(ev[8], st));
    read_counters(ctx);
    for (int i = 0; i < 8; i++) ctx->stats.ms_stage[i] += elapsed_ms(ctx, ev[i], ev[i + 1]);
    record_counts(ctx, pl, dump);
    SceneHost &s = get_scene(ctx, id);
    s.counted = true;
    s.last_lines = nl;
    s.last_fills = ctx->w->h_counters[C_FILLS];
    s.last_alpha = ctx->w->h_counters[C_ALPHA];
    s.last_entries = ctx->w->h_counters[C_ENTRIES];
    s.dumped = dump;
}

// Batched: every scene in one pass, no host round trip inside; the launch sequence is
// captured once per (scene set, capacities, flags) and replayed from a HIP graph.  Split in
// three so that the request loop can keep launches of several workspaces in flight:
// batched_prepare (capacities), batched_enqueue (asynchronous), batched_check (after the
// launch's counters have arrived: accept, or grow and go again).
struct BatchedLaunch {
    Plan pl;
    uint32_t flags = 0;
    bool dump = false, use_graph = true;
    uint64_t plc = 0, pf = 0, pa = 0;  // requested chord / fill / alpha-tile capacities
    int attempt = 0;
};

static void batched_prepare(pfh_ctx *ctx, const int32_t *ids, int32_t n, uint32_t flags, BatchedLaunch &bl) {
    bl = BatchedLaunch{};
    bl.flags = flags;
    bl.dump = (flags & PFH_DUMP) != 0;
    bl.use_graph = (flags & PFH_NO_GRAPH) == 0;
    Plan &pl = bl.pl;
    pl.ids.assign(ids, ids + n);
    write_tables(ctx, pl, (flags & PFH_LOAD_DST) != 0);
    // Capacities: entries are bounded by tiles exactly; chords, fills and alpha tiles come
    // from the predictor, else from the last exact counts, else from a count pass read back.
    bool known = true;
    for (int32_t id : pl.ids) {
        SceneHost &s = get_scene(ctx, id);
        if ((flags & PFH_PREDICT) && !ctx->predictor.empty()) {
            bl.plc += (uint64_t)std::ceil(s.pred_lines * ctx->predictor_safety) + 2 * s.n_breaks + 64;
            bl.pf += (uint64_t)std::ceil(s.pred_fills * ctx->predictor_safety) + 64;
            bl.pa += (uint64_t)std::ceil(s.pred_alpha * ctx->predictor_safety) + 16;
        } else if (s.counted) {
            bl.plc += s.last_lines;
            bl.pf += s.last_fills;
            bl.pa += s.last_alpha;
        } else {
            known = false;
        }
    }
    pl.cap_e = clamp_cap(std::max<uint64_t>(pl.T, 1));
    if (!known) {
        // size from an eager count pass (read back), as the plain pipeline does
        Workspace &w = *ctx->w;
        Caps held = held_caps(w, bl.dump);
        pl.cap_l = held.l;
        pl.cap_f = held.f;
        pl.cap_a = held.a;
        ensure_sizes(w, pl, bl.dump);
        upload_tables(w, pl);
        BatchDev b = batch_dev(w, pl, bl.dump);
        stage_clear(w, b);
        HIP_TRY(launch_flatten(b, false, flatten_segs(pl), pl.sched.flatten_take, w.stream));
        HIP_TRY(scan_total(b.seg_cnt, b.n_segs, b.counters, C_LINES, w.scan_tmp.as<uint32_t>(), w.stream));
        read_counters(ctx);
        bl.plc = w.h_counters[C_LINES];
        pl.cap_l = clamp_cap(std::max<uint64_t>(bl.plc, pl.cap_l));
        ensure_sizes(w, pl, bl.dump);
        b = batch_dev(w, pl, bl.dump);
        stage_clear(w, b);
        stage_count(w, b, pl);
        read_counters(ctx);
        bl.pf = w.h_counters[C_FILLS];
        bl.pa = w.h_counters[C_ALPHA];
    }
}

// Enqueues one attempt on the current workspace's stream and the copy of its counters;
// returns without waiting.  ev1 marks the end of the attempt.
static void batched_enqueue(pfh_ctx *ctx, BatchedLaunch &bl) {
    Workspace &w = *ctx->w;
    Plan &pl = bl.pl;
    Caps held = held_caps(w, bl.dump);
    pl.cap_l = clamp_cap(std::max<uint64_t>(bl.plc, held.l));
    pl.cap_f = clamp_cap(std::max<uint64_t>(bl.pf, held.f));
    pl.cap_a = clamp_cap(std::max<uint64_t>(bl.pa, held.a));
    ensure_sizes(w, pl, bl.dump);
    // ensure() rounds to 2^k: use the full allocation, then size the buffers that depend
    // on it (scan scratch) for exactly those capacities, so that a repeat call asks for
    // nothing new and replays its graph
    for (int fix = 0; fix < 4; fix++) {
        held = held_caps(w, bl.dump);
        if (fix > 0 && held.l == pl.cap_l && held.f == pl.cap_f && held.a == pl.cap_a) break;
        pl.cap_l = held.l;
        pl.cap_f = held.f;
        pl.cap_a = held.a;
        ensure_sizes(w, pl, bl.dump);
    }
    upload_tables(w, pl);
    BatchDev b = batch_dev(w, pl, bl.dump);
    HIP_TRY(hipEventRecord(w.ev0, w.stream));
    if (bl.use_graph) {
        GraphKey key{pl.ids,
                     bl.flags & ~(uint32_t)PFH_PREDICT,
                     pl.cap_f,
                     pl.cap_a,
                     pl.cap_e,
                     pl.cap_l,
                     {pl.sched.fill_grid, pl.sched.composite_block, pl.sched.column_block, pl.sched.xcd_order,
                      pl.sched.flatten_segs, pl.sched.flatten_take, pl.sched.column_rows, pl.sched.flatten_deep}};
        if (w.graphs_generation != g_alloc_generation) {
            // some buffer moved since these graphs were captured: never replay them
            w.drop_graphs();
            w.graphs_generation = g_alloc_generation;
        }
        auto it = w.graphs.find(key);
        if (it == w.graphs.end()) {
            if (w.graphs.size() > 32) w.drop_graphs();
            CapturedGraph cg;
            try {
                hipGraph_t graph;
                HIP_TRY(hipStreamBeginCapture(w.stream, hipStreamCaptureModeThreadLocal));
                try {
                    stage_clear(w, b);
                    stage_count(w, b, pl);
                    stage_emit(w, b, pl);
                    stage_finish(w, b, pl);
                } catch (...) {
                    hipGraph_t dummy = nullptr;
                    (void)hipStreamEndCapture(w.stream, &dummy);
                    if (dummy) (void)hipGraphDestroy(dummy);
                    throw;
                }
                HIP_TRY(hipStreamEndCapture(w.stream, &graph));
                hipError_t ie = hipGraphInstantiate(&cg.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HIP_TRY(ie);
            } catch (...) {
                Workspace::destroy_graph(cg);
                throw;
            }
            it = w.graphs.emplace(key, cg).first;
        } else {
            // the sort's result buffer parity depends only on the key width
            w.last_sorted_alt = ((pl.key_bits + 7) / 8) % 2 == 1;
        }
        HIP_TRY(hipGraphLaunch(it->second.exec, w.stream));
    } else {
        stage_clear(w, b);
        stage_count(w, b, pl);
        stage_emit(w, b, pl);
        stage_finish(w, b, pl);
    }
    HIP_TRY(hipMemcpyAsync(w.h_counters, w.counters.ptr, 4 * C_COUNT, hipMemcpyDeviceToHost, w.stream));
    HIP_TRY(hipEventRecord(w.ev1, w.stream));
}

// After the attempt's counters have arrived (ev1 complete): true when every capacity held;
// otherwise the requested capacities grow to the exact counts for another attempt.
static bool batched_check(pfh_ctx *ctx, BatchedLaunch &bl) {
    Workspace &w = *ctx->w;
    Plan &pl = bl.pl;
    uint32_t nl = w.h_counters[C_LINES], nf = w.h_counters[C_FILLS], na = w.h_counters[C_ALPHA];
    if (nl <= pl.cap_l && nf <= pl.cap_f && na <= pl.cap_a) return true;
    // under-prediction: a retry with the exact counts, never a wrong pixel.  When the
    // chord buffer overflowed, the fill and alpha counts are partial: keep growing.
    ctx->stats.retries++;
    bl.plc = std::max<uint64_t>(bl.plc, nl);
    bl.pf = std::max<uint64_t>(bl.pf, nl > pl.cap_l ? 2ull * nf : nf);
    bl.pa = std::max<uint64_t>(bl.pa, nl > pl.cap_l ? 2ull * na : na);
    if (++bl.attempt >= 4) throw PfhError{PFH_E_OVERFLOW, "capacity still exceeded after retries"};
    return false;
}

static void batched_accept(pfh_ctx *ctx, BatchedLaunch &bl) {
    Workspace &w = *ctx->w;
    ctx->stats.ms_total = elapsed_ms(ctx, w.ev0, w.ev1);
    for (double &m : ctx->stats.ms_stage) m = 0;
    ctx->stats.ms_stage[6] = elapsed_ms(ctx, w.kev[0], w.kev[1]);
    ctx->stats.ms_stage[7] = elapsed_ms(ctx, w.kev[1], w.kev[2]);
    record_counts(ctx, bl.pl, bl.dump);
    for (int32_t id : bl.pl.ids) {
        SceneHost &s = get_scene(ctx, id);
        s.dumped = bl.dump;
    }
}

static void rasterize_batched(pfh_ctx *ctx, const int32_t *ids, int32_t n, uint32_t flags) {
    BatchedLaunch bl;
    batched_prepare(ctx, ids, n, flags, bl);
    while (true) {
        batched_enqueue(ctx, bl);
        ctx->stats.host_syncs++;
        HIP_TRY(hipEventSynchronize(ctx->w->ev1));
        if (batched_check(ctx, bl)) break;
    }
    batched_accept(ctx, bl);
}


// ---------------------------------------------------------------------------
// Request loop: continuous batching over several workspaces ("slots")
// ---------------------------------------------------------------------------
//
// Requests name resident scenes.  Each free slot takes the longest prefix-compatible set of
// queued requests whose predicted device bytes fit the memory budget beside the launches
// already in flight, launches them as one batched pass on its own stream and returns to the
// caller at once; requests that arrive meanwhile queue and join the next launch of whichever
// slot frees first, so work of different launches overlaps on the device.  A scene is in at
// most one launch at a time (its framebuffer is the launch's output).

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t id = 0;
    int32_t scene = 0;
    uint64_t bytes = 0;  // predicted device bytes of this scene's share of a launch
    double t_submit = 0, t_start = 0;
};

struct Slot {
    Workspace ws;
    bool busy = false;
    std::vector<Request> reqs;
    BatchedLaunch bl;
    uint64_t reserved = 0;
};

struct Server {
    std::vector<std::unique_ptr<Slot>> slots;
    std::deque<Request> queue;
    std::deque<int64_t> done;  // completed, not yet returned by pfh_server_poll
    std::vector<double> latency_ms;
    std::vector<int32_t> scene_busy;  // per scene id: in a launch
    uint64_t budget = 0, reserved = 0, peak_reserved = 0;
    int32_t max_batch = 64;
    uint32_t flags = 0;
    int64_t next_id = 1, submitted = 0, completed = 0, launches = 0, retries = 0, launched_requests = 0;
    int64_t max_concurrent = 0;
};

// Device bytes one scene adds to a launch, from its resident sizes and its predicted (or last
// exact) chord / fill / alpha-tile counts; per-element sizes are those ensure_sizes asks for.
static uint64_t scene_launch_bytes(const pfh_ctx *ctx, const SceneHost &s) {
    double L, F, A;
    if (!ctx->predictor.empty() && s.pred_lines >= 0) {
        L = s.pred_lines * ctx->predictor_safety + 2.0 * s.n_breaks;  // stream points
        F = s.pred_fills * ctx->predictor_safety;
        A = s.pred_alpha * ctx->predictor_safety;
    } else if (s.counted) {
        L = s.last_lines;
        F = s.last_fills;
        A = s.last_alpha;
    } else {  // unknown: generous bounds from the resident sizes
        L = 16.0 * s.n_segs;
        F = 2.0 * s.n_tiles + L;
        A = s.n_tiles;
    }
    const double T = s.n_tiles, W = s.n_cols, S = s.n_segs, FB = (double)s.fb_w * s.fb_h;
    const double per_tile = 4 + 4 + 8 + 1 + 4 * 4 + 8;  // counts, offsets, status, sort, records
    double b = T * per_tile + 4 * W + 16 * s.n_paths + 8 * S + 12 * FB + 12 * L + 8 * F + (16 + 4 + 256) * A;
    return (uint64_t)b + 4096;
}

static Server &server_of(pfh_ctx *ctx) {
    if (!ctx->server) throw PfhError{PFH_E_STATE, "request loop not started (pfh_server_start)"};
    return *ctx->server;
}

// Graphs of every workspace that rasterize scene `id`, after each workspace's stream is idle
// (the request loop refuses to release a scene it has queued or in flight).
static void drop_graphs_of_scene(pfh_ctx *ctx, int32_t id) {
    ctx->main.drop_graphs_of(id);
    if (ctx->server)
        for (auto &sl : ctx->server->slots) sl->ws.drop_graphs_of(id);
}

// One turn of the loop: retire finished launches (or re-launch them after an overflow),
// then fill free slots from the queue.
static void server_pump(pfh_ctx *ctx) {
    Server &sv = server_of(ctx);
    struct Restore {
        pfh_ctx *c;
        ~Restore() { c->w = &c->main; }
    } restore{ctx};
    for (auto &slp : sv.slots) {
        Slot &sl = *slp;
        if (!sl.busy) continue;
        hipError_t q = hipEventQuery(sl.ws.ev1);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) throw PfhError{PFH_E_HIP, std::string("request launch: ") + hipGetErrorString(q)};
        ctx->w = &sl.ws;
        if (!batched_check(ctx, sl.bl)) {
            sv.retries++;
            batched_enqueue(ctx, sl.bl);
            continue;
        }
        double t = now_s();
        for (const Request &r : sl.reqs) {
            sv.latency_ms.push_back((t - r.t_submit) * 1e3);
            sv.done.push_back(r.id);
            sv.scene_busy[r.scene] = 0;
        }
        sv.completed += (int64_t)sl.reqs.size();
        sl.reqs.clear();
        sv.reserved -= sl.reserved;
        sl.reserved = 0;
        sl.busy = false;
    }
    ctx->w = &ctx->main;
    for (auto &slp : sv.slots) {
        Slot &sl = *slp;
        if (sl.busy || sv.queue.empty()) continue;
        bool others_busy = false;
        for (auto &o : sv.slots) others_busy |= o->busy;
        std::vector<Request> batch;
        uint64_t bytes = 0;
        for (auto it = sv.queue.begin(); it != sv.queue.end() && (int32_t)batch.size() < sv.max_batch;) {
            bool fits = sv.reserved + bytes + it->bytes <= sv.budget || (batch.empty() && !others_busy);
            bool dup = sv.scene_busy[it->scene] != 0;
            for (const Request &r : batch) dup |= r.scene == it->scene;
            if (!fits) break;  // keep arrival order for admission
            if (dup) {
                ++it;
                continue;
            }
            bytes += it->bytes;
            batch.push_back(*it);
            it = sv.queue.erase(it);
        }
        if (batch.empty()) break;
        std::vector<int32_t> ids;
        double t = now_s();
        for (Request &r : batch) {
            ids.push_back(r.scene);
            r.t_start = t;
            sv.scene_busy[r.scene] = 1;
        }
        ctx->w = &sl.ws;
        batched_prepare(ctx, ids.data(), (int32_t)ids.size(), sv.flags | PFH_BATCHED, sl.bl);
        batched_enqueue(ctx, sl.bl);
        ctx->w = &ctx->main;
        sl.reqs = std::move(batch);
        sl.reserved = bytes;
        sl.busy = true;
        sv.reserved += bytes;
        sv.peak_reserved = std::max(sv.peak_reserved, sv.reserved);
        sv.launches++;
        sv.launched_requests += (int64_t)ids.size();
        int64_t busy = 0;
        for (auto &o : sv.slots) busy += o->busy;
        sv.max_concurrent = std::max(sv.max_concurrent, busy);
    }
}

static bool server_idle(const Server &sv) {
    if (!sv.queue.empty()) return false;
    for (auto &sl : sv.slots)
        if (sl->busy) return false;
    return true;
}

static void server_destroy(pfh_ctx *ctx) {
    if (!ctx->server) return;
    for (auto &sl : ctx->server->slots) sl->ws.destroy();
    ctx->server.reset();
}

static double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0;
    size_t k = (size_t)std::min<double>(v.size() - 1, std::floor(q * (v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + k, v.end());
    return v[k];
}

// A failed HIP call also leaves its status as the thread's last error, which the launch helpers
// read back after each launch: clear it on the way out so that a later, unrelated call does not
// report it again.
static int run_guarded(pfh_ctx *ctx, const std::function<void()> &f) {
    try {
        f();
        return PFH_OK;
    } catch (const PfhError &e) {
        if (ctx) ctx->err = e.msg;
        (void)hipGetLastError();
        return e.code;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        (void)hipGetLastError();
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
        ctx->main.create();
        ctx->fixed.create(1);  // plain mode: one scene per launch
        ctx->fixed.use_arena(scratch_on_device(device), PFH_SCRATCH_BYTES);
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
    server_destroy(ctx);
    ctx->main.destroy();
    ctx->fixed.destroy();
    for (auto &s : ctx->scenes) {
        if (!s) continue;
        for (DevBuf *b : {&s->d_features, &s->d_pts, &s->d_segs, &s->d_seg_path, &s->d_path_rect, &s->d_path_tile_off,
                          &s->d_path_col_off, &s->d_col_path, &s->d_path_rgba, &s->d_path_rule, &s->d_path_seg_off, &s->d_bin_paths, &s->d_out_owned, &s->d_path_clip,
                          &s->d_path_level, &s->d_path_paint, &s->d_paints, &s->d_ramps})
            b->release();
    }
    for (DevBuf *b : {&ctx->mlp_params, &ctx->mlp_x, &ctx->mlp_h, &ctx->mlp_out}) b->release();
    delete ctx;
}

const char *pfh_last_error(const pfh_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pfh_scene_upload(pfh_ctx *ctx, const pfh_scene_desc *desc, int32_t *scene_id) {
    if (!ctx || !scene_id) return PFH_E_ARG;
    return run_guarded(ctx, [&] { *scene_id = scene_upload(ctx, desc); });
}

int pfh_gradient_compile(const pfh_gradient *g, const float *stop_offsets, const uint32_t *stop_rgba,
                         int32_t ramp_row, uint32_t *ramp, float *record) {
    if (!g || !ramp || !record || ramp_row < 0 || ramp_row >= RAMP_TEXELS || g->stop_count < 0 || g->stop_start < 0 ||
        (g->stop_count > 0 && (!stop_offsets || !stop_rgba)))
        return PFH_E_ARG;
    try {
        PaintRec rec{};
        compile_gradient(*g, stop_offsets, stop_rgba, nullptr, ramp_row, ramp, rec);
        const float f[16] = {rec.m11, rec.m21, rec.m12, rec.m22, rec.m13, rec.m23, rec.from_x, rec.from_y,
                             rec.vec_x, rec.vec_y, rec.r0, rec.r1, rec.uv_x, rec.uv_y, (float)rec.kind, (float)rec.wrap};
        std::memcpy(record, f, sizeof(f));
        return PFH_OK;
    } catch (const PfhError &e) {
        return e.code;
    } catch (...) {
        return PFH_E_ARG;
    }
}

int pfh_stroke_outlines(pfh_ctx *ctx, const pfh_outline *in, const int32_t *path_contour_start, int32_t n_paths,
                        const pfh_stroke_style *styles, const float *dashes, int32_t n_dashes, pfh_outline_buf *out) {
    if (!ctx || !in || !out || n_paths < 0 || n_dashes < 0 || (n_paths > 0 && (!styles || !path_contour_start)))
        return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        if (in->n_contours < 0 || in->n_points < 0 || !in->contour_start || (in->n_points && (!in->points || !in->point_flags)) ||
            (in->n_contours && !in->contour_closed))
            throw PfhError{PFH_E_ARG, "null array"};
        if (in->contour_start[0] != 0 || in->contour_start[in->n_contours] != in->n_points ||
            (n_paths > 0 && (path_contour_start[0] != 0 || path_contour_start[n_paths] != in->n_contours)))
            throw PfhError{PFH_E_ARG, "contour / path ranges do not cover the arrays"};
        for (int32_t c = 0; c < in->n_contours; c++)
            if (in->contour_start[c + 1] < in->contour_start[c]) throw PfhError{PFH_E_ARG, "contour_start not monotone"};
        for (int32_t p = 0; p < n_paths; p++)
            if (path_contour_start[p + 1] < path_contour_start[p])
                throw PfhError{PFH_E_ARG, "path_contour_start not monotone"};
        HIP_TRY(hipSetDevice(ctx->device));
        HostOutlines so = dash_and_stroke(ctx, reinterpret_cast<const float2 *>(in->points), in->point_flags,
                                          in->contour_start, in->contour_closed, path_contour_start, n_paths, styles,
                                          dashes, n_dashes);
        out->n_points = (int64_t)so.pts.size();
        out->n_contours = (int64_t)so.closed.size();
        if (out->n_points > out->cap_points || out->n_contours > out->cap_contours)
            throw PfhError{PFH_E_LIMIT, "output capacity too small for the stroked outlines"};
        if ((out->n_points && (!out->points || !out->point_flags)) || !out->contour_start ||
            (out->n_contours && !out->contour_closed) || !out->path_contour_start)
            throw PfhError{PFH_E_ARG, "null output array"};
        if (out->n_points) {
            std::memcpy(out->points, so.pts.data(), 8 * so.pts.size());
            std::memcpy(out->point_flags, so.flags.data(), so.flags.size());
        }
        std::memcpy(out->contour_start, so.contour_start.data(), 4 * so.contour_start.size());
        if (out->n_contours) std::memcpy(out->contour_closed, so.closed.data(), so.closed.size());
        std::memcpy(out->path_contour_start, so.path_contour_start.data(), 4 * so.path_contour_start.size());
    });
}

int pfh_scene_release(pfh_ctx *ctx, int32_t scene_id) {
    if (!ctx) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        SceneHost &s = get_scene(ctx, scene_id);
        if (ctx->server) {
            Server &sv = *ctx->server;
            bool queued = false;
            for (const Request &r : sv.queue) queued |= r.scene == scene_id;
            if (queued || ((size_t)scene_id < sv.scene_busy.size() && sv.scene_busy[scene_id]))
                throw PfhError{PFH_E_STATE, "scene has requests in the request loop"};
        }
        HIP_TRY(hipStreamSynchronize(ctx->w->stream));
        drop_graphs_of_scene(ctx, scene_id);
        for (DevBuf *b : {&s.d_features, &s.d_pts, &s.d_segs, &s.d_seg_path, &s.d_path_rect, &s.d_path_tile_off, &s.d_path_col_off,
                          &s.d_col_path, &s.d_path_rgba, &s.d_path_rule, &s.d_path_seg_off, &s.d_bin_paths, &s.d_out_owned, &s.d_path_clip,
                          &s.d_path_level, &s.d_path_paint, &s.d_paints, &s.d_ramps})
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
            HIP_TRY(hipEventRecord(ctx->w->ev0, ctx->w->stream));
            for (int32_t i = 0; i < n; i++) rasterize_plain(ctx, scene_ids[i], flags);
            HIP_TRY(hipEventRecord(ctx->w->ev1, ctx->w->stream));
            HIP_TRY(hipEventSynchronize(ctx->w->ev1));
            ctx->stats.ms_total = elapsed_ms(ctx, ctx->w->ev0, ctx->w->ev1);
        }
    });
}

int pfh_rasterize_noalloc(pfh_ctx *ctx, const int32_t *scene_ids, int32_t n, uint32_t flags) {
    if (!ctx || (n > 0 && !scene_ids) || n < 0) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        if (flags & ~(uint32_t)PFH_LOAD_DST)
            throw PfhError{PFH_E_ARG, "pfh_rasterize_noalloc takes PFH_PLAIN, optionally with PFH_LOAD_DST"};
        HIP_TRY(hipSetDevice(ctx->device));
        for (int32_t i = 0; i < n; i++) get_scene(ctx, scene_ids[i]);
        if (n == 0) return;
        struct Restore {
            pfh_ctx *c;
            ~Restore() {
                // a launch that stopped part way leaves kernels in flight on the scratch: they end
                // before the next call lays the scratch out again
                (void)hipStreamSynchronize(c->fixed.stream);
                c->w = &c->main;
            }
        } restore{ctx};
        ctx->w = &ctx->fixed;
        ctx->stats.host_syncs = 0;
        ctx->stats.paths = 0;
        for (double &m : ctx->stats.ms_stage) m = 0;
        double total = 0;
        for (int32_t i = 0; i < n; i++) {
            ctx->stats.paths += get_scene(ctx, scene_ids[i]).n_paths;
            ctx->fixed.reset_arena();
            rasterize_plain(ctx, scene_ids[i], flags);
            total += elapsed_ms(ctx, ctx->fixed.stage_ev[0], ctx->fixed.stage_ev[8]);
        }
        ctx->stats.ms_total = total;
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
        // Captured graphs stay valid: kernels read the target through the scene table, which
        // every launch copies to the device before it starts.
        s.d_out = reinterpret_cast<uint8_t *>(device_rgba8);
        s.out_pitch = pitch_bytes;
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
        HIP_TRY(hipStreamSynchronize(ctx->w->stream));
        HIP_TRY(hipMemcpy2D(dst, (size_t)pitch_bytes, s.d_out, (size_t)s.out_pitch, (size_t)s.width * 4,
                            (size_t)s.height, hipMemcpyDeviceToHost));
    });
}

int pfh_dump_counts(pfh_ctx *ctx, int32_t scene_id, int64_t *counts) {
    if (!ctx || !counts) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        SceneHost &s = get_scene(ctx, scene_id);
        if (!s.dumped) throw PfhError{PFH_E_STATE, "scene was not rasterized with PFH_DUMP"};
        HIP_TRY(hipStreamSynchronize(ctx->w->stream));
        uint32_t f0 = 0, f1 = 0, a0 = 0, a1 = 0, e0 = 0, e1 = 0, l0 = 0, l1 = 0;
        uint32_t sb = s.seg_base, tb = s.tile_base, nfb = (uint32_t)(s.fb_w * s.fb_h);
        HIP_TRY(hipMemcpy(&l0, ctx->w->seg_cnt.as<uint32_t>() + sb, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&l1, ctx->w->seg_cnt.as<uint32_t>() + sb + s.n_segs, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&f0, ctx->w->line_cnt.as<uint32_t>() + l0, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&f1, ctx->w->line_cnt.as<uint32_t>() + l1, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&a0, ctx->w->tile_alpha_idx.as<uint32_t>() + tb, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&a1, ctx->w->tile_alpha_idx.as<uint32_t>() + tb + s.n_tiles, 4, hipMemcpyDeviceToHost));
        // a scene's entries keep their index range through the sort (keys are grouped by scene)
        HIP_TRY(hipMemcpy(&e0, ctx->w->path_ent.as<uint2>() + s.path_base, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&e1, ctx->w->path_ent.as<uint2>() + s.path_base + s.n_paths, 4, hipMemcpyDeviceToHost));
        counts[0] = s.n_tiles;
        counts[1] = s.n_cols;
        counts[2] = f1 - f0;
        counts[3] = e1 - e0;
        counts[4] = nfb;
        counts[5] = s.n_paths;
        // alpha tiles after clip resolution (a clip path's tiles hold two mask rows each)
        std::vector<uint8_t> st(s.n_tiles);
        if (s.n_tiles) HIP_TRY(hipMemcpy(st.data(), ctx->w->tile_st.as<uint8_t>() + tb, s.n_tiles, hipMemcpyDeviceToHost));
        int64_t na = 0;
        for (uint8_t v : st) na += (v == TILE_ALPHA || v == TILE_CLIP_MASK) ? 1 : 0;
        (void)a0;
        (void)a1;
        counts[6] = na;
        counts[7] = ctx->w->h_counters[C_GUARD];
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
        HIP_TRY(hipStreamSynchronize(ctx->w->stream));
        uint32_t sb = s.seg_base, tb = s.tile_base, cb = s.col_base, fbb = s.fb_base;
        uint32_t nfb = (uint32_t)(s.fb_w * s.fb_h);
        auto d2h = [](void *dst, const void *src, size_t bytes) {
            if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        };
        // chords of a path start at its first segment's chord offset; fills of a chord at
        // that chord's scanned fill offset
        std::vector<uint32_t> seg_line(s.n_segs + 1);
        d2h(seg_line.data(), ctx->w->seg_cnt.as<uint32_t>() + sb, 4 * (s.n_segs + 1));
        uint32_t l0 = seg_line[0], l1 = seg_line[s.n_segs];
        std::vector<uint32_t> line_fill(l1 - l0 + 1);
        d2h(line_fill.data(), ctx->w->line_cnt.as<uint32_t>() + l0, 4 * (size_t)(l1 - l0 + 1));
        for (uint32_t p = 0; p < s.n_paths; p++) std::memcpy(path_rect + 4 * p, &s.path_rect[p], 16);
        for (uint32_t p = 0; p <= s.n_paths; p++) {
            tile_off[p] = s.path_tile_off[p];
            col_off[p] = s.path_col_off[p];
            uint32_t first_line = seg_line[s.path_seg_off[p]] - l0;
            fill_off[p] = (int64_t)line_fill[first_line] - line_fill[0];
        }
        d2h(tile_backdrop, ctx->w->tile_bd.as<int8_t>() + tb, s.n_tiles);
        // a tile draws a mask when it has fills, or (clipped) when it took its clip tile's
        std::vector<uint8_t> tst(s.n_tiles + 1);
        d2h(tst.data(), ctx->w->tile_st.as<uint8_t>() + tb, s.n_tiles);
        for (uint32_t t = 0; t < s.n_tiles; t++) tile_alpha[t] = (tst[t] == TILE_ALPHA || tst[t] == TILE_CLIP_MASK) ? 1 : 0;
        d2h(col_sum, ctx->w->col_bd.as<int32_t>() + cb, 4 * s.n_cols);
        uint32_t nf = line_fill[l1 - l0] - line_fill[0];
        std::vector<uint4> df(nf);
        d2h(df.data(), ctx->w->dump_fills.as<uint4>() + line_fill[0], 16 * (size_t)nf);
        uint32_t p = 0;
        for (uint32_t i = 0; i < nf; i++) {
            while (p + 1 <= s.n_paths && (int64_t)i >= fill_off[p + 1]) p++;
            fill_seg[4 * i + 0] = (uint16_t)(df[i].x & 0xffff);
            fill_seg[4 * i + 1] = (uint16_t)(df[i].x >> 16);
            fill_seg[4 * i + 2] = (uint16_t)(df[i].y & 0xffff);
            fill_seg[4 * i + 3] = (uint16_t)(df[i].y >> 16);
            fill_tile[i] = df[i].z - tb - s.path_tile_off[p];
        }
        // a scene's entries keep their index range through the sort (keys are grouped by scene)
        uint32_t ebase = 0, eend = 0;
        d2h(&ebase, ctx->w->path_ent.as<uint2>() + s.path_base, 4);
        d2h(&eend, ctx->w->path_ent.as<uint2>() + s.path_base + s.n_paths, 4);
        uint32_t ne = eend - ebase;
        std::vector<uint32_t> sv(ne);
        const uint32_t *sorted = ctx->w->last_sorted_alt ? ctx->w->vals_alt.as<uint32_t>() : ctx->w->vals.as<uint32_t>();
        d2h(sv.data(), sorted + ebase, 4 * (size_t)ne);
        uint32_t elo = 0xffffffffu, ehi = 0;
        for (uint32_t v : sv) {
            elo = std::min(elo, v);
            ehi = std::max(ehi, v + 1);
        }
        if (ne) {
            std::vector<uint2> meta(ehi - elo);
            d2h(meta.data(), ctx->w->entry_meta.as<uint2>() + elo, 8 * (size_t)(ehi - elo));
            for (uint32_t i = 0; i < ne; i++) {
                entry_fbtile[i] = meta[sv[i] - elo].y;
                entry_path[i] = meta[sv[i] - elo].x;
            }
        }
        d2h(zbuf, ctx->w->zbuf.as<uint32_t>() + fbb, 4 * (size_t)nfb);
        fb_rect[0] = s.fb_x0;
        fb_rect[1] = s.fb_y0;
        fb_rect[2] = s.fb_x0 + s.fb_w;
        fb_rect[3] = s.fb_y0 + s.fb_h;
        if (masks) {
            // the mask row of each alpha tile in (path, row, column) order: its own, or for a
            // tile that took its clip tile's mask, that row (a clip path's own tiles: the first
            // of their two rows); the scene's rows are [a0, a1)
            std::vector<uint32_t> aidx(s.n_tiles + 1), cai(s.n_clip ? s.n_tiles : 0);
            d2h(aidx.data(), ctx->w->tile_alpha_idx.as<uint32_t>() + tb, 4 * (s.n_tiles + 1));
            if (s.n_clip) d2h(cai.data(), ctx->w->tile_clip_ai.as<uint32_t>() + tb, 4 * s.n_tiles);
            const uint32_t a0 = aidx[0], a1 = aidx[s.n_tiles];
            std::vector<uint8_t> rows(256ull * (a1 - a0));
            if (a1 > a0) d2h(rows.data(), ctx->w->masks.as<uint8_t>() + 256ull * a0, rows.size());
            uint64_t k = 0;
            for (uint32_t t = 0; t < s.n_tiles; t++) {
                if (!tile_alpha[t]) continue;
                const uint32_t row = (tst[t] == TILE_CLIP_MASK ? cai[t] : aidx[t]) - a0;
                if (row >= a1 - a0) throw PfhError{PFH_E_STATE, "mask row outside the scene's rows"};
                // device masks are column-major (byte 16 x + y); the dump is row-major (16 y + x)
                const uint8_t *m = rows.data() + 256ull * row;
                uint8_t *o = masks + 256 * k++;
                for (int xx = 0; xx < 16; xx++)
                    for (int yy = 0; yy < 16; yy++) o[16 * yy + xx] = m[16 * xx + yy];
            }
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
        HIP_TRY(hipStreamSynchronize(ctx->w->stream));
        std::vector<uint32_t> seg_line(s.n_segs + 1);
        HIP_TRY(hipMemcpy(seg_line.data(), ctx->w->seg_cnt.as<uint32_t>() + s.seg_base, 4 * (s.n_segs + 1),
                          hipMemcpyDeviceToHost));
        uint32_t l0 = seg_line[0], l1 = seg_line[s.n_segs];
        std::vector<uint32_t> line_fill(l1 - l0 + 1);
        HIP_TRY(hipMemcpy(line_fill.data(), ctx->w->line_cnt.as<uint32_t>() + l0, 4 * (size_t)(l1 - l0 + 1),
                          hipMemcpyDeviceToHost));
        std::vector<uint32_t> fill_off(s.n_tiles + 1);
        HIP_TRY(hipMemcpy(fill_off.data(), ctx->w->tile_fill_off.as<uint32_t>() + s.tile_base, 4 * (s.n_tiles + 1),
                          hipMemcpyDeviceToHost));
        for (uint32_t p = 0; p < s.n_paths; p++) {
            uint32_t a = seg_line[s.path_seg_off[p]] - l0, e = seg_line[s.path_seg_off[p + 1]] - l0;
            lines[p] = (int32_t)(e - a - 2 * s.path_breaks[p]);  // chords: the stream's points less two per chain
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
                            ctx->w->stream));
        HIP_TRY(hipStreamSynchronize(ctx->w->stream));
        HIP_TRY(hipMemcpy(out, d.out.ptr, 4 * n * 3, hipMemcpyDeviceToHost));
        for (DevBuf *bf : {&d.params, &d.x, &d.h, &d.out}) bf->release();
    });
}

namespace {
struct MlpDev {
    DevBuf params, m, v, grad, loss, x, t, h, out, scratch;
};
static void mlp_grad_dev(pfh_ctx *ctx, MlpDev &d, const float *x, const float *t, int64_t n) {
    HIP_TRY(mlp_forward(d.params.as<float>(), x, n, d.h.as<float>(), d.out.as<float>(), ctx->w->stream));
    HIP_TRY(mlp_backward(d.params.as<float>(), x, n, d.h.as<float>(), d.out.as<float>(), t, d.grad.as<float>(),
                         d.loss.as<float>(), d.scratch.as<float>(), ctx->w->stream));
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
        HIP_TRY(hipStreamSynchronize(ctx->w->stream));
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
                                 ++step, ctx->w->stream));
            }
            if (loss_history) {
                HIP_TRY(mlp_forward(d.params.as<float>(), d.x.as<float>(), n, d.h.as<float>(), d.out.as<float>(),
                                    ctx->w->stream));
                HIP_TRY(mlp_backward(d.params.as<float>(), d.x.as<float>(), n, d.h.as<float>(), d.out.as<float>(),
                                     d.t.as<float>(), d.grad.as<float>(), d.loss.as<float>(), d.scratch.as<float>(),
                                     ctx->w->stream));
                HIP_TRY(hipStreamSynchronize(ctx->w->stream));
                HIP_TRY(hipMemcpy(&loss_history[ep], d.loss.ptr, 4, hipMemcpyDeviceToHost));
            }
        }
        HIP_TRY(hipStreamSynchronize(ctx->w->stream));
        HIP_TRY(hipMemcpy(params, d.params.ptr, 4 * np, hipMemcpyDeviceToHost));
        for (DevBuf *bf : {&d.params, &d.m, &d.v, &d.grad, &d.loss, &d.x, &d.t, &d.h, &d.out, &d.scratch})
            bf->release();
    });
}

static bool valid_schedule(const pfh_schedule &s) {
    auto blk = [](int32_t v) { return v == 64 || v == 128 || v == 256; };
    return s.fill_grid >= 1 && s.fill_grid <= (1 << 20) && blk(s.composite_block) && blk(s.column_block) &&
           s.xcd_order >= 0 && s.xcd_order <= 3 && s.flatten_segs >= 0 && s.flatten_segs <= (1 << 16) &&
           s.flatten_segs % 64 == 0 && s.flatten_take >= 0 && s.flatten_take <= 64 && s.column_rows >= 0 &&
           s.column_rows <= 64 && s.flatten_deep >= 0 && s.flatten_deep <= 12;
}

int pfh_get_schedule(pfh_ctx *ctx, pfh_schedule *out) {
    if (!ctx || !out) return PFH_E_ARG;
    out->fill_grid = (int32_t)ctx->sched.fill_grid;
    out->composite_block = (int32_t)ctx->sched.composite_block;
    out->column_block = (int32_t)ctx->sched.column_block;
    out->xcd_order = (int32_t)ctx->sched.xcd_order;
    out->flatten_segs = (int32_t)ctx->sched.flatten_segs;
    out->flatten_take = (int32_t)ctx->sched.flatten_take;
    out->column_rows = (int32_t)ctx->sched.column_rows;
    out->flatten_deep = (int32_t)ctx->sched.flatten_deep;
    return PFH_OK;
}

int pfh_set_schedule(pfh_ctx *ctx, const pfh_schedule *in) {
    if (!ctx || !in) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        if (!valid_schedule(*in)) throw PfhError{PFH_E_ARG, "invalid schedule"};
        ctx->sched.fill_grid = (uint32_t)in->fill_grid;
        ctx->sched.composite_block = (uint32_t)in->composite_block;
        ctx->sched.column_block = (uint32_t)in->column_block;
        ctx->sched.xcd_order = (uint32_t)in->xcd_order;
        ctx->sched.flatten_segs = (uint32_t)in->flatten_segs;
        ctx->sched.flatten_take = in->flatten_take ? (uint32_t)in->flatten_take : 8u;
        ctx->sched.column_rows = (uint32_t)in->column_rows;
        ctx->sched.flatten_deep = (uint32_t)in->flatten_deep;
    });
}

// Coordinate descent over the schedule, one parameter at a time from the current schedule,
// each candidate timed as the median device time of `reps` graph replays of the batched
// launch of these scenes (capture excluded).  Keeps the fastest (a candidate has to be 1 %
// faster than the best so far); results never change.
int pfh_autotune(pfh_ctx *ctx, const int32_t *scene_ids, int32_t n, uint32_t flags, int32_t reps, pfh_schedule *out,
                 double *ms_start, double *ms_tuned) {
    if (!ctx || !scene_ids || n <= 0 || reps < 1 || reps > 100 || !out) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        HIP_TRY(hipSetDevice(ctx->device));
        for (int32_t i = 0; i < n; i++) get_scene(ctx, scene_ids[i]);
        const uint32_t f = (flags & (PFH_PREDICT | PFH_LOAD_DST)) | PFH_BATCHED;
        auto measure = [&]() {
            rasterize_batched(ctx, scene_ids, n, f);  // capture (or re-capture) and settle sizes
            std::vector<double> t;
            for (int32_t r = 0; r < reps; r++) {
                rasterize_batched(ctx, scene_ids, n, f);
                t.push_back(ctx->stats.ms_total);
            }
            std::sort(t.begin(), t.end());
            return t[t.size() / 2];
        };
        Schedule best = ctx->sched;
        double best_ms = measure();
        if (ms_start) *ms_start = best_ms;
        const std::vector<uint32_t> cand[8] = {{1024, 2048, 4096, 8192, 16384, 32768, 65536}, {64, 128, 256}, {64, 128, 256},
                                              {0, 1, 2, 3}, {0, 64, 128, 256, 512, 1024, 2048}, {1, 4, 8, 16, 32},
                                              {0, 8, 16, 32, 64}, {0, 3, 4, 5, 6}};
        for (int param = 0; param < 8; param++) {
            for (uint32_t v : cand[param]) {
                Schedule trial = best;
                uint32_t *field = param == 0 ? &trial.fill_grid
                                  : param == 1 ? &trial.composite_block
                                  : param == 2 ? &trial.column_block
                                  : param == 3 ? &trial.xcd_order
                                  : param == 4 ? &trial.flatten_segs
                                  : param == 5 ? &trial.flatten_take
                                  : param == 6 ? &trial.column_rows
                                               : &trial.flatten_deep;
                if (*field == v) continue;
                *field = v;
                ctx->sched = trial;
                const double ms = measure();
                // a candidate must beat the best by more than replay-to-replay noise (~0.5 %)
                if (ms < best_ms * 0.99) {
                    best_ms = ms;
                    best = trial;
                }
            }
        }
        ctx->sched = best;
        out->fill_grid = (int32_t)best.fill_grid;
        out->composite_block = (int32_t)best.composite_block;
        out->column_block = (int32_t)best.column_block;
        out->xcd_order = (int32_t)best.xcd_order;
        out->flatten_segs = (int32_t)best.flatten_segs;
        out->flatten_take = (int32_t)best.flatten_take;
        out->column_rows = (int32_t)best.column_rows;
        out->flatten_deep = (int32_t)best.flatten_deep;
        if (ms_tuned) *ms_tuned = best_ms;
    });
}

int pfh_server_start(pfh_ctx *ctx, int32_t slots, int64_t memory_budget_bytes, int32_t max_batch, uint32_t flags) {
    if (!ctx) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        if (ctx->server) throw PfhError{PFH_E_STATE, "request loop already started"};
        if (slots < 1 || slots > 3) throw PfhError{PFH_E_ARG, "slots must be 1..3 (one HIP stream each)"};
        if (memory_budget_bytes <= 0 || max_batch < 1) throw PfhError{PFH_E_ARG, "budget and max_batch must be positive"};
        if (flags & (PFH_DUMP | PFH_LOAD_DST)) throw PfhError{PFH_E_ARG, "PFH_DUMP / PFH_LOAD_DST are not served"};
        HIP_TRY(hipSetDevice(ctx->device));
        auto sv = std::make_unique<Server>();
        sv->budget = (uint64_t)memory_budget_bytes;
        sv->max_batch = max_batch;
        sv->flags = flags & (PFH_PREDICT | PFH_NO_GRAPH);
        for (int32_t i = 0; i < slots; i++) {
            sv->slots.push_back(std::make_unique<Slot>());
            sv->slots.back()->ws.create();
        }
        ctx->server = std::move(sv);
    });
}

int pfh_server_submit(pfh_ctx *ctx, int32_t scene_id, int64_t *request_id) {
    if (!ctx || !request_id) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        Server &sv = server_of(ctx);
        SceneHost &s = get_scene(ctx, scene_id);
        if (sv.scene_busy.size() < ctx->scenes.size()) sv.scene_busy.resize(ctx->scenes.size(), 0);
        Request r;
        r.id = sv.next_id++;
        r.scene = scene_id;
        r.bytes = scene_launch_bytes(ctx, s);
        r.t_submit = now_s();
        sv.queue.push_back(r);
        sv.submitted++;
        *request_id = r.id;
    });
}

int pfh_server_poll(pfh_ctx *ctx, double wait_ms, int64_t *done_ids, int32_t capacity, int32_t *n_done) {
    if (!ctx || !n_done || capacity < 0 || (capacity > 0 && !done_ids)) return PFH_E_ARG;
    *n_done = 0;
    return run_guarded(ctx, [&] {
        Server &sv = server_of(ctx);
        HIP_TRY(hipSetDevice(ctx->device));
        const double until = now_s() + std::max(0.0, wait_ms) * 1e-3;
        while (true) {
            server_pump(ctx);
            if (!sv.done.empty() || server_idle(sv) || now_s() >= until) break;
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        int32_t k = 0;
        while (k < capacity && !sv.done.empty()) {
            done_ids[k++] = sv.done.front();
            sv.done.pop_front();
        }
        *n_done = k;
    });
}

int pfh_server_stats_get(pfh_ctx *ctx, pfh_server_stats *out) {
    if (!ctx || !out) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        Server &sv = server_of(ctx);
        pfh_server_stats st{};
        st.submitted = sv.submitted;
        st.completed = sv.completed;
        st.queued = (int64_t)sv.queue.size();
        for (auto &sl : sv.slots) st.in_flight += (int64_t)sl->reqs.size();
        st.launches = sv.launches;
        st.retries = sv.retries;
        st.max_concurrent_launches = sv.max_concurrent;
        st.mean_batch = sv.launches ? (double)sv.launched_requests / (double)sv.launches : 0.0;
        st.latency_ms_p50 = percentile(sv.latency_ms, 0.50);
        st.latency_ms_p99 = percentile(sv.latency_ms, 0.99);
        double sum = 0, mx = 0;
        for (double v : sv.latency_ms) {
            sum += v;
            mx = std::max(mx, v);
        }
        st.latency_ms_mean = sv.latency_ms.empty() ? 0.0 : sum / (double)sv.latency_ms.size();
        st.latency_ms_max = mx;
        st.reserved_bytes = (int64_t)sv.reserved;
        st.peak_reserved_bytes = (int64_t)sv.peak_reserved;
        st.budget_bytes = (int64_t)sv.budget;
        for (auto &sl : sv.slots) st.device_bytes += (int64_t)sl->ws.device_bytes();
        *out = st;
    });
}

int pfh_server_stop(pfh_ctx *ctx) {
    if (!ctx) return PFH_E_ARG;
    return run_guarded(ctx, [&] {
        if (!ctx->server) return;
        HIP_TRY(hipSetDevice(ctx->device));
        while (!server_idle(*ctx->server)) {
            server_pump(ctx);
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        server_destroy(ctx);
    });
}

}  // extern "C"
