// ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this library, and only as the checker.
// The HIP product never links or calls it.
//
// A CPU restatement of the reference's hybrid-mode (D3D9-level) tiling path,
// written from the reference sources, plus an FP64 coverage and composite
// model of the same fills.  Every function cites the reference file:line it
// follows (paths relative to the reference's pathfinder/ directory).
//
// Float semantics follow the reference's Linux/x86 build (SSE, -msse4.1,
// PATHFINDER_ENABLE_SIMD): plain IEEE single-precision mul/add/div with no FMA
// contraction (build with -ffp-contract=off), F32x4::round = round-half-even
// (common/f32x4.h:64-66), _mm_min/max_ps operand order for clamp.
//
// Pinning (see tests/): the tiler restatement is checked bit-for-bit against
// fixtures written by the reference's own compiled Tiler (oracle/_ref); the
// FP64 coverage model against closed forms (rectangles, half-pixel edges,
// triangle areas, polygon area sums); the composite against closed forms
// (single opaque layer, n identical layers -> 1-(1-a)^n).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

struct V2 {
    float x, y;
};

static inline V2 add(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
static inline V2 sub(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
static inline V2 muls(V2 a, float s) { return {a.x * s, a.y * s}; }

// _mm_max_ps(a, b) / _mm_min_ps(a, b): the second operand wins on NaN or equality.
static inline float sse_max(float a, float b) { return a > b ? a : b; }
static inline float sse_min(float a, float b) { return a < b ? a : b; }

enum Kind { K_LINE = 1, K_QUAD = 2, K_CUBIC = 3 };

struct Seg {
    int kind;
    V2 p0, c0, c1, p1; // baseline (p0 -> p1), ctrl (c0, c1); quad uses c0 only
};

struct RectFo {
    float l = 1, t = 1, r = 0, b = 0; // default-invalid, common/math/rect.h:132-135
    bool valid() const { return l <= r && t <= b; }
};

struct OFill {
    uint16_t x0, y0, x1, y1;
    uint32_t tile; // tile-local index inside the path's tile rect
};

struct PathTiling {
    int32_t rect[4] = {0, 0, 0, 0}; // l, t, r, b in tiles
    std::vector<int8_t> backdrop;   // per tile, after prepare_tiles
    std::vector<uint8_t> alpha;     // per tile: has an alpha (mask) tile
    std::vector<int8_t> delta;      // per tile: local winding change before propagation
    std::vector<int32_t> col;       // per column: running sum (above-rect part during tiling, final after)
    std::vector<int32_t> col_above; // per column: above-rect sum before propagation
    std::vector<OFill> fills;       // emission order
    uint32_t lines = 0;             // chords handed to process_line_segment
    uint32_t guard_hits = 0;        // loop guards tripped (reference would not terminate)
    // Clip resolution of a draw path with a clip path (tiler.cpp:385-431), per tile:
    std::vector<uint8_t> combine;   // 1: both tiles have masks; the clip combine job applies comb_bd
    std::vector<int8_t> comb_bd;    // the draw tile's own backdrop, handed to the combine job
    std::vector<int64_t> clip_src;  // global tile index of the clip tile whose mask is used, or -1
};

struct Scene {
    int n_paths;
    const int32_t *path_contour_start;
    const uint8_t *path_fill_rule;
    const int32_t *contour_start;
    const uint8_t *contour_closed;
    const float *points;
    const uint8_t *point_flags;
    RectFo view_box;
};

// core/scene.cpp:170-179 — the view box's right edge is rounded up to a multiple of 16;
// common/math/basic.h:93-108 round_up.
static RectFo scene_view_box(const float *vb) {
    RectFo r{vb[0], vb[1], vb[2], vb[3]};
    int number = (int)std::ceil(vb[2]);
    int multiple = 16;
    int remainder = std::abs(number) % multiple;
    int rounded = number;
    if (remainder != 0) rounded = number < 0 ? -(std::abs(number) - remainder) : number + multiple - remainder;
    r.r = (float)rounded;
    return r;
}

// core/data/contour.cpp:110-173 SegmentsIter::get_next(force_closed = true), driven by
// core/d3d9/tiler.cpp:345-367 generate_fills (break on SegmentKind::None).
static void contour_segments(const V2 *pts, const uint8_t *flags, int n, bool closed, std::vector<Seg> &out) {
    if (n == 0) return; // tiler.cpp:349-351
    int head = 0;
    bool has_next = true;
    while (has_next) {
        Seg s{};
        s.kind = 0;
        if (head >= n || flags[head] != 0) break;
        s.p0 = pts[head];
        if (head + 1 < n) {
            if (flags[head + 1] == 0) {
                s.p1 = pts[head + 1];
                s.kind = K_LINE;
                head += 1;
            } else if (head + 2 < n) {
                if (flags[head + 1] == 1 && flags[head + 2] == 0) {
                    s.c0 = pts[head + 1];
                    s.p1 = pts[head + 2];
                    s.kind = K_QUAD;
                    head += 2;
                } else if (head + 3 < n) {
                    if (flags[head + 1] == 1 && flags[head + 2] == 2 && flags[head + 3] == 0) {
                        s.c0 = pts[head + 1];
                        s.c1 = pts[head + 2];
                        s.p1 = pts[head + 3];
                        s.kind = K_CUBIC;
                        head += 3;
                    }
                }
            }
        } else {
            // contour.cpp:155-170: close unless first and last are within FLOAT_EPSILON (1e-4).
            if (n > 1) {
                V2 d = sub(pts[0], pts[n - 1]);
                float len = std::sqrt(d.x * d.x + d.y * d.y);
                if (!(len <= 0.0001f)) {
                    s.p0 = pts[n - 1];
                    s.p1 = pts[0];
                    s.kind = K_LINE;
                }
            }
            has_next = false;
        }
        if (s.kind == 0) break;
        out.push_back(s);
    }
}

struct Builder {
    RectFo view_box;       // scene view box (rounded right edge)
    int32_t tl, tt, tr, tb; // tile rect
    PathTiling *pt;

    int width() const { return tr - tl; }
    int height() const { return tb - tt; }

    // core/d3d9/object_builder.cpp:19-64 ObjectBuilder::add_fill
    void add_fill(V2 from, V2 to, int tx, int ty) {
        // rect.h:313-316 contains_point(Vec2I): max edge excluded.
        if (!(tl <= tx && tx <= tr - 1 && tt <= ty && ty <= tb - 1)) return;
        float ox = (float)tx * 16.0f, oy = (float)ty * 16.0f;
        float v[4] = {(from.x - ox) * 256.0f, (from.y - oy) * 256.0f, (to.x - ox) * 256.0f, (to.y - oy) * 256.0f};
        uint16_t q[4];
        for (int i = 0; i < 4; i++) {
            float c = sse_min(sse_max(v[i], 0.0f), 4095.0f); // F32x4::clamp = max(min).min(max)
            c = std::nearbyint(c);                            // _mm_round_ps NEAREST_INT: half to even
            q[i] = (uint16_t)c;
        }
        if (q[0] == q[2]) return; // object_builder.cpp:46-48
        int local = (tx - tl) + width() * (ty - tt);
        pt->alpha[local] = 1; // get_or_allocate_alpha_tile_index (:72-94)
        pt->fills.push_back(OFill{q[0], q[1], q[2], q[3], (uint32_t)local});
    }

    // core/d3d9/object_builder.cpp:96-114 adjust_alpha_tile_backdrop (int8 tile delta, int32 column)
    void adjust_backdrop(int tx, int ty, int8_t delta) {
        int ox = tx - tl, oy = ty - tt;
        if (ox < 0 || ox >= width() || oy >= height()) return;
        if (oy < 0) {
            pt->col[ox] += delta;
            return;
        }
        int local = oy * width() + ox;
        pt->delta[local] = (int8_t)(pt->delta[local] + delta);
    }

    // core/d3d9/tiler.cpp:46-62 compute_outcode
    static uint8_t outcode(V2 p, const RectFo &r) {
        uint8_t f = 0;
        if (p.x < r.l) f |= 1;
        else if (p.x > r.r) f |= 2;
        if (p.y < r.t) f |= 4;
        else if (p.y > r.b) f |= 8;
        return f;
    }

    // common/math/basic.h:45-53 clamp + lerp (t clamped to [0, 1])
    static float lerp(float a, float b, float t) {
        t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
        return a + (b - a) * t;
    }

    // core/d3d9/tiler.cpp:65-125 clip_line_segment_to_rect (Cohen–Sutherland)
    bool clip(V2 &from, V2 &to, const RectFo &r) {
        uint8_t of = outcode(from, r), ot = outcode(to, r);
        for (int iter = 0; iter < 64; iter++) {
            if (of == 0 && ot == 0) return true;
            if ((of & ot) != 0) return false;
            bool clip_from = of > ot;
            V2 pnt = clip_from ? from : to;
            uint8_t oc = clip_from ? of : ot;
            if (oc & 1) pnt = V2{r.l, lerp(from.y, to.y, (r.l - from.x) / (to.x - from.x))};
            else if (oc & 2) pnt = V2{r.r, lerp(from.y, to.y, (r.r - from.x) / (to.x - from.x))};
            else if (oc & 4) pnt = V2{lerp(from.x, to.x, (r.t - from.y) / (to.y - from.y)), r.t};
            else if (oc & 8) pnt = V2{lerp(from.x, to.x, (r.b - from.y) / (to.y - from.y)), r.b};
            if (clip_from) {
                from = pnt;
                of = outcode(pnt, r);
            } else {
                to = pnt;
                ot = outcode(pnt, r);
            }
        }
        pt->guard_hits++;
        return false;
    }

    // core/d3d9/tiler.cpp:131-279 process_line_segment (Amanatides–Woo tile walk)
    void process_line(V2 from, V2 to) {
        if (!std::isfinite(from.x) || !std::isfinite(from.y) || !std::isfinite(to.x) || !std::isfinite(to.y)) return;
        RectFo cb = view_box;
        cb.t = -std::numeric_limits<float>::infinity(); // :144
        if (!clip(from, to, cb)) return;

        const float inv = 1.0f / 16.0f;
        int fx = (int32_t)std::floor(from.x * inv), fy = (int32_t)std::floor(from.y * inv);
        int tx = (int32_t)std::floor(to.x * inv), ty = (int32_t)std::floor(to.y * inv);
        V2 vec = sub(to, from);
        int stx = vec.x < 0 ? -1 : 1, sty = vec.y < 0 ? -1 : 1;
        V2 cross{((float)fx + (vec.x >= 0 ? 1.0f : 0.0f)) * 16.0f, ((float)fy + (vec.y >= 0 ? 1.0f : 0.0f)) * 16.0f};
        V2 tmax{(cross.x - from.x) / vec.x, (cross.y - from.y) / vec.y};
        V2 tdelta{std::fabs(16.0f / vec.x), std::fabs(16.0f / vec.y)};
        V2 cur = from;
        int cx = fx, cy = fy;
        int last = 0; // 0 none, 1 X, 2 Y
        int64_t guard = (int64_t)std::abs((int64_t)tx - fx) + std::abs((int64_t)ty - fy) + 2;
        for (int64_t it = 0;; it++) {
            if (it > guard) {
                pt->guard_hits++;
                break;
            }
            int next;
            if (tmax.x < tmax.y) next = 1;
            else if (tmax.x > tmax.y) next = 2;
            else next = stx > 0 ? 1 : 2;
            float nt = next == 1 ? tmax.x : tmax.y;
            nt = (1.0f < nt) ? 1.0f : nt; // std::min(next_t, 1.0f)
            if (cx == tx && cy == ty) next = 0;
            V2 npos = add(from, muls(vec, nt)); // LineSegmentF::sample
            add_fill(cur, npos, cx, cy);
            V2 corner{(float)cx * 16.0f, (float)cy * 16.0f};
            if (sty < 0 && next == 2) add_fill(npos, corner, cx, cy);
            else if (sty > 0 && last == 2) add_fill(corner, cur, cx, cy);
            if (stx < 0 && last == 1) adjust_backdrop(cx, cy, 1);
            else if (stx > 0 && next == 1) adjust_backdrop(cx, cy, -1);
            if (next == 1) {
                tmax.x += tdelta.x;
                cx += stx;
            } else if (next == 2) {
                tmax.y += tdelta.y;
                cy += sty;
            } else {
                break;
            }
            cur = npos;
            last = next;
        }
    }

    // core/data/segment.cpp:12-20 is_flat_cubic
    static bool flat_cubic(const Seg &s, float tol) {
        float u0 = 3.0f * s.c0.x - s.p0.x - s.p0.x - s.p1.x;
        float u1 = 3.0f * s.c0.y - s.p0.y - s.p0.y - s.p1.y;
        float u2 = 3.0f * s.c1.x - s.p1.x - s.p1.x - s.p0.x;
        float u3 = 3.0f * s.c1.y - s.p1.y - s.p1.y - s.p0.y;
        u0 = u0 * u0;
        u1 = u1 * u1;
        u2 = u2 * u2;
        u3 = u3 * u3;
        float a = sse_max(u0, u2), b = sse_max(u1, u3);
        return a + b <= tol;
    }

    // core/data/segment.cpp:22-36 is_flat_quadratic
    static bool flat_quad(const Seg &s, float tol) {
        V2 mid = muls(add(s.p0, s.p1), 0.5f);
        V2 dev = sub(s.c0, mid);
        return dev.x * dev.x + dev.y * dev.y <= tol * 0.25f;
    }

    static inline V2 lerp2(V2 a, V2 b, float t) { return add(a, muls(sub(b, a), t)); }

    // core/data/segment.cpp:42-100 split_cubic at t = 0.5 (F32x4 lane algebra expanded)
    static void split_cubic(const Seg &s, Seg &a, Seg &b) {
        const float t = 0.5f;
        V2 p01 = lerp2(s.p0, s.c0, t), p12 = lerp2(s.c0, s.c1, t), p23 = lerp2(s.c1, s.p1, t);
        V2 p012 = lerp2(p01, p12, t), p123 = lerp2(p12, p23, t);
        V2 p0123 = lerp2(p012, p123, t);
        a = Seg{K_CUBIC, s.p0, p01, p012, p0123};
        b = Seg{K_CUBIC, p0123, p123, p23, s.p1};
    }

    // core/data/segment.cpp:102-126 split_quadratic at t = 0.5
    static void split_quad(const Seg &s, Seg &a, Seg &b) {
        const float t = 0.5f;
        V2 qa = lerp2(s.p0, s.c0, t), qb = lerp2(s.c0, s.p1, t), qc = lerp2(qa, qb, t);
        a = Seg{K_QUAD, s.p0, qa, qa, qc};
        b = Seg{K_QUAD, qc, qb, qb, s.p1};
    }

    // core/d3d9/tiler.cpp:284-315 process_segment (depth-first, first half first)
    void process_segment(const Seg &s, int depth) {
        const float tol = 1.0f; // FLATTENING_TOLERANCE, tiler.cpp:15
        if (s.kind == K_LINE) {
            pt->lines++;
            process_line(s.p0, s.p1);
            return;
        }
        bool flat = s.kind == K_QUAD ? flat_quad(s, tol) : flat_cubic(s, tol);
        // Depth guard: a finite curve with coordinates below ~2^30 px is flat long before
        // depth 16 (the flatness measure shrinks 16x per halving).
        if (flat || depth >= 16) {
            if (!flat) pt->guard_hits++;
            pt->lines++;
            process_line(s.p0, s.p1);
            return;
        }
        Seg a, b;
        if (s.kind == K_QUAD) split_quad(s, a, b);
        else split_cubic(s, a, b);
        process_segment(a, depth + 1);
        process_segment(b, depth + 1);
    }

    // core/d3d9/tiler.cpp:369-439 prepare_tiles (no clip path)
    void prepare_tiles() {
        int w = width(), h = height();
        pt->col_above = pt->col;
        for (int i = 0; i < w * h; i++) {
            int column = i % w;
            int8_t d = pt->delta[i];
            pt->backdrop[i] = (int8_t)pt->col[column];
            pt->col[column] += d;
        }
    }
};

// Outline bounds: union of contour bounds, each the min/max of its points
// (core/data/contour.cpp:73-81 push_point + rect.h:319-325 union_rect, path.cpp:22-38).
static RectFo path_bounds(const Scene &sc, int p) {
    RectFo b;
    for (int c = sc.path_contour_start[p]; c < sc.path_contour_start[p + 1]; c++) {
        for (int i = sc.contour_start[c]; i < sc.contour_start[c + 1]; i++) {
            float x = sc.points[2 * i], y = sc.points[2 * i + 1];
            if (!b.valid()) {
                b = RectFo{x, y, x, y};
            } else {
                b.l = std::min(b.l, x);
                b.t = std::min(b.t, y);
                b.r = std::max(b.r, x);
                b.b = std::max(b.b, y);
            }
        }
    }
    return b;
}

// rect.h:279-291 intersects / intersection
static RectFo intersect(const RectFo &a, const RectFo &o) {
    if (a.l > o.r || a.r < o.l || a.t > o.b || a.b < o.t) return RectFo{};
    return RectFo{std::max(a.l, o.l), std::max(a.t, o.t), std::min(a.r, o.r), std::min(a.b, o.b)};
}

// core/d3d9/tiler.cpp:317-343 Tiler ctor + generate_tiles, core/data/built_path.cpp:8-34,
// core/data/data.h:53-55 round_rect_out_to_tile_bounds.
// destructive: a draw path of a destructive blend mode tiles the whole view box (built_path.cpp:21).
static void tile_path(const Scene &sc, int p, PathTiling &pt, bool destructive = false) {
    RectFo b = destructive ? sc.view_box : intersect(path_bounds(sc, p), sc.view_box);
    float s = 1.0f / 16.0f;
    pt.rect[0] = (int32_t)std::floor(b.l * s);
    pt.rect[1] = (int32_t)std::floor(b.t * s);
    pt.rect[2] = (int32_t)std::ceil(b.r * s);
    pt.rect[3] = (int32_t)std::ceil(b.b * s);
    int w = pt.rect[2] - pt.rect[0], h = pt.rect[3] - pt.rect[1];
    if (w < 0) w = 0;
    if (h < 0) h = 0;
    pt.backdrop.assign((size_t)w * h, 0);
    pt.alpha.assign((size_t)w * h, 0);
    pt.delta.assign((size_t)w * h, 0);
    pt.col.assign(w, 0);
    Builder bld;
    bld.view_box = sc.view_box;
    bld.tl = pt.rect[0];
    bld.tt = pt.rect[1];
    bld.tr = pt.rect[2];
    bld.tb = pt.rect[3];
    bld.pt = &pt;
    std::vector<Seg> segs;
    // With no tile columns nothing can land: add_fill culls every tile and
    // adjust_alpha_tile_backdrop ignores x outside [0, width).
    int first_contour = w > 0 ? sc.path_contour_start[p] : sc.path_contour_start[p + 1];
    for (int c = first_contour; c < sc.path_contour_start[p + 1]; c++) {
        int c0 = sc.contour_start[c], c1 = sc.contour_start[c + 1];
        segs.clear();
        contour_segments((const V2 *)(sc.points + 2 * c0), sc.point_flags + c0, c1 - c0, sc.contour_closed[c] != 0,
                         segs);
        for (const Seg &sg : segs) {
            // Any non-finite point makes every flattened piece non-finite, and those are
            // dropped by process_line_segment (tiler.cpp:133-136).
            bool finite = std::isfinite(sg.p0.x) && std::isfinite(sg.p0.y) && std::isfinite(sg.p1.x) &&
                          std::isfinite(sg.p1.y);
            if (sg.kind != K_LINE) finite = finite && std::isfinite(sg.c0.x) && std::isfinite(sg.c0.y);
            if (sg.kind == K_CUBIC) finite = finite && std::isfinite(sg.c1.x) && std::isfinite(sg.c1.y);
            if (finite) bld.process_segment(sg, 0);
        }
    }
    bld.prepare_tiles();
}

// core/d3d9/tiler.cpp:385-431 (the clip part of prepare_tiles), applied after the draw path's
// own propagation, which it does not change: the column sums add the tile deltas read before it.
// The clip tile at the same tile coordinates decides: both masks -> a combine job (the draw tile's
// backdrop moves into the job, the tile keeps its mask); a solid draw tile over a clip mask -> the
// tile takes the clip's mask and backdrop; a blank clip tile, or none, -> the tile is culled.
// A clip path with a clip of its own (nesting, undefined in the hybrid tiler) is resolved against
// it the same way before anything it clips is (the level structure of core/d3d11/scene_builder.cpp:
// 92-133 and shaders/d3d11/propagate.comp:133-176), so `cp` always holds resolved tiles.
// "Solid" is backdrop != 0 in the reference (own_rules 0).  With own_rules (ClipPath::fill_rule
// honoured, which the reference never reads) a tile without a mask is solid where its path's own
// rule covers: an even-odd path's tile with an even backdrop is not (pt_eo / cp_eo: the rules).
static bool covers(int bd, int even_odd, int own_rules) { return bd != 0 && !(own_rules && even_odd && bd % 2 == 0); }

static void resolve_clip(PathTiling &pt, const PathTiling &cp, int64_t clip_tile_base, int pt_eo, int cp_eo,
                         int own_rules) {
    int w = std::max(pt.rect[2] - pt.rect[0], 0), h = std::max(pt.rect[3] - pt.rect[1], 0);
    int cw = std::max(cp.rect[2] - cp.rect[0], 0), ch = std::max(cp.rect[3] - cp.rect[1], 0);
    pt.combine.assign((size_t)w * h, 0);
    pt.comb_bd.assign((size_t)w * h, 0);
    pt.clip_src.assign((size_t)w * h, -1);
    for (int i = 0; i < w * h; i++) {
        int cx = pt.rect[0] + i % w - cp.rect[0], cy = pt.rect[1] + i / w - cp.rect[1];
        if (cx < 0 || cy < 0 || cx >= cw || cy >= ch) {  // outside the clip path's rect
            pt.alpha[i] = 0;
            pt.backdrop[i] = 0;
            continue;
        }
        int ci = cy * cw + cx;
        bool clip_alpha = cp.alpha[ci] != 0, draw_alpha = pt.alpha[i] != 0;
        if (clip_alpha && draw_alpha) {
            pt.combine[i] = 1;
            pt.comb_bd[i] = pt.backdrop[i];
            pt.clip_src[i] = clip_tile_base + ci;
            pt.backdrop[i] = 0;
        } else if (clip_alpha && covers(pt.backdrop[i], pt_eo, own_rules)) {
            pt.alpha[i] = 1;
            pt.backdrop[i] = cp.backdrop[ci];
            pt.clip_src[i] = clip_tile_base + ci;
        } else if (!clip_alpha && !covers(cp.backdrop[ci], cp_eo, own_rules)) {
            pt.alpha[i] = 0;
            pt.backdrop[i] = 0;
        }
    }
}

struct Result {
    int n_draw = 0;
    std::vector<PathTiling> paths;
    int32_t fb_rect[4];
    std::vector<uint32_t> zbuf;
    std::vector<uint32_t> entry_fbtile, entry_path;
    std::vector<int64_t> entry_tile; // global canonical tile index
    double tiling_s = 0;
};

// GLSL mod / fill rules: shaders/d3d11/fill.comp:134-140, shaders/d3d9/tile.frag:580-585.
static double apply_rule(double c, int even_odd) {
    if (!even_odd) return std::min(std::fabs(c), 1.0);
    double m = c - 2.0 * std::floor(c / 2.0);
    double v = 1.0 - std::fabs(1.0 - m);
    return std::min(std::max(v, 0.0), 1.0);
}

// Exact average of clamp(u, 0, 1) for u linear between ua and ub.
static double avg_clamp01(double ua, double ub) {
    double a = std::min(ua, ub), b = std::max(ua, ub);
    if (b - a <= 0.0) return std::min(std::max(a, 0.0), 1.0);
    double full = b > 1.0 ? b - std::max(a, 1.0) : 0.0;
    double lo = std::max(a, 0.0), hi = std::min(b, 1.0);
    double mid = hi > lo ? (hi - lo) * (hi + lo) * 0.5 : 0.0;
    return (full + mid) / (b - a);
}

} // namespace

extern "C" {

// FP64 signed area of one fill (tile-local 8.8 fixed point) over pixel (px, py) of its tile:
// sign(from.x - to.x) * ∫_{[px,px+1] ∩ [xmin,xmax]} clamp(py + 1 - y(x), 0, 1) dx — the quantity
// the area LUT approximates in shaders/d3d9/fill.frag:25-45 and shaders/d3d11/fill.comp:51-71.
double pfo_fill_area(const uint16_t *seg, int px, int py) {
    double x0 = seg[0] / 256.0, y0 = seg[1] / 256.0, x1 = seg[2] / 256.0, y1 = seg[3] / 256.0;
    if (x0 == x1) return 0.0;
    double sign = x0 > x1 ? 1.0 : -1.0;
    double xl = std::min(x0, x1), xr = std::max(x0, x1);
    double yl = x0 < x1 ? y0 : y1, yr = x0 < x1 ? y1 : y0;
    double lo = std::max((double)px, xl), hi = std::min((double)px + 1.0, xr);
    if (hi <= lo) return 0.0;
    double slope = (yr - yl) / (xr - xl);
    double ya = yl + (lo - xl) * slope, yb = yl + (hi - xl) * slope;
    double ua = py + 1.0 - ya, ub = py + 1.0 - yb;
    return sign * (hi - lo) * avg_clamp01(ua, ub);
}

// FP64 coverage of alpha tiles.  fills grouped by alpha tile: tile a owns
// fill_seg[fill_off[a] .. fill_off[a+1]).  out: n_alpha * 256 doubles, row-major
// 16x16, after the fill rule.
void pfo_coverage(int64_t n_alpha, const int64_t *fill_off, const uint16_t *fill_seg, const int8_t *backdrop,
                  const uint8_t *even_odd, double *out) {
    for (int64_t a = 0; a < n_alpha; a++) {
        for (int py = 0; py < 16; py++) {
            for (int px = 0; px < 16; px++) {
                double c = backdrop[a];
                for (int64_t f = fill_off[a]; f < fill_off[a + 1]; f++) c += pfo_fill_area(fill_seg + 4 * f, px, py);
                out[a * 256 + py * 16 + px] = apply_rule(c, even_odd[a]);
            }
        }
    }
}

// Raw (pre-rule) FP64 coverage sums, for closed-form tests of the area model.
void pfo_coverage_raw(int64_t n_alpha, const int64_t *fill_off, const uint16_t *fill_seg, const int8_t *backdrop,
                      double *out) {
    for (int64_t a = 0; a < n_alpha; a++)
        for (int py = 0; py < 16; py++)
            for (int px = 0; px < 16; px++) {
                double c = backdrop[a];
                for (int64_t f = fill_off[a]; f < fill_off[a + 1]; f++) c += pfo_fill_area(fill_seg + 4 * f, px, py);
                out[a * 256 + py * 16 + px] = c;
            }
}

double pfo_apply_rule(double c, int even_odd) {
    return apply_rule(c, even_odd);
}

// Tiles a whole scene (all draw paths).  zmode 0 = reference z-buffer rule
// (core/d3d9/scene_builder.cpp:57-79: any opaque solid tile with backdrop != 0
// occludes, even an even-odd tile with even backdrop); zmode 1 = the fixed rule
// (such even-odd tiles cover nothing and never occlude, shaders/d3d11/propagate.comp:190-198).
//
// Clip paths (Scene::clip_paths, core/scene.h) are the last n_clip paths; path_clip[p] (NULL: none)
// names the clip path of draw path p by its index among them, or -1.  Clip paths are tiled like draw
// paths (core/d3d9/scene_builder.cpp:200-225) but draw nothing themselves.  clip_rule 0 resolves
// clips as the reference does; 1 under each path's own fill rule (resolve_clip's own_rules).
// path_blend (NULL: every path src-over) holds BlendMode values (core/paint/effects.h:56-123): a
// destructive mode (clear, copy, src-in, dest-in, src-out, dest-atop: effects.h:234-243) tiles the
// view box, and only src-over and clear occlude (blend_mode_occludes_backdrop, effects.h:127-160;
// core/data/built_path.cpp:40).
static bool blend_destructive(int m) { return m == 0 || m == 1 || m == 2 || m == 6 || m == 3 || m == 9; }
static bool blend_occludes(int m) { return m == 4 || m == 0; }

void *pfo_run(int n_paths, const int32_t *path_contour_start, const uint8_t *path_fill_rule,
              const uint8_t *path_opaque, const int32_t *contour_start, const uint8_t *contour_closed,
              const float *points, const uint8_t *point_flags, const float *view_box, int zmode, int n_clip,
              const int32_t *path_clip, int clip_rule, const uint8_t *path_blend) {
    auto *res = new Result();
    Scene sc{n_paths, path_contour_start, path_fill_rule, contour_start, contour_closed, points, point_flags,
             scene_view_box(view_box)};
    res->paths.resize(n_paths);
    const int n_draw = n_paths - n_clip;
    res->n_draw = n_draw;
    // Nesting level of each clip path: 0 without a clip of its own, else its clip's level + 1.
    // A chain longer than 4 clip paths, or a cycle, is refused (no result).
    std::vector<int> level(n_clip, 0);
    for (int q = 0; q < n_clip && path_clip; q++) {
        int k = 0;
        for (int c = path_clip[n_draw + q]; c >= 0; c = path_clip[n_draw + c]) {
            if (++k > 3 || c >= n_clip) {
                delete res;
                return nullptr;
            }
        }
        level[q] = k;
    }
    auto t0 = std::chrono::steady_clock::now();
    for (int p = 0; p < n_paths; p++)
        tile_path(sc, p, res->paths[p], path_blend && p < n_draw && blend_destructive(path_blend[p]));
    std::vector<int64_t> tile_base(n_paths + 1, 0);
    for (int p = 0; p < n_paths; p++) tile_base[p + 1] = tile_base[p] + (int64_t)res->paths[p].backdrop.size();
    for (int lev = 1; lev <= 3 && path_clip; lev++) {
        for (int q = 0; q < n_clip; q++) {
            if (level[q] != lev) continue;
            int g = n_draw + path_clip[n_draw + q];
            resolve_clip(res->paths[n_draw + q], res->paths[g], tile_base[g], path_fill_rule[n_draw + q],
                         path_fill_rule[g], clip_rule);
        }
    }
    for (int p = 0; path_clip && p < n_draw; p++) {
        if (path_clip[p] < 0) continue;
        int g = n_draw + path_clip[p];
        resolve_clip(res->paths[p], res->paths[g], tile_base[g], path_fill_rule[p], path_fill_rule[g], clip_rule);
    }
    res->tiling_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    // core/d3d9/scene_builder.cpp:18-95 — one batch (solid paints only), z over the view box.
    float s = 1.0f / 16.0f;
    res->fb_rect[0] = (int32_t)std::floor(sc.view_box.l * s);
    res->fb_rect[1] = (int32_t)std::floor(sc.view_box.t * s);
    res->fb_rect[2] = (int32_t)std::ceil(sc.view_box.r * s);
    res->fb_rect[3] = (int32_t)std::ceil(sc.view_box.b * s);
    int fw = res->fb_rect[2] - res->fb_rect[0], fh = res->fb_rect[3] - res->fb_rect[1];
    res->zbuf.assign((size_t)fw * fh, 0);
    std::vector<std::vector<std::pair<uint32_t, int64_t>>> lists((size_t)fw * fh);
    for (int p = 0; p < n_draw; p++) {
        PathTiling &pt = res->paths[p];
        int w = pt.rect[2] - pt.rect[0], h = pt.rect[3] - pt.rect[1];
        if (w < 0) w = 0;
        if (h < 0) h = 0;
        for (int i = 0; i < w * h; i++) {
            int tx = pt.rect[0] + i % w, ty = pt.rect[1] + i / w;
            if (!pt.alpha[i] && pt.backdrop[i] == 0) continue;
            int64_t fb = (int64_t)(ty - res->fb_rect[1]) * fw + (tx - res->fb_rect[0]);
            lists[fb].push_back({(uint32_t)p, tile_base[p] + i});
            if (!path_opaque[p] || pt.alpha[i] || (path_blend && !blend_occludes(path_blend[p]))) continue;
            if (zmode == 1 && path_fill_rule[p] && (pt.backdrop[i] % 2) == 0) continue;
            res->zbuf[fb] = std::max(res->zbuf[fb], (uint32_t)p);
        }
    }
    for (size_t fb = 0; fb < lists.size(); fb++) {
        for (auto &e : lists[fb]) {
            if ((int32_t)e.first < (int32_t)res->zbuf[fb]) continue; // shaders/d3d9/tile.vert:114
            res->entry_fbtile.push_back((uint32_t)fb);
            res->entry_path.push_back(e.first);
            res->entry_tile.push_back(e.second);
        }
    }
    return res;
}

// counts: tiles, columns, fills, entries, fb tiles, paths, alpha tiles, guard hits
void pfo_counts(void *h, int64_t *counts) {
    auto *r = (Result *)h;
    int64_t t = 0, c = 0, f = 0, a = 0, g = 0;
    for (auto &p : r->paths) {
        t += (int64_t)p.backdrop.size();
        c += (int64_t)p.col.size();
        f += (int64_t)p.fills.size();
        for (auto v : p.alpha) a += v;
        g += p.guard_hits;
    }
    counts[0] = t;
    counts[1] = c;
    counts[2] = f;
    counts[3] = (int64_t)r->entry_path.size();
    counts[4] = (int64_t)r->zbuf.size();
    counts[5] = (int64_t)r->paths.size();
    counts[6] = a;
    counts[7] = g;
}

void pfo_path_lines(void *h, int32_t *out) {
    auto *r = (Result *)h;
    for (size_t p = 0; p < r->paths.size(); p++) out[p] = (int32_t)r->paths[p].lines;
}

double pfo_tiling_seconds(void *h) {
    return ((Result *)h)->tiling_s;
}

void pfo_fetch(void *h, int32_t *path_rect, int64_t *tile_off, int64_t *col_off, int64_t *fill_off,
               int8_t *tile_backdrop, uint8_t *tile_alpha, int32_t *col_sum, uint16_t *fill_seg, uint32_t *fill_tile,
               uint32_t *entry_fbtile, uint32_t *entry_path, uint32_t *zbuf, int32_t *fb_rect, int8_t *tile_delta,
               int32_t *col_above, int64_t *entry_tile) {
    auto *r = (Result *)h;
    int64_t t = 0, c = 0, f = 0;
    tile_off[0] = col_off[0] = fill_off[0] = 0;
    for (size_t p = 0; p < r->paths.size(); p++) {
        auto &pt = r->paths[p];
        memcpy(path_rect + 4 * p, pt.rect, 16);
        for (size_t i = 0; i < pt.backdrop.size(); i++) {
            tile_backdrop[t + i] = pt.backdrop[i];
            tile_alpha[t + i] = pt.alpha[i];
            if (tile_delta) tile_delta[t + i] = pt.delta[i];
        }
        for (size_t i = 0; i < pt.col.size(); i++) {
            col_sum[c + i] = pt.col[i];
            if (col_above) col_above[c + i] = pt.col_above[i];
        }
        for (size_t i = 0; i < pt.fills.size(); i++) {
            const OFill &o = pt.fills[i];
            uint16_t *d = fill_seg + 4 * (f + i);
            d[0] = o.x0;
            d[1] = o.y0;
            d[2] = o.x1;
            d[3] = o.y1;
            fill_tile[f + i] = o.tile;
        }
        t += (int64_t)pt.backdrop.size();
        c += (int64_t)pt.col.size();
        f += (int64_t)pt.fills.size();
        tile_off[p + 1] = t;
        col_off[p + 1] = c;
        fill_off[p + 1] = f;
    }
    size_t ne = r->entry_path.size();
    if (ne) {
        memcpy(entry_fbtile, r->entry_fbtile.data(), ne * 4);
        memcpy(entry_path, r->entry_path.data(), ne * 4);
        if (entry_tile) memcpy(entry_tile, r->entry_tile.data(), ne * 8);
    }
    if (!r->zbuf.empty()) memcpy(zbuf, r->zbuf.data(), r->zbuf.size() * 4);
    memcpy(fb_rect, r->fb_rect, 16);
}

// Per global tile: clip combine flag, the backdrop handed to the combine job, and the clip
// tile whose mask a draw tile uses (-1 where none).
void pfo_fetch_clip(void *h, uint8_t *combine, int8_t *comb_bd, int64_t *clip_src) {
    auto *r = (Result *)h;
    int64_t t = 0;
    for (auto &pt : r->paths) {
        size_t n = pt.backdrop.size();
        for (size_t i = 0; i < n; i++) {
            bool has = !pt.combine.empty();
            combine[t + i] = has ? pt.combine[i] : 0;
            comb_bd[t + i] = has ? pt.comb_bd[i] : 0;
            clip_src[t + i] = has ? pt.clip_src[i] : -1;
        }
        t += (int64_t)n;
    }
}

// out[i] = the fill rule even_odd[i] applied to c[i].
void pfo_apply_rule_array(int64_t n, const double *c, const uint8_t *even_odd, double *out) {
    for (int64_t i = 0; i < n; i++) out[i] = apply_rule(c[i], even_odd[i]);
}

void pfo_free(void *h) {
    delete (Result *)h;
}

// FP64 premultiplied src-over composite of solid colours (u8/255 exact) over a
// transparent clear, following shaders/d3d9/tile.frag:590-653 (mask alpha, then
// premultiply) and BlendState::from_over (gpu/base.h:83-93).  Entries must be in
// composite order (grouped by framebuffer tile, ascending draw order).
//   entry_coverage_index: alpha index of the entry's tile, or -1 for a solid tile
//   coverage: n_alpha * 256 doubles after the fill rule
// out: H * W * 4 doubles (premultiplied RGBA in [0, 1]).
void pfo_composite(int64_t n_entries, const uint32_t *entry_fbtile, const uint32_t *entry_path,
                   const int64_t *entry_alpha, const int8_t *entry_backdrop, const uint8_t *path_even_odd,
                   const uint32_t *path_rgba, const double *coverage, int fb_w, int fb_h, int width, int height,
                   double *out, int fb_x0, int fb_y0) {
    std::fill(out, out + (size_t)width * height * 4, 0.0);
    for (int64_t e = 0; e < n_entries; e++) {
        uint32_t fb = entry_fbtile[e];
        // the framebuffer tile's position on the target: its index in the view box's tile grid,
        // from the grid's origin (fb_x0, fb_y0)
        int tx = (int)(fb % (uint32_t)fb_w) + fb_x0, ty = (int)(fb / (uint32_t)fb_w) + fb_y0;
        uint32_t p = entry_path[e];
        uint32_t c = path_rgba[p];
        double r = (c & 0xff) / 255.0, g = ((c >> 8) & 0xff) / 255.0, b = ((c >> 16) & 0xff) / 255.0,
               a = ((c >> 24) & 0xff) / 255.0;
        double solid_cov = apply_rule((double)entry_backdrop[e], path_even_odd[p]);
        for (int py = 0; py < 16; py++) {
            int y = ty * 16 + py;
            if (y < 0) continue;
            if (y >= height) break;
            for (int px = 0; px < 16; px++) {
                int x = tx * 16 + px;
                if (x < 0) continue;
                if (x >= width) break;
                double cov = entry_alpha[e] >= 0 ? coverage[entry_alpha[e] * 256 + py * 16 + px] : solid_cov;
                double sa = a * cov;
                double *d = out + ((size_t)y * width + x) * 4;
                double k = 1.0 - sa;
                d[0] = r * sa + d[0] * k;
                d[1] = g * sa + d[1] * k;
                d[2] = b * sa + d[2] * k;
                d[3] = sa + d[3] * k;
            }
        }
    }
    (void)fb_h;
}

// A CPU raster half for the timed baseline (bench.py's cpu_baseline leg): the hybrid path's fill
// (shaders/d3d9/fill.frag:25-45: per alpha-tile pixel, the box-filtered signed area of the tile's
// fills, accumulated on the backdrop, then the fill rule of shaders/d3d9/tile.frag:568-590) and
// its per-layer src-over composite (shaders/d3d9/tile.frag, BlendState::from_over,
// gpu/base.h:83-93) into RGBA8, in float on `threads` host threads: alpha tiles split by path,
// framebuffer tiles split by row.  Scenes without clip paths.  Writes RGBA8 [height][width] and
// returns the wall seconds of fill + composite.
double pfo_raster_cpu(void *h, const uint8_t *path_even_odd, const uint32_t *path_rgba, int width, int height,
                      int threads, uint8_t *out) {
    auto *r = (Result *)h;
    const int n_paths = (int)r->paths.size();
    std::vector<int64_t> tile_base(n_paths + 1, 0), alpha_base(n_paths + 1, 0);
    for (int p = 0; p < n_paths; p++) {
        tile_base[p + 1] = tile_base[p] + (int64_t)r->paths[p].backdrop.size();
        int64_t a = 0;
        for (auto v : r->paths[p].alpha) a += v;
        alpha_base[p + 1] = alpha_base[p] + a;
    }
    const int64_t n_alpha = alpha_base[n_paths];
    std::vector<float> mask((size_t)n_alpha * 256);
    std::vector<int64_t> tile_alpha((size_t)tile_base[n_paths], -1);
    for (int p = 0; p < n_paths; p++) {
        int64_t a = alpha_base[p];
        for (size_t i = 0; i < r->paths[p].alpha.size(); i++)
            if (r->paths[p].alpha[i]) tile_alpha[tile_base[p] + (int64_t)i] = a++;
    }
    if (threads < 1) threads = 1;
    auto t0 = std::chrono::steady_clock::now();
    auto fill = [&](int first) {
        std::vector<float> acc;
        for (int p = first; p < n_paths; p += threads) {
            const PathTiling &pt = r->paths[p];
            const int64_t a0 = alpha_base[p];
            acc.assign((size_t)(alpha_base[p + 1] - a0) * 256, 0.0f);
            for (const OFill &f : pt.fills) {
                const int64_t a = tile_alpha[tile_base[p] + f.tile];
                if (a < 0) continue;
                float *t = acc.data() + (size_t)(a - a0) * 256;
                const float x0 = f.x0 / 256.0f, y0 = f.y0 / 256.0f, x1 = f.x1 / 256.0f, y1 = f.y1 / 256.0f;
                const float sign = x0 > x1 ? 1.0f : -1.0f;
                const float xl = std::min(x0, x1), xr = std::max(x0, x1);
                const float yl = x0 < x1 ? y0 : y1, yr = x0 < x1 ? y1 : y0;
                const float slope = (yr - yl) / (xr - xl);
                for (int c = (int)xl; c < 16 && (float)c < xr; c++) {
                    const float lo = std::max((float)c, xl), hi = std::min((float)c + 1.0f, xr);
                    if (hi <= lo) continue;
                    const float ya = yl + (lo - xl) * slope, yb = yl + (hi - xl) * slope;
                    const float ymin = std::min(ya, yb), ymax = std::max(ya, yb);
                    const float w = sign * (hi - lo);
                    const int r0 = std::max(0, (int)ymin), r1 = std::min(15, (int)ymax);
                    for (int row = r0; row <= r1; row++) {
                        // mean of clamp(row + 1 - y, 0, 1) over the column's stretch of the fill
                        const float ua = row + 1.0f - ya, ub = row + 1.0f - yb;
                        const float lo_u = std::min(ua, ub), hi_u = std::max(ua, ub);
                        float m;
                        if (hi_u - lo_u <= 0.0f) {
                            m = std::min(std::max(lo_u, 0.0f), 1.0f);
                        } else {
                            const float full = hi_u > 1.0f ? hi_u - std::max(lo_u, 1.0f) : 0.0f;
                            const float l = std::max(lo_u, 0.0f), u = std::min(hi_u, 1.0f);
                            m = (full + (u > l ? (u - l) * (u + l) * 0.5f : 0.0f)) / (hi_u - lo_u);
                        }
                        t[row * 16 + c] += w * m;
                    }
                    for (int row = r1 + 1; row < 16; row++) t[row * 16 + c] += w;
                }
            }
            for (size_t i = 0; i < pt.alpha.size(); i++) {
                const int64_t a = tile_alpha[tile_base[p] + (int64_t)i];
                if (a < 0) continue;
                const float *t = acc.data() + (size_t)(a - a0) * 256;
                float *m = mask.data() + (size_t)a * 256;
                for (int k = 0; k < 256; k++) m[k] = (float)apply_rule((double)(pt.backdrop[i] + t[k]), path_even_odd[p]);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; i++) pool.emplace_back(fill, i);
    for (auto &t : pool) t.join();
    pool.clear();
    // entries are grouped by framebuffer tile, in draw order within one
    const int fw = r->fb_rect[2] - r->fb_rect[0];
    const size_t ne = r->entry_path.size();
    std::vector<size_t> fb_start((size_t)std::max(0, fw * (r->fb_rect[3] - r->fb_rect[1])) + 1, ne);
    for (size_t e = ne; e-- > 0;) fb_start[r->entry_fbtile[e]] = e;
    for (size_t i = fb_start.size() - 1; i-- > 0;) fb_start[i] = std::min(fb_start[i], fb_start[i + 1]);
    const int fh = r->fb_rect[3] - r->fb_rect[1];
    auto composite = [&](int first) {
        float px[16][16][4];
        for (int ty = first; ty < fh; ty += threads) {
            for (int tx = 0; tx < fw; tx++) {
                const size_t fb = (size_t)ty * fw + tx;
                std::memset(px, 0, sizeof(px));
                for (size_t e = fb_start[fb]; e < fb_start[fb + 1]; e++) {
                    const uint32_t p = r->entry_path[e];
                    const uint32_t c = path_rgba[p];
                    const float cr = (c & 255) / 255.0f, cg = ((c >> 8) & 255) / 255.0f,
                                cb = ((c >> 16) & 255) / 255.0f, ca = (c >> 24) / 255.0f;
                    const int64_t a = tile_alpha[r->entry_tile[e]];
                    const PathTiling &pt = r->paths[p];
                    const float solid = (float)apply_rule(
                        (double)pt.backdrop[r->entry_tile[e] - tile_base[p]], path_even_odd[p]);
                    const float *m = a >= 0 ? mask.data() + (size_t)a * 256 : nullptr;
                    for (int k = 0; k < 256; k++) {
                        const float sa = ca * (m ? m[k] : solid), inv = 1.0f - sa;
                        float *d = px[k >> 4][k & 15];
                        d[0] = cr * sa + d[0] * inv;
                        d[1] = cg * sa + d[1] * inv;
                        d[2] = cb * sa + d[2] * inv;
                        d[3] = sa + d[3] * inv;
                    }
                }
                for (int yy = 0; yy < 16; yy++) {
                    const int y = (ty + r->fb_rect[1]) * 16 + yy;
                    if (y < 0 || y >= height) continue;
                    for (int xx = 0; xx < 16; xx++) {
                        const int x = (tx + r->fb_rect[0]) * 16 + xx;
                        if (x < 0 || x >= width) continue;
                        for (int ch = 0; ch < 4; ch++)
                            out[((size_t)y * width + x) * 4 + ch] =
                                (uint8_t)std::lrint(std::min(std::max(px[yy][xx][ch], 0.0f), 1.0f) * 255.0f);
                    }
                }
            }
        }
    };
    for (int i = 0; i < threads; i++) pool.emplace_back(composite, i);
    for (auto &t : pool) t.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

} // extern "C"
