// Stroke-to-fill on the GPU: the outline whose nonzero fill is the stroke of an outline, as
// Canvas::stroke_path makes it before a stroked path reaches the scene (core/canvas.cpp:272-301,
// OutlineStrokeToFill::offset, core/stroke.cpp:124-167), for many paths at once.
//
// Work split:
//   host    — each input contour becomes one or two strands of segments in the order the
//             reference offsets them: forward (SegmentsIter order with the closing line,
//             core/stroke.cpp:28-52) and backward (the closing line reversed, then the segments
//             from the tail, core/stroke.cpp:54-122); O(points) bookkeeping;
//   k_offset_pieces<COUNT/EMIT> — one thread per segment: the offset recursion of
//             Segment::offset (core/stroke.cpp:502-541): offset the control polygon, check 9
//             samples against the tolerance, else split at t = 0.5 and recurse (explicit stack,
//             depth <= 16); each leaf is one add_to_contour call (a piece and its join point);
//   k_assemble — one thread per output contour: the leaves of its strands in order with the
//             joins before them (core/stroke.cpp:251-305, 469-480), the caps of open contours
//             (core/stroke.cpp:197-249) and the closing join (core/stroke.cpp:169-195), into a
//             region sized by an upper bound (a leaf adds at most 4 points plus a 16-point round
//             join, a cap or the closing join at most 16);
//   k_compact — the regions' points packed into the output outline.
// Every float operation keeps the reference's order (the file is built without FMA
// contraction), so the stroked outlines are bit-identical to the reference's.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pfh {

namespace stroke {

constexpr float FLOAT_EPSILON = 0.0001f;     // common/math/basic.h:11
constexpr float GEOMETRIC_EPSILON = 0.001f;  // common/math/basic.h:14
constexpr float STROKE_TOL = 0.1f;           // core/stroke.cpp:16
constexpr uint32_t SAMPLE_COUNT = 8;         // core/stroke.cpp:22
constexpr uint32_t MAX_DEPTH = 16;           // core/stroke.cpp:520

enum : uint32_t { K_NONE = 0, K_LINE = 1, K_QUAD = 2, K_CUBIC = 3 };
enum : int32_t { JOIN_MITER = 0, JOIN_BEVEL = 1, JOIN_ROUND = 2 };
enum : int32_t { CAP_BUTT = 0, CAP_SQUARE = 1, CAP_ROUND = 2 };

struct V {
    float x, y;
};
__device__ __forceinline__ V mk(float x, float y) { return V{x, y}; }
__device__ __forceinline__ V operator+(V a, V b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ V operator-(V a, V b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ V operator-(V a) { return {-a.x, -a.y}; }
__device__ __forceinline__ V operator*(V a, V b) { return {a.x * b.x, a.y * b.y}; }
__device__ __forceinline__ V operator*(V a, float s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ V operator/(V a, float s) { return {a.x / s, a.y / s}; }
__device__ __forceinline__ V add_s(V a, float s) { return {a.x + s, a.y + s}; }
__device__ __forceinline__ V yx(V a) { return {a.y, a.x}; }
__device__ __forceinline__ float sq_len(V a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ float len(V a) { return sqrtf(sq_len(a)); }
__device__ __forceinline__ V normalize(V a) { return a / len(a); }
__device__ __forceinline__ V vsqrt(V a) { return {sqrtf(a.x), sqrtf(a.y)}; }
__device__ __forceinline__ V vmax(V a, V b) { return {a.x < b.x ? b.x : a.x, a.y < b.y ? b.y : a.y}; }
__device__ __forceinline__ bool finite(V a) { return isfinite(a.x) && isfinite(a.y); }

// LineSegmentF (core/data/line_segment.cpp)
struct Line {
    V from, to;
    __device__ V vector() const { return to - from; }
    __device__ V sample(float t) const { return from + vector() * t; }
    __device__ float square_length() const { return sq_len(vector()); }
    __device__ bool valid() const { return finite(from) && finite(to); }
    __device__ Line offset(float d) const {  // :98-104
        const V v = vector();
        if (v.x == 0.0f && v.y == 0.0f) return *this;
        const V n = normalize(yx(v)) * mk(-d, d);
        return Line{from + n, to + n};
    }
    __device__ bool intersection_t(const Line &o, float &out) const {  // :86-96 with mat2.h
        const V p = vector(), q = o.vector();
        const float m0 = q.x, m1 = q.y, m2 = -p.x, m3 = -p.y;
        const float det = m0 * m3 - m1 * m2;
        if (fabsf(det) < FLOAT_EPSILON) return false;
        const float s = 1.0f / det;
        const float i1 = (m1 * -1.0f) * s, i3 = (m0 * 1.0f) * s;
        const V d = from - o.from;
        out = i1 * d.x + i3 * d.y;
        return true;
    }
};

// Transform2 (common/math/transform2.h): column-major 2x2 matrix and a translation
struct Xform {
    float m[4];
    V t;
    __device__ static Xform identity() { return Xform{{1.0f, 0.0f, 0.0f, 1.0f}, {0.0f, 0.0f}}; }
    __device__ V apply(V p) const { return V{m[0] * p.x + m[2] * p.y, m[1] * p.x + m[3] * p.y} + t; }
    __device__ Xform mul(const Xform &o) const {
        Xform r;
        r.m[0] = m[0] * o.m[0] + m[2] * o.m[1];
        r.m[1] = m[1] * o.m[0] + m[3] * o.m[1];
        r.m[2] = m[0] * o.m[2] + m[2] * o.m[3];
        r.m[3] = m[1] * o.m[2] + m[3] * o.m[3];
        r.t = apply(o.t);
        return r;
    }
    __device__ static Xform scale(V s) { return Xform{{s.x, 0.0f, 0.0f, s.y}, {0.0f, 0.0f}}; }
    __device__ static Xform rotation(V u) {
        return Xform{{1.0f * u.x, 1.0f * u.y, -1.0f * u.y, 1.0f * u.x}, {0.0f, 0.0f}};
    }
    __device__ Xform translate(V v) const {  // Transform2::translate: from_translation(v) * this
        Xform tr = identity();
        tr.t = v;
        return tr.mul(*this);
    }
};

// UnitVector (common/math/unit_vector.cpp)
__device__ __forceinline__ V rotate_by(V a, V o) { return {a.x * o.x - a.y * o.y, a.y * o.x + a.x * o.y}; }
__device__ __forceinline__ V rev_rotate_by(V a, V o) { return {a.x * o.x + a.y * o.y, a.y * o.x - a.x * o.y}; }
__device__ __forceinline__ V halve_angle(V a) { return vsqrt(vmax(mk(0.5f, 0.5f) * add_s(mk(a.x, -a.x), 1.0f), mk(0, 0))); }

// Segment (core/data/segment.cpp)
struct Seg {
    Line base, ctrl;
    uint32_t kind;

    __device__ bool valid() const {
        if (kind == K_LINE) return base.valid();
        if (kind == K_QUAD || kind == K_CUBIC) return base.valid() && ctrl.valid();
        return false;
    }
    __device__ void split_cubic(float t, Seg &a, Seg &b) const {  // :42-100
        if (t <= 0.0f) {
            const V f = base.from;
            a.base = Line{f, f};
            a.ctrl = Line{f, f};
            b.base = base;
            b.ctrl = ctrl;
        } else if (t >= 1.0f) {
            a.base = base;
            a.ctrl = ctrl;
            const V e = base.to;
            b.base = Line{e, e};
            b.ctrl = Line{e, e};
        } else {
            const V p0 = base.from, p3 = base.to, p1 = ctrl.from, p2 = ctrl.to;
            const V p01 = p0 + (p1 - p0) * t, p12 = p1 + (p2 - p1) * t;
            const V p23 = p2 + (p3 - p2) * t;
            const V p012 = p01 + (p12 - p01) * t, p123 = p12 + (p23 - p12) * t;
            const V p0123 = p012 + (p123 - p012) * t;
            a.base = Line{p0, p0123};
            a.ctrl = Line{p01, p012};
            b.base = Line{p0123, p3};
            b.ctrl = Line{p123, p23};
        }
        a.kind = b.kind = K_CUBIC;
    }
    __device__ void split(float t, Seg &a, Seg &b) const {  // :126-137, :102-126
        if (kind == K_LINE) {
            const V mid = base.from + (base.to - base.from) * t;
            a.base = Line{base.from, mid};
            b.base = Line{mid, base.to};
            a.ctrl = b.ctrl = Line{mk(0, 0), mk(0, 0)};
            a.kind = b.kind = K_LINE;
        } else if (kind == K_QUAD) {
            const V p0 = base.from, p1 = ctrl.from, p2 = base.to;
            const V qa = p0 + (p1 - p0) * t, qb = p1 + (p2 - p1) * t;
            const V c = qa + (qb - qa) * t;
            a = Seg{Line{p0, c}, Line{qa, qa}, K_QUAD};
            b = Seg{Line{c, p2}, Line{qb, qb}, K_QUAD};
        } else {
            split_cubic(t, a, b);
        }
    }
    __device__ V sample(float t) const {
        Seg a, b;
        split(t, a, b);
        return a.base.to;
    }
    __device__ Seg transform(const Xform &x) const {
        return Seg{Line{x.apply(base.from), x.apply(base.to)}, Line{x.apply(ctrl.from), x.apply(ctrl.to)}, kind};
    }
    // core/stroke.cpp:394-467
    __device__ Seg offset_once(float d) const {
        if (kind == K_LINE) return Seg{base.offset(d), Line{mk(0, 0), mk(0, 0)}, K_LINE};
        if (kind == K_QUAD) {
            const Line s0 = Line{base.from, ctrl.from}.offset(d), s1 = Line{ctrl.from, base.to}.offset(d);
            float t;
            const V c = s0.intersection_t(s1, t) ? s0.sample(t) : s0.to + (s1.from - s0.to) * 0.5f;
            return Seg{Line{s0.from, s1.to}, Line{c, c}, K_QUAD};
        }
        if (base.from.x == ctrl.from.x && base.from.y == ctrl.from.y) {
            const Line s0 = Line{base.from, ctrl.to}.offset(d), s1 = Line{ctrl.to, base.to}.offset(d);
            float t;
            const V c = s0.intersection_t(s1, t) ? s0.sample(t) : s0.to + (s1.from - s0.to) * 0.5f;
            return Seg{Line{s0.from, s1.to}, Line{s0.from, c}, K_CUBIC};
        }
        if (ctrl.to.x == base.to.x && ctrl.to.y == base.to.y) {
            const Line s0 = Line{base.from, ctrl.from}.offset(d), s1 = Line{ctrl.from, base.to}.offset(d);
            float t;
            const V c = s0.intersection_t(s1, t) ? s0.sample(t) : s0.to + (s1.from - s0.to) * 0.5f;
            return Seg{Line{s0.from, s1.to}, Line{c, s1.to}, K_CUBIC};
        }
        const Line s0 = Line{base.from, ctrl.from}.offset(d), s1 = Line{ctrl.from, ctrl.to}.offset(d);
        const Line s2 = Line{ctrl.to, base.to}.offset(d);
        V c0, c1;
        float t0, t1;
        if (s0.intersection_t(s1, t0) && s1.intersection_t(s2, t1)) {
            c0 = s0.sample(t0);
            c1 = s1.sample(t1);
        } else {
            c0 = s0.to + (s1.from - s0.to) * 0.5f;
            c1 = s1.to + (s2.from - s1.to) * 0.5f;
        }
        return Seg{Line{s0.from, s2.to}, Line{c0, c1}, K_CUBIC};
    }
    // core/stroke.cpp:481-500
    __device__ bool within_tolerance(const Seg &other, float d) const {
        float lo = fabsf(d) - STROKE_TOL, hi = fabsf(d) + STROKE_TOL;
        lo = lo <= 0 ? 0.0f : lo * lo;
        hi = hi <= 0 ? 0.0f : hi * hi;
        for (uint32_t k = 0; k < SAMPLE_COUNT + 1; k++) {
            const float t = (float)k / (float)SAMPLE_COUNT;
            const float s2 = sq_len(sample(t) - other.sample(t));
            if (s2 < lo || s2 > hi) return false;
        }
        return true;
    }
};

// core/data/segment.cpp:213-222
__device__ Seg quarter_circle_arc() {
    const float r2 = sqrtf(2.0f);
    const V p0 = mk(r2 * 0.5f, r2 * 0.5f);
    const V p1 = mk(-r2 / 6.0f + 4.0f / 3.0f, 7.0f * r2 / 6.0f - 4.0f / 3.0f);
    const V flip = mk(1.0f, -1.0f);
    return Seg{Line{p0 * flip, p0}, Line{p1 * flip, p1}, K_CUBIC};
}

// core/data/segment.cpp:224-256
__device__ Seg arc_from_cos(float c) {
    if ((double)c >= 1.0 - (double)GEOMETRIC_EPSILON) return Seg{Line{mk(1, 0), mk(1, 0)}, Line{mk(0, 0), mk(0, 0)}, K_LINE};
    const V r = vsqrt(add_s(mk(c, -c), 1.0f) * mk(0.5f, 0.5f));
    const V p3 = r * mk(1.0f, -1.0f), p0 = r * mk(1.0f, 1.0f);
    const float p1x = 4.0f - p0.x, p1y = (1.0f - p0.x) * (3.0f - p0.x) / p0.y;
    const V p2 = mk(p1x, -p1y) * (1.0f / 3.0f), p1 = mk(p1x, p1y) * (1.0f / 3.0f);
    return Seg{Line{p3, p0}, Line{p2, p1}, K_CUBIC};
}

}  // namespace stroke

// One segment of a strand, in the order the reference offsets it.
struct StrokeSegment {
    float2 p0, c0, c1, p1;
    uint32_t kind;    // K_LINE / K_QUAD / K_CUBIC
    int32_t join;     // the join of its pieces (the strand's first segment: bevel)
    float distance;   // -radius (the reference offsets by the negated radius)
    uint32_t pad;
};

// A piece of offset outline and the point its join is made around (add_to_contour's arguments).
struct StrokeLeaf {
    float2 p0, c0, c1, p1;
    float2 join_point;
    uint32_t kind;
    uint32_t pad;
};

// An output contour: its strands' segment ranges and what surrounds them.
struct StrokeContour {
    uint32_t seg0, seg1;     // forward strand [seg0, seg1) (first strand)
    uint32_t seg2, seg3;     // second strand (an open contour's backward strand), empty otherwise
    uint32_t closed;         // 1: closed input contour (closing join, no caps); 0: open (caps)
    int32_t cap, join;
    float width, miter;
    float2 input_first;      // the input contour's first point (the closing join's point)
    uint32_t pad;
};

// A path to dash: its contours [c0, c1) of the input and its pattern [d0, d1) of the dash lengths.
struct DashPath {
    uint32_t c0, c1;
    uint32_t d0, d1;
    float offset;
    uint32_t pad;
};
constexpr uint32_t DASH_MAX_STEPS = 1u << 24;  // a walk this long marks the path as over-long

namespace stroke {

// One add_to_contour call (core/stroke.cpp:469-480) in leaf form.
__device__ __forceinline__ StrokeLeaf make_leaf(const Seg &piece, V join_point) {
    StrokeLeaf l;
    l.p0 = make_float2(piece.base.from.x, piece.base.from.y);
    l.c0 = make_float2(piece.ctrl.from.x, piece.ctrl.from.y);
    l.c1 = make_float2(piece.ctrl.to.x, piece.ctrl.to.y);
    l.p1 = make_float2(piece.base.to.x, piece.base.to.y);
    l.join_point = make_float2(join_point.x, join_point.y);
    l.kind = piece.kind;
    l.pad = 0;
    return l;
}

}  // namespace stroke

// Segment::offset (core/stroke.cpp:502-541) per segment, depth first, first half first: COUNT
// stores the number of leaves, EMIT writes them from leaf_off[i].
template <bool EMIT>
__global__ __launch_bounds__(128) void k_offset_pieces(const StrokeSegment *segs, uint32_t n, uint32_t *leaf_cnt,
                                                       const uint32_t *leaf_off, StrokeLeaf *leaves) {
    using namespace stroke;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const StrokeSegment sg = segs[i];
    Seg s{Line{mk(sg.p0.x, sg.p0.y), mk(sg.p1.x, sg.p1.y)}, Line{mk(sg.c0.x, sg.c0.y), mk(sg.c1.x, sg.c1.y)}, sg.kind};
    const float d = sg.distance;
    Seg stack[MAX_DEPTH + 1];
    uint32_t sdepth[MAX_DEPTH + 1];
    uint32_t sp = 0, depth = 0, count = 0;
    const uint32_t base = EMIT ? leaf_off[i] : 0u;
    while (true) {
        bool leaf_done = false;
        if (!s.valid()) {
            leaf_done = true;  // an invalid (sub)segment adds nothing
        } else {
            const V join_point = s.base.from;
            bool emit = false;
            Seg piece = s;
            if (s.base.square_length() < STROKE_TOL * STROKE_TOL || depth >= MAX_DEPTH) {
                emit = true;  // added as it is
            } else if (s.kind == K_LINE) {
                piece = s.offset_once(d);
                emit = true;
            } else {
                const Seg cand = s.offset_once(d);
                if (s.within_tolerance(cand, d)) {
                    piece = cand;
                    emit = true;
                }
            }
            if (emit) {
                if (EMIT) leaves[base + count] = make_leaf(piece, join_point);
                count++;
                leaf_done = true;
            } else {
                Seg a, b;
                s.split(0.5f, a, b);
                stack[sp] = b;
                sdepth[sp] = depth + 1;
                sp++;
                s = a;
                depth++;
            }
        }
        if (leaf_done) {
            if (sp == 0) break;
            sp--;
            s = stack[sp];
            depth = sdepth[sp];
        }
    }
    if (!EMIT) leaf_cnt[i] = count;
}

namespace stroke {

// The output contour under construction, written into its region.
struct Out {
    float2 *pts;
    uint8_t *flags;
    uint32_t n, cap;
    bool write;
    V first0, first1;  // its first two points (the closing join's final segment)
    V last2, last1;    // its last two points

    __device__ void push(V p, uint8_t f) {
        if (write && n < cap) {
            pts[n] = make_float2(p.x, p.y);
            flags[n] = f;
        }
        if (n == 0) first0 = p;
        if (n == 1) first1 = p;
        last2 = last1;
        last1 = p;
        n++;
    }
    __device__ V at(uint32_t k) const { return k < cap ? mk(pts[k].x, pts[k].y) : mk(0, 0); }
    // core/data/contour.cpp:11-38
    __device__ void push_segment(const Seg &s) {
        if (s.kind == K_NONE || !s.valid()) return;
        push(s.base.from, 0);
        if (s.kind != K_LINE) {
            push(s.ctrl.from, 1);
            if (s.kind != K_QUAD) push(s.ctrl.to, 2);
        }
        push(s.base.to, 0);
    }
    __device__ bool might_need_join(int32_t join) const { return n >= 2 && (join == JOIN_MITER || join == JOIN_ROUND); }
    // core/stroke.cpp:307-348, clockwise
    __device__ void push_arc_from_unit_chord(const Xform &x, const Line &chord) {
        V vec = chord.from;
        const V end = chord.to;
        const Xform dir = Xform::identity();
        for (int k = 0; k < 4; k++) {
            V sweep = rev_rotate_by(end, vec);
            const bool last_piece = sweep.x >= -FLOAT_EPSILON && sweep.y >= -FLOAT_EPSILON;
            Seg s;
            if (!last_piece) {
                sweep = mk(0.0f, 1.0f);
                s = quarter_circle_arc();
            } else {
                s = arc_from_cos(sweep.x);
            }
            const V half = halve_angle(sweep);
            const Xform rot = Xform::rotation(rotate_by(half, vec));
            s = s.transform(x.mul(dir).mul(rot));
            push_segment(s);
            if (last_piece) break;
            vec = rotate_by(vec, sweep);
        }
    }
    // core/stroke.cpp:251-305
    __device__ void add_join(float distance, int32_t join, V join_point, const Line &next, float miter_limit) {
        const Line prev{last2, last1};
        if (prev.square_length() < FLOAT_EPSILON || next.square_length() < FLOAT_EPSILON) return;
        if (join == JOIN_MITER) {
            float t;
            if (prev.intersection_t(next, t)) {
                if (t < -FLOAT_EPSILON) return;
                const V m = prev.sample(t);
                const float threshold = miter_limit * distance;
                if (sq_len(m - join_point) > threshold * threshold) return;
                push(m, 0);
            }
        } else if (join == JOIN_ROUND) {
            if (sq_len(prev.to - join_point) == 0 || sq_len(next.to - join_point) == 0) return;
            const float scale = fabsf(distance);
            const Xform x = Xform::scale(mk(scale, scale)).translate(join_point);
            const Line chord{normalize(prev.to - join_point), normalize(next.to - join_point)};
            push_arc_from_unit_chord(x, chord);
        }
    }
    // core/stroke.cpp:469-480
    __device__ void add_leaf(const StrokeLeaf &l, float distance, int32_t join, float miter) {
        const Seg piece{Line{mk(l.p0.x, l.p0.y), mk(l.p1.x, l.p1.y)}, Line{mk(l.c0.x, l.c0.y), mk(l.c1.x, l.c1.y)}, l.kind};
        if (might_need_join(join)) {
            const V p3 = piece.base.from, p4 = piece.kind == K_LINE ? piece.base.to : piece.ctrl.from;
            add_join(distance, join, mk(l.join_point.x, l.join_point.y), Line{p4, p3}, miter);
        }
        push_segment(piece);
    }
    // core/stroke.cpp:197-249 (reads back the points written so far: only in the writing pass)
    __device__ void add_cap(int32_t cap, float width) {
        if (cap == CAP_BUTT || n < 2) return;
        const V p1 = last1;
        V p0;
        uint32_t i0 = n - 2;
        while (true) {
            p0 = i0 + 2 == n ? last2 : at(i0);
            if (sq_len(p1 - p0) > FLOAT_EPSILON) break;
            if (i0 == 0) return;
            i0--;
        }
        const V g = normalize(p1 - p0);
        if (cap == CAP_SQUARE) {
            const V off = g * (width * 0.5f);
            const V p2 = p1 + off, p3 = p2 + yx(g) * mk(-width, width), p4 = p3 - off;
            push(p2, 0);
            push(p3, 0);
            push(p4, 0);
        } else {
            const float scale = width * 0.5f;
            const V off = yx(g) * mk(-1.0f, 1.0f);
            const V tr = p1 + off * (width * 0.5f);
            const Xform x = Xform::scale(mk(scale, scale)).translate(tr);
            push_arc_from_unit_chord(x, Line{-off, off});
        }
    }
};

}  // namespace stroke

// One thread per output contour: its leaves in order with their joins, the caps of an open
// contour between and after its strands, the closing join of a closed one.  Writes into
// [region_off[c], region_off[c] + region_cap) and stores the contour's point count.
__global__ __launch_bounds__(64) void k_assemble(const StrokeContour *contours, uint32_t n, const uint32_t *leaf_off,
                                                 const StrokeLeaf *leaves, const StrokeSegment *segs,
                                                 const uint32_t *region_off, float2 *region_pts,
                                                 uint8_t *region_flags, uint32_t *point_cnt) {
    using namespace stroke;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const StrokeContour sc = contours[c];
    Out o;
    o.pts = region_pts + region_off[c];
    o.flags = region_flags + region_off[c];
    o.cap = region_off[c + 1] - region_off[c];
    o.n = 0;
    o.write = true;
    o.first0 = o.first1 = o.last2 = o.last1 = mk(0, 0);
    auto strand = [&](uint32_t s0, uint32_t s1) {
        for (uint32_t s = s0; s < s1; s++) {
            const StrokeSegment sg = segs[s];
            for (uint32_t k = leaf_off[s]; k < leaf_off[s + 1]; k++) o.add_leaf(leaves[k], sg.distance, sg.join, sc.miter);
        }
    };
    strand(sc.seg0, sc.seg1);
    if (!sc.closed) o.add_cap(sc.cap, sc.width);
    strand(sc.seg2, sc.seg3);
    if (!sc.closed) o.add_cap(sc.cap, sc.width);
    if (sc.closed && o.might_need_join(sc.join)) {  // push_stroked_contour's closing join
        const Line final_segment{o.first1, o.first0};
        o.add_join(sc.width * 0.5f, sc.join, mk(sc.input_first.x, sc.input_first.y), final_segment, sc.miter);
    }
    point_cnt[c] = o.n;
}

// Region points packed at the contours' scanned offsets.
__global__ __launch_bounds__(256) void k_compact(const uint32_t *region_off, const uint32_t *out_off, uint32_t n,
                                                 const float2 *region_pts, const uint8_t *region_flags,
                                                 float2 *out_pts, uint8_t *out_flags) {
    const uint32_t c = blockIdx.x;
    if (c >= n) return;
    const uint32_t src = region_off[c], dst = out_off[c], cnt = out_off[c + 1] - out_off[c];
    for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) {
        out_pts[dst + k] = region_pts[src + k];
        out_flags[dst + k] = region_flags[src + k];
    }
}

// ---------------------------------------------------------------------------
// Dashes (OutlineDash, core/dash.cpp:9-126), before stroking (core/canvas.cpp:284-289)
// ---------------------------------------------------------------------------
//
// A thread per dashed path: the dash state (pattern index, distance left, the dash being built)
// carries over the path's contours, as in the reference, so a path is one sequential walk.  The
// count pass sizes each path's dashed outline; the emit pass writes it at the offsets the host scanned.

namespace stroke {

// Segment::arc_length (core/data/segment.cpp:168-187): a line's length; for a curve, the mean of its
// chord and its control polygon's length, a quadratic raised to a cubic first (:141-151).
__device__ __forceinline__ float arc_length(const Seg &s) {
    if (s.kind == K_LINE) return len(s.base.vector());
    if (s.kind != K_QUAD && s.kind != K_CUBIC) return 0.0f;
    V c0 = s.ctrl.from, c1 = s.ctrl.to;
    if (s.kind == K_QUAD) {
        const V p1_2 = s.ctrl.from + s.ctrl.from;
        c0 = (s.base.from + p1_2) / 3.0f;
        c1 = (p1_2 + s.base.to) / 3.0f;
    }
    const float chord = len(s.base.vector());
    const float cont_net = len(s.base.from - c0) + len(c1 - c0) + len(s.base.to - c1);
    return (cont_net + chord) * 0.5f;
}

// SegmentsIter::get_next (core/data/contour.cpp:110-173) without force_closed.
struct SegIter {
    const float2 *p;
    const uint8_t *f;
    int n;
    bool closed;
    int head;
    bool has_next;
    __device__ V at(int i) const { return mk(p[i].x, p[i].y); }
    __device__ Seg next() {
        Seg s{Line{mk(0, 0), mk(0, 0)}, Line{mk(0, 0), mk(0, 0)}, K_NONE};
        if (head >= n || f[head] != 0) return s;
        s.base.from = at(head);
        if (head + 1 < n) {
            if (f[head + 1] == 0) {
                s.base.to = at(head + 1);
                s.kind = K_LINE;
                head += 1;
            } else if (head + 2 < n) {
                if (f[head + 1] == 1 && f[head + 2] == 0) {
                    s.ctrl.from = at(head + 1);
                    s.base.to = at(head + 2);
                    s.kind = K_QUAD;
                    head += 2;
                } else if (head + 3 < n) {
                    if (f[head + 1] == 1 && f[head + 2] == 2 && f[head + 3] == 0) {
                        s.ctrl.from = at(head + 1);
                        s.ctrl.to = at(head + 2);
                        s.base.to = at(head + 3);
                        s.kind = K_CUBIC;
                        head += 3;
                    }
                }
            }
        } else {
            // the closing line unless the contour ends where it starts (Vec2::approx_eq, 1e-4)
            if (closed && n > 1 && !(len(at(0) - at(n - 1)) <= FLOAT_EPSILON)) {
                s.base.from = at(n - 1);
                s.base.to = at(0);
                s.kind = K_LINE;
            }
            has_next = false;
        }
        return s;
    }
};

}  // namespace stroke

template <bool EMIT>
__global__ __launch_bounds__(64) void k_dash(const float2 *pts, const uint8_t *flags, const int32_t *contour_start,
                                             const uint8_t *closed, const DashPath *paths, uint32_t n,
                                             const float *dashes, uint32_t *cnt, const uint32_t *off, float2 *out_pts,
                                             uint8_t *out_flags, uint32_t *out_cstart) {
    using namespace stroke;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const DashPath dp = paths[p];
    const float *D = dashes + dp.d0;
    const uint32_t nd = dp.d1 - dp.d0;
    // DashState::DashState (core/dash.cpp:9-29); the host checked the pattern: finite, >= 0, total > 0
    float total = 0.0f;
    for (uint32_t i = 0; i < nd; i++) total += D[i];
    float offset = fmodf(dp.offset, total);
    uint32_t idx = 0;
    while (idx < nd) {
        const float d = D[idx];
        if (offset < d) break;
        offset -= d;
        idx += 1;
    }
    // Rounding in the subtractions can walk past the last dash (the reference then indexes one past
    // its pattern); the pattern restarts instead.
    if (idx >= nd) idx = 0;
    float left = offset;
    const uint32_t pbase = EMIT ? off[2 * p] : 0u, cbase = EMIT ? off[2 * p + 1] : 0u;
    const uint32_t pcap = EMIT ? off[2 * p + 2] - pbase : 0xFFFFFFFFu, ccap = EMIT ? off[2 * p + 3] - cbase : 0xFFFFFFFFu;
    uint32_t npts = 0, ncont = 0, cur = 0, steps = 0;
    bool over = false;
    auto push_point = [&](V v, uint8_t f) {
        if (EMIT && npts < pcap) {
            out_pts[pbase + npts] = make_float2(v.x, v.y);
            out_flags[pbase + npts] = f;
        }
        npts++;
        cur++;
    };
    // Contour::push_segment (core/data/contour.cpp:11-41): every point, its start point included
    auto push_segment = [&](const Seg &s) {
        if (!s.valid()) return;
        push_point(s.base.from, 0);
        if (s.kind != K_LINE) {
            push_point(s.ctrl.from, 1);
            if (s.kind != K_QUAD) push_point(s.ctrl.to, 2);
        }
        push_point(s.base.to, 0);
    };
    // Outline::push_contour (core/data/path.cpp:24-38) skips an empty contour
    auto push_contour = [&]() {
        if (cur > 0) {
            if (EMIT && ncont < ccap) out_cstart[cbase + ncont] = pbase + npts - cur;
            ncont++;
        }
        cur = 0;
    };
    for (uint32_t c = dp.c0; c < dp.c1 && !over; c++) {
        const int a = contour_start[c];
        SegIter it{pts + a, flags + a, contour_start[c + 1] - a, closed[c] != 0, 0, true};
        Seg queued{Line{mk(0, 0), mk(0, 0)}, Line{mk(0, 0), mk(0, 0)}, K_NONE};
        bool queued_none = true;
        // ContourDash::dash (core/dash.cpp:74-124)
        while (true) {
            if (++steps > DASH_MAX_STEPS) {
                over = true;
                break;
            }
            if (queued_none) {
                if (!it.has_next) break;
                queued = it.next();
                if (queued.kind == K_NONE) break;
                queued_none = false;
            }
            Seg seg = queued;
            float distance = left;
            const float t = distance / arc_length(seg);
            if (t < 1.0f) {
                Seg s0, s1;
                seg.split(t, s0, s1);
                seg = s0;
                queued = s1;
                queued_none = false;
            } else {
                distance = arc_length(seg);
                queued_none = true;
            }
            if (idx % 2 == 0) push_segment(seg);
            left -= distance;
            if (left < FLOAT_EPSILON) {
                if (idx % 2 == 0) push_contour();
                idx += 1;
                if (idx == nd) idx = 0;
                left = D[idx];
            }
        }
    }
    if (idx % 2 == 0) push_contour();  // OutlineDash::into_outline (core/dash.cpp:57-63)
    if (!EMIT) {
        cnt[2 * p] = over ? 0xFFFFFFFFu : npts;
        cnt[2 * p + 1] = ncont;
    }
}

hipError_t launch_dash(const float2 *pts, const uint8_t *flags, const int32_t *contour_start, const uint8_t *closed,
                       const DashPath *paths, uint32_t n, const float *dashes, uint32_t *cnt, const uint32_t *off,
                       float2 *out_pts, uint8_t *out_flags, uint32_t *out_cstart, bool emit, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 grid((n + 63) / 64);
    if (emit)
        hipLaunchKernelGGL(k_dash<true>, grid, dim3(64), 0, s, pts, flags, contour_start, closed, paths, n, dashes, cnt,
                           off, out_pts, out_flags, out_cstart);
    else
        hipLaunchKernelGGL(k_dash<false>, grid, dim3(64), 0, s, pts, flags, contour_start, closed, paths, n, dashes, cnt,
                           off, out_pts, out_flags, out_cstart);
    return hipGetLastError();
}

hipError_t launch_offset_pieces(const StrokeSegment *segs, uint32_t n, uint32_t *leaf_cnt, const uint32_t *leaf_off,
                                StrokeLeaf *leaves, bool emit, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 g((n + 127) / 128);
    if (emit) hipLaunchKernelGGL(k_offset_pieces<true>, g, dim3(128), 0, s, segs, n, leaf_cnt, leaf_off, leaves);
    else hipLaunchKernelGGL(k_offset_pieces<false>, g, dim3(128), 0, s, segs, n, leaf_cnt, leaf_off, leaves);
    return hipGetLastError();
}

hipError_t launch_assemble(const StrokeContour *contours, uint32_t n, const uint32_t *leaf_off, const StrokeLeaf *leaves,
                           const StrokeSegment *segs, const uint32_t *region_off, float2 *region_pts,
                           uint8_t *region_flags, uint32_t *point_cnt, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_assemble, dim3((n + 63) / 64), dim3(64), 0, s, contours, n, leaf_off, leaves, segs,
                       region_off, region_pts, region_flags, point_cnt);
    return hipGetLastError();
}

hipError_t launch_compact(const uint32_t *region_off, const uint32_t *out_off, uint32_t n, const float2 *region_pts,
                          const uint8_t *region_flags, float2 *out_pts, uint8_t *out_flags, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact, dim3(n), dim3(256), 0, s, region_off, out_off, n, region_pts, region_flags, out_pts,
                       out_flags);
    return hipGetLastError();
}

}  // namespace pfh
