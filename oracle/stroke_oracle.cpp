// ORACLE — test infrastructure only (tests/ and the checker legs of __graft_entry__.smoke() and
// bench.py).  The product's stroke-to-fill runs on the GPU (pfh_stroke.hip); this is the CPU
// restatement it is checked against, itself pinned bit for bit to fixtures the reference's own
// compiled OutlineStrokeToFill wrote (tests/golden/stroke_fixtures.npz).
//
// Stroke-to-fill: an outline and a stroke style in, the outline whose nonzero fill is the stroke
// out, as Canvas::stroke_path does before a stroked path reaches the scene (core/canvas.cpp:272-301):
// OutlineStrokeToFill::offset (core/stroke.cpp:124-167).  Every float operation keeps the
// reference's order: offset curves per segment with the tolerance check and midpoint recursion
// (core/stroke.cpp:502-541), miter / bevel / round joins (core/stroke.cpp:251-305), butt / square /
// round caps (core/stroke.cpp:197-249), round arcs as cubic quarter arcs (core/stroke.cpp:307-348,
// core/data/segment.cpp:213-256).  Built without FMA contraction.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>


// The outline and style layouts of the library's pfh_outline / pfh_stroke_style (include/pfh.h).
struct pfo_outline {
    int32_t n_contours, n_points;
    const float *points;
    const uint8_t *point_flags;
    const int32_t *contour_start;
    const uint8_t *contour_closed;
};
struct pfo_stroke_style {
    float line_width;
    int32_t line_cap;
    int32_t line_join;
    float miter_limit;
};

namespace {

constexpr float FLOAT_EPSILON = 0.0001f;     // common/math/basic.h:11
constexpr float GEOMETRIC_EPSILON = 0.001f;  // common/math/basic.h:14
constexpr float STROKE_TOL = 0.1f;           // core/stroke.cpp:16
constexpr uint32_t SAMPLE_COUNT = 8;         // core/stroke.cpp:22
constexpr uint32_t MAX_OFFSET_DEPTH = 16;    // core/stroke.cpp:520

enum Flag : uint8_t { ON_CURVE = 0, CONTROL_0 = 1, CONTROL_1 = 2 };
enum Kind : uint8_t { K_NONE = 0, K_LINE, K_QUAD, K_CUBIC };
enum Join : int32_t { JOIN_MITER = 0, JOIN_BEVEL = 1, JOIN_ROUND = 2 };
enum Cap : int32_t { CAP_BUTT = 0, CAP_SQUARE = 1, CAP_ROUND = 2 };

struct V {
    float x = 0, y = 0;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y}; }
inline V operator-(V a) { return {-a.x, -a.y}; }
inline V operator*(V a, V b) { return {a.x * b.x, a.y * b.y}; }
inline V operator*(V a, float s) { return {a.x * s, a.y * s}; }
inline V operator/(V a, float s) { return {a.x / s, a.y / s}; }
inline V add_s(V a, float s) { return {a.x + s, a.y + s}; }
inline V yx(V a) { return {a.y, a.x}; }
inline float sq_len(V a) { return a.x * a.x + a.y * a.y; }
inline float len(V a) { return std::sqrt(sq_len(a)); }
inline V normalize(V a) { return a / len(a); }
inline V vsqrt(V a) { return {std::sqrt(a.x), std::sqrt(a.y)}; }
// std::max(a, b): b only when a < b (a NaN operand keeps a)
inline V vmax(V a, V b) { return {a.x < b.x ? b.x : a.x, a.y < b.y ? b.y : a.y}; }
inline bool approx_eq(V a, V b, float eps) { return len(a - b) <= eps; }
inline bool finite(V a) { return std::isfinite(a.x) && std::isfinite(a.y); }

struct Line {
    V from, to;
    V vector() const { return to - from; }
    V sample(float t) const { return from + vector() * t; }
    float square_length() const { return sq_len(vector()); }
    bool valid() const { return finite(from) && finite(to); }
    // core/data/line_segment.cpp:98-104: moved along the normal (a zero vector stays put)
    Line offset(float d) const {
        const V v = vector();
        if (v.x == 0.0f && v.y == 0.0f) return *this;
        const V n = normalize(yx(v)) * V{-d, d};
        return {from + n, to + n};
    }
    // core/data/line_segment.cpp:86-96 with common/math/mat2.h (det, adjugate * (1 / det))
    bool intersection_t(const Line &o, float &out) const {
        const V p = vector(), q = o.vector();
        const float m0 = q.x, m1 = q.y, m2 = -p.x, m3 = -p.y;
        const float det = m0 * m3 - m1 * m2;
        if (std::fabs(det) < FLOAT_EPSILON) return false;
        const float s = 1.0f / det;
        const float i1 = (m1 * -1.0f) * s, i3 = (m0 * 1.0f) * s;
        const V d = from - o.from;
        out = i1 * d.x + i3 * d.y;
        return true;
    }
};

// 2x2 matrix (column-major m11, m21, m12, m22) and translation (common/math/transform2.h)
struct Xform {
    float m[4] = {1, 0, 0, 1};
    V t;
    V apply(V p) const { return V{m[0] * p.x + m[2] * p.y, m[1] * p.x + m[3] * p.y} + t; }
    Xform operator*(const Xform &o) const {
        Xform r;
        r.m[0] = m[0] * o.m[0] + m[2] * o.m[1];
        r.m[1] = m[1] * o.m[0] + m[3] * o.m[1];
        r.m[2] = m[0] * o.m[2] + m[2] * o.m[3];
        r.m[3] = m[1] * o.m[2] + m[3] * o.m[3];
        r.t = apply(o.t);
        return r;
    }
    static Xform scale(V s) {
        Xform r;
        r.m[0] = s.x;
        r.m[1] = 0.0f;
        r.m[2] = 0.0f;
        r.m[3] = s.y;
        return r;
    }
    static Xform translation(V v) {
        Xform r;
        r.t = v;
        return r;
    }
    static Xform rotation(V u) {  // Mat2::from_rotation_vector
        Xform r;
        r.m[0] = 1.0f * u.x;
        r.m[1] = 1.0f * u.y;
        r.m[2] = -1.0f * u.y;
        r.m[3] = 1.0f * u.x;
        return r;
    }
    // Transform2::translate: the translation applied after this transform
    Xform translate(V v) const { return translation(v) * *this; }
};

// common/math/unit_vector.cpp
inline V rotate_by(V a, V o) { return {a.x * o.x - a.y * o.y, a.y * o.x + a.x * o.y}; }
inline V rev_rotate_by(V a, V o) { return {a.x * o.x + a.y * o.y, a.y * o.x - a.x * o.y}; }
inline V halve_angle(V a) { return vsqrt(vmax(V{0.5f, 0.5f} * add_s(V{a.x, -a.x}, 1.0f) , V{0, 0})); }

struct Seg {
    Line base, ctrl;
    Kind kind = K_NONE;

    static Seg line(Line b) { return Seg{b, Line{}, K_LINE}; }
    static Seg quad(Line b, V c) { return Seg{b, Line{c, c}, K_QUAD}; }
    static Seg cubic(Line b, Line c) { return Seg{b, c, K_CUBIC}; }

    bool valid() const {
        if (kind == K_LINE) return base.valid();
        if (kind == K_QUAD || kind == K_CUBIC) return base.valid() && ctrl.valid();
        return false;
    }

    // core/data/segment.cpp:42-100
    void split_cubic(float t, Seg &a, Seg &b) const {
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
    // core/data/segment.cpp:102-126
    void split_quad(float t, Seg &a, Seg &b) const {
        const V p0 = base.from, p1 = ctrl.from, p2 = base.to;
        const V qa = p0 + (p1 - p0) * t, qb = p1 + (p2 - p1) * t;
        const V c = qa + (qb - qa) * t;
        a = Seg{Line{p0, c}, Line{qa, qa}, K_QUAD};
        b = Seg{Line{c, p2}, Line{qb, qb}, K_QUAD};
    }
    // core/data/segment.cpp:126-137 (a line splits through LineSegmentF::split)
    void split(float t, Seg &a, Seg &b) const {
        if (kind == K_LINE) {
            const V mid = base.from + (base.to - base.from) * t;
            a.base = Line{base.from, mid};
            b.base = Line{mid, base.to};
            a.kind = b.kind = K_LINE;
        } else if (kind == K_QUAD) {
            split_quad(t, a, b);
        } else {
            split_cubic(t, a, b);
        }
    }
    V sample(float t) const {
        Seg a, b;
        split(t, a, b);
        return a.base.to;
    }
    Seg transform(const Xform &x) const {
        return Seg{Line{x.apply(base.from), x.apply(base.to)}, Line{x.apply(ctrl.from), x.apply(ctrl.to)}, kind};
    }

    // Offset control polygon: each leg moved along its normal, inner controls at the legs'
    // intersections (core/stroke.cpp:394-467).
    Seg offset_once(float d) const {
        if (kind == K_LINE) return Seg::line(base.offset(d));
        auto meet = [](const Line &s0, const Line &s1) {
            float t;
            return s0.intersection_t(s1, t) ? s0.sample(t) : s0.to + (s1.from - s0.to) * 0.5f;
        };
        if (kind == K_QUAD) {
            const Line s0 = Line{base.from, ctrl.from}.offset(d), s1 = Line{ctrl.from, base.to}.offset(d);
            return Seg::quad(Line{s0.from, s1.to}, meet(s0, s1));
        }
        if (base.from.x == ctrl.from.x && base.from.y == ctrl.from.y) {
            const Line s0 = Line{base.from, ctrl.to}.offset(d), s1 = Line{ctrl.to, base.to}.offset(d);
            const V c = meet(s0, s1);
            return Seg::cubic(Line{s0.from, s1.to}, Line{s0.from, c});
        }
        if (ctrl.to.x == base.to.x && ctrl.to.y == base.to.y) {
            const Line s0 = Line{base.from, ctrl.from}.offset(d), s1 = Line{ctrl.from, base.to}.offset(d);
            const V c = meet(s0, s1);
            return Seg::cubic(Line{s0.from, s1.to}, Line{c, s1.to});
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
        return Seg::cubic(Line{s0.from, s2.to}, Line{c0, c1});
    }

    // core/stroke.cpp:481-500: the offset curve stays within STROKE_TOL of |d| at 9 samples
    bool within_tolerance(const Seg &other, float d) const {
        float lo = std::fabs(d) - STROKE_TOL, hi = std::fabs(d) + STROKE_TOL;
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
Seg quarter_circle_arc() {
    const V p0{std::sqrt(2.0f) * 0.5f, std::sqrt(2.0f) * 0.5f};
    const V p1{-std::sqrt(2.0f) / 6.0f + 4.0f / 3.0f, 7.0f * std::sqrt(2.0f) / 6.0f - 4.0f / 3.0f};
    const V flip{1.0f, -1.0f};
    return Seg{Line{p0 * flip, p0}, Line{p1 * flip, p1}, K_CUBIC};
}

// core/data/segment.cpp:224-256 (DeVeneza's cubic for an arc of at most 90 degrees)
Seg arc_from_cos(float c) {
    if ((double)c >= 1.0 - (double)GEOMETRIC_EPSILON) return Seg::line(Line{V{1, 0}, V{1, 0}});
    const V term{c, -c};
    const V r = vsqrt(add_s(term, 1.0f) * V{0.5f, 0.5f});
    const V p3 = r * V{1.0f, -1.0f}, p0 = r * V{1.0f, 1.0f};
    const float p1x = 4.0f - p0.x, p1y = (1.0f - p0.x) * (3.0f - p0.x) / p0.y;
    const V p2 = V{p1x, -p1y} * (1.0f / 3.0f), p1 = V{p1x, p1y} * (1.0f / 3.0f);
    return Seg::cubic(Line{p3, p0}, Line{p2, p1});
}

struct Contour {
    std::vector<V> pts;
    std::vector<uint8_t> flags;
    bool closed = false;

    void push(V p, uint8_t f) {
        pts.push_back(p);
        flags.push_back(f);
    }
    V last(size_t i) const { return pts[pts.size() - i]; }
    // core/data/contour.cpp:11-38: every point of the segment, its start point included
    void push_segment(const Seg &s) {
        if (s.kind == K_NONE || !s.valid()) return;
        push(s.base.from, ON_CURVE);
        if (s.kind != K_LINE) {
            push(s.ctrl.from, CONTROL_0);
            if (s.kind != K_QUAD) push(s.ctrl.to, CONTROL_1);
        }
        push(s.base.to, ON_CURVE);
    }
    bool might_need_join(int32_t join) const { return pts.size() >= 2 && (join == JOIN_MITER || join == JOIN_ROUND); }

    // core/stroke.cpp:307-348, clockwise
    void push_arc_from_unit_chord(const Xform &x, const Line &chord) {
        V vec = chord.from;
        const V end = chord.to;
        const Xform dir;  // clockwise: identity
        for (int i = 0; i < 4; i++) {
            V sweep = rev_rotate_by(end, vec);
            const bool last_piece = sweep.x >= -FLOAT_EPSILON && sweep.y >= -FLOAT_EPSILON;
            Seg s;
            if (!last_piece) {
                sweep = V{0.0f, 1.0f};
                s = quarter_circle_arc();
            } else {
                s = arc_from_cos(sweep.x);
            }
            const V half = halve_angle(sweep);
            const Xform rot = Xform::rotation(rotate_by(half, vec));
            s = s.transform((x * dir) * rot);
            push_segment(s);
            if (last_piece) break;
            vec = rotate_by(vec, sweep);
        }
    }

    // core/stroke.cpp:251-305
    void add_join(float distance, int32_t join, V join_point, const Line &next, float miter_limit) {
        const Line prev{last(2), last(1)};
        if (prev.square_length() < FLOAT_EPSILON || next.square_length() < FLOAT_EPSILON) return;
        if (join == JOIN_MITER) {
            float t;
            if (prev.intersection_t(next, t)) {
                if (t < -FLOAT_EPSILON) return;
                const V m = prev.sample(t);
                const float threshold = miter_limit * distance;
                if (sq_len(m - join_point) > threshold * threshold) return;
                push(m, ON_CURVE);
            }
        } else if (join == JOIN_ROUND) {
            if (sq_len(prev.to - join_point) == 0 || sq_len(next.to - join_point) == 0) return;
            const float scale = std::fabs(distance);
            const Xform x = Xform::scale(V{scale, scale}).translate(join_point);
            const Line chord{normalize(prev.to - join_point), normalize(next.to - join_point)};
            push_arc_from_unit_chord(x, chord);
        }
    }
};

// core/stroke.cpp:469-480
void add_to_contour(const Seg &s, float distance, int32_t join, V join_point, float miter_limit, Contour &c) {
    if (c.might_need_join(join)) {
        const V p3 = s.base.from, p4 = s.kind == K_LINE ? s.base.to : s.ctrl.from;
        c.add_join(distance, join, join_point, Line{p4, p3}, miter_limit);
    }
    c.push_segment(s);
}

// core/stroke.cpp:502-541
void offset_segment(const Seg &s, float distance, int32_t join, float miter_limit, Contour &c, uint32_t depth = 0) {
    if (!s.valid()) return;
    const V join_point = s.base.from;
    if (s.base.square_length() < STROKE_TOL * STROKE_TOL || depth >= MAX_OFFSET_DEPTH) {
        add_to_contour(s, distance, join, join_point, miter_limit, c);
        return;
    }
    if (s.kind == K_LINE) {
        add_to_contour(s.offset_once(distance), distance, join, join_point, miter_limit, c);
        return;
    }
    const Seg cand = s.offset_once(distance);
    if (s.within_tolerance(cand, distance)) {
        add_to_contour(cand, distance, join, join_point, miter_limit, c);
        return;
    }
    Seg a, b;
    s.split(0.5f, a, b);
    offset_segment(a, distance, join, miter_limit, c, depth + 1);
    offset_segment(b, distance, join, miter_limit, c, depth + 1);
}

struct InContour {
    const V *pts;
    const uint8_t *flags;
    int n;
    bool closed;
};

// core/stroke.cpp:28-52: the contour's segments in order (SegmentsIter without force_closed,
// core/data/contour.cpp:110-173), the first with a bevel join
void offset_forward(const InContour &in, float radius, int32_t join, float miter, Contour &out) {
    int head = 0, index = -1;
    while (true) {
        if (head >= in.n || in.flags[head] != ON_CURVE) break;
        Seg s;
        s.base.from = in.pts[head];
        bool end = false;
        if (head + 1 < in.n) {
            if (in.flags[head + 1] == ON_CURVE) {
                s.base.to = in.pts[head + 1];
                s.kind = K_LINE;
                head += 1;
            } else if (head + 2 < in.n) {
                if (in.flags[head + 1] == CONTROL_0 && in.flags[head + 2] == ON_CURVE) {
                    s.ctrl.from = in.pts[head + 1];
                    s.base.to = in.pts[head + 2];
                    s.kind = K_QUAD;
                    head += 2;
                } else if (head + 3 < in.n && in.flags[head + 1] == CONTROL_0 && in.flags[head + 2] == CONTROL_1 &&
                           in.flags[head + 3] == ON_CURVE) {
                    s.ctrl.from = in.pts[head + 1];
                    s.ctrl.to = in.pts[head + 2];
                    s.base.to = in.pts[head + 3];
                    s.kind = K_CUBIC;
                    head += 3;
                }
            }
        } else {
            if (in.closed && in.n > 1 && !approx_eq(in.pts[0], in.pts[in.n - 1], FLOAT_EPSILON)) {
                s.base.from = in.pts[in.n - 1];
                s.base.to = in.pts[0];
                s.kind = K_LINE;
            }
            end = true;
        }
        index++;
        if (s.kind == K_NONE) break;
        offset_segment(s, -radius, index == 0 ? JOIN_BEVEL : join, miter, out);
        if (end) break;
    }
}

// core/stroke.cpp:54-122: the closing line (reversed) first, then the segments backwards
void offset_backward(const InContour &in, float radius, int32_t join, float miter, Contour &out) {
    int tail = in.n - 1, index = -1;
    if (in.closed && in.n > 1 && !approx_eq(in.pts[0], in.pts[in.n - 1], FLOAT_EPSILON)) {
        index++;
        offset_segment(Seg::line(Line{in.pts[0], in.pts[in.n - 1]}), -radius, index == 0 ? JOIN_BEVEL : join, miter,
                       out);
    }
    while (tail >= 0) {
        if (in.flags[tail] != ON_CURVE) break;
        Seg s;
        if (tail >= 3 && in.flags[tail - 1] == CONTROL_1 && in.flags[tail - 2] == CONTROL_0 &&
            in.flags[tail - 3] == ON_CURVE) {
            s.base.from = in.pts[tail];
            s.ctrl.from = in.pts[tail - 1];
            s.ctrl.to = in.pts[tail - 2];
            s.base.to = in.pts[tail - 3];
            s.kind = K_CUBIC;
            tail -= 3;
        } else if (tail >= 2 && in.flags[tail - 1] == CONTROL_0 && in.flags[tail - 2] == ON_CURVE) {
            s.base.from = in.pts[tail];
            s.ctrl.from = in.pts[tail - 1];
            s.base.to = in.pts[tail - 2];
            s.kind = K_QUAD;
            tail -= 2;
        } else if (tail >= 1 && in.flags[tail - 1] == ON_CURVE) {
            s.base.from = in.pts[tail];
            s.base.to = in.pts[tail - 1];
            s.kind = K_LINE;
            tail -= 1;
        } else {
            break;
        }
        index++;
        offset_segment(s, -radius, index == 0 ? JOIN_BEVEL : join, miter, out);
    }
}

// core/stroke.cpp:197-249
void add_cap(Contour &c, const pfo_stroke_style &st) {
    if (st.line_cap == CAP_BUTT || c.pts.size() < 2) return;
    const float width = st.line_width;
    const V p1 = c.last(1);
    V p0;
    size_t i0 = c.pts.size() - 2;
    while (true) {
        p0 = c.pts[i0];
        if (sq_len(p1 - p0) > FLOAT_EPSILON) break;
        if (i0 == 0) return;
        i0--;
    }
    const V g = normalize(p1 - p0);
    if (st.line_cap == CAP_SQUARE) {
        const V off = g * (width * 0.5f);
        const V p2 = p1 + off, p3 = p2 + yx(g) * V{-width, width}, p4 = p3 - off;
        c.push(p2, ON_CURVE);
        c.push(p3, ON_CURVE);
        c.push(p4, ON_CURVE);
    } else {
        const float scale = width * 0.5f;
        const V off = yx(g) * V{-1.0f, 1.0f};
        const V tr = p1 + off * (width * 0.5f);
        const Xform x = Xform::scale(V{scale, scale}).translate(tr);
        c.push_arc_from_unit_chord(x, Line{-off, off});
    }
}

// core/stroke.cpp:169-195
void push_stroked(std::vector<Contour> &out, Contour &c, const InContour &in, bool closed,
                  const pfo_stroke_style &st) {
    if (closed && c.might_need_join(st.line_join)) {
        const Line final_segment{c.pts[1], c.pts[0]};
        c.add_join(st.line_width * 0.5f, st.line_join, in.pts[0], final_segment, st.miter_limit);
    }
    c.closed = true;
    out.push_back(c);
}

// core/data/segment.cpp:168-187: the length estimate the dasher walks by (a line's length; for a
// curve, the mean of its chord and its control polygon's length, a quadratic raised to a cubic first,
// core/data/segment.cpp:141-151)
float arc_length(const Seg &s) {
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

// SegmentsIter (core/data/contour.cpp:107-177) with force_closed = false: a contour's segments in
// order, then its closing line when it is closed and does not end where it starts (within 1e-4);
// a malformed point pattern ends the iteration with a None segment.
struct SegIter {
    const V *p;
    const uint8_t *f;
    int n;
    bool closed;
    int head = 0;
    bool has_next = true;
    Seg next() {
        Seg s;
        if (head >= n || f[head] != ON_CURVE) return s;
        s.base.from = p[head];
        if (head + 1 < n) {
            if (f[head + 1] == ON_CURVE) {
                s.base.to = p[head + 1];
                s.kind = K_LINE;
                head += 1;
            } else if (head + 2 < n) {
                if (f[head + 1] == CONTROL_0 && f[head + 2] == ON_CURVE) {
                    s.ctrl.from = p[head + 1];
                    s.base.to = p[head + 2];
                    s.kind = K_QUAD;
                    head += 2;
                } else if (head + 3 < n) {
                    if (f[head + 1] == CONTROL_0 && f[head + 2] == CONTROL_1 && f[head + 3] == ON_CURVE) {
                        s.ctrl.from = p[head + 1];
                        s.ctrl.to = p[head + 2];
                        s.base.to = p[head + 3];
                        s.kind = K_CUBIC;
                        head += 3;
                    }
                }
            }
        } else {
            if (closed && n > 1 && !approx_eq(p[0], p[n - 1], FLOAT_EPSILON)) {
                s.base.from = p[n - 1];
                s.base.to = p[0];
                s.kind = K_LINE;
            }
            has_next = false;
        }
        return s;
    }
};

// OutlineDash (core/dash.cpp:9-126): DashState carried across the contours of the outline, each
// contour walked by ContourDash::dash; the last dash is pushed by into_outline when it is on.
// Returns false when the walk exceeds max_steps (a pattern of dashes far shorter than the outline).
bool dash_outline(const pfo_outline &in, const float *dashes, int nd, float offset, std::vector<Contour> &out,
                  uint64_t max_steps) {
    float total = 0;
    for (int i = 0; i < nd; i++) total += dashes[i];
    offset = std::fmod(offset, total);
    int idx = 0;
    while (idx < nd) {
        const float d = dashes[idx];
        if (offset < d) break;
        offset -= d;
        idx += 1;
    }
    // Rounding in the subtractions can walk past the last dash (the reference then indexes one past
    // its pattern); the pattern restarts instead.
    if (idx >= nd) idx = 0;
    float left = offset;
    Contour cur;
    uint64_t steps = 0;
    const V *P = reinterpret_cast<const V *>(in.points);
    for (int c = 0; c < in.n_contours; c++) {
        const int a = in.contour_start[c];
        SegIter it{P + a, in.point_flags + a, in.contour_start[c + 1] - a, in.contour_closed[c] != 0};
        Seg queued;
        bool queued_none = true;
        while (true) {
            if (++steps > max_steps) return false;
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
            if (idx % 2 == 0) cur.push_segment(seg);
            left -= distance;
            if (left < FLOAT_EPSILON) {
                if (idx % 2 == 0) {
                    if (!cur.pts.empty()) out.push_back(cur);  // Outline::push_contour skips empty ones
                    cur = Contour{};
                }
                idx += 1;
                if (idx == nd) idx = 0;
                left = dashes[idx];
            }
        }
    }
    if (idx % 2 == 0 && !cur.pts.empty()) out.push_back(cur);
    return true;
}

}  // namespace

// Dashes of an outline (OutlineDash::dash + into_outline, core/dash.cpp:48-126): open contours.
// The pattern must be finite, non-negative and of positive total, the offset finite and
// non-negative.  0, -1 bad input, -4 the walk ran past 2^24 steps, -5 a capacity was too small
// (n_points_out / n_contours_out hold the sizes).
extern "C" int pfo_dash_outline(const pfo_outline *in, const float *dashes, int32_t n_dashes, float offset,
                                float *points_out, uint8_t *flags_out, int64_t cap_points, int32_t *contour_start_out,
                                uint8_t *contour_closed_out, int64_t cap_contours, int64_t *n_points_out,
                                int64_t *n_contours_out) {
    if (!in || !n_points_out || !n_contours_out || in->n_contours < 0 || in->n_points < 0 || n_dashes < 1 || !dashes)
        return -1;
    if ((in->n_points && (!in->points || !in->point_flags)) || !in->contour_start ||
        (in->n_contours && !in->contour_closed))
        return -1;
    if (in->contour_start[0] != 0 || in->contour_start[in->n_contours] != in->n_points) return -1;
    float total = 0;
    for (int i = 0; i < n_dashes; i++) {
        if (!std::isfinite(dashes[i]) || dashes[i] < 0.0f) return -1;
        total += dashes[i];
    }
    if (!(total > 0.0f) || !std::isfinite(total) || !std::isfinite(offset) || offset < 0.0f) return -1;
    try {
        std::vector<Contour> out;
        if (!dash_outline(*in, dashes, n_dashes, offset, out, 1ull << 24)) return -4;
        int64_t np = 0;
        for (const Contour &k : out) np += (int64_t)k.pts.size();
        *n_points_out = np;
        *n_contours_out = (int64_t)out.size();
        if (np > cap_points || (int64_t)out.size() > cap_contours) return -5;
        if ((np && (!points_out || !flags_out)) || !contour_start_out || (!out.empty() && !contour_closed_out))
            return -1;
        int64_t o = 0;
        contour_start_out[0] = 0;
        for (size_t k = 0; k < out.size(); k++) {
            std::memcpy(points_out + 2 * o, out[k].pts.data(), out[k].pts.size() * sizeof(V));
            std::memcpy(flags_out + o, out[k].flags.data(), out[k].flags.size());
            o += (int64_t)out[k].pts.size();
            contour_start_out[k + 1] = (int32_t)o;
            contour_closed_out[k] = 0;
        }
        return 0;
    } catch (...) {
        return -3;
    }
}

extern "C" int pfo_stroke_outline(const pfo_outline *in, const pfo_stroke_style *style, float *points_out,
                                  uint8_t *flags_out, int64_t cap_points, int32_t *contour_start_out,
                                  uint8_t *contour_closed_out, int64_t cap_contours, int64_t *n_points_out,
                                  int64_t *n_contours_out) {
    if (!in || !style || !n_points_out || !n_contours_out || in->n_contours < 0 || in->n_points < 0)
        return -1;
    if ((in->n_points && (!in->points || !in->point_flags)) || !in->contour_start ||
        (in->n_contours && !in->contour_closed))
        return -1;
    if (in->contour_start[0] != 0 || in->contour_start[in->n_contours] != in->n_points) return -1;
    for (int c = 0; c < in->n_contours; c++)
        if (in->contour_start[c + 1] < in->contour_start[c]) return -1;
    if (style->line_cap < CAP_BUTT || style->line_cap > CAP_ROUND || style->line_join < JOIN_MITER ||
        style->line_join > JOIN_ROUND || !(style->line_width >= 0.0f))
        return -1;
    try {
        std::vector<Contour> out;
        const V *P = reinterpret_cast<const V *>(in->points);
        const float radius = style->line_width * 0.5f;
        for (int c = 0; c < in->n_contours; c++) {
            const int a = in->contour_start[c];
            const InContour ic{P + a, in->point_flags + a, in->contour_start[c + 1] - a, in->contour_closed[c] != 0};
            Contour k;
            offset_forward(ic, radius, style->line_join, style->miter_limit, k);
            if (ic.closed) {
                push_stroked(out, k, ic, true, *style);
                k = Contour{};
            } else {
                add_cap(k, *style);
            }
            offset_backward(ic, radius, style->line_join, style->miter_limit, k);
            if (!ic.closed) add_cap(k, *style);
            push_stroked(out, k, ic, ic.closed, *style);
        }
        int64_t np = 0;
        for (const Contour &k : out) np += (int64_t)k.pts.size();
        *n_points_out = np;
        *n_contours_out = (int64_t)out.size();
        if (np > cap_points || (int64_t)out.size() > cap_contours) return -5;
        if ((np && (!points_out || !flags_out)) || !contour_start_out || (!out.empty() && !contour_closed_out))
            return -1;
        int64_t o = 0;
        contour_start_out[0] = 0;
        for (size_t k = 0; k < out.size(); k++) {
            const Contour &ct = out[k];
            std::memcpy(points_out + 2 * o, ct.pts.data(), ct.pts.size() * sizeof(V));
            std::memcpy(flags_out + o, ct.flags.data(), ct.flags.size());
            o += (int64_t)ct.pts.size();
            contour_start_out[k + 1] = (int32_t)o;
            contour_closed_out[k] = ct.closed ? 1 : 0;
        }
        return 0;
    } catch (...) {
        return -3;
    }
}
