// Test infrastructure only: a CPU restatement of the reference's gradient paints, the checker of
// the library's gradient ramps, records and composite.  Nothing in the product links it.
//
//   pfo_half_value       the metadata texture's float-to-half conversion (third_party/umHalf.inl:49-111)
//   pfo_gradient_ramp    Gradient::add + Gradient::sample + GradientTileBuilder::allocate
//                        (core/paint/gradient.cpp:7-56, 86-123)
//   pfo_gradient_record  Palette::calculate_texture_transforms (core/paint/palette.cpp:398-440),
//                        PaintMetadata::filter (core/paint/paint.cpp:45-65), through the half texture
//   pfo_paint_colour     the tile shader's colour for one pixel (shaders/d3d9/tile.vert:65-77,
//                        tile.frag:296-323, 590-653): the colour texture coordinate and the radial
//                        filter in float, the bilinear ramp sample in double
//   pfo_composite_paint  pfo_composite (hybrid_oracle.cpp) with gradient paths
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// The half's value: exponent bits below 1 give zero (2^-24 and below) or a denormal, the mantissa
// keeps its top 10 bits (no rounding), past 2^15 an infinity.
double half_value(float v) {
    uint32_t bits;
    std::memcpy(&bits, &v, 4);
    const bool neg = (bits >> 31) != 0;
    const int exp = (int)((bits >> 23) & 0xff);
    const uint32_t mant = bits & 0x7fffff;
    double mag;
    if (exp == 0xff) {
        mag = mant ? NAN : INFINITY;
    } else if (exp == 0) {
        mag = 0.0;
    } else {
        const int e = exp - 127;
        if (e < -24) {
            mag = 0.0;
        } else if (e < -14) {
            // denormal half: the implicit one and the mantissa shifted down, in units of 2^-24
            const int shift = -14 - e;
            const uint32_t frac = (1024u >> shift) + (mant >> (13 + shift));
            mag = std::ldexp((double)frac, -24);
        } else if (e > 15) {
            mag = INFINITY;
        } else {
            mag = std::ldexp(1.0 + (double)(mant >> 13) / 1024.0, e);
        }
    }
    return neg ? -mag : mag;
}

struct Stop {
    float offset;
    uint32_t rgba;
};

uint8_t chan(uint32_t c, int k) { return (uint8_t)(c >> (8 * k)); }

uint32_t sample(const std::vector<Stop> &stops, float t) {
    if (stops.empty()) return 0;
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    size_t i = 0;
    while (i < stops.size() && (stops[i].offset < t || stops[i].offset == 0)) i++;
    const size_t hi = i < stops.size() - 1 ? i : stops.size() - 1;
    const size_t lo = hi > 0 ? hi - 1 : hi;
    const float denom = stops[hi].offset - stops[lo].offset;
    if (denom == 0) return stops[lo].rgba;
    float r = (t - stops[lo].offset) / denom;
    if (r > 1.0f) r = 1.0f;
    if (r < 0.0f) r = 0.0f;
    uint32_t out = 0;
    const float inv = 1.f / 255.f;
    for (int k = 0; k < 4; k++) {
        const float a = (float)chan(stops[lo].rgba, k) * inv, b = (float)chan(stops[hi].rgba, k) * inv;
        const float m = a + (b - a) * r;
        out |= (uint32_t)(uint8_t)(m * 255.f) << (8 * k);
    }
    return out;
}

// shaders/d3d9/tile.frag:444-583: B(dest, src) of composite op `op` per channel (dest, src in
// [0, 1]), through HSL for hue (12), saturation (13), color (14), luminosity (15).
double dodge(double d, double s) { return d == 0.0 ? 0.0 : (s == 1.0 ? 1.0 : d / (1.0 - s)); }
double screen(double d, double s) { return d + s - d * s; }
double hard_light(double d, double s) { return s <= 0.5 ? d * 2.0 * s : screen(d, 2.0 * s - 1.0); }
double soft_light(double d, double s) {
    const double dark = d <= 0.25 ? ((16.0 * d - 12.0) * d + 4.0) * d : std::sqrt(d);
    return d + (s * 2.0 - 1.0) * (s <= 0.5 ? d * (1.0 - d) : dark - d);
}
double safe_div(double n, double d) { return d != 0.0 ? n / d : 0.0; }
// https://en.wikipedia.org/wiki/HSL_and_HSV: hue in radians (pi/3 per sector), saturation c / v
void rgb_to_hsl(const double *c, double *hsl) {
    const double v = std::max(std::max(c[0], c[1]), c[2]), mn = std::min(std::min(c[0], c[1]), c[2]);
    const double ch = v - mn;
    double sector, a, b;
    if (c[0] == v) { sector = 0.0; a = c[1]; b = c[2]; }
    else if (c[1] == v) { sector = 2.0; a = c[2]; b = c[0]; }
    else { sector = 4.0; a = c[0]; b = c[1]; }
    hsl[0] = (M_PI / 3.0) * safe_div(sector * ch + a - b, ch);
    hsl[1] = safe_div(ch, v);
    hsl[2] = 0.5 * (mn + v);
}
void hsl_to_rgb(const double *hsl, double *c) {
    const double a = hsl[1] * std::min(hsl[2], 1.0 - hsl[2]);
    const double n[3] = {0.0, 8.0, 4.0};
    for (int i = 0; i < 3; i++) {
        const double t = n[i] + hsl[0] * (6.0 / M_PI);
        const double k = t - 12.0 * std::floor(t / 12.0);
        c[i] = hsl[2] - std::min(std::max(std::min(k - 3.0, 9.0 - k), -1.0), 1.0) * a;
    }
}
void composite_op(const double *d, const double *s, int op, double *out) {
    if (op >= 12) {
        double hd[3], hs[3], h[3];
        rgb_to_hsl(d, hd);
        rgb_to_hsl(s, hs);
        h[0] = (op == 12 || op == 14) ? hs[0] : hd[0];
        h[1] = (op == 13 || op == 14) ? hs[1] : hd[1];
        h[2] = op == 15 ? hs[2] : hd[2];
        hsl_to_rgb(h, out);
        return;
    }
    for (int i = 0; i < 3; i++) {
        const double dc = d[i], sc = s[i];
        double r = sc;
        switch (op) {
            case 1: r = dc * sc; break;
            case 2: r = screen(dc, sc); break;
            case 3: r = hard_light(sc, dc); break;
            case 4: r = std::min(dc, sc); break;
            case 5: r = std::max(dc, sc); break;
            case 6: r = dodge(dc, sc); break;
            case 7: r = 1.0 - dodge(1.0 - dc, 1.0 - sc); break;
            case 8: r = hard_light(dc, sc); break;
            case 9: r = soft_light(dc, sc); break;
            case 10: r = std::fabs(dc - sc); break;
            case 11: r = dc + sc - 2.0 * dc * sc; break;
        }
        out[i] = r;
    }
}

}  // namespace

extern "C" {

// One composite op on single colours (for tests): out = B(dest, src) of `op`.
void pfo_composite_op(const double *dest, const double *src, int op, double *out) { composite_op(dest, src, op, out); }


double pfo_half_value(float v) { return half_value(v); }

// Stops in the order given, each inserted ahead of the first stop whose offset is not below its own.
void pfo_gradient_ramp(const float *offsets, const uint32_t *rgba, int n, uint32_t *ramp) {
    std::vector<Stop> stops;
    for (int i = 0; i < n; i++) {
        auto it = stops.begin();
        while (it != stops.end() && it->offset < offsets[i]) ++it;
        stops.insert(it, Stop{offsets[i], rgba[i]});
    }
    for (int x = 0; x < 256; x++) ramp[x] = sample(stops, ((float)x + 0.5f) / 256.0f);
}

// rec: m11 m21 m12 m22 m13 m23, line from x y, vector x y, radii, uv origin x y, kind (1 linear,
// 2 radial), wrap -- every value as the half texture holds it.  xform: the radial gradient's
// transform (m11 m21 m12 m22 m13 m23).  Returns 0, or -1 for a radial transform without inverse.
int pfo_gradient_record(int kind, int wrap, const float *line, const float *radii, const float *xform, int row,
                        float *rec) {
    float m[6];
    for (int i = 0; i < 16; i++) rec[i] = 0.0f;
    const float page = 1.f / 256.f;
    if (kind == 0) {
        // t = (p - from) . d / |d|^2 as a transform: m0 = d / |d|^2, offset m0 . (-from); v the row centre
        const float dx = line[2] - line[0], dy = line[3] - line[1];
        const float len2 = dx * dx + dy * dy;
        const float m0x = dx / len2, m0y = dy / len2;
        m[0] = m0x;
        m[1] = 0.0f;
        m[2] = m0y;
        m[3] = 0.0f;
        m[4] = m0x * -line[0] + m0y * -line[1];
        m[5] = ((float)row + (float)(row + 1)) * 0.5f * page;
    } else {
        const float det = xform[0] * xform[3] - xform[1] * xform[2];
        if (det == 0.0f) return -1;
        const float inv = 1.0f / det;
        m[0] = xform[3] * inv;
        m[1] = -xform[1] * inv;
        m[2] = -xform[2] * inv;
        m[3] = xform[0] * inv;
        m[4] = -(m[0] * xform[4] + m[2] * xform[5]);
        m[5] = -(m[1] * xform[4] + m[3] * xform[5]);
        const float p[8] = {line[0], line[1], line[2] - line[0], line[3] - line[1], radii[0], radii[1], 0.0f,
                            (float)row * page + page * 0.5f};
        for (int i = 0; i < 8; i++) rec[6 + i] = (float)half_value(p[i]);
    }
    for (int i = 0; i < 6; i++) rec[i] = (float)half_value(m[i]);
    rec[14] = kind == 0 ? 1.0f : 2.0f;
    rec[15] = (float)wrap;
    return 0;
}

// The paint's un-premultiplied colour at pixel centre (px, py), channels in [0, 1]; 0 where the
// paint leaves the pixel transparent (a radial discriminant <= 0, a non-finite coordinate).
int pfo_paint_colour(const float *rec, const uint32_t *ramp, float px, float py, double *out) {
    const float u = rec[0] * px + rec[2] * py + rec[4];
    float s = u;
    if (rec[14] == 2.0f) {
        const float v = rec[1] * px + rec[3] * py + rec[5];
        const float fx = rec[6], fy = rec[7], cx = rec[8], cy = rec[9], r0 = rec[10], r1 = rec[11];
        const float dpx = u - fx, dpy = v - fy, dr = r1 - r0;
        const float a = cx * cx + cy * cy - dr * dr;
        const float b = dpx * cx + dpy * cy + r0 * dr;
        const float c = dpx * dpx + dpy * dpy - r0 * r0;
        const float disc = b * b - a * c;
        if (!(disc > 0.0f)) return 0;
        const float q = std::sqrt(disc);
        float t0 = (q + b) / a, t1 = (b - q) / a;
        if (t0 > t1) std::swap(t0, t1);
        s = rec[12] + (t0 >= 0.0f ? t0 : t1);
    }
    if (!std::isfinite(s)) return 0;
    const bool repeat = rec[15] != 0.0f;
    const double su = repeat ? (double)s - std::floor((double)s) : std::fmin(std::fmax((double)s, -1.0), 2.0);
    const double cxy = su * 256.0 - 0.5;
    const double fl = std::floor(cxy), f = cxy - fl;
    long i0 = (long)fl, i1 = i0 + 1;
    if (repeat) {
        i0 = ((i0 % 256) + 256) % 256;
        i1 = ((i1 % 256) + 256) % 256;
    } else {
        i0 = i0 < 0 ? 0 : (i0 > 255 ? 255 : i0);
        i1 = i1 < 0 ? 0 : (i1 > 255 ? 255 : i1);
    }
    for (int k = 0; k < 4; k++) {
        const double a = chan(ramp[i0], k), b = chan(ramp[i1], k);
        out[k] = (a + (b - a) * f) / 255.0;
    }
    return 1;
}

// FP64 premultiplied composite (pfo_composite's model) where a path with path_paint[p] >= 0 takes
// its colour per pixel from paint path_paint[p] (records [16 floats], ramps [256 texels]).
double pfo_apply_rule(double c, int even_odd);

// path_op (NULL: none): per path the tile shader's composite op (COMBINER_CTRL_COMPOSITE_*, 0 over),
// applied as composite() does (shaders/d3d9/tile.frag:585-598) against the colour composited so far.
void pfo_composite_paint(int64_t n_entries, const uint32_t *entry_fbtile, const uint32_t *entry_path,
                         const int64_t *entry_alpha, const int8_t *entry_backdrop, const uint8_t *path_even_odd,
                         const uint32_t *path_rgba, const int32_t *path_paint, const float *recs,
                         const uint32_t *ramps, const double *coverage, int fb_w, int width, int height,
                         double *out, const uint8_t *path_op, int fb_x0, int fb_y0) {
    std::fill(out, out + (size_t)width * height * 4, 0.0);
    for (int64_t e = 0; e < n_entries; e++) {
        const uint32_t fb = entry_fbtile[e];
        // its position on the target: the index in the view box's tile grid from the grid's origin
        const int tx = (int)(fb % (uint32_t)fb_w) + fb_x0, ty = (int)(fb / (uint32_t)fb_w) + fb_y0;
        const uint32_t p = entry_path[e];
        const int32_t paint = path_paint ? path_paint[p] : -1;
        const int op = path_op ? path_op[p] : 0;
        const uint32_t c = path_rgba[p];
        double col[4] = {(c & 0xff) / 255.0, ((c >> 8) & 0xff) / 255.0, ((c >> 16) & 0xff) / 255.0,
                         ((c >> 24) & 0xff) / 255.0};
        const double solid_cov = pfo_apply_rule((double)entry_backdrop[e], path_even_odd[p]);
        for (int py = 0; py < 16; py++) {
            const int y = ty * 16 + py;
            if (y < 0) continue;
            if (y >= height) break;
            for (int px = 0; px < 16; px++) {
                const int x = tx * 16 + px;
                if (x < 0) continue;
                if (x >= width) break;
                if (paint >= 0 && !pfo_paint_colour(recs + 16 * (size_t)paint, ramps + 256 * (size_t)paint,
                                                    (float)x + 0.5f, (float)y + 0.5f, col)) {
                    if (op == 0) continue;
                    col[0] = col[1] = col[2] = col[3] = 0.0;  // transparent where the paint leaves out
                }
                const double cov = entry_alpha[e] >= 0 ? coverage[entry_alpha[e] * 256 + py * 16 + px] : solid_cov;
                const double sa = col[3] * cov;
                double *d = out + ((size_t)y * width + x) * 4;
                if (op != 0) {
                    double bl[3];
                    composite_op(d, col, op, bl);
                    for (int ch = 0; ch < 3; ch++)
                        d[ch] = sa * (1.0 - d[3]) * col[ch] + sa * d[3] * bl[ch] + (1.0 - sa) * d[ch];
                    d[3] = 1.0;
                    continue;
                }
                const double k = 1.0 - sa;
                d[0] = col[0] * sa + d[0] * k;
                d[1] = col[1] * sa + d[1] * k;
                d[2] = col[2] * sa + d[2] * k;
                d[3] = sa + d[3] * k;
            }
        }
    }
}

}  // extern "C"
