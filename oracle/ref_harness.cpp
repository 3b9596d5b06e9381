// Test infrastructure only: drives the reference's own hybrid-mode CPU tiler
// (core/d3d9/tiler.cpp, object_builder.cpp, scene_builder.cpp), compiled from
// its unchanged sources by oracle/build_ref.sh, so that the oracle
// restatement can be checked against it.  Nothing in the product links this.
//
// Exports the same canonical form as oracle/hybrid_oracle.cpp (pfo_*):
//   per path: tile rect, per-tile (backdrop, alpha flag), per-column final
//   backdrop sum, fills in emission order with their tile-local index;
//   per framebuffer tile: draw entries in composite order after the z-cull.
#include <cstdint>
#include <cstring>
#include <chrono>
#include <thread>
#include <string>
#include <vector>

#include "core/d3d9/scene_builder.h"
#include "core/d3d9/tiler.h"
#include "core/data/built_path.h"
#include "core/scene.h"
#include "core/stroke.h"
#include "core/dash.h"
#include "core/path2d.h"

#include "core/paint/gradient.h"
#include "core/paint/texture_allocator.h"
#include <umHalf.h>

// The reference's vendored SVG parser (third_party/nanosvg.h), compiled here as core/svg.cpp
// compiles it.
#define NANOSVG_IMPLEMENTATION
#include <nanosvg.h>

namespace Pathfinder {
// Defined with external linkage in core/d3d9/scene_builder.cpp:18 (not in a header).
std::vector<DrawTileBatchD3D9> build_tile_batches_for_draw_path_display_item(const Scene &scene,
                                                                             const std::vector<BuiltDrawPath> &built_paths,
                                                                             Range draw_path_range);
} // namespace Pathfinder

using namespace Pathfinder;

namespace {

struct HarnessBuilder : public SceneBuilderD3D9 {
    void attach(Scene *s) {
        scene = s;
    }
};

struct RefResult {
    std::vector<int32_t> path_rect;
    std::vector<int64_t> tile_off, col_off, fill_off;
    std::vector<int8_t> tile_backdrop;
    std::vector<uint8_t> tile_alpha;
    std::vector<int32_t> col_sum;
    std::vector<uint16_t> fill_seg;
    std::vector<uint32_t> fill_tile;
    std::vector<uint32_t> entry_fbtile, entry_path;
    std::vector<uint32_t> zbuf;
    int32_t fb_rect[4] = {0, 0, 0, 0};
    std::vector<uint8_t> combine;  // per tile: the reference queued a clip combine job for it
    std::vector<int8_t> comb_bd;   // that job's dest_backdrop
};

Outline make_outline(const float *pts, const uint8_t *pflags, const int32_t *contour_start, const uint8_t *closed,
                     int c0, int c1) {
    Outline outline;
    for (int c = c0; c < c1; c++) {
        Contour contour;
        for (int i = contour_start[c]; i < contour_start[c + 1]; i++) {
            PointFlag flag = pflags[i] == 0 ? PointFlag::ON_CURVE_POINT
                                            : (pflags[i] == 1 ? PointFlag::CONTROL_POINT_0 : PointFlag::CONTROL_POINT_1);
            contour.push_point(Vec2F(pts[2 * i], pts[2 * i + 1]), flag, true);
        }
        contour.closed = closed[c] != 0;
        outline.push_contour(contour);
    }
    return outline;
}

} // namespace

extern "C" {

// Tiles every draw path of one scene with the reference Tiler.
// threads > 1 reproduces SceneBuilderD3D9::build_paths_on_cpu's interleaved
// split (core/d3d9/scene_builder.cpp:170-194); the canonical output does not
// depend on it.  Returns an opaque handle, or nullptr.
//
// Clip paths: the last n_clip paths are Scene::clip_paths, tiled first as in
// SceneBuilderD3D9::build_paths_on_cpu (core/d3d9/scene_builder.cpp:137-166, 200-225);
// path_clip[p] (NULL: none) is draw path p's clip path index among them, or -1.  A clipped
// draw path is tiled twice: without its clip for its fills and their tiles (clip resolution
// changes only tiles, tiler.cpp:385-431, and can drop a tile's alpha id), and with it for the
// resolved tiles and the clip combine jobs.  Scenes with clip paths are tiled on one thread.
void *pfref_run(int n_paths, const int32_t *path_contour_start, const uint8_t *path_fill_rule,
                const uint8_t *path_opaque, const int32_t *contour_start, const uint8_t *contour_closed,
                const float *points, const uint8_t *point_flags, const float *view_box, int threads,
                double *elapsed_s, int n_clip, const int32_t *path_clip, const uint8_t *path_blend) {
    auto *res = new RefResult();
    Scene scene(0, RectF(view_box[0], view_box[1], view_box[2], view_box[3]));
    HarnessBuilder builder;
    builder.attach(&scene);
    for (auto &n : builder.next_alpha_tile_indices) n = 0;
    RectF vb = scene.get_view_box();

    std::vector<Outline> outlines(n_paths);
    for (int p = 0; p < n_paths; p++) {
        outlines[p] = make_outline(points, point_flags, contour_start, contour_closed, path_contour_start[p],
                                   path_contour_start[p + 1]);
    }

    const int n_draw = n_paths - n_clip;
    std::vector<BuiltDrawPath> built(n_paths);
    std::vector<std::vector<Fill>> fills(n_paths);
    std::vector<BuiltPath> built_clips(n_clip);
    for (int q = 0; q < n_clip; q++) {
        int g = n_draw + q;
        TilingPathInfo info{};
        info.type = TilingPathInfo::Type::Clip;
        FillRule rule = path_fill_rule[g] ? FillRule::EvenOdd : FillRule::Winding;
        Tiler tiler(builder, (uint32_t)q, outlines[g], rule, vb, nullptr, {}, info);
        tiler.generate_tiles();
        built_clips[q] = tiler.object_builder.built_path;
        built[g].path = tiler.object_builder.built_path;
        fills[g] = std::move(tiler.object_builder.fills);
    }
    std::vector<BuiltPath> resolved(n_draw);
    if (n_clip > 0) threads = 1;
    auto work = [&](int begin, int stride) {
        for (int p = begin; p < n_draw; p += stride) {
            TilingPathInfo info{};
            info.type = TilingPathInfo::Type::Draw;
            info.info.paint_id = 0;
            // path_blend (NULL: src-over): BlendMode values, as DrawPath::blend_mode
            const BlendMode mode = path_blend ? (BlendMode)path_blend[p] : BlendMode::SrcOver;
            info.info.blend_mode = mode;
            info.info.fill_rule = path_fill_rule[p] ? FillRule::EvenOdd : FillRule::Winding;
            Tiler tiler(builder, (uint32_t)p, outlines[p], info.info.fill_rule, vb, nullptr, {}, info);
            tiler.generate_tiles();
            built[p].path = tiler.object_builder.built_path;
            // BuiltDrawPath's rule (core/data/built_path.cpp:40), the paint's opacity given
            built[p].occludes = path_opaque[p] != 0 && blend_mode_occludes_backdrop(mode);
            built[p].blend_mode = mode;
            built[p].mask_fill_rule = info.info.fill_rule;
            fills[p] = std::move(tiler.object_builder.fills);
            if (path_clip && path_clip[p] >= 0) {
                auto clip_id = std::make_shared<uint32_t>((uint32_t)path_clip[p]);
                Tiler clipped(builder, (uint32_t)p, outlines[p], info.info.fill_rule, vb, clip_id, built_clips, info);
                clipped.generate_tiles();
                resolved[p] = clipped.object_builder.built_path;
            }
        }
    };
    auto t0 = std::chrono::steady_clock::now();
    if (threads <= 1) {
        work(0, 1);
    } else {
        std::vector<std::thread> pool;
        for (int i = 0; i < threads; i++) pool.emplace_back(work, i, threads);
        for (auto &t : pool) t.join();
    }
    auto t1 = std::chrono::steady_clock::now();
    if (elapsed_s) *elapsed_s = std::chrono::duration<double>(t1 - t0).count();

    int64_t tsum = 0, csum = 0, fsum = 0;
    res->tile_off.push_back(0);
    res->col_off.push_back(0);
    res->fill_off.push_back(0);
    for (int p = 0; p < n_paths; p++) {
        const BuiltPath &bp = built[p].path;
        const RectI &r = bp.data.tiles.rect;
        res->path_rect.insert(res->path_rect.end(), {r.left, r.top, r.right, r.bottom});
        const auto &tiles = bp.data.tiles.data;
        const bool clipped = p < n_draw && path_clip && path_clip[p] >= 0;
        const auto &out_tiles = clipped ? resolved[p].data.tiles.data : tiles;
        for (size_t i = 0; i < out_tiles.size(); i++) {
            res->tile_backdrop.push_back(out_tiles[i].backdrop);
            res->tile_alpha.push_back(out_tiles[i].alpha_tile_id.is_valid() ? 1 : 0);
            uint8_t comb = 0;
            int8_t cbd = 0;
            if (clipped && resolved[p].data.clip_tiles) {
                const auto &ct = resolved[p].data.clip_tiles->data[i];
                if (ct.dest_tile_id.is_valid() && ct.src_tile_id.is_valid()) {
                    comb = 1;
                    cbd = (int8_t)ct.dest_backdrop;
                }
            }
            res->combine.push_back(comb);
            res->comb_bd.push_back(cbd);
        }
        for (auto v : bp.data.backdrops) res->col_sum.push_back(v);
        for (const Fill &f : fills[p]) {
            res->fill_seg.insert(res->fill_seg.end(), {f.line_segment.from_x, f.line_segment.from_y,
                                                       f.line_segment.to_x, f.line_segment.to_y});
            uint32_t local = 0xffffffffu;
            for (size_t i = 0; i < tiles.size(); i++) {
                if (tiles[i].alpha_tile_id.value == f.link) {
                    local = (uint32_t)i;
                    break;
                }
            }
            res->fill_tile.push_back(local);
        }
        tsum += (int64_t)tiles.size();
        csum += (int64_t)bp.data.backdrops.size();
        fsum += (int64_t)fills[p].size();
        res->tile_off.push_back(tsum);
        res->col_off.push_back(csum);
        res->fill_off.push_back(fsum);
    }

    // Draw order and z-buffer through the reference's own batch builder.
    if (n_draw > 0) {
        std::vector<BuiltDrawPath> draws(built.begin(), built.begin() + n_draw);
        for (int p = 0; p < n_draw; p++)
            if (path_clip && path_clip[p] >= 0) draws[p].path = resolved[p];
        auto batches = build_tile_batches_for_draw_path_display_item(scene, draws, Range(0, (uint64_t)n_draw));
        if (!batches.empty()) {
            const auto &batch = batches[0];
            const RectI &zr = batch.z_buffer_data.rect;
            res->fb_rect[0] = zr.left;
            res->fb_rect[1] = zr.top;
            res->fb_rect[2] = zr.right;
            res->fb_rect[3] = zr.bottom;
            res->zbuf.assign(batch.z_buffer_data.data.begin(), batch.z_buffer_data.data.end());
            int w = zr.width();
            int64_t nfb = (int64_t)zr.width() * zr.height();
            // Stable grouping by framebuffer tile keeps batch (draw) order, which is
            // the order the reference rasterizes and blends the tile instances in.
            std::vector<std::vector<uint32_t>> lists(nfb);
            for (const auto &t : batch.tiles) {
                int64_t fb = (int64_t)(t.tile_y - zr.top) * w + (t.tile_x - zr.left);
                if ((int32_t)t.path_id < (int32_t)batch.z_buffer_data.data[fb]) continue; // tile.vert:114
                lists[fb].push_back(t.path_id);
            }
            for (int64_t fb = 0; fb < nfb; fb++) {
                for (auto pid : lists[fb]) {
                    res->entry_fbtile.push_back((uint32_t)fb);
                    res->entry_path.push_back(pid);
                }
            }
        }
    }
    return res;
}

// Wall time of tiling every draw path `repeats` times with the reference Tiler
// (no canonical output): the CPU baseline.
double pfref_time(int n_paths, const int32_t *path_contour_start, const uint8_t *path_fill_rule,
                  const int32_t *contour_start, const uint8_t *contour_closed, const float *points,
                  const uint8_t *point_flags, const float *view_box, int threads, int repeats) {
    Scene scene(0, RectF(view_box[0], view_box[1], view_box[2], view_box[3]));
    HarnessBuilder builder;
    builder.attach(&scene);
    RectF vb = scene.get_view_box();
    std::vector<Outline> outlines(n_paths);
    for (int p = 0; p < n_paths; p++)
        outlines[p] = make_outline(points, point_flags, contour_start, contour_closed, path_contour_start[p],
                                   path_contour_start[p + 1]);
    std::vector<BuiltPath> built(n_paths);
    std::vector<size_t> nfills(n_paths);
    auto work = [&](int begin, int stride) {
        for (int p = begin; p < n_paths; p += stride) {
            TilingPathInfo info{};
            info.type = TilingPathInfo::Type::Draw;
            info.info.blend_mode = BlendMode::SrcOver;
            info.info.fill_rule = path_fill_rule[p] ? FillRule::EvenOdd : FillRule::Winding;
            Tiler tiler(builder, (uint32_t)p, outlines[p], info.info.fill_rule, vb, nullptr, {}, info);
            tiler.generate_tiles();
            built[p] = tiler.object_builder.built_path;
            nfills[p] = tiler.object_builder.fills.size();
        }
    };
    auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < repeats; r++) {
        for (auto &n : builder.next_alpha_tile_indices) n = 0;
        std::vector<std::thread> pool;
        for (int i = 0; i < threads; i++) pool.emplace_back(work, i, threads);
        for (auto &t : pool) t.join();
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// counts[0..5] = tiles, columns, fills, entries, fb tiles, paths
void pfref_counts(void *h, int64_t *counts) {
    auto *r = (RefResult *)h;
    counts[0] = (int64_t)r->tile_backdrop.size();
    counts[1] = (int64_t)r->col_sum.size();
    counts[2] = (int64_t)r->fill_tile.size();
    counts[3] = (int64_t)r->entry_path.size();
    counts[4] = (int64_t)r->zbuf.size();
    counts[5] = (int64_t)r->path_rect.size() / 4;
}

void pfref_fetch(void *h, int32_t *path_rect, int64_t *tile_off, int64_t *col_off, int64_t *fill_off,
                 int8_t *tile_backdrop, uint8_t *tile_alpha, int32_t *col_sum, uint16_t *fill_seg,
                 uint32_t *fill_tile, uint32_t *entry_fbtile, uint32_t *entry_path, uint32_t *zbuf,
                 int32_t *fb_rect) {
    auto *r = (RefResult *)h;
    auto cp = [](void *dst, const void *src, size_t bytes) {
        if (bytes) memcpy(dst, src, bytes);
    };
    cp(path_rect, r->path_rect.data(), r->path_rect.size() * 4);
    cp(tile_off, r->tile_off.data(), r->tile_off.size() * 8);
    cp(col_off, r->col_off.data(), r->col_off.size() * 8);
    cp(fill_off, r->fill_off.data(), r->fill_off.size() * 8);
    cp(tile_backdrop, r->tile_backdrop.data(), r->tile_backdrop.size());
    cp(tile_alpha, r->tile_alpha.data(), r->tile_alpha.size());
    cp(col_sum, r->col_sum.data(), r->col_sum.size() * 4);
    cp(fill_seg, r->fill_seg.data(), r->fill_seg.size() * 2);
    cp(fill_tile, r->fill_tile.data(), r->fill_tile.size() * 4);
    cp(entry_fbtile, r->entry_fbtile.data(), r->entry_fbtile.size() * 4);
    cp(entry_path, r->entry_path.data(), r->entry_path.size() * 4);
    cp(zbuf, r->zbuf.data(), r->zbuf.size() * 4);
    cp(fb_rect, r->fb_rect, 16);
}

// Per tile: whether the reference queued a clip combine job for it, and the job's dest_backdrop.
void pfref_fetch_clip(void *h, uint8_t *combine, int8_t *comb_bd) {
    auto *r = (RefResult *)h;
    if (!r->combine.empty()) {
        memcpy(combine, r->combine.data(), r->combine.size());
        memcpy(comb_bd, r->comb_bd.data(), r->comb_bd.size());
    }
}

void pfref_free(void *h) {
    delete (RefResult *)h;
}

// The reference's stroke-to-fill (OutlineStrokeToFill::offset + into_outline, core/stroke.cpp:124-195)
// on one outline; fetch with pfref_stroke_counts / pfref_stroke_fetch, free with pfref_stroke_free.
struct RefStroke {
    std::vector<float> points;
    std::vector<uint8_t> flags;
    std::vector<int32_t> contour_start{0};
    std::vector<uint8_t> closed;
};

void *pfref_stroke(const float *points, const uint8_t *point_flags, const int32_t *contour_start,
                   const uint8_t *contour_closed, int n_contours, float width, int cap, int join, float miter_limit) {
    Outline in = make_outline(points, point_flags, contour_start, contour_closed, 0, n_contours);
    StrokeStyle style;
    style.line_width = width;
    style.line_cap = (LineCap)cap;
    style.line_join = (LineJoin)join;
    style.miter_limit = miter_limit;
    OutlineStrokeToFill stroker(in, style);
    stroker.offset();
    Outline out = stroker.into_outline();
    auto *r = new RefStroke();
    for (const Contour &c : out.contours) {
        for (size_t i = 0; i < c.points.size(); i++) {
            r->points.push_back(c.points[i].x);
            r->points.push_back(c.points[i].y);
            r->flags.push_back((uint8_t)c.flags[i]);
        }
        r->contour_start.push_back((int32_t)(r->flags.size()));
        r->closed.push_back(c.closed ? 1 : 0);
    }
    return r;
}

void pfref_stroke_counts(void *h, int64_t *counts) {
    auto *r = (RefStroke *)h;
    counts[0] = (int64_t)r->flags.size();
    counts[1] = (int64_t)r->closed.size();
}

void pfref_stroke_fetch(void *h, float *points, uint8_t *flags, int32_t *contour_start, uint8_t *closed) {
    auto *r = (RefStroke *)h;
    if (!r->flags.empty()) {
        memcpy(points, r->points.data(), r->points.size() * sizeof(float));
        memcpy(flags, r->flags.data(), r->flags.size());
    }
    memcpy(contour_start, r->contour_start.data(), r->contour_start.size() * sizeof(int32_t));
    if (!r->closed.empty()) memcpy(closed, r->closed.data(), r->closed.size());
}

void pfref_stroke_free(void *h) {
    delete (RefStroke *)h;
}

// The reference's dasher (OutlineDash::dash + into_outline, core/dash.cpp:48-126) on one outline, then,
// when width > 0, its stroke-to-fill of the dashed outline (the order of Canvas::stroke_path,
// core/canvas.cpp:283-295); fetch with pfref_stroke_counts / pfref_stroke_fetch, free with
// pfref_stroke_free.
void *pfref_dash(const float *points, const uint8_t *point_flags, const int32_t *contour_start,
                 const uint8_t *contour_closed, int n_contours, const float *dashes, int n_dashes, float offset,
                 float width, int cap, int join, float miter_limit) {
    Outline in = make_outline(points, point_flags, contour_start, contour_closed, 0, n_contours);
    OutlineDash dasher(in, std::vector<float>(dashes, dashes + n_dashes), offset);
    dasher.dash();
    Outline out = dasher.into_outline();
    if (width > 0.0f) {
        StrokeStyle style;
        style.line_width = width;
        style.line_cap = (LineCap)cap;
        style.line_join = (LineJoin)join;
        style.miter_limit = miter_limit;
        OutlineStrokeToFill stroker(out, style);
        stroker.offset();
        out = stroker.into_outline();
    }
    auto *r = new RefStroke();
    for (const Contour &c : out.contours) {
        for (size_t i = 0; i < c.points.size(); i++) {
            r->points.push_back(c.points[i].x);
            r->points.push_back(c.points[i].y);
            r->flags.push_back((uint8_t)c.flags[i]);
        }
        r->contour_start.push_back((int32_t)(r->flags.size()));
        r->closed.push_back(c.closed ? 1 : 0);
    }
    return r;
}

// The reference's SVG ingest (SvgScene::SvgScene, core/svg.cpp:128-206): nanosvg parses the text
// ("px", 96 dpi) into shapes whose paths are cubic chains; each shape becomes a Path2d (move_to,
// cubic_to, close_path), filled when its fill paint is visible (Canvas::fill_path,
// core/canvas.cpp:265-270) and stroked when its stroke paint is visible and its width > 0, dashed
// first with offset 0 when it has a dash array (Canvas::stroke_path, core/canvas.cpp:272-301).
// The Canvas itself is not linked (it pulls in the GPU renderers); these are its steps for an
// identity transform, no clip, no shadow.  Per draw path: the outline before stroking, the style,
// and the outline the scene receives.
struct RefSvg {
    float size[2] = {0, 0};
    std::vector<int32_t> path_i32;  // kind (0 fill, 1 stroke), fill rule, paint type, rgba, cap, join, dash start, count
    std::vector<float> path_f32;    // width, miter limit
    // gradient paints: per draw path an index or -1; per gradient kind, wrap, stop start, stop count
    // and line (4), radii (2), transform (6); stops (offset, rgba) in the order added
    std::vector<int32_t> path_grad, grad_i32, stop_rgba;
    std::vector<float> grad_f32, stop_off;
    std::vector<float> dashes;
    RefStroke src, out;
    std::vector<int32_t> src_pcs{0}, out_pcs{0};
};

static void append_outline(RefStroke &dst, std::vector<int32_t> &pcs, const Outline &o) {
    for (const Contour &c : o.contours) {
        for (size_t i = 0; i < c.points.size(); i++) {
            dst.points.push_back(c.points[i].x);
            dst.points.push_back(c.points[i].y);
            dst.flags.push_back((uint8_t)c.flags[i]);
        }
        dst.contour_start.push_back((int32_t)dst.flags.size());
        dst.closed.push_back(c.closed ? 1 : 0);
    }
    pcs.push_back((int32_t)dst.closed.size());
}

// convert_nsvg_paint's gradient branch (core/svg.cpp:75-121): a linear gradient's ends through the
// path and gradient transforms, a radial one's centre-to-focus line with radii (0, r) and the
// transforms kept as its transform; pad -> clamp, repeat -> repeat, reflect left at clamp.
static int32_t add_gradient(RefSvg &r, const NSVGpaint &p) {
    const NSVGgradient *ng = p.gradient;
    auto gradient_xform = Transform2((float *)ng->xform2);
    auto path_xform = Transform2((float *)ng->xform3);
    Gradient g;
    if (p.type == NSVG_PAINT_LINEAR_GRADIENT) {
        Vec2F from = {ng->x1, ng->y1}, to = {ng->x2, ng->y2};
        from = path_xform * gradient_xform * from;
        to = path_xform * gradient_xform * to;
        g = Gradient::linear(LineSegmentF(from, to));
    } else {
        Vec2F from = {ng->cx, ng->cy}, to = {ng->fx, ng->fy};
        g = Gradient::radial(LineSegmentF(from, to), Vec2F(0.0, ng->r));
        std::get<GradientRadial>(g.geometry).transform = path_xform * gradient_xform;
    }
    const int32_t wrap = ng->spread == NSVG_SPREAD_REPEAT ? 1 : 0;
    const int32_t start = (int32_t)r.stop_off.size();
    for (int i = 0; i < ng->nstops; i++) {
        r.stop_off.push_back(ng->stops[i].offset);
        r.stop_rgba.push_back((int32_t)ng->stops[i].color);
    }
    float f[12] = {0};
    int32_t kind = 0;
    if (auto *lin = std::get_if<GradientLinear>(&g.geometry)) {
        f[0] = lin->line.from().x, f[1] = lin->line.from().y, f[2] = lin->line.to().x, f[3] = lin->line.to().y;
        f[6] = 1, f[9] = 1;
    } else {
        const auto &rad = std::get<GradientRadial>(g.geometry);
        kind = 1;
        f[0] = rad.line.from().x, f[1] = rad.line.from().y, f[2] = rad.line.to().x, f[3] = rad.line.to().y;
        f[4] = rad.radii.x, f[5] = rad.radii.y;
        f[6] = rad.transform.m11(), f[7] = rad.transform.m21(), f[8] = rad.transform.m12();
        f[9] = rad.transform.m22(), f[10] = rad.transform.m13(), f[11] = rad.transform.m23();
    }
    r.grad_i32.insert(r.grad_i32.end(), {kind, wrap, start, ng->nstops});
    r.grad_f32.insert(r.grad_f32.end(), f, f + 12);
    return (int32_t)(r.grad_i32.size() / 4) - 1;
}

void *pfref_svg(const char *text, int64_t len) {
    std::string copy(text, (size_t)len);  // nsvgParse edits its input
    NSVGimage *image = nsvgParse((char *)copy.data(), "px", 96);
    if (!image) return nullptr;
    auto *r = new RefSvg();
    r->size[0] = image->width;
    r->size[1] = image->height;
    // convert_nsvg_paint (core/svg.cpp:64-124) for visibility: a colour paint is visible when its
    // alpha is not 0 (ColorU::is_visible); a gradient's base colour is opaque white
    auto paint_type = [](const NSVGpaint &p) { return (int32_t)p.type; };
    // a colour paint's rgba (a gradient paint's union holds its gradient instead: 0)
    auto colour_of = [](const NSVGpaint &p) { return p.type == NSVG_PAINT_COLOR ? (int32_t)p.color : 0; };
    auto visible = [](const NSVGpaint &p) {
        if (p.type == NSVG_PAINT_COLOR) return (p.color >> 24) != 0;
        return p.type == NSVG_PAINT_LINEAR_GRADIENT || p.type == NSVG_PAINT_RADIAL_GRADIENT;
    };
    for (NSVGshape *sh = image->shapes; sh; sh = sh->next) {
        Path2d path;
        for (NSVGpath *np = sh->paths; np; np = np->next) {
            path.move_to(np->pts[0], np->pts[1]);
            for (int i = 0; i < np->npts - 3; i += 3) {
                const float *q = &np->pts[i * 2];
                path.cubic_to(q[2], q[3], q[4], q[5], q[6], q[7]);
            }
            if (np->closed) path.close_path();
        }
        if (visible(sh->fill)) {
            Outline o = path.into_outline();
            append_outline(r->src, r->src_pcs, o);
            append_outline(r->out, r->out_pcs, o);
            const int32_t rule = sh->fillRule == NSVG_FILLRULE_EVENODD ? 1 : 0;
            r->path_i32.insert(r->path_i32.end(), {0, rule, paint_type(sh->fill), colour_of(sh->fill), 0, 0, 0, 0});
            r->path_grad.push_back(sh->fill.type == NSVG_PAINT_COLOR ? -1 : add_gradient(*r, sh->fill));
            r->path_f32.insert(r->path_f32.end(), {0.0f, 0.0f});
        }
        const float width = sh->strokeWidth;
        if (visible(sh->stroke) && width > 0.0f) {
            const int32_t cap = sh->strokeLineCap == NSVG_CAP_ROUND ? 2 : sh->strokeLineCap == NSVG_CAP_SQUARE ? 1 : 0;
            const int32_t join =
                sh->strokeLineJoin == NSVG_JOIN_ROUND ? 2 : sh->strokeLineJoin == NSVG_JOIN_BEVEL ? 1 : 0;
            Outline o = path.into_outline();
            append_outline(r->src, r->src_pcs, o);
            const int32_t d0 = (int32_t)r->dashes.size(), dn = sh->strokeDashCount;
            if (dn > 0) {
                std::vector<float> pattern(sh->strokeDashArray, sh->strokeDashArray + dn);
                r->dashes.insert(r->dashes.end(), pattern.begin(), pattern.end());
                OutlineDash dasher(o, pattern, 0);
                dasher.dash();
                o = dasher.into_outline();
            }
            StrokeStyle style;
            style.line_width = width;
            style.line_cap = (LineCap)cap;
            style.line_join = (LineJoin)join;
            style.miter_limit = sh->miterLimit;
            OutlineStrokeToFill stroker(o, style);
            stroker.offset();
            append_outline(r->out, r->out_pcs, stroker.into_outline());
            r->path_i32.insert(r->path_i32.end(),
                               {1, 0, paint_type(sh->stroke), colour_of(sh->stroke), cap, join, d0, dn});
            r->path_grad.push_back(sh->stroke.type == NSVG_PAINT_COLOR ? -1 : add_gradient(*r, sh->stroke));
            r->path_f32.insert(r->path_f32.end(), {width, sh->miterLimit});
        }
    }
    nsvgDelete(image);
    return r;
}

// The reference's gradient ramp: a Gradient with the stops added in the order given
// (Gradient::add_color_stop), sampled into a texture row by GradientTileBuilder::allocate
// (core/paint/gradient.cpp:86-123); writes its 256 texels as r | g << 8 | b << 16 | a << 24.
void pfref_gradient_ramp(const float *offsets, const uint32_t *rgba, int n, uint32_t *out) {
    Gradient g = Gradient::linear(LineSegmentF(Vec2F(0, 0), Vec2F(1, 0)));
    for (int i = 0; i < n; i++) g.add_color_stop(ColorU(rgba[i]), offsets[i]);
    TextureAllocator allocator;
    GradientTileBuilder builder;
    auto loc = builder.allocate(g, allocator, {});
    const auto &tile = builder.tiles.back();
    for (int x = 0; x < (int)GRADIENT_TILE_LENGTH; x++) {
        const ColorU c = tile.texels[loc.rect.min_y() * GRADIENT_TILE_LENGTH + x];
        out[x] = (uint32_t)c.r_ | (uint32_t)c.g_ << 8 | (uint32_t)c.b_ << 16 | (uint32_t)c.a_ << 24;
    }
}

// The metadata texture's float-to-half conversion (third_party/umHalf, as core/renderer.cpp:167-250
// uses it): the half's bits.
void pfref_half_bits(const float *in, int64_t n, uint16_t *out) {
    for (int64_t i = 0; i < n; i++) {
        half h = in[i];
        uint16_t bits;
        memcpy(&bits, &h, 2);
        out[i] = bits;
    }
}

// Transform2::inverse (common/math/transform2.cpp:32-36) of m11, m21, m12, m22, m13, m23.
void pfref_transform_inverse(const float *t, float *out) {
    Transform2 x(Mat2(t[0], t[1], t[2], t[3]), Vec2F(t[4], t[5]));
    Transform2 i = x.inverse();
    const float v[6] = {i.m11(), i.m21(), i.m12(), i.m22(), i.m13(), i.m23()};
    memcpy(out, v, sizeof(v));
}

// [paths, source points, source contours, output points, output contours, dash lengths]
void pfref_svg_counts(void *h, int64_t *counts) {
    auto *r = (RefSvg *)h;
    counts[0] = (int64_t)r->path_f32.size() / 2;
    counts[1] = (int64_t)r->src.flags.size();
    counts[2] = (int64_t)r->src.closed.size();
    counts[3] = (int64_t)r->out.flags.size();
    counts[4] = (int64_t)r->out.closed.size();
    counts[5] = (int64_t)r->dashes.size();
}

void pfref_svg_fetch(void *h, float *size, int32_t *path_i32, float *path_f32, float *dashes, int32_t *src_pcs,
                     float *src_points, uint8_t *src_flags, int32_t *src_cs, uint8_t *src_closed, int32_t *out_pcs,
                     float *out_points, uint8_t *out_flags, int32_t *out_cs, uint8_t *out_closed) {
    auto *r = (RefSvg *)h;
    size[0] = r->size[0];
    size[1] = r->size[1];
    auto cp = [](void *dst, const void *src, size_t n) {
        if (n) memcpy(dst, src, n);
    };
    cp(path_i32, r->path_i32.data(), r->path_i32.size() * 4);
    cp(path_f32, r->path_f32.data(), r->path_f32.size() * 4);
    cp(dashes, r->dashes.data(), r->dashes.size() * 4);
    cp(src_pcs, r->src_pcs.data(), r->src_pcs.size() * 4);
    pfref_stroke_fetch(&r->src, src_points, src_flags, src_cs, src_closed);
    cp(out_pcs, r->out_pcs.data(), r->out_pcs.size() * 4);
    pfref_stroke_fetch(&r->out, out_points, out_flags, out_cs, out_closed);
}

// [gradients, stops]
void pfref_svg_gradient_counts(void *h, int64_t *counts) {
    auto *r = (RefSvg *)h;
    counts[0] = (int64_t)r->grad_i32.size() / 4;
    counts[1] = (int64_t)r->stop_off.size();
}

void pfref_svg_gradient_fetch(void *h, int32_t *path_grad, int32_t *grad_i32, float *grad_f32, float *stop_off,
                              int32_t *stop_rgba) {
    auto *r = (RefSvg *)h;
    auto cp = [](void *dst, const void *src, size_t n) {
        if (n) memcpy(dst, src, n);
    };
    cp(path_grad, r->path_grad.data(), r->path_grad.size() * 4);
    cp(grad_i32, r->grad_i32.data(), r->grad_i32.size() * 4);
    cp(grad_f32, r->grad_f32.data(), r->grad_f32.size() * 4);
    cp(stop_off, r->stop_off.data(), r->stop_off.size() * 4);
    cp(stop_rgba, r->stop_rgba.data(), r->stop_rgba.size() * 4);
}

void pfref_svg_free(void *h) {
    delete (RefSvg *)h;
}

} // extern "C"
