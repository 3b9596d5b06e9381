// Drives the reference-side binding end to end: a Scene built through the reference's own API
// (Scene, Path2d, Paint, ClipPath with a nested clip, a stroke made by the reference's
// OutlineStrokeToFill as Canvas::stroke_path does, linear and radial gradient paints), packed by
// SceneBuilderHip and drawn by
// RendererHip on a DeviceHip / QueueHip into a TextureHip, read back with
// CommandEncoder::read_texture.
//   check pack <desc file>                 packs the scene (no device needed)
//   check draw <desc file> <image file>    also draws it and reads the image back
// The desc file holds the packed pfh_scene_desc (int32 header, then its arrays); the image file
// the RGBA8 pixels.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "core/data/path.h"
#include "core/paint/gradient.h"
#include "core/path2d.h"
#include "core/scene.h"
#include "core/stroke.h"
#include "hip_gpu.h"
#include "renderer_hip.h"

using namespace Pathfinder;

namespace {

constexpr int W = 300, H = 220;

std::shared_ptr<Scene> make_scene() {
    auto scene = std::make_shared<Scene>(0, RectF(0, 0, W, H));
    const uint32_t red = scene->push_paint(Paint::from_color(ColorU(230, 40, 40, 255)));
    const uint32_t blue = scene->push_paint(Paint::from_color(ColorU(40, 80, 220, 160)));
    const uint32_t green = scene->push_paint(Paint::from_color(ColorU(30, 180, 60, 200)));
    // clip paths: an outer rect, and a circle-ish cubic blob inside it clipped by the rect
    ClipPath outer;
    {
        Path2d p;
        p.move_to(20.5f, 15.25f);
        p.line_to(270.75f, 15.25f);
        p.line_to(270.75f, 200.5f);
        p.line_to(20.5f, 200.5f);
        p.close_path();
        outer.outline = p.into_outline();
        outer.fill_rule = FillRule::Winding;
    }
    const uint32_t outer_id = scene->push_clip_path(outer);
    ClipPath inner;
    {
        Path2d p;
        p.move_to(240.0f, 110.0f);
        p.cubic_to(240.0f, 170.0f, 60.0f, 190.0f, 40.0f, 110.0f);
        p.cubic_to(20.0f, 30.0f, 240.0f, 50.0f, 240.0f, 110.0f);
        p.close_path();
        inner.outline = p.into_outline();
        inner.fill_rule = FillRule::EvenOdd;
        inner.clip_path = std::make_shared<uint32_t>(outer_id);  // nested: clipped by the outer rect
    }
    const uint32_t inner_id = scene->push_clip_path(inner);
    // a filled star (even-odd) under the inner clip
    {
        Path2d p;
        p.move_to(150.0f, 10.0f);
        p.line_to(205.0f, 190.0f);
        p.line_to(55.0f, 75.0f);
        p.line_to(245.0f, 75.0f);
        p.line_to(95.0f, 190.0f);
        p.close_path();
        DrawPath dp;
        dp.outline = p.into_outline();
        dp.paint = (uint16_t)red;
        dp.fill_rule = FillRule::EvenOdd;
        dp.clip_path = std::make_shared<uint32_t>(inner_id);
        scene->push_draw_path(dp);
    }
    // a quadratic blob, unclipped, nonzero, multiplied onto what lies below (DrawPath::blend_mode)
    {
        Path2d p;
        p.move_to(10.0f, 120.0f);
        p.quadratic_to(80.0f, 20.0f, 150.0f, 120.0f);
        p.quadratic_to(80.0f, 215.0f, 10.0f, 120.0f);
        p.close_path();
        DrawPath dp;
        dp.outline = p.into_outline();
        dp.paint = (uint16_t)blue;
        dp.blend_mode = BlendMode::Multiply;
        scene->push_draw_path(dp);
    }
    // a stroked open cubic polyline (round caps and joins), as Canvas::stroke_path makes it
    {
        Path2d p;
        p.move_to(30.0f, 30.0f);
        p.cubic_to(90.0f, 0.0f, 120.0f, 90.0f, 170.0f, 60.0f);
        p.cubic_to(220.0f, 30.0f, 250.0f, 140.0f, 280.0f, 180.0f);
        StrokeStyle style;
        style.line_width = 9.5f;
        style.line_cap = LineCap::Round;
        style.line_join = LineJoin::Round;
        const Outline line = p.into_outline();  // the stroker keeps a reference to its input
        OutlineStrokeToFill stroker(line, style);
        stroker.offset();
        DrawPath dp;
        dp.outline = stroker.into_outline();
        dp.paint = (uint16_t)green;
        dp.fill_rule = FillRule::Winding;
        dp.clip_path = std::make_shared<uint32_t>(outer_id);
        scene->push_draw_path(dp);
    }
    // a band with a linear gradient (two of its stops at one offset, one translucent)
    {
        Gradient g = Gradient::linear(LineSegmentF(Vec2F(40.0f, 180.0f), Vec2F(260.0f, 200.0f)));
        g.add_color_stop(ColorU(255, 200, 0, 255), 0.0f);
        g.add_color_stop(ColorU(0, 120, 255, 255), 0.5f);
        g.add_color_stop(ColorU(255, 0, 120, 200), 0.5f);
        g.add_color_stop(ColorU(20, 20, 20, 255), 1.0f);
        const uint32_t paint = scene->push_paint(Paint::from_gradient(g));
        Path2d p;
        p.move_to(40.0f, 160.0f);
        p.line_to(260.0f, 160.0f);
        p.line_to(260.0f, 212.0f);
        p.line_to(40.0f, 212.0f);
        p.close_path();
        DrawPath dp;
        dp.outline = p.into_outline();
        dp.paint = (uint16_t)paint;
        scene->push_draw_path(dp);
    }
    // a blob with a repeating radial gradient in a rotated, scaled gradient space
    {
        Gradient g = Gradient::radial(LineSegmentF(Vec2F(0.0f, 0.0f), Vec2F(6.0f, -4.0f)), Vec2F(4.0f, 40.0f));
        g.wrap = GradientWrap::Repeat;
        std::get<GradientRadial>(g.geometry).transform =
            Transform2(Mat2(0.8f, 0.6f, -0.6f, 0.8f) * 1.25f, Vec2F(205.0f, 80.0f));
        g.add_color_stop(ColorU(250, 250, 250, 255), 0.0f);
        g.add_color_stop(ColorU(120, 0, 160, 255), 0.7f);
        g.add_color_stop(ColorU(0, 90, 40, 255), 1.0f);
        const uint32_t paint = scene->push_paint(Paint::from_gradient(g));
        Path2d p;
        p.move_to(160.0f, 80.0f);
        p.cubic_to(160.0f, 30.0f, 250.0f, 30.0f, 250.0f, 80.0f);
        p.cubic_to(250.0f, 130.0f, 160.0f, 130.0f, 160.0f, 80.0f);
        p.close_path();
        DrawPath dp;
        dp.outline = p.into_outline();
        dp.paint = (uint16_t)paint;
        scene->push_draw_path(dp);
    }
    return scene;
}

void write_desc(const pfh_scene_desc &d, const char *path) {
    FILE *f = std::fopen(path, "wb");
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    const int32_t head[7] = {d.width, d.height, d.n_paths, d.n_contours, d.n_points, d.n_clip_paths, 0};
    std::fwrite(head, 4, 7, f);
    std::fwrite(d.view_box, 4, 4, f);
    std::fwrite(d.points, 8, d.n_points, f);
    std::fwrite(d.point_flags, 1, d.n_points, f);
    std::fwrite(d.contour_start, 4, d.n_contours + 1, f);
    std::fwrite(d.contour_closed, 1, d.n_contours, f);
    std::fwrite(d.path_contour_start, 4, d.n_paths + 1, f);
    std::fwrite(d.path_fill_rule, 1, d.n_paths, f);
    std::fwrite(d.path_rgba, 4, d.n_paths, f);
    std::fwrite(d.path_clip, 4, d.n_paths, f);
    // gradient paints: counts, then per path its gradient (-1: none), the records and the stops
    const int32_t gh[2] = {d.path_gradient ? d.n_gradients : 0, d.path_gradient ? d.n_gradient_stops : 0};
    std::fwrite(gh, 4, 2, f);
    if (d.path_gradient) {
        std::fwrite(d.path_gradient, 4, d.n_paths, f);
        std::fwrite(d.gradients, sizeof(pfh_gradient), d.n_gradients, f);
        std::fwrite(d.gradient_stop_offsets, 4, d.n_gradient_stops, f);
        std::fwrite(d.gradient_stop_rgba, 4, d.n_gradient_stops, f);
        const int32_t has_ramps = d.gradient_ramps ? 1 : 0;
        std::fwrite(&has_ramps, 4, 1, f);
        if (d.gradient_ramps) std::fwrite(d.gradient_ramps, 4, (size_t)d.n_gradients * 256, f);
    }
    // blend modes: a flag, then per path its BlendMode
    const int32_t has_blend = d.path_blend ? 1 : 0;
    std::fwrite(&has_blend, 4, 1, f);
    if (d.path_blend) std::fwrite(d.path_blend, 1, d.n_paths, f);
    std::fclose(f);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: check pack <desc> | check draw <desc> <image>\n");
        return 2;
    }
    try {
        const std::string mode = argv[1];
        auto scene = make_scene();
        auto builder = std::make_shared<SceneBuilderHip>();
        builder->build(scene.get(), nullptr);
        write_desc(builder->desc(), argv[2]);
        if (mode == "pack") return 0;
        if (mode != "draw" || argc < 4) return 2;
        auto device = std::make_shared<DeviceHip>(0);
        auto queue = std::make_shared<QueueHip>(0);
        auto renderer = std::make_shared<RendererHip>(device, queue, 0);
        auto dest = device->create_texture({{W, H}, TextureFormat::Rgba8Unorm}, "destination");
        renderer->set_dest_texture(dest);
        renderer->draw(builder, true);
        std::vector<uint8_t> pixels((size_t)W * H * 4);
        auto encoder = device->create_command_encoder("read back");
        encoder->read_texture(dest, {}, pixels.data());
        queue->submit(encoder, renderer->fence);
        FILE *f = std::fopen(argv[3], "wb");
        if (!f) return 3;
        std::fwrite(pixels.data(), 1, pixels.size(), f);
        std::fclose(f);
        std::printf("drew %d paths into a %dx%d TextureHip\n", builder->desc().n_paths, W, H);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "check: %s\n", e.what());
        return 1;
    }
}
