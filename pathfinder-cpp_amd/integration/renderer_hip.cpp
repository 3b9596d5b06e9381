// SceneBuilderHip / RendererHip (see renderer_hip.h).
#include "renderer_hip.h"

#include <algorithm>
#include <stdexcept>
#include <variant>

#include "common/logger.h"
#include "hip_gpu.h"

namespace Pathfinder {

void SceneBuilderHip::build(Scene *scene_, Renderer *) {
    scene = scene_;
    pts_.clear();
    flags_.clear();
    closed_.clear();
    rule_.clear();
    rgba_.clear();
    clip_.clear();
    grad_of_path_.clear();
    blend_.clear();
    grads_.clear();
    ramps_.clear();
    cstart_.assign(1, 0);
    pstart_.assign(1, 0);
    auto add = [&](const Outline &outline, FillRule fill_rule, uint32_t rgba, const std::shared_ptr<uint32_t> &clip) {
        for (const Contour &c : outline.contours) {
            for (size_t i = 0; i < c.points.size(); i++) {
                pts_.push_back(c.points[i].x);
                pts_.push_back(c.points[i].y);
                flags_.push_back((uint8_t)c.flags[i]);  // PointFlag: 0 on-curve, 1 / 2 control points
            }
            cstart_.push_back((int32_t)flags_.size());
            closed_.push_back(c.closed ? 1 : 0);
        }
        pstart_.push_back((int32_t)closed_.size());
        rule_.push_back(fill_rule == FillRule::EvenOdd ? PFH_FILL_EVEN_ODD : PFH_FILL_NONZERO);
        rgba_.push_back(rgba);
        clip_.push_back(clip ? (int32_t)*clip : -1);
    };
    // a gradient paint (Paint::from_gradient: its overlay holds the Gradient) as a pfh_gradient whose ramp
    // is sampled here with the Gradient's own sample(), as GradientTileBuilder::allocate does
    // (core/paint/gradient.cpp:115-119; the stops themselves are private to Gradient)
    auto gradient_of = [&](const Paint &paint) -> int32_t {
        const auto overlay = paint.get_overlay();
        if (!overlay || !std::holds_alternative<Gradient>(overlay->contents)) return -1;
        const Gradient &g = std::get<Gradient>(overlay->contents);
        pfh_gradient r{};
        r.wrap = g.wrap == GradientWrap::Repeat ? PFH_WRAP_REPEAT : PFH_WRAP_CLAMP;
        r.transform[0] = r.transform[3] = 1.0f;
        if (const auto *lin = std::get_if<GradientLinear>(&g.geometry)) {
            r.kind = PFH_GRADIENT_LINEAR;
            r.line[0] = lin->line.from().x, r.line[1] = lin->line.from().y;
            r.line[2] = lin->line.to().x, r.line[3] = lin->line.to().y;
        } else {
            const auto &rad = std::get<GradientRadial>(g.geometry);
            r.kind = PFH_GRADIENT_RADIAL;
            r.line[0] = rad.line.from().x, r.line[1] = rad.line.from().y;
            r.line[2] = rad.line.to().x, r.line[3] = rad.line.to().y;
            r.radii[0] = rad.radii.x, r.radii[1] = rad.radii.y;
            const float t[6] = {rad.transform.m11(), rad.transform.m21(), rad.transform.m12(),
                                rad.transform.m22(), rad.transform.m13(), rad.transform.m23()};
            std::copy(t, t + 6, r.transform);
        }
        r.stop_start = 0;
        r.stop_count = -1;
        for (uint32_t x = 0; x < GRADIENT_TILE_LENGTH; x++) {
            const ColorU c = g.sample(((float)x + 0.5f) / GRADIENT_TILE_LENGTH);
            ramps_.push_back((uint32_t)c.r_ | (uint32_t)c.g_ << 8 | (uint32_t)c.b_ << 16 | (uint32_t)c.a_ << 24);
        }
        grads_.push_back(r);
        return (int32_t)grads_.size() - 1;
    };
    for (const DrawPath &dp : scene->draw_paths) {
        const Paint paint = scene->palette.get_paint(dp.paint);
        const ColorU c = paint.get_base_color();
        add(dp.outline, dp.fill_rule, (uint32_t)c.r_ | (uint32_t)c.g_ << 8 | (uint32_t)c.b_ << 16 | (uint32_t)c.a_ << 24,
            dp.clip_path);
        grad_of_path_.push_back(gradient_of(paint));
        blend_.push_back((uint8_t)dp.blend_mode);  // BlendMode, the order of PFH_BLEND_*
    }
    for (const ClipPath &cp : scene->clip_paths) {
        add(cp.outline, cp.fill_rule, 0, cp.clip_path);  // never drawn
        grad_of_path_.push_back(-1);
        blend_.push_back((uint8_t)BlendMode::SrcOver);
    }
    const RectF vb = scene->get_view_box();
    desc_ = pfh_scene_desc{};
    desc_.width = (int32_t)std::ceil(vb.width());
    desc_.height = (int32_t)std::ceil(vb.height());
    desc_.view_box[0] = vb.left;
    desc_.view_box[1] = vb.top;
    desc_.view_box[2] = vb.right;
    desc_.view_box[3] = vb.bottom;
    desc_.n_paths = (int32_t)rule_.size();
    desc_.n_contours = (int32_t)closed_.size();
    desc_.n_points = (int32_t)flags_.size();
    desc_.points = pts_.data();
    desc_.point_flags = flags_.data();
    desc_.contour_start = cstart_.data();
    desc_.contour_closed = closed_.data();
    desc_.path_contour_start = pstart_.data();
    desc_.path_fill_rule = rule_.data();
    desc_.path_rgba = rgba_.data();
    desc_.n_clip_paths = (int32_t)scene->clip_paths.size();
    desc_.path_clip = clip_.data();
    desc_.clip_rule = clip_rule_;
    desc_.path_blend = blend_.data();
    desc_.path_stroke = nullptr;
    desc_.stroke_dashes = nullptr;
    desc_.n_stroke_dashes = 0;
    if (!grads_.empty()) {
        desc_.path_gradient = grad_of_path_.data();
        desc_.gradients = grads_.data();
        desc_.n_gradients = (int32_t)grads_.size();
        desc_.gradient_ramps = ramps_.data();
    }
}

RendererHip::RendererHip(const std::shared_ptr<Device> &device, const std::shared_ptr<Queue> &queue, int hip_device)
    : Renderer(device, queue) {
    if (pfh_create(hip_device, &ctx_) != PFH_OK) throw std::runtime_error("pfh_create failed");
}

RendererHip::~RendererHip() {
    if (scene_id_ >= 0) pfh_scene_release(ctx_, scene_id_);
    pfh_destroy(ctx_);
}

bool RendererHip::check(int rc) {
    if (rc == PFH_OK) return true;
    Logger::error(pfh_last_error(ctx_));  // the reference's error style: log and return
    return false;
}

void RendererHip::draw(const std::shared_ptr<SceneBuilder> &scene_builder, bool clear_dst_texture) {
    auto *builder = dynamic_cast<SceneBuilderHip *>(scene_builder.get());
    if (!builder) {
        Logger::error("RendererHip draws scenes built by a SceneBuilderHip");
        return;
    }
    clear_dest_texture = clear_dst_texture;
    if (scene_id_ >= 0) pfh_scene_release(ctx_, scene_id_);
    scene_id_ = -1;
    // the canvas is the destination texture's size (Canvas::set_dst_texture); the view box only
    // bounds the tiling (its right edge is rounded up to whole tiles, core/scene.cpp:170-179)
    pfh_scene_desc desc = builder->desc();
    auto *t = dynamic_cast<TextureHip *>(dest_.get());
    if (t) {
        if (t->get_format() != TextureFormat::Rgba8Unorm) {
            Logger::error("the destination must be an RGBA8 texture");
            return;
        }
        desc.width = t->get_size().x;
        desc.height = t->get_size().y;
    }
    if (!check(pfh_scene_upload(ctx_, &desc, &scene_id_))) return;
    if (t && !check(pfh_scene_set_target(ctx_, scene_id_, t->device_ptr(), (int64_t)t->pitch()))) return;
    check(pfh_rasterize(ctx_, &scene_id_, 1, clear_dst_texture ? PFH_PLAIN : PFH_LOAD_DST));
}

}  // namespace Pathfinder
