// The reference-side binding of libpfh.so: a SceneBuilder and a Renderer beside the D3D9 / D3D11
// ones (core/scene_builder.h:16-32, core/renderer.h:58-139), selected by a RenderMode::Hip branch in
// Canvas::Canvas (core/canvas.cpp:159-171).  Compiled against the reference's headers.
#pragma once

#include <memory>
#include <vector>

#include "core/renderer.h"
#include "core/scene.h"
#include "core/scene_builder.h"
#include "pfh.h"

namespace Pathfinder {

// Packs Scene::draw_paths, then Scene::clip_paths, into a pfh_scene_desc: outlines as they are
// (Canvas has applied transforms, strokes and dashes), solid paints as RGBA8, gradient paints as
// pfh_gradient records, fill rules, and each path's clip path (DrawPath::clip_path /
// ClipPath::clip_path) as an index among the clip paths.
class SceneBuilderHip : public SceneBuilder {
public:
    void build(Scene *scene_, Renderer *renderer) override;
    const pfh_scene_desc &desc() const { return desc_; }
    // PFH_CLIP_RULE_ABS (the reference's clip coverage, the default) or PFH_CLIP_RULE_OWN
    // (ClipPath::fill_rule honoured), for the scenes built after the call.
    void set_clip_rule(int32_t rule) { clip_rule_ = rule; }

private:
    pfh_scene_desc desc_{};
    int32_t clip_rule_ = PFH_CLIP_RULE_ABS;
    std::vector<float> pts_;
    std::vector<uint8_t> flags_, closed_, rule_, blend_;
    std::vector<int32_t> cstart_, pstart_, clip_, grad_of_path_;
    std::vector<uint32_t> rgba_, ramps_;
    std::vector<pfh_gradient> grads_;
};

// Renderer over the C ABI.  The destination is a texture of the device (a TextureHip, whose device
// memory the rasterizer writes directly, Canvas::set_dst_texture) or, without one, the context's
// own image.
class RendererHip : public Renderer {
public:
    RendererHip(const std::shared_ptr<Device> &device, const std::shared_ptr<Queue> &queue, int hip_device = 0);
    ~RendererHip() override;

    void set_up_pipelines() override {}  // the kernels are in libpfh.so
    std::shared_ptr<Texture> get_dest_texture() override { return dest_; }
    void set_dest_texture(const std::shared_ptr<Texture> &texture) override { dest_ = texture; }
    void draw(const std::shared_ptr<SceneBuilder> &scene_builder, bool clear_dst_texture) override;

    pfh_ctx *context() const { return ctx_; }
    int32_t scene_id() const { return scene_id_; }

protected:
    TextureFormat mask_texture_format() const override { return TextureFormat::Rgba8Unorm; }

private:
    bool check(int rc);
    pfh_ctx *ctx_ = nullptr;
    int32_t scene_id_ = -1;
    std::shared_ptr<Texture> dest_;
};

}  // namespace Pathfinder
