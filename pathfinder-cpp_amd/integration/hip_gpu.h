// The HIP-backed subset of the reference's GPU abstraction that a RendererHip needs: a Device, a
// Queue, textures, buffers, a command encoder and a fence (gpu/device.h:24-147, gpu/queue.h,
// gpu/texture.h, gpu/buffer.h, gpu/command_encoder.h:158-273, gpu/fence.h).  Renderer's
// constructor (core/renderer.cpp:13-37) and Palette::build_paint_info (core/paint/palette.cpp:
// 137-170) use only resource creation and the data-transfer commands (write / read buffer and
// texture) submitted through a queue, so those are implemented on HIP memory; render and compute
// pipelines, render passes and shader modules do not exist on this device (the kernels are in
// libpfh.so) and their factories fail with an error.
//
// Compiled against the reference's headers (a maintainer adds it as pathfinder/gpu/hip/).
#pragma once

#include <hip/hip_runtime_api.h>

#include <memory>
#include <string>
#include <vector>

#include "gpu/device.h"
#include "gpu/queue.h"

namespace Pathfinder {

// Device memory of width x height texels; rows `pitch` bytes apart.
class TextureHip : public Texture {
public:
    TextureHip(const TextureDescriptor &desc, const std::string &label);
    ~TextureHip() override;
    void *device_ptr() const { return ptr_; }
    size_t pitch() const { return pitch_; }

private:
    void *ptr_ = nullptr;
    size_t pitch_ = 0;
};

// Device memory (DeviceLocal) or host memory (HostVisibleAndCoherent: staging, mapped).
class BufferHip : public Buffer {
public:
    BufferHip(const BufferDescriptor &desc, const std::string &label);
    ~BufferHip() override;
    void upload_via_mapping(size_t data_size, size_t offset, const void *data) override;
    void download_via_mapping(size_t data_size, size_t offset, void *data) override;
    void *device_ptr() const { return dev_; }
    uint8_t *host_ptr() { return host_.data(); }

private:
    void *dev_ = nullptr;
    std::vector<uint8_t> host_;
};

class SamplerHip : public Sampler {
public:
    explicit SamplerHip(const SamplerDescriptor &d) : Sampler(d) {}
};

// Records the reference's commands (CommandEncoder's own write_* / read_* recorders) and runs the
// data-transfer ones on a HIP stream when its queue submits it.
class CommandEncoderHip : public CommandEncoder {
public:
    CommandEncoderHip(const std::shared_ptr<Device> &device, const std::string &label);
    // Runs every recorded command in order on `stream`, waits, then invokes the callbacks.
    void execute(hipStream_t stream);

protected:
    bool prepare() override;
};

class DeviceHip : public Device {
public:
    explicit DeviceHip(int hip_device);
    int hip_device() const { return hip_device_; }

    std::shared_ptr<Framebuffer> create_framebuffer(const std::shared_ptr<RenderPass> &render_pass,
                                                    const std::shared_ptr<Texture> &texture,
                                                    const std::string &label) override;
    std::shared_ptr<Buffer> create_buffer(const BufferDescriptor &desc, const std::string &label) override;
    std::shared_ptr<Texture> create_texture(const TextureDescriptor &desc, const std::string &label) override;
    std::shared_ptr<Sampler> create_sampler(SamplerDescriptor descriptor) override;
    std::shared_ptr<CommandEncoder> create_command_encoder(const std::string &label) override;
    std::shared_ptr<DescriptorSetLayout> create_descriptor_set_layout(
        const std::vector<DescriptorLayout> &descriptors) override;
    std::shared_ptr<DescriptorSet> create_descriptor_set(std::shared_ptr<DescriptorSetLayout> layout) override;
    std::shared_ptr<RenderPass> create_render_pass(TextureFormat format, AttachmentLoadOp load_op,
                                                   const std::string &label) override;
    std::shared_ptr<RenderPass> create_swap_chain_render_pass(TextureFormat format, AttachmentLoadOp load_op) override;
    std::shared_ptr<ShaderModule> create_shader_module(const std::shared_ptr<Shader> &shader,
                                                       const std::string &label) override;
    std::shared_ptr<ShaderModule> create_shader_module(const std::vector<char> &source_code, ShaderStage shader_stage,
                                                       const std::string &label) override;
    std::shared_ptr<RenderPipeline> create_render_pipeline(
        const std::shared_ptr<ShaderModule> &vert_shader_module,
        const std::shared_ptr<ShaderModule> &frag_shader_module,
        const std::vector<VertexInputAttributeDescription> &attribute_descriptions, BlendState blend_state,
        const std::shared_ptr<DescriptorSetLayout> &descriptor_set_layout, TextureFormat target_format,
        const std::string &label) override;
    std::shared_ptr<ComputePipeline> create_compute_pipeline(
        const std::shared_ptr<ShaderModule> &comp_shader_module,
        const std::shared_ptr<DescriptorSetLayout> &descriptor_set_layout, const std::string &label) override;
    std::shared_ptr<Fence> create_fence(const std::string &label) override;
    void *map_staging(const StagingAllocation &allocation) override;
    size_t get_aligned_uniform_size(size_t original_size) override;

protected:
    std::shared_ptr<Buffer> create_staging_buffer(size_t size) override;

private:
    int hip_device_ = 0;
};

// Submits run synchronously on the queue's stream (the reference's queues block on a fence too,
// gpu/vk/queue.cpp:29-34).
class QueueHip : public Queue {
public:
    explicit QueueHip(int hip_device);
    ~QueueHip() override;
    void submit(const std::shared_ptr<CommandEncoder> &encoder, const std::shared_ptr<Fence> &fence) override;
    void wait_idle() override;

private:
    hipStream_t stream_ = nullptr;
};

}  // namespace Pathfinder
