// HIP-backed Device / Queue / Texture / Buffer / CommandEncoder (see hip_gpu.h).
#include "hip_gpu.h"

#include <cstring>
#include <stdexcept>

#include "common/logger.h"

namespace Pathfinder {

namespace {
void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

TextureHip::TextureHip(const TextureDescriptor &desc, const std::string &label) : Texture(desc) {
    label_ = label;
    pitch_ = (size_t)desc.size.x * get_pixel_size(desc.format);
    const size_t bytes = pitch_ * (size_t)desc.size.y;
    if (bytes) hip_check(hipMalloc(&ptr_, bytes), "hipMalloc(texture)");
}

TextureHip::~TextureHip() {
    if (ptr_) (void)hipFree(ptr_);
}

BufferHip::BufferHip(const BufferDescriptor &desc, const std::string &label) : Buffer(desc) {
    label_ = label;
    if (desc.property == MemoryProperty::HostVisibleAndCoherent) host_.resize(desc.size);
    else if (desc.size) hip_check(hipMalloc(&dev_, desc.size), "hipMalloc(buffer)");
}

BufferHip::~BufferHip() {
    if (dev_) (void)hipFree(dev_);
}

void BufferHip::upload_via_mapping(size_t data_size, size_t offset, const void *data) {
    if (dev_) hip_check(hipMemcpy((uint8_t *)dev_ + offset, data, data_size, hipMemcpyHostToDevice), "upload");
    else std::memcpy(host_.data() + offset, data, data_size);
}

void BufferHip::download_via_mapping(size_t data_size, size_t offset, void *data) {
    if (dev_) hip_check(hipMemcpy(data, (uint8_t *)dev_ + offset, data_size, hipMemcpyDeviceToHost), "download");
    else std::memcpy(data, host_.data() + offset, data_size);
}

CommandEncoderHip::CommandEncoderHip(const std::shared_ptr<Device> &device, const std::string &label) {
    device_ = device;
    label_ = label;
}

bool CommandEncoderHip::prepare() {
    if (submitted_) return false;
    prepared_ = true;
    return true;
}

void CommandEncoderHip::execute(hipStream_t stream) {
    if (!prepare()) return;
    for (const Command &cmd : commands_) {
        switch (cmd.type) {
            case CommandType::WriteBuffer: {
                const auto &a = cmd.args.write_buffer;
                auto *dst = static_cast<BufferHip *>(a.buffer);
                auto *src = static_cast<BufferHip *>(a.staging_buffer);
                if (dst->device_ptr())
                    hip_check(hipMemcpyAsync((uint8_t *)dst->device_ptr() + a.offset, src->host_ptr() + a.staging_offset,
                                             a.data_size, hipMemcpyHostToDevice, stream),
                              "write_buffer");
                else
                    std::memcpy(dst->host_ptr() + a.offset, src->host_ptr() + a.staging_offset, a.data_size);
            } break;
            case CommandType::ReadBuffer: {
                // into the staging allocation; the recorder's callback copies it on to a.data
                const auto &a = cmd.args.read_buffer;
                auto *src = static_cast<BufferHip *>(a.buffer);
                auto *staging = static_cast<BufferHip *>(a.staging_buffer);
                if (src->device_ptr())
                    hip_check(hipMemcpyAsync(staging->host_ptr() + a.staging_offset,
                                             (uint8_t *)src->device_ptr() + a.offset, a.data_size,
                                             hipMemcpyDeviceToHost, stream),
                              "read_buffer");
                else
                    std::memcpy(staging->host_ptr() + a.staging_offset, src->host_ptr() + a.offset, a.data_size);
            } break;
            case CommandType::WriteTexture: {
                const auto &a = cmd.args.write_texture;
                auto *t = static_cast<TextureHip *>(a.texture);
                auto *src = static_cast<BufferHip *>(a.staging_buffer);
                const size_t px = get_pixel_size(t->get_format());
                hip_check(hipMemcpy2DAsync((uint8_t *)t->device_ptr() + a.offset_y * t->pitch() + a.offset_x * px,
                                           t->pitch(), src->host_ptr() + a.staging_offset, a.width * px,
                                           a.width * px, a.height, hipMemcpyHostToDevice, stream),
                          "write_texture");
            } break;
            case CommandType::ReadTexture: {
                // into the staging allocation; the recorder's callback (CommandEncoder::read_texture,
                // gpu/command_encoder.cpp:317-355) copies it on to a.data once the queue has run it
                const auto &a = cmd.args.read_texture;
                auto *t = static_cast<TextureHip *>(a.texture);
                auto *staging = static_cast<BufferHip *>(a.staging_buffer);
                const size_t px = get_pixel_size(t->get_format());
                hip_check(hipMemcpy2DAsync(staging->host_ptr() + a.staging_offset, a.width * px,
                                           (uint8_t *)t->device_ptr() + a.offset_y * t->pitch() + a.offset_x * px,
                                           t->pitch(), a.width * px, a.height, hipMemcpyDeviceToHost, stream),
                          "read_texture");
            } break;
            default:
                Logger::error("command not supported by the HIP device (its kernels run in libpfh.so)");
                break;
        }
    }
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    submitted_ = true;
    invoke_callbacks();
}

DeviceHip::DeviceHip(int hip_device) : Device(1), hip_device_(hip_device) {
    hip_check(hipSetDevice(hip_device), "hipSetDevice");
}

namespace {
[[noreturn]] void unsupported(const char *what) {
    throw std::runtime_error(std::string(what) + " does not exist on the HIP device (its kernels run in libpfh.so)");
}
}  // namespace

std::shared_ptr<Framebuffer> DeviceHip::create_framebuffer(const std::shared_ptr<RenderPass> &,
                                                           const std::shared_ptr<Texture> &, const std::string &) {
    unsupported("a framebuffer");
}
std::shared_ptr<Buffer> DeviceHip::create_buffer(const BufferDescriptor &desc, const std::string &label) {
    return std::make_shared<BufferHip>(desc, label);
}
std::shared_ptr<Texture> DeviceHip::create_texture(const TextureDescriptor &desc, const std::string &label) {
    return std::make_shared<TextureHip>(desc, label);
}
std::shared_ptr<Sampler> DeviceHip::create_sampler(SamplerDescriptor descriptor) {
    return std::make_shared<SamplerHip>(descriptor);
}
std::shared_ptr<CommandEncoder> DeviceHip::create_command_encoder(const std::string &label) {
    return std::make_shared<CommandEncoderHip>(shared_from_this(), label);
}
std::shared_ptr<DescriptorSetLayout> DeviceHip::create_descriptor_set_layout(const std::vector<DescriptorLayout> &) {
    unsupported("a descriptor set layout");
}
std::shared_ptr<DescriptorSet> DeviceHip::create_descriptor_set(std::shared_ptr<DescriptorSetLayout>) {
    unsupported("a descriptor set");
}
std::shared_ptr<RenderPass> DeviceHip::create_render_pass(TextureFormat, AttachmentLoadOp, const std::string &) {
    unsupported("a render pass");
}
std::shared_ptr<RenderPass> DeviceHip::create_swap_chain_render_pass(TextureFormat, AttachmentLoadOp) {
    unsupported("a swap chain");
}
std::shared_ptr<ShaderModule> DeviceHip::create_shader_module(const std::shared_ptr<Shader> &, const std::string &) {
    unsupported("a shader module");
}
std::shared_ptr<ShaderModule> DeviceHip::create_shader_module(const std::vector<char> &, ShaderStage,
                                                              const std::string &) {
    unsupported("a shader module");
}
std::shared_ptr<RenderPipeline> DeviceHip::create_render_pipeline(const std::shared_ptr<ShaderModule> &,
                                                                  const std::shared_ptr<ShaderModule> &,
                                                                  const std::vector<VertexInputAttributeDescription> &,
                                                                  BlendState, const std::shared_ptr<DescriptorSetLayout> &,
                                                                  TextureFormat, const std::string &) {
    unsupported("a render pipeline");
}
std::shared_ptr<ComputePipeline> DeviceHip::create_compute_pipeline(const std::shared_ptr<ShaderModule> &,
                                                                    const std::shared_ptr<DescriptorSetLayout> &,
                                                                    const std::string &) {
    unsupported("a compute pipeline");
}
std::shared_ptr<Fence> DeviceHip::create_fence(const std::string &label) {
    auto f = std::make_shared<Fence>();
    f->label = label;
    return f;
}
void *DeviceHip::map_staging(const StagingAllocation &allocation) {
    return static_cast<BufferHip *>(allocation.buffer.get())->host_ptr() + allocation.offset;
}
size_t DeviceHip::get_aligned_uniform_size(size_t original_size) { return (original_size + 255) & ~(size_t)255; }
std::shared_ptr<Buffer> DeviceHip::create_staging_buffer(size_t size) {
    return std::make_shared<BufferHip>(
        BufferDescriptor{BufferType::Storage, size, MemoryProperty::HostVisibleAndCoherent}, "staging");
}

QueueHip::QueueHip(int hip_device) {
    hip_check(hipSetDevice(hip_device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
}
QueueHip::~QueueHip() {
    if (stream_) (void)hipStreamDestroy(stream_);
}
void QueueHip::submit(const std::shared_ptr<CommandEncoder> &encoder, const std::shared_ptr<Fence> &) {
    auto *e = dynamic_cast<CommandEncoderHip *>(encoder.get());
    if (!e) {
        Logger::error("QueueHip takes encoders of a DeviceHip");
        return;
    }
    e->execute(stream_);
}
void QueueHip::wait_idle() { (void)hipStreamSynchronize(stream_); }

}  // namespace Pathfinder
