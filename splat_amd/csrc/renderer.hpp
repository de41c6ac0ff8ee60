// C++ host interface of the HIP rasterizer, mirroring splat-tui's Renderer
// trait (src/render/mod.rs:86-121) for C++ callers. The C ABI in
// include/splat_hip.h is the drop-in boundary; this is the idiomatic layer
// above it.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "splat_hip.h"

namespace splat {

// Frame (src/render/mod.rs:44-70): 8-bit RGB, row-major.
struct Frame {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> pixels;  // width*height*3
    // Resizes like Frame::resize; sides above SP_MAX_DIM throw ("pixel limit"),
    // where the reference panics.
    void resize(uint32_t w, uint32_t h);
};

using View = sp_view;
using RenderStats = sp_stats;

// Renderer trait (src/render/mod.rs:86-97).
class Renderer {
public:
    virtual ~Renderer() = default;
    virtual const char *name() const = 0;
    virtual RenderStats render(const View &view, uint32_t width, uint32_t height, const float background[3],
                               Frame &frame) = 0;
};

// Backend (src/render/mod.rs:99-113), restricted to what this library provides.
enum class Backend { Hip };

// HIP backend over the C ABI.
class HipRenderer final : public Renderer {
public:
    HipRenderer(const sp_scene &scene, const sp_options *opt = nullptr);
    ~HipRenderer() override;
    HipRenderer(const HipRenderer &) = delete;
    HipRenderer &operator=(const HipRenderer &) = delete;
    const char *name() const override { return "hip"; }
    RenderStats render(const View &view, uint32_t width, uint32_t height, const float background[3],
                       Frame &frame) override;
    sp_handle *handle() { return h_; }

private:
    sp_handle *h_ = nullptr;
};

// render::create (src/render/mod.rs:115-121): throws std::runtime_error where the
// reference returns Err (no device, scene too large).
std::unique_ptr<Renderer> create(Backend backend, const sp_scene &scene);

}  // namespace splat
