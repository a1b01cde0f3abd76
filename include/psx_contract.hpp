// psx_contract.hpp - C++ host mirror of Psycles' render contract over psx.h.
//
// Keeps the class shapes of include/psycles/contract/render.h:136-181 so a
// host written against RendererBackend / RenderSession / OutputSink can drive
// the MI355X render path unchanged:
//
//   psx::RendererBackend backend{0};
//   auto compilation = backend.compile_scene(scene_desc);   // SceneCompilation
//   auto session = backend.create_session(*compilation.scene, settings);
//   session->render_samples({first, count, offset, total}, sink);  // bool
//
// Error behaviour matches the reference: compile failures return a null scene
// plus diagnostics, invalid settings return a null session, render_samples
// returns false on cancel / bad range / changed total, and nothing throws
// except std::runtime_error for device failures (which the reference reports
// through Luisa exceptions).
#pragma once

#include <atomic>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "psx.h"

namespace psx {

struct SampleRange {
    uint32_t first = 0;
    uint32_t count = 0;
    uint32_t offset = 0;
    uint32_t total = 0;
};

struct PassTile {
    uint32_t pass = PSX_PASS_COMBINED;
    uint32_t channels = 4;
    uint32_t window_x = 0, window_y = 0, window_width = 0, window_height = 0;
    uint32_t full_width = 0, full_height = 0;
    std::span<const float> pixels;  // valid only during OutputSink::write
};

class OutputSink {
  public:
    virtual ~OutputSink() noexcept = default;
    virtual void begin(const psx_render_settings &settings) = 0;
    virtual void write(const PassTile &tile) = 0;
    virtual void end(bool cancelled) = 0;
};

struct RenderDiagnostic {
    std::string message;
};

class CompiledScene {
  public:
    explicit CompiledScene(psx_scene *handle) : handle_(handle) {}
    ~CompiledScene() { psx_scene_destroy(handle_); }
    CompiledScene(const CompiledScene &) = delete;
    CompiledScene &operator=(const CompiledScene &) = delete;
    psx_scene *handle() const noexcept { return handle_; }
    // parameter-only rebind (MaterialLibrary::update, rebind branch)
    bool update_parameters(uint32_t material, std::span<const float> block) {
        return psx_scene_update_parameters(handle_, material, block.data(),
                                           static_cast<uint32_t>(block.size())) == PSX_OK;
    }

  private:
    psx_scene *handle_;
};

struct SceneCompilation {
    std::unique_ptr<CompiledScene> scene;
    std::vector<RenderDiagnostic> diagnostics;
    bool ok() const noexcept { return scene != nullptr; }
};

class RenderSession {
  public:
    explicit RenderSession(psx_session *handle) : handle_(handle) {}
    ~RenderSession() { psx_session_destroy(handle_); }
    RenderSession(const RenderSession &) = delete;
    RenderSession &operator=(const RenderSession &) = delete;

    void reset(const psx_render_settings &settings) {
        if (psx_session_reset(handle_, &settings) != PSX_OK) throw std::runtime_error("psx_session_reset failed");
    }

    bool render_samples(const SampleRange &range, OutputSink &output) {
        psx_sink sink{};
        sink.user = &output;
        sink.begin = [](void *user, const psx_render_settings *s) { static_cast<OutputSink *>(user)->begin(*s); };
        sink.write = [](void *user, uint32_t pass, uint32_t channels, uint32_t wx, uint32_t wy, uint32_t ww,
                        uint32_t wh, uint32_t fw, uint32_t fh, const float *pixels) {
            PassTile tile{pass, channels, wx, wy, ww, wh, fw, fh,
                          std::span<const float>{pixels, static_cast<size_t>(ww) * wh * channels}};
            static_cast<OutputSink *>(user)->write(tile);
        };
        sink.end = [](void *user, int cancelled) { static_cast<OutputSink *>(user)->end(cancelled != 0); };
        const int rc = psx_session_render(handle_, range.first, range.count, range.offset, range.total, &sink);
        if (rc == PSX_ERROR_DEVICE || rc == PSX_ERROR_OUT_OF_MEMORY || rc == PSX_ERROR_NO_DEVICE)
            throw std::runtime_error("psx_session_render: device failure");
        return rc == PSX_OK;
    }

    void cancel() noexcept { psx_session_cancel(handle_); }
    psx_session *handle() const noexcept { return handle_; }

  private:
    psx_session *handle_;
};

class RendererBackend {
  public:
    explicit RendererBackend(int device, const psx_backend_options *options = nullptr) {
        if (psx_backend_create(device, options, &handle_) != PSX_OK)
            throw std::runtime_error("psx_backend_create failed");
    }
    ~RendererBackend() { psx_backend_destroy(handle_); }
    RendererBackend(const RendererBackend &) = delete;
    RendererBackend &operator=(const RendererBackend &) = delete;

    SceneCompilation compile_scene(const psx_scene_desc &desc) {
        SceneCompilation result;
        psx_scene *scene = nullptr;
        const int rc = psx_scene_compile(
            handle_, &desc,
            [](void *user, const char *msg) {
                static_cast<std::vector<RenderDiagnostic> *>(user)->push_back(RenderDiagnostic{msg});
            },
            &result.diagnostics, &scene);
        if (rc == PSX_ERROR_DEVICE || rc == PSX_ERROR_OUT_OF_MEMORY)
            throw std::runtime_error("psx_scene_compile: device failure");
        if (rc == PSX_OK) result.scene = std::make_unique<CompiledScene>(scene);
        return result;
    }

    std::unique_ptr<RenderSession> create_session(const CompiledScene &scene, const psx_render_settings &settings) {
        psx_session *session = nullptr;
        if (psx_session_create(scene.handle(), &settings, &session) != PSX_OK) return nullptr;
        return std::make_unique<RenderSession>(session);
    }

  private:
    psx_backend *handle_ = nullptr;
};

}  // namespace psx
