/* C++ host side over the C ABI of btrc_hip.h, header only.
 *
 * btrc_hip::WavefrontPathTracer mirrors Btrc's Renderer interface for the
 * wavefront tracer (core/renderer.h:13-59: set_scene / set_camera /
 * set_film / commit / render, stop via stop_async, core/renderer.cpp:38-48)
 * with WavefrontPathTracer::Params and their defaults
 * (builtin/renderer/wavefront.h:13-27). Failures throw BtrcException
 * carrying the context's message, as the reference's throw_on_error does
 * (utils/cuda/error.h:10-30). Scenes and cameras are passed as the C ABI's
 * descriptors; serialising Btrc's Scene into them is the job of the
 * renderer plugin that hosts this class. */
#ifndef BTRC_HIP_HPP
#define BTRC_HIP_HPP

#include "btrc_hip.h"

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace btrc_hip {

class BtrcException : public std::runtime_error
{
public:
    explicit BtrcException(const std::string &msg, int status = BTRC_ERR_INVALID)
        : std::runtime_error(msg), status_(status)
    {
    }
    int status() const { return status_; }

private:
    int status_;
};

// one device context (one HIP stream); not copyable
class Context
{
public:
    explicit Context(int device = 0)
    {
        const int rc = btrc_hip_create(device, &ctx_);
        if(rc != BTRC_OK)
            throw BtrcException("btrc_hip_create failed on device " + std::to_string(device) + " (status " +
                                std::to_string(rc) + ")", rc);
    }
    ~Context()
    {
        if(ctx_)
            btrc_hip_destroy(ctx_);
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;

    btrc_ctx *get() const { return ctx_; }

    void check(int rc) const
    {
        if(rc != BTRC_OK)
            throw BtrcException(btrc_hip_last_error(ctx_), rc);
    }

private:
    btrc_ctx *ctx_ = nullptr;
};

struct RenderResult
{
    int width = 0, height = 0;
    std::vector<float> color;  // RGBA per pixel, value / weight
    std::vector<float> albedo; // RGBA, when Params::albedo
    std::vector<float> normal; // RGBA, when Params::normal
    btrc_render_stats stats{};
    bool stopped = false;      // ended by stop_async
};

// Reporter (core/reporter.h:7-36): override what you need; calls come from
// the rendering thread during WavefrontPathTracer::render
class Reporter
{
public:
    virtual ~Reporter() = default;
    virtual void new_stage(const char *name) { (void)name; }
    virtual void complete_stage() {}
    virtual void progress(float percentage) { (void)percentage; }
};

class WavefrontPathTracer
{
public:
    // WavefrontPathTracer::Params, wavefront.h:13-27
    struct Params
    {
        bool tile = false;
        int spp = 128;
        int min_depth = 5;
        int max_depth = 10;
        float rr_threshold = 0.1f;
        float rr_cont_prob = 0.5f;
        int state_count = 1000000;
        bool albedo = false;
        bool normal = false;
        int sampler = BTRC_SAMPLER_HALTON;
    };

    explicit WavefrontPathTracer(Context &ctx) : ctx_(ctx) {}
    WavefrontPathTracer(Context &ctx, const Params &p) : ctx_(ctx), params_(p) {}

    void set_scene(const btrc_scene_desc &scene) { scene_ = &scene; dirty_ = true; }
    void set_camera(const btrc_camera_desc &camera) { camera_ = camera; has_camera_ = true; dirty_ = true; }
    void set_film(int width, int height) { width_ = width; height_ = height; dirty_ = true; }
    void set_film_filter(const btrc_filter_desc &filter) { filter_ = filter; dirty_ = true; }
    const Params &params() const { return params_; }

    // Renderer::set_reporter; the reporter must outlive the renders it sees
    // (nullptr removes it)
    void set_reporter(Reporter *reporter)
    {
        btrc_reporter r{ reporter, nullptr, nullptr, nullptr };
        if(reporter)
        {
            r.new_stage = [](void *u, const char *name) { static_cast<Reporter *>(u)->new_stage(name); };
            r.complete_stage = [](void *u) { static_cast<Reporter *>(u)->complete_stage(); };
            r.progress = [](void *u, float pct) { static_cast<Reporter *>(u)->progress(pct); };
        }
        ctx_.check(btrc_hip_set_reporter(ctx_.get(), reporter ? &r : nullptr));
    }

    // the "aces" post processor on the last render (ACESToneMap::process)
    std::vector<float> tonemapped(float exposure = 1.0f)
    {
        std::vector<float> rgba(static_cast<size_t>(width_) * static_cast<size_t>(height_) * 4);
        ctx_.check(btrc_hip_tonemap_readback(ctx_.get(), exposure, rgba.data()));
        return rgba;
    }

    // Renderer::commit: the scene, the camera (its aspect from the film,
    // main.cpp:52) and the filter go to the device
    void commit()
    {
        if(!scene_)
            throw BtrcException("commit before set_scene", BTRC_ERR_STATE);
        if(!has_camera_)
            throw BtrcException("commit before set_camera", BTRC_ERR_STATE);
        ctx_.check(btrc_hip_upload_scene(ctx_.get(), scene_));
        btrc_camera_desc cam = camera_;
        cam.w_over_h = static_cast<float>(width_) / static_cast<float>(height_);
        ctx_.check(btrc_hip_set_camera(ctx_.get(), &cam));
        ctx_.check(btrc_hip_set_filter(ctx_.get(), &filter_));
        dirty_ = false;
    }

    RenderResult render()
    {
        if(dirty_)
            commit();
        btrc_render_params p{};
        p.width = width_;
        p.height = height_;
        p.spp = params_.spp;
        p.min_depth = params_.min_depth;
        p.max_depth = params_.max_depth;
        p.rr_threshold = params_.rr_threshold;
        p.rr_cont_prob = params_.rr_cont_prob;
        p.state_count = params_.state_count;
        p.tile = params_.tile ? 1 : 0;
        p.sampler = params_.sampler;
        p.albedo = params_.albedo ? 1 : 0;
        p.normal = params_.normal ? 1 : 0;
        RenderResult r;
        r.width = width_;
        r.height = height_;
        const int rc = btrc_hip_render(ctx_.get(), &p, &r.stats);
        if(rc == BTRC_ERR_STOPPED)
        {
            r.stopped = true; // like should_stop() in wavefront.cpp:308-315
            return r;
        }
        ctx_.check(rc);
        const size_t n = static_cast<size_t>(width_) * static_cast<size_t>(height_) * 4;
        r.color.resize(n);
        ctx_.check(btrc_hip_film_readback(ctx_.get(), r.color.data(), nullptr));
        if(params_.albedo || params_.normal)
        {
            if(params_.albedo)
                r.albedo.resize(n);
            if(params_.normal)
                r.normal.resize(n);
            ctx_.check(btrc_hip_aov_readback(ctx_.get(), params_.albedo ? r.albedo.data() : nullptr,
                                             params_.normal ? r.normal.data() : nullptr));
        }
        return r;
    }

    // Renderer::stop_async: thread-safe; the running render returns stopped
    void stop_async() { btrc_hip_request_stop(ctx_.get()); }

private:
    Context &ctx_;
    Params params_;
    const btrc_scene_desc *scene_ = nullptr;
    btrc_camera_desc camera_{};
    bool has_camera_ = false;
    btrc_filter_desc filter_{ BTRC_FILTER_BOX, 0.0f, 0.0f };
    int width_ = 512, height_ = 512;
    bool dirty_ = true;
};

} // namespace btrc_hip

#endif /* BTRC_HIP_HPP */
