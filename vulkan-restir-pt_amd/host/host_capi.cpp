// C entry points of the host layer (include/restirpt_host.h).
#include "Passes.h"
#include "restirpt_host.h"
#include <cstring>
#include <exception>

struct rt_host_scene { std::unique_ptr<Scene> scene; };
struct rt_pipeline {
    std::unique_ptr<DeviceScene> scene;
    std::unique_ptr<FramePipeline> p;
};
struct rt_renderer {
    std::unique_ptr<DeviceScene> scene;
    std::unique_ptr<Renderer> r;
    std::unique_ptr<LearnedTarget> learned;
    std::unique_ptr<LearnedTarget> proposal;
};

namespace rti {
void setError(const std::string& msg);  // the calling thread's rt_last_error() message (librestirpt.so)
}

namespace {
rt_status fail(rt_status st, const std::string& m) { rti::setError(m); return st; }
}  // namespace

extern "C" {

rt_status rt_scenegen(uint32_t config, uint32_t seed, rt_host_scene** out) {
    if (!out) return RT_ERR_INVALID_ARGUMENT;
    try {
        auto s = generateScene(config, seed);
        if (!s) return RT_ERR_INVALID_ARGUMENT;
        *out = new rt_host_scene{std::move(s)};
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_ERR_INVALID_ARGUMENT, e.what());
    }
}

rt_status rt_scene_load(const char* path, rt_host_scene** out) {
    if (!path || !out) return fail(RT_ERR_INVALID_ARGUMENT, "null argument");
    try {
        *out = new rt_host_scene{loadSceneFile(path)};
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_ERR_INVALID_ARGUMENT, e.what());
    }
}

rt_status rt_host_scene_film_size(const rt_host_scene* s, uint32_t* width, uint32_t* height) {
    if (!s || !width || !height) return fail(RT_ERR_INVALID_ARGUMENT, "null argument");
    *width = s->scene->filmSize[0];
    *height = s->scene->filmSize[1];
    return RT_OK;
}

rt_status rt_host_scene_destroy(rt_host_scene* s) {
    delete s;
    return RT_OK;
}

rt_status rt_host_scene_arrays(const rt_host_scene* s, rt_scene_arrays* out) {
    if (!s || !out) return RT_ERR_INVALID_ARGUMENT;
    *out = s->scene->arrays();
    return RT_OK;
}

// uCamera of frame f (src/Renderer.cpp:358-368 with accumulation off): the
// frame's matrices, lastProjView of frame f-1, frameIndex 0 (clear flag on
// frame 0, whose previous reservoirs hold nothing), seed of frame f.
static rt_camera deviceCamera(const Scene& S, uint32_t f, uint32_t width, uint32_t height) {
    auto camAt = [&](uint32_t k) {
        Camera c;
        c.setFilmSize(width, height);
        c.setPlanes(0.001f, 200.0f);  // src/Renderer.cpp:124
        c.setFOV(S.fov);
        c.setPos(S.cameraPath[k].pos);
        c.setAngle(S.cameraPath[k].angle);
        c.setLens(S.lensRadius, S.focalDist);
        return c;
    };
    Camera last = camAt(f > 0 ? f - 1 : 0);
    Camera cur = camAt(f);
    cur.nextFrame(S.cameraPath[f].seed);
    rt_camera c, l;
    cur.toDevice(c);
    last.toDevice(l);
    std::memcpy(c.lastProjView, l.projView, 64);
    c.frameIndex = (f == 0) ? 0x80000000u : 0u;
    return c;
}

rt_status rt_host_scene_camera(const rt_host_scene* s, uint32_t frame, uint32_t width, uint32_t height, rt_camera out[2]) {
    if (!s || !out || width == 0 || height == 0) return RT_ERR_INVALID_ARGUMENT;
    const Scene& S = *s->scene;
    if (frame >= S.cameraPath.size()) return RT_ERR_INVALID_ARGUMENT;
    out[0] = deviceCamera(S, frame, width, height);
    out[1] = deviceCamera(S, frame > 0 ? frame - 1 : 0, width, height);  // uPrevCamera = last frame's uCamera
    return RT_OK;
}

rt_status rt_light_table_build(const float* distribution, uint32_t n, rt_light_table_entry* out) {
    if ((!distribution && n) || !out) return RT_ERR_INVALID_ARGUMENT;
    DiscreteSampler1D d;
    d.build(std::vector<float>(distribution, distribution + n));
    std::memcpy(out, d.binomDistribs.data(), sizeof(rt_light_table_entry) * (n + 1));
    return RT_OK;
}

rt_status rt_renderer_create(rt_device_scene* scene, uint32_t width, uint32_t height, rt_renderer** out) {
    if (!scene || !out || width == 0 || height == 0) return RT_ERR_INVALID_ARGUMENT;
    try {
        rt_renderer* r = new rt_renderer();
        r->scene = std::make_unique<DeviceScene>(scene);  // borrowed: the caller keeps ownership
        r->r = std::make_unique<Renderer>(*r->scene, width, height);
        *out = r;
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_ERR_HIP, e.what());
    }
}

rt_status rt_renderer_destroy(rt_renderer* r) {
    if (!r) return RT_ERR_INVALID_ARGUMENT;
    r->r.reset();
    r->learned.reset();
    delete r;
    return RT_OK;
}

rt_status rt_renderer_set_settings(rt_renderer* r, const rt_gris_settings* s) {
    if (!r || !s) return RT_ERR_INVALID_ARGUMENT;
    r->r->gris.settings = *s;
    r->r->maxDepthNaive = s->maxDepth;
    return RT_OK;
}

rt_status rt_renderer_frame(rt_renderer* r, const rt_camera cams[2], uint32_t method, hipStream_t stream) {
    if (!r || !cams || method > 1) return RT_ERR_INVALID_ARGUMENT;
    try {
        r->r->drawFrame(stream, cams, method == 0 ? Renderer::ResampledPT : Renderer::Naive);
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_ERR_HIP, e.what());
    }
}

rt_status rt_renderer_frame_methods(rt_renderer* r, const rt_camera cams[2], uint32_t directMethod,
                                    uint32_t indirectMethod, hipStream_t stream) {
    if (!r || !cams) return RT_ERR_INVALID_ARGUMENT;
    if (directMethod > Renderer::MethodResampled || indirectMethod > Renderer::MethodResampledPT)
        return fail(RT_ERR_INVALID_ARGUMENT, "direct method must be 0..2 and indirect method 0..3");
    try {
        r->r->drawFrame(stream, cams, directMethod, indirectMethod);
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_ERR_HIP, e.what());
    }
}

rt_status rt_renderer_set_di_settings(rt_renderer* r, const rt_di_settings* s) {
    if (!r || !s || s->shiftType > 2 || s->sampleType > 2) return RT_ERR_INVALID_ARGUMENT;
    r->r->restirDI.settings = *s;
    return RT_OK;
}

rt_status rt_renderer_di_bindings(const rt_renderer* r, rt_di_reservoirs* out) {
    if (!r || !out) return RT_ERR_INVALID_ARGUMENT;
    *out = r->r->diBindings();
    return RT_OK;
}

rt_status rt_renderer_bindings(const rt_renderer* r, rt_frame* out) {
    if (!r || !out) return RT_ERR_INVALID_ARGUMENT;
    *out = r->r->bindings();
    return RT_OK;
}

rt_status rt_renderer_flip(rt_renderer* r) {
    if (!r) return RT_ERR_INVALID_ARGUMENT;
    r->r->flip();
    return RT_OK;
}

rt_status rt_renderer_read_output(const rt_renderer* r, uint32_t which, float* host, hipStream_t stream) {
    if (!r || !host || which > 1) return RT_ERR_INVALID_ARGUMENT;
    const float* src = which == 0 ? r->r->directOutput : r->r->indirectOutput;
    size_t bytes = (size_t)r->r->width * r->r->height * 16;
    if (hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, stream) != hipSuccess) return RT_ERR_HIP;
    if (hipStreamSynchronize(stream) != hipSuccess) return RT_ERR_HIP;
    return RT_OK;
}

rt_status rt_renderer_screenshot(rt_renderer* r, const rt_post_settings* settings, const char* path, hipStream_t stream) {
    if (!r || !settings || !path) return RT_ERR_INVALID_ARGUMENT;
    try {
        r->r->screenshot(stream, *settings, path);
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_ERR_INVALID_ARGUMENT, e.what());
    }
}

rt_status rt_renderer_reset_reservoirs(rt_renderer* r, hipStream_t stream) {
    if (!r) return RT_ERR_INVALID_ARGUMENT;
    try {
        r->r->resetReservoirs(stream);
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_ERR_HIP, e.what());
    }
}

rt_status rt_pipeline_create(rt_device_scene* scene, uint32_t width, uint32_t height, rt_pipeline** out) {
    if (!scene || !out || width == 0 || height == 0) return RT_ERR_INVALID_ARGUMENT;
    try {
        auto p = std::make_unique<rt_pipeline>();
        p->scene = std::make_unique<DeviceScene>(scene);  // borrowed: the caller keeps ownership
        p->p = std::make_unique<FramePipeline>(*p->scene, width, height);
        *out = p.release();
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_ERR_HIP, e.what());
    }
}

rt_status rt_pipeline_destroy(rt_pipeline* p) {
    if (!p) return RT_ERR_INVALID_ARGUMENT;
    p->p.reset();
    delete p;
    return RT_OK;
}

rt_status rt_pipeline_set_settings(rt_pipeline* p, const rt_gris_settings* s) {
    if (!p || !s) return RT_ERR_INVALID_ARGUMENT;
    if (s->targetFunction != 0u || s->guidedProposal != 0u)
        return fail(RT_ERR_INVALID_ARGUMENT, "rt_pipeline: the luminance target and BSDF sampling only");
    p->p->settings = *s;
    return RT_OK;
}

rt_status rt_pipeline_set_trace_groups(rt_pipeline* p, uint32_t workgroups) {
    if (!p) return RT_ERR_INVALID_ARGUMENT;
    p->p->traceGroups = workgroups;
    return RT_OK;
}

rt_status rt_pipeline_frame(rt_pipeline* p, const rt_camera cams[2], hipStream_t traceStream, hipStream_t reuseStream) {
    if (!p || !cams) return RT_ERR_INVALID_ARGUMENT;
    try {
        p->p->drawFrame(traceStream, reuseStream, cams);
        return RT_OK;
    } catch (const std::invalid_argument& e) {
        return fail(RT_ERR_INVALID_ARGUMENT, e.what());
    } catch (const std::exception& e) {
        return fail(RT_ERR_HIP, e.what());
    }
}

rt_status rt_pipeline_bindings(const rt_pipeline* p, rt_frame* out) {
    if (!p || !out) return RT_ERR_INVALID_ARGUMENT;
    *out = p->p->bindings();
    return RT_OK;
}

rt_status rt_pipeline_set_direct_method(rt_pipeline* p, uint32_t directMethod) {
    if (!p) return RT_ERR_INVALID_ARGUMENT;
    if (directMethod > Renderer::MethodResampled)
        return fail(RT_ERR_INVALID_ARGUMENT, "rt_pipeline: direct light 0 (None), 1 (Naive) or 2 (ReSTIR DI)");
    try {
        p->p->setDirectMethod(directMethod);
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_ERR_HIP, e.what());
    }
}

rt_status rt_pipeline_set_di_settings(rt_pipeline* p, const rt_di_settings* s) {
    if (!p || !s || s->shiftType > 2 || s->sampleType > 2) return RT_ERR_INVALID_ARGUMENT;
    p->p->diSettings = *s;
    return RT_OK;
}

rt_status rt_pipeline_set_direct_stream(rt_pipeline* p, hipStream_t stream) {
    if (!p) return RT_ERR_INVALID_ARGUMENT;
    p->p->directStream = stream;
    return RT_OK;
}

rt_status rt_pipeline_di_bindings(const rt_pipeline* p, rt_di_reservoirs* out) {
    if (!p || !out) return RT_ERR_INVALID_ARGUMENT;
    *out = p->p->diBindings();
    return RT_OK;
}

rt_status rt_pipeline_read_output(const rt_pipeline* p, uint32_t which, float* host, hipStream_t stream) {
    if (!p || !host || which > 1) return RT_ERR_INVALID_ARGUMENT;
    const size_t bytes = (size_t)p->p->width * p->p->height * 16;
    const float* src = which == 0 ? p->p->directOutput : p->p->indirectOutput;
    if (hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, stream) != hipSuccess) return RT_ERR_HIP;
    if (hipStreamSynchronize(stream) != hipSuccess) return RT_ERR_HIP;
    return RT_OK;
}

rt_status rt_pipeline_reset_reservoirs(rt_pipeline* p, hipStream_t stream) {
    if (!p) return RT_ERR_INVALID_ARGUMENT;
    try {
        p->p->resetReservoirs(stream);
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_ERR_HIP, e.what());
    }
}

rt_status rt_renderer_enable_learned(rt_renderer* r, uint32_t seed, uint32_t log2TableSize, hipStream_t stream,
                                     rt_mlp** mlp) {
    if (!r || log2TableSize < 10 || log2TableSize > 24) return RT_ERR_INVALID_ARGUMENT;
    try {
        r->r->gris.learned = nullptr;
        r->learned = std::make_unique<LearnedTarget>(*r->scene, seed, stream, log2TableSize);
        r->r->gris.learned = r->learned->mlp;
        if (mlp) *mlp = r->learned->mlp;
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_ERR_HIP, e.what());
    }
}

rt_status rt_renderer_enable_proposal(rt_renderer* r, uint32_t seed, uint32_t log2TableSize, hipStream_t stream,
                                      rt_mlp** mlp) {
    if (!r || log2TableSize < 10 || log2TableSize > 24) return fail(RT_ERR_INVALID_ARGUMENT, "invalid argument");
    try {
        r->r->gris.proposal = nullptr;
        r->r->allocateProposal();
        r->proposal = std::make_unique<LearnedTarget>(*r->scene, seed, stream, log2TableSize, true);
        r->r->gris.proposal = r->proposal->mlp;
        if (mlp) *mlp = r->proposal->mlp;
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_ERR_HIP, e.what());
    }
}

}  // extern "C"
