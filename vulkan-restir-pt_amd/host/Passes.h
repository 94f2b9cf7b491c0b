// Render passes of the host layer, keeping the reference's pass surface:
//   GBufferPass  (src/GBufferPass.h:22-69)  images depthNormal[2],
//                albedoMatId[2], motionVector; render() for a frame parity
//   RayTracing   (src/RayTracing.h:27-48)   one per-pixel pass bound to a
//                kernel entry of the C ABI; execute() with the frame bindings
//                and a push-constant block
//   GRISReSTIR   (src/GRISReSTIR.h:7-37)    Settings, createPipeline(),
//                render() = path trace, temporal reuse, spatial reuse
//   TestReSTIR   (src/TestReSTIR.h:7-38)    the ReSTIR DI method: Settings,
//                createPipeline(), render() = path generation, temporal
//                reuse, spatial reuse
//   Renderer     headless stand-in for src/Renderer.cpp's frame loop: owns the
//                ray images and reservoirs, binds the frame parity and flips it
#pragma once
#include "Scene.h"
#include <functional>

class GBufferPass {
public:
    static constexpr size_t DepthNormalBytes = 16, AlbedoMatIdBytes = 8, MotionVectorBytes = 4;

    GBufferPass(uint32_t width, uint32_t height);
    ~GBufferPass();
    GBufferPass(const GBufferPass&) = delete;
    GBufferPass& operator=(const GBufferPass&) = delete;

    void render(hipStream_t stream, uint32_t curFrame, const DeviceScene& scene, const rt_camera& camera);

    float* depthNormal[2] = {nullptr, nullptr};
    uint32_t* albedoMatId[2] = {nullptr, nullptr};
    uint16_t* motionVector = nullptr;
    uint32_t width, height;
};

struct RayTracing {
    enum Mode : uint32_t { RayQuery = 0, Undefined };
    using Entry = std::function<rt_status(const rt_device_scene*, const rt_camera*, const rt_frame*, const void*, hipStream_t)>;

    void createPipeline(Entry entry) { mEntry = std::move(entry); mode = mEntry ? RayQuery : Undefined; }
    void execute(hipStream_t stream, const DeviceScene& scene, const rt_camera cams[2], const rt_frame& frame,
                 const void* pushConstant = nullptr);

    Mode mode = Undefined;

private:
    Entry mEntry;
};

class GRISReSTIR {
public:
    enum ShiftType : uint32_t { Reconnection = 0, Replay = 1, Hybrid = 2 };
    using Settings = rt_gris_settings;

    void createPipeline();
    // With settings.targetFunction = 1 the frame is: path trace, learned refresh,
    // temporal reuse, spatial reuse, one training step of the network. With
    // settings.guidedProposal = 1 the proposal's refresh precedes the path trace and
    // its training step follows the spatial reuse (the frame binds its lobe planes).
    void render(hipStream_t stream, const DeviceScene& scene, const rt_camera cams[2], const rt_frame& frame);

    // src/GRISReSTIR.h:29 {Hybrid, 1, false, true, 20}; then the kernel constants
    // MaxTracingDepth = 15, ResampleNum = 3, ResampleRadius = 20; luminance target; the reference's 1/M weights
    Settings settings = {Hybrid, 1.0f, 0u, 1u, 20u, 15u, 3u, 20.0f, 0u, 0u};
    rt_mlp* learned = nullptr;   // borrowed; required when settings.targetFunction = 1
    rt_mlp* proposal = nullptr;  // borrowed (rt_guide_create); required when settings.guidedProposal = 1

private:
    RayTracing mPathTracePass, mTemporalReusePass, mSpatialReusePass;
};

class TestReSTIR {
public:
    enum ShiftType : uint32_t { Reconnection = 0, Replay = 1, Hybrid = 2 };
    enum SampleType : uint32_t { Light = 0, BSDF = 1, Both = 2 };
    using Settings = rt_di_settings;

    void createPipeline();
    // reservoirs: descriptor bindings 7/8/9 of the frame being rendered
    void render(hipStream_t stream, const DeviceScene& scene, const rt_camera cams[2], const rt_frame& frame,
                const rt_di_reservoirs& reservoirs);

    Settings settings = {Reconnection, Light, 0u, 1u};  // src/TestReSTIR.h:31

private:
    RayTracing mPathTracePass, mTemporalReusePass, mSpatialReusePass;
    rt_di_reservoirs mBound{};
};

// Owns a learned network of a scene: the learned target function, or with
// proposal = true the learned vMF-mixture proposal; network bounds are the scene
// box padded by 10%.
class LearnedTarget {
public:
    LearnedTarget(const DeviceScene& scene, uint32_t seed, hipStream_t stream, uint32_t log2TableSize = 19,
                  bool proposal = false);
    ~LearnedTarget();
    LearnedTarget(const LearnedTarget&) = delete;
    LearnedTarget& operator=(const LearnedTarget&) = delete;
    rt_mlp* mlp = nullptr;
};

class Renderer {
public:
    enum Method : uint32_t { ResampledPT = 0, Naive = 1 };
    // Renderer::Settings directMethod / indirectMethod values (src/Renderer.h:25-31):
    // direct None, Naive, ResampledDI; indirect None, Naive, ResampledGI, ResampledPT
    enum RayTracingMethod : uint32_t { MethodNone = 0, MethodNaive = 1, MethodResampled = 2, MethodResampledPT = 3 };

    Renderer(const DeviceScene& scene, uint32_t width, uint32_t height);
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    // G-buffer pass, then the indirect (and for Naive also direct) method; flips the parity.
    void drawFrame(hipStream_t stream, const rt_camera cams[2], Method method);
    // G-buffer pass, the direct method, the indirect method (src/Renderer.cpp:440-470); flips the parity.
    void drawFrame(hipStream_t stream, const rt_camera cams[2], uint32_t directMethod, uint32_t indirectMethod);
    rt_frame bindings() const;
    rt_di_reservoirs diBindings() const;
    rt_gi_reservoirs giBindings() const;
    void flip() { mCurFrame ^= 1u; }
    // Resets the GRIS and DI reservoir buffers (the GUI's clearReservoir, src/Renderer.cpp:567).
    void resetReservoirs(hipStream_t stream);
    // PostProcessFrag + Renderer::writeScreenshot (src/Renderer.cpp:758-793): the current outputs
    // post-processed and written to path as a PNG.
    void screenshot(hipStream_t stream, const rt_post_settings& post, const std::string& path);
    // Allocates the learned proposal's frame buffers (two lobe-plane parities, the training
    // records), bound by bindings() from then on.
    void allocateProposal();

    GRISReSTIR gris;
    TestReSTIR restirDI;
    uint32_t maxDepthNaive = 15;
    float* directOutput = nullptr;
    float* indirectOutput = nullptr;
    void* grisReservoir[2] = {nullptr, nullptr};
    void* grisReservoirTemp = nullptr;
    void* diReservoir[2] = {nullptr, nullptr};
    void* diReservoirTemp = nullptr;
    void* giReservoir[2] = {nullptr, nullptr};
    void* guideLobes[2] = {nullptr, nullptr};
    void* guideSamples = nullptr;
    uint32_t width, height;

private:
    const DeviceScene& mScene;
    GBufferPass mGBuffer;
    RayTracing mNaiveDI, mNaiveGI;
    uint32_t mCurFrame = 0;
};

// GRIS frames with consecutive frames overlapped on two streams (no reference
// counterpart; the reference records one frame after the other). Frame k + 1's
// G-buffer and path trace depend only on its camera, so they run on the trace
// stream beside frame k's temporal and spatial reuse on the reuse stream: the
// path trace's latency-bound tail and the reuse passes fill each other's idle
// CUs. Buffers: three G-buffers (a frame reads its own and the previous one
// while the next is written), two motion-vector images, the path trace's
// reservoirs and the final reservoirs apart (two each), one Temp. Frame k:
//   trace stream:  wait for frame k - 2's temporal reuse; G-buffer; path trace
//   reuse stream:  wait for frame k's path trace; temporal reuse; spatial reuse
// Every frame computes exactly what Renderer::drawFrame(directMethod,
// ResampledPT) computes for it. The direct light: MethodNaive runs on the trace
// stream after the G-buffer; MethodResampled (ReSTIR DI) runs its path
// generation there and its temporal and spatial reuse on the reuse stream as
// soon as the generation is done (beside the path trace), with its reservoirs
// buffered like the GRIS ones (allocated when the method is chosen). The
// luminance target only (the learned forms' networks are read by one frame's
// refresh while the previous frame trains them); either resampling weights.
class FramePipeline {
public:
    FramePipeline(const DeviceScene& scene, uint32_t width, uint32_t height);
    ~FramePipeline();
    FramePipeline(const FramePipeline&) = delete;
    FramePipeline& operator=(const FramePipeline&) = delete;

    void drawFrame(hipStream_t traceStream, hipStream_t reuseStream, const rt_camera cams[2]);
    // The bindings a sequential renderer presents after the last enqueued frame: its G-buffer,
    // its final reservoirs (grisReservoir), the previous frame's (grisReservoirPrev), the outputs.
    rt_frame bindings() const;
    // Resets every reservoir buffer and restarts the frame count (the next frame is frame 0).
    void resetReservoirs(hipStream_t stream);

    GRISReSTIR::Settings settings = {GRISReSTIR::Hybrid, 1.0f, 0u, 1u, 20u, 15u, 3u, 20.0f, 0u, 0u, 0u};
    // workgroups of the persistent path trace (rt_gris_path_trace_groups; 0: all the device holds)
    uint32_t traceGroups = 0;
    // Renderer::RayTracingMethod of the direct light: MethodNone, MethodNaive or MethodResampled
    // (allocates the ReSTIR DI reservoirs the first time)
    void setDirectMethod(uint32_t method);
    uint32_t directMethod() const { return mDirect; }
    // ReSTIR DI's final reservoirs after the last enqueued frame (reservoir) and the frame before
    // (reservoirPrev), as Renderer::diBindings gives them before the flip
    rt_di_reservoirs diBindings() const;
    TestReSTIR::Settings diSettings = {TestReSTIR::Reconnection, TestReSTIR::Light, 0u, 1u};
    // the stream of ReSTIR DI's temporal and spatial reuse (null: the reuse stream)
    hipStream_t directStream = nullptr;
    float* directOutput = nullptr;
    float* indirectOutput = nullptr;
    uint32_t width, height;

private:
    rt_frame frameOf(uint64_t k) const;
    const DeviceScene& mScene;
    float* mDepthNormal[3] = {nullptr, nullptr, nullptr};
    uint32_t* mAlbedoMatId[3] = {nullptr, nullptr, nullptr};
    uint16_t* mMotion[2] = {nullptr, nullptr};
    void* mTraced[2] = {nullptr, nullptr};  // path-trace reservoirs
    void* mFinal[2] = {nullptr, nullptr};   // reservoirs after spatial reuse
    void* mTemp = nullptr;
    hipEvent_t mTracedEv[3] = {nullptr, nullptr, nullptr};
    hipEvent_t mTemporalEv[3] = {nullptr, nullptr, nullptr};
    hipEvent_t mDiGenEv[3] = {nullptr, nullptr, nullptr};
    hipEvent_t mDiDoneEv[3] = {nullptr, nullptr, nullptr};
    void* mDiTraced[2] = {nullptr, nullptr};  // ReSTIR DI: path generation's reservoirs
    void* mDiFinal[2] = {nullptr, nullptr};   // after spatial reuse
    void* mDiTemp = nullptr;
    uint32_t mDirect = 0;
    uint64_t mFrame = 0;
};
