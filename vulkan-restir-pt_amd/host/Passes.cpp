// Render passes of the host layer over the C ABI.
#include "Passes.h"
#include "restirpt_host.h"
#include <vector>

namespace {
template <class T>
T* deviceAlloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    if (hipMemset(p, 0, bytes) != hipSuccess) throw std::runtime_error("hipMemset failed");
    return reinterpret_cast<T*>(p);
}
}  // namespace

// ---- GBufferPass: src/GBufferPass.cpp:69-123 (createResource), :22-56 (render) ----
GBufferPass::GBufferPass(uint32_t w, uint32_t h) : width(w), height(h) {
    size_t n = (size_t)w * h;
    for (int j = 0; j < 2; j++) {
        depthNormal[j] = deviceAlloc<float>(n * DepthNormalBytes);
        albedoMatId[j] = deviceAlloc<uint32_t>(n * AlbedoMatIdBytes);
    }
    motionVector = deviceAlloc<uint16_t>(n * MotionVectorBytes);
}
GBufferPass::~GBufferPass() {
    for (int j = 0; j < 2; j++) { (void)hipFree(depthNormal[j]); (void)hipFree(albedoMatId[j]); }
    (void)hipFree(motionVector);
}
void GBufferPass::render(hipStream_t stream, uint32_t curFrame, const DeviceScene& scene, const rt_camera& cam) {
    rtThrow(rt_gbuffer_trace(scene.handle(), &cam, depthNormal[curFrame], albedoMatId[curFrame], motionVector, stream),
            "GBufferPass::render");
}

// ---- RayTracing: src/RayTracing.cpp:137-147 ----
void RayTracing::execute(hipStream_t stream, const DeviceScene& scene, const rt_camera cams[2], const rt_frame& frame,
                         const void* pushConstant) {
    if (mode != RayQuery || !mEntry) throw std::runtime_error("RayTracing::execute: pipeline not created");
    rtThrow(mEntry(scene.handle(), cams, &frame, pushConstant, stream), "RayTracing::execute");
}

// ---- GRISReSTIR: src/GRISReSTIR.cpp:9-67 ----
void GRISReSTIR::createPipeline() {
    auto bind = [](rt_status (*fn)(const rt_device_scene*, const rt_camera*, const rt_frame*, const rt_gris_settings*,
                                   hipStream_t)) {
        return [fn](const rt_device_scene* s, const rt_camera* c, const rt_frame* f, const void* pc, hipStream_t st) {
            return fn(s, c, f, static_cast<const rt_gris_settings*>(pc), st);
        };
    };
    mPathTracePass.createPipeline(bind(&rt_gris_path_trace));
    mTemporalReusePass.createPipeline(bind(&rt_gris_temporal));
    mSpatialReusePass.createPipeline(bind(&rt_gris_spatial));
}
// Passes are stream ordered, which stands in for the reference's pipeline barriers.
void GRISReSTIR::render(hipStream_t stream, const DeviceScene& scene, const rt_camera cams[2], const rt_frame& frame) {
    const bool learnedTarget = settings.targetFunction == 1u;
    const bool guided = settings.guidedProposal == 1u;
    if (learnedTarget && !learned) throw std::runtime_error("GRISReSTIR: the learned target needs a network");
    if (guided && !proposal) throw std::runtime_error("GRISReSTIR: the learned proposal needs a network");
    if (guided) rtThrow(rt_gris_guide_refresh(scene.handle(), proposal, cams, &frame, stream), "rt_gris_guide_refresh");
    mPathTracePass.execute(stream, scene, cams, frame, &settings);
    if (learnedTarget)
        rtThrow(rt_gris_learned_refresh(scene.handle(), learned, cams, &frame, stream), "rt_gris_learned_refresh");
    mTemporalReusePass.execute(stream, scene, cams, frame, &settings);
    mSpatialReusePass.execute(stream, scene, cams, frame, &settings);
    if (learnedTarget) rtThrow(rt_mlp_train_frame(learned, nullptr, stream), "rt_mlp_train_frame");
    if (guided) rtThrow(rt_gris_guide_train(scene.handle(), proposal, cams, &frame, nullptr, stream), "rt_gris_guide_train");
}

// ---- TestReSTIR: src/TestReSTIR.cpp:9-46 ----
void TestReSTIR::createPipeline() {
    const rt_di_reservoirs* bound = &mBound;
    auto bind = [bound](rt_status (*fn)(const rt_device_scene*, const rt_camera*, const rt_frame*,
                                        const rt_di_reservoirs*, const rt_di_settings*, hipStream_t)) {
        return [fn, bound](const rt_device_scene* s, const rt_camera* c, const rt_frame* f, const void* pc,
                           hipStream_t st) { return fn(s, c, f, bound, static_cast<const rt_di_settings*>(pc), st); };
    };
    mPathTracePass.createPipeline(bind(&rt_di_path_gen));
    mTemporalReusePass.createPipeline(bind(&rt_di_temporal));
    mSpatialReusePass.createPipeline(bind(&rt_di_spatial));
}
void TestReSTIR::render(hipStream_t stream, const DeviceScene& scene, const rt_camera cams[2], const rt_frame& frame,
                        const rt_di_reservoirs& reservoirs) {
    mBound = reservoirs;
    mPathTracePass.execute(stream, scene, cams, frame, &settings);
    mTemporalReusePass.execute(stream, scene, cams, frame, &settings);
    mSpatialReusePass.execute(stream, scene, cams, frame, &settings);
}

LearnedTarget::LearnedTarget(const DeviceScene& scene, uint32_t seed, hipStream_t stream, uint32_t log2TableSize,
                             bool proposal) {
    rt_scene_info info{};
    rtThrow(rt_device_scene_info(scene.handle(), &info), "rt_device_scene_info");
    rt_mlp_config c{};
    c.levels = 16; c.featuresPerLevel = 2; c.log2TableSize = log2TableSize; c.baseResolution = 16;
    c.maxResolution = 2048; c.hiddenWidth = 64; c.hiddenLayers = 2; c.seed = seed;
    c.learningRate = 1e-2f; c.beta1 = 0.9f; c.beta2 = 0.99f; c.epsilon = 1e-15f;
    for (int a = 0; a < 3; a++) {
        float ext = std::max(info.sceneHi[a] - info.sceneLo[a], 1e-3f);
        c.boundsLo[a] = info.sceneLo[a] - 0.1f * ext;
        c.boundsHi[a] = info.sceneHi[a] + 0.1f * ext;
    }
    if (proposal) rtThrow(rt_guide_create(&c, stream, &mlp), "rt_guide_create");
    else rtThrow(rt_mlp_create(&c, stream, &mlp), "rt_mlp_create");
}
LearnedTarget::~LearnedTarget() {
    if (mlp) rt_mlp_destroy(mlp);
}

void Renderer::resetReservoirs(hipStream_t stream) {
    for (int j = 0; j < 2; j++) rtThrow(rt_gris_reservoir_reset(grisReservoir[j], width, height, stream), "reset");
    rtThrow(rt_gris_reservoir_reset(grisReservoirTemp, width, height, stream), "reset");
    for (int j = 0; j < 2; j++) rtThrow(rt_di_reservoir_reset(diReservoir[j], width, height, stream), "reset");
    rtThrow(rt_di_reservoir_reset(diReservoirTemp, width, height, stream), "reset");
    for (int j = 0; j < 2; j++) rtThrow(rt_gi_reservoir_reset(giReservoir[j], width, height, stream), "reset");
}

// ---- Renderer: src/Renderer.cpp:193-257 (createRayImage), 319-353 (bindings), 400-569 ----
Renderer::Renderer(const DeviceScene& scene, uint32_t w, uint32_t h)
    : width(w), height(h), mScene(scene), mGBuffer(w, h) {
    size_t n = (size_t)w * h;
    directOutput = deviceAlloc<float>(n * 16);
    indirectOutput = deviceAlloc<float>(n * 16);
    size_t rb = rt_gris_reservoir_bytes(w, h);
    for (int j = 0; j < 2; j++) {
        grisReservoir[j] = deviceAlloc<uint8_t>(rb);
        rtThrow(rt_gris_reservoir_reset(grisReservoir[j], w, h, 0), "rt_gris_reservoir_reset");
    }
    grisReservoirTemp = deviceAlloc<uint8_t>(rb);
    rtThrow(rt_gris_reservoir_reset(grisReservoirTemp, w, h, 0), "rt_gris_reservoir_reset");
    const size_t db = rt_di_reservoir_bytes(w, h);
    for (int j = 0; j < 2; j++) {
        diReservoir[j] = deviceAlloc<uint8_t>(db);
        rtThrow(rt_di_reservoir_reset(diReservoir[j], w, h, 0), "rt_di_reservoir_reset");
    }
    diReservoirTemp = deviceAlloc<uint8_t>(db);
    rtThrow(rt_di_reservoir_reset(diReservoirTemp, w, h, 0), "rt_di_reservoir_reset");
    for (int j = 0; j < 2; j++) {
        giReservoir[j] = deviceAlloc<uint8_t>(rt_gi_reservoir_bytes(w, h));
        rtThrow(rt_gi_reservoir_reset(giReservoir[j], w, h, 0), "rt_gi_reservoir_reset");
    }
    if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize failed");
    gris.createPipeline();
    restirDI.createPipeline();
    uint32_t* maxDepth = &maxDepthNaive;
    mNaiveDI.createPipeline([maxDepth](const rt_device_scene* s, const rt_camera* c, const rt_frame* f, const void*,
                                       hipStream_t st) { return rt_pt_plain(s, c, f, *maxDepth, 1u, st); });
    mNaiveGI.createPipeline([maxDepth](const rt_device_scene* s, const rt_camera* c, const rt_frame* f, const void*,
                                       hipStream_t st) { return rt_pt_plain(s, c, f, *maxDepth, 2u, st); });
}
Renderer::~Renderer() {
    (void)hipFree(directOutput); (void)hipFree(indirectOutput);
    (void)hipFree(grisReservoir[0]); (void)hipFree(grisReservoir[1]); (void)hipFree(grisReservoirTemp);
    (void)hipFree(diReservoir[0]); (void)hipFree(diReservoir[1]); (void)hipFree(diReservoirTemp);
    (void)hipFree(giReservoir[0]); (void)hipFree(giReservoir[1]);
    (void)hipFree(guideLobes[0]); (void)hipFree(guideLobes[1]); (void)hipFree(guideSamples);
}
void Renderer::allocateProposal() {
    if (guideSamples) return;
    const size_t lobeBytes = rt_guide_lobe_bytes(width, height);
    for (int j = 0; j < 2; j++) guideLobes[j] = deviceAlloc<uint8_t>(lobeBytes);
    guideSamples = deviceAlloc<uint8_t>((size_t)width * height * 16);
}
void Renderer::screenshot(hipStream_t stream, const rt_post_settings& post, const std::string& path) {
    const size_t bytes = (size_t)width * height * 4;
    uint8_t* dev = deviceAlloc<uint8_t>(bytes);
    std::vector<uint8_t> host(bytes);
    const rt_frame f = bindings();
    rt_status st = rt_post_process(&f, &post, dev, stream);
    hipError_t e = st == RT_OK ? hipMemcpyAsync(host.data(), dev, bytes, hipMemcpyDeviceToHost, stream) : hipSuccess;
    if (e == hipSuccess && st == RT_OK) e = hipStreamSynchronize(stream);
    (void)hipFree(dev);
    rtThrow(st, "rt_post_process");
    if (e != hipSuccess) throw std::runtime_error("Renderer::screenshot: copy failed");
    rtThrow(rt_write_png(path.c_str(), host.data(), width, height), "rt_write_png");
}
rt_gi_reservoirs Renderer::giBindings() const {
    rt_gi_reservoirs r{};
    r.reservoir = giReservoir[mCurFrame];
    r.reservoirPrev = giReservoir[mCurFrame ^ 1u];
    return r;
}
rt_di_reservoirs Renderer::diBindings() const {
    rt_di_reservoirs r{};
    r.reservoir = diReservoir[mCurFrame];
    r.reservoirPrev = diReservoir[mCurFrame ^ 1u];
    r.reservoirTemp = diReservoirTemp;
    return r;
}
rt_frame Renderer::bindings() const {
    rt_frame f{};
    uint32_t thisFrame = mCurFrame, lastFrame = mCurFrame ^ 1u;
    f.width = width; f.height = height;
    f.directOutput = directOutput;
    f.indirectOutput = indirectOutput;
    f.depthNormal = mGBuffer.depthNormal[thisFrame];
    f.depthNormalPrev = mGBuffer.depthNormal[lastFrame];
    f.albedoMatId = mGBuffer.albedoMatId[thisFrame];
    f.albedoMatIdPrev = mGBuffer.albedoMatId[lastFrame];
    f.motionVector = mGBuffer.motionVector;
    f.grisReservoir = grisReservoir[thisFrame];
    f.grisReservoirPrev = grisReservoir[lastFrame];
    f.grisReservoirTemp = grisReservoirTemp;
    f.guideLobes = guideLobes[thisFrame];
    f.guideLobesPrev = guideLobes[lastFrame];
    f.guideSamples = guideSamples;
    return f;
}
void Renderer::drawFrame(hipStream_t stream, const rt_camera cams[2], Method method) {
    mGBuffer.render(stream, mCurFrame, mScene, cams[0]);
    rt_frame f = bindings();
    if (method == ResampledPT) {
        gris.render(stream, mScene, cams, f);
    } else {
        mNaiveDI.execute(stream, mScene, cams, f);
        mNaiveGI.execute(stream, mScene, cams, f);
    }
    flip();
}
void Renderer::drawFrame(hipStream_t stream, const rt_camera cams[2], uint32_t directMethod, uint32_t indirectMethod) {
    if (directMethod > MethodResampled || indirectMethod > MethodResampledPT)
        throw std::invalid_argument("Renderer::drawFrame: unsupported direct or indirect method");
    mGBuffer.render(stream, mCurFrame, mScene, cams[0]);
    const rt_frame f = bindings();
    if (directMethod == MethodNaive) mNaiveDI.execute(stream, mScene, cams, f);
    else if (directMethod == MethodResampled) restirDI.render(stream, mScene, cams, f, diBindings());
    if (indirectMethod == MethodNaive) mNaiveGI.execute(stream, mScene, cams, f);
    else if (indirectMethod == MethodResampled) {
        const rt_gi_reservoirs gi = giBindings();
        rtThrow(rt_gi_temporal(mScene.handle(), cams, &f, &gi, stream), "rt_gi_temporal");
    } else if (indirectMethod == MethodResampledPT) gris.render(stream, mScene, cams, f);
    flip();
}

// ---- FramePipeline ----
FramePipeline::FramePipeline(const DeviceScene& scene, uint32_t w, uint32_t h) : width(w), height(h), mScene(scene) {
    const size_t n = (size_t)w * h;
    for (int j = 0; j < 3; j++) {
        mDepthNormal[j] = deviceAlloc<float>(n * GBufferPass::DepthNormalBytes);
        mAlbedoMatId[j] = deviceAlloc<uint32_t>(n * GBufferPass::AlbedoMatIdBytes);
    }
    for (int j = 0; j < 2; j++) mMotion[j] = deviceAlloc<uint16_t>(n * GBufferPass::MotionVectorBytes);
    const size_t rb = rt_gris_reservoir_bytes(w, h);
    for (int j = 0; j < 2; j++) {
        mTraced[j] = deviceAlloc<uint8_t>(rb);
        mFinal[j] = deviceAlloc<uint8_t>(rb);
    }
    mTemp = deviceAlloc<uint8_t>(rb);
    directOutput = deviceAlloc<float>(n * 16);
    indirectOutput = deviceAlloc<float>(n * 16);
    for (int j = 0; j < 3; j++) {
        if (hipEventCreateWithFlags(&mTracedEv[j], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&mTemporalEv[j], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&mDiGenEv[j], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&mDiDoneEv[j], hipEventDisableTiming) != hipSuccess)
            throw std::runtime_error("hipEventCreate failed");
    }
    resetReservoirs(0);
    if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize failed");
}
FramePipeline::~FramePipeline() {
    (void)hipDeviceSynchronize();
    for (int j = 0; j < 3; j++) {
        (void)hipFree(mDepthNormal[j]); (void)hipFree(mAlbedoMatId[j]);
        if (mTracedEv[j]) (void)hipEventDestroy(mTracedEv[j]);
        if (mTemporalEv[j]) (void)hipEventDestroy(mTemporalEv[j]);
        if (mDiGenEv[j]) (void)hipEventDestroy(mDiGenEv[j]);
        if (mDiDoneEv[j]) (void)hipEventDestroy(mDiDoneEv[j]);
    }
    for (int j = 0; j < 2; j++) {
        (void)hipFree(mMotion[j]); (void)hipFree(mTraced[j]); (void)hipFree(mFinal[j]);
        (void)hipFree(mDiTraced[j]); (void)hipFree(mDiFinal[j]);
    }
    (void)hipFree(mDiTemp);
    (void)hipFree(mTemp); (void)hipFree(directOutput); (void)hipFree(indirectOutput);
}
void FramePipeline::resetReservoirs(hipStream_t stream) {
    // frames still in flight on the caller's streams read and write these buffers
    if (mFrame && hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize failed");
    for (int j = 0; j < 2; j++) {
        rtThrow(rt_gris_reservoir_reset(mTraced[j], width, height, stream), "reset");
        rtThrow(rt_gris_reservoir_reset(mFinal[j], width, height, stream), "reset");
    }
    rtThrow(rt_gris_reservoir_reset(mTemp, width, height, stream), "reset");
    if (mDiTemp) {
        for (int j = 0; j < 2; j++) {
            rtThrow(rt_di_reservoir_reset(mDiTraced[j], width, height, stream), "reset");
            rtThrow(rt_di_reservoir_reset(mDiFinal[j], width, height, stream), "reset");
        }
        rtThrow(rt_di_reservoir_reset(mDiTemp, width, height, stream), "reset");
    }
    // the next frames run on other streams
    if (hipStreamSynchronize(stream) != hipSuccess) throw std::runtime_error("hipStreamSynchronize failed");
    mFrame = 0;
}
// Bindings of frame k: its G-buffer and the previous one, its motion vectors, the path trace's
// reservoirs as grisReservoir and the previous frame's final reservoirs as grisReservoirPrev.
rt_frame FramePipeline::frameOf(uint64_t k) const {
    rt_frame f{};
    const uint32_t g = (uint32_t)(k % 3u), gp = (uint32_t)((k + 2u) % 3u), p = (uint32_t)(k & 1u);
    f.width = width; f.height = height;
    f.directOutput = directOutput;
    f.indirectOutput = indirectOutput;
    f.depthNormal = mDepthNormal[g]; f.depthNormalPrev = mDepthNormal[gp];
    f.albedoMatId = mAlbedoMatId[g]; f.albedoMatIdPrev = mAlbedoMatId[gp];
    f.motionVector = mMotion[p];
    f.grisReservoir = mTraced[p];
    f.grisReservoirPrev = mFinal[p ^ 1u];
    f.grisReservoirTemp = mTemp;
    return f;
}
void FramePipeline::setDirectMethod(uint32_t method) {
    if (method > Renderer::MethodResampled)
        throw std::invalid_argument("FramePipeline: direct light None, Naive or ReSTIR DI");
    if (method == mDirect) return;
    // the two methods write the direct output from different streams: let the frames in
    // flight finish before the next one switches
    if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize failed");
    if (method == Renderer::MethodResampled && !mDiTemp) {
        const size_t db = rt_di_reservoir_bytes(width, height);
        for (int j = 0; j < 2; j++) {
            mDiTraced[j] = deviceAlloc<uint8_t>(db);
            mDiFinal[j] = deviceAlloc<uint8_t>(db);
            rtThrow(rt_di_reservoir_reset(mDiTraced[j], width, height, 0), "rt_di_reservoir_reset");
            rtThrow(rt_di_reservoir_reset(mDiFinal[j], width, height, 0), "rt_di_reservoir_reset");
        }
        mDiTemp = deviceAlloc<uint8_t>(db);
        rtThrow(rt_di_reservoir_reset(mDiTemp, width, height, 0), "rt_di_reservoir_reset");
        if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize failed");
    }
    mDirect = method;
}
rt_di_reservoirs FramePipeline::diBindings() const {
    const uint64_t k = mFrame ? mFrame - 1u : 0u;
    rt_di_reservoirs r{};
    r.reservoir = mDiFinal[k & 1u];
    r.reservoirPrev = mDiFinal[(k & 1u) ^ 1u];
    r.reservoirTemp = mDiTemp;
    return r;
}
void FramePipeline::drawFrame(hipStream_t traceStream, hipStream_t reuseStream, const rt_camera cams[2]) {
    if (settings.targetFunction != 0u || settings.guidedProposal != 0u)
        throw std::invalid_argument("FramePipeline: the luminance target and BSDF sampling only");
    if (cams[0].filmSize[0] != width || cams[0].filmSize[1] != height)
        throw std::invalid_argument("FramePipeline: camera film size differs from the pipeline's");

    const uint64_t k = mFrame;
    const uint32_t e = (uint32_t)(k % 3u);
    // frame k overwrites what frame k - 2's temporal reuse read (the G-buffer three frames
    // back as its previous one, the path-trace reservoirs and motion vectors of its parity),
    // the DI reuse of frame k - 2 included when it runs on a stream of its own
    if (k >= 2 && (hipStreamWaitEvent(traceStream, mTemporalEv[(k - 2) % 3u], 0) != hipSuccess ||
                   hipStreamWaitEvent(traceStream, mDiDoneEv[(k - 2) % 3u], 0) != hipSuccess))
        throw std::runtime_error("hipStreamWaitEvent failed");
    hipStream_t diStream = directStream ? directStream : reuseStream;
    rt_frame f = frameOf(k);
    rtThrow(rt_gbuffer_trace(mScene.handle(), &cams[0], const_cast<float*>(f.depthNormal),
                             const_cast<uint32_t*>(f.albedoMatId), const_cast<uint16_t*>(f.motionVector), traceStream),
            "rt_gbuffer_trace");
    // the direct light: Renderer's naive pass (rt_pt_plain, direct part, its maxDepthNaive =
    // settings.maxDepth), or ReSTIR DI's path generation, whose reuse the DI stream (the
    // direct stream, else the reuse stream) starts as soon as it is done
    rt_di_reservoirs di{};
    if (mDirect == Renderer::MethodNaive)
        rtThrow(rt_pt_plain(mScene.handle(), cams, &f, settings.maxDepth, 1u, traceStream), "rt_pt_plain");
    if (mDirect == Renderer::MethodResampled) {
        di.reservoir = mDiTraced[k & 1u];
        di.reservoirPrev = mDiFinal[(k & 1u) ^ 1u];
        di.reservoirTemp = mDiTemp;
        rtThrow(rt_di_path_gen(mScene.handle(), cams, &f, &di, &diSettings, traceStream), "rt_di_path_gen");
        // after this frame's generation and the frame before's DI reuse (which may have run on
        // another stream if the direct stream changed)
        if (hipEventRecord(mDiGenEv[e], traceStream) != hipSuccess ||
            hipStreamWaitEvent(diStream, mDiGenEv[e], 0) != hipSuccess ||
            (k >= 1 && hipStreamWaitEvent(diStream, mDiDoneEv[(k - 1) % 3u], 0) != hipSuccess))
            throw std::runtime_error("hipEventRecord failed");
        rtThrow(rt_di_temporal(mScene.handle(), cams, &f, &di, &diSettings, diStream), "rt_di_temporal");
        di.reservoir = mDiFinal[k & 1u];  // spatial reuse reads Temp and writes the final reservoirs
        rtThrow(rt_di_spatial(mScene.handle(), cams, &f, &di, &diSettings, diStream), "rt_di_spatial");
        if (hipEventRecord(mDiDoneEv[e], diStream) != hipSuccess) throw std::runtime_error("hipEventRecord failed");
    }
    rtThrow(rt_gris_path_trace_groups(mScene.handle(), cams, &f, &settings, traceGroups, traceStream),
            "rt_gris_path_trace_groups");
    if (hipEventRecord(mTracedEv[e], traceStream) != hipSuccess ||
        hipStreamWaitEvent(reuseStream, mTracedEv[e], 0) != hipSuccess)
        throw std::runtime_error("hipEventRecord failed");
    rtThrow(rt_gris_temporal(mScene.handle(), cams, &f, &settings, reuseStream), "rt_gris_temporal");
    if (hipEventRecord(mTemporalEv[e], reuseStream) != hipSuccess) throw std::runtime_error("hipEventRecord failed");
    // spatial reuse reads Temp and writes the frame's final reservoirs
    f.grisReservoir = mFinal[k & 1u];
    rtThrow(rt_gris_spatial(mScene.handle(), cams, &f, &settings, reuseStream), "rt_gris_spatial");
    mFrame++;
}
rt_frame FramePipeline::bindings() const {
    const uint64_t k = mFrame ? mFrame - 1u : 0u;
    rt_frame f = frameOf(k);
    f.grisReservoir = mFinal[k & 1u];
    f.grisReservoirPrev = mFinal[(k & 1u) ^ 1u];
    return f;
}
