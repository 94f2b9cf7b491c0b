// Batched views behind the rt_batch_* entry points of include/restirpt.h:
// view descriptors are staged in pinned host memory (two buffers, reused
// only after the copy of two frames ago has completed), copied to the device
// on the caller's stream, and the batch kernels of rt_passes.hip run over
// them. Each view owns a pixel queue of the batch.
//
// Views are grouped by their form (target function, resampling weights,
// learned proposal): the descriptors of one form are contiguous, its G-buffer
// and reuse passes are one launch each, and one persistent path-trace launch
// covers all views. The learned networks' per-view steps run around the
// batched passes in the order of a single view's frame: proposal refresh
// before the path trace, target refresh after it, training after spatial reuse.
//
// With tuning.graph the frame's launches are recorded and replayed as one HIP
// graph (rti::KernelChain): the descriptors are still copied on the stream
// before it, so a frame with the views' shapes unchanged only rewrites the
// kernel nodes' arguments. With tuning.shading = 1 the path trace runs as the
// wavefront variant with material-sorted shading, whose path state (one slot
// per pixel of the batch) the batch allocates on the first such frame and
// grows when a frame has more pixels.
#include "rt_internal.h"
#include <algorithm>
#include <cstring>
#include <vector>

namespace rti {
size_t batchDescBytes();
uint32_t batchViewTiles(const rt_frame& f);
void batchFillDesc(void* dst, const rt_device_scene* scene, const rt_camera* cams, const rt_frame& f,
                   const rt_gris_settings& set, uint32_t tileBase, rtd::WorkQueue* queue, uint32_t queueShift,
                   uint32_t pixBase);
size_t wavefrontBytes(uint32_t n);
void wavefrontCarve(void* base, uint32_t n, WavefrontBuffers& wb);
hipError_t launchBatchWavefront(const void* descs, uint32_t numViews, uint32_t totalPixels, uint32_t maxDepth,
                                const WavefrontBuffers& wb, uint32_t groups, hipStream_t stream);
hipError_t batchOccupancy(uint32_t perCU[2][kBatchWaveVariants], uint32_t& cus);
hipError_t launchBatchGBuffer(const void* descs, uint32_t numViews, uint32_t totalTiles, uint32_t xcdSwizzle,
                              hipStream_t stream);
hipError_t launchBatchPathTrace(const void* descs, uint32_t numViews, uint32_t persistentGroups, bool guided,
                                uint32_t waveVariant, hipStream_t stream);
hipError_t launchBatchReuse(const void* descs, uint32_t numViews, uint32_t totalTiles, uint32_t xcdSwizzle,
                            bool learned, bool unbiased, bool guided, hipStream_t stream);
hipError_t launchLearnedRefresh(const rt_device_scene*, rt_mlp*, const rt_camera*, const rt_frame&, hipStream_t);
hipError_t mlpGradients(rt_mlp* m, const float* pos, const float* target, uint32_t n, float* loss, hipStream_t st);
hipError_t mlpAdam(rt_mlp* m, hipStream_t st);
hipError_t launchGuideRefresh(const rt_mlp* m, const rt_camera* cams, const rt_frame& f, hipStream_t stream);
hipError_t launchGuideGather(rt_mlp* m, const rt_camera* cams, const rt_frame& f, hipStream_t stream);
}  // namespace rti

struct rt_batch {
    uint32_t maxViews = 0;
    size_t descBytes = 0;
    void* staging[2] = {nullptr, nullptr};   // pinned host descriptors
    hipEvent_t copied[2] = {nullptr, nullptr};
    void* descs = nullptr;                   // device descriptors
    rtd::WorkQueue* queues = nullptr;        // one per view
    uint64_t frames = 0;
    rt_batch_tuning tuning{0u, 16u, 0u, 0u, 0u, 0u};
    uint32_t perCU[2][rti::kBatchWaveVariants] = {};  // resident path-trace workgroups per CU [guided][variant]
    uint32_t cus = 1;
    std::vector<uint32_t> order;             // views sorted by form (stable)
    rti::KernelChain chain;                  // the frame's graph (tuning.graph)
    std::vector<rti::KernelLaunch> recorded;
    void* wfMem = nullptr;                   // wavefront path state (tuning.shading = 1)
    rti::WavefrontBuffers wf;
};

namespace {

rt_status invalid(const char* m) { rti::setError(m); return RT_ERR_INVALID_ARGUMENT; }

constexpr uint64_t kMaxTiles = 1ull << 25;   // tiles of one form per frame (2^31 pixels)
constexpr uint32_t kMaxChunk = 4096;          // tiles per XCD chunk

// Path-trace queue tiles of 2^shift x 2^(6-shift) pixels: 8x8, 16x4, 64x1.
constexpr uint32_t kQueueShift[3] = {3u, 4u, 6u};

// The batched path-trace variant of a tuning's `waves` (0: the default, 4).
uint32_t waveVariant(uint32_t waves) {
    for (uint32_t w = 0; w < rti::kBatchWaveVariants; w++)
        if ((int)waves == rti::kBatchWaves[w]) return w;
    return 1u;
}

rt_status checkTuning(const rt_batch_tuning& t) {
    if (t.xcdSwizzle > kMaxChunk) return invalid("xcdSwizzle (tiles per XCD chunk) must be <= 4096");
    if (t.waves != 0u && t.waves != 3u && t.waves != 4u && t.waves != 5u) return invalid("waves must be 0, 3, 4 or 5");
    if (t.queueTile > 2u) return invalid("queueTile must be 0 (8x8), 1 (16x4) or 2 (64x1)");
    if (t.graph > 1u) return invalid("graph must be 0 or 1");
    if (t.shading > 1u) return invalid("shading must be 0 (persistent) or 1 (wavefront, material-sorted)");
    return RT_OK;
}

void freeBatch(rt_batch* b) {
    for (int k = 0; k < 2; k++) {
        if (b->copied[k]) (void)hipEventDestroy(b->copied[k]);
        if (b->staging[k]) (void)hipHostFree(b->staging[k]);
    }
    (void)hipFree(b->descs);
    (void)hipFree(b->queues);
    (void)hipFree(b->wfMem);
    delete b;
}

rt_status checkView(const rt_view& v) {
    const rt_frame& f = v.frame;
    if (!v.scene) return invalid("view scene is null");
    if (f.width == 0 || f.height == 0 || f.width >= 32768 || f.height >= 32768)
        return invalid("view frame size must be in [1, 32767]");
    if (v.cams[0].filmSize[0] != f.width || v.cams[0].filmSize[1] != f.height)
        return invalid("view camera filmSize differs from its frame size");
    if (!f.depthNormal || !f.albedoMatId || !f.motionVector || !f.depthNormalPrev || !f.albedoMatIdPrev ||
        !f.grisReservoir || !f.grisReservoirPrev || !f.grisReservoirTemp || !f.indirectOutput)
        return invalid("view frame bindings are incomplete");
    const rt_gris_settings& s = v.settings;
    if (s.maxDepth == 0 || s.maxDepth > 255 || s.shiftType > 2 || s.spatialNeighbors > 64 || s.targetFunction > 1 ||
        s.unbiasedReuse > 1 || s.guidedProposal > 1)
        return invalid("invalid view settings");
    if (s.unbiasedReuse && s.spatialNeighbors > 8) return invalid("unbiasedReuse supports at most 8 spatialNeighbors");
    if (s.targetFunction && (!v.target || v.target->outputs != (uint32_t)rtm::kOut))
        return invalid("a view with the learned target needs its radiance network (rt_mlp_create) as target");
    if (s.guidedProposal) {
        if (!v.guide || v.guide->outputs != (uint32_t)rtm::kGuideOut)
            return invalid("a view with the learned proposal needs its proposal network (rt_guide_create) as guide");
        if (!f.guideLobes || !f.guideLobesPrev || !f.guideSamples)
            return invalid("guidedProposal needs guideLobes, guideLobesPrev and guideSamples");
    }
    return RT_OK;
}

// Form of a view: bit 0 learned target, bit 1 unbiased reuse, bit 2 learned proposal.
uint32_t formOf(const rt_gris_settings& s) {
    return (s.targetFunction ? 1u : 0u) | (s.unbiasedReuse ? 2u : 0u) | (s.guidedProposal ? 4u : 0u);
}

struct Group { uint32_t first, count, form; uint64_t tiles; };

// The frame's kernel launches (launched, or recorded while a recorder is
// installed): per form G-buffer, proposal refreshes, one persistent path trace,
// target refreshes, per form temporal + spatial reuse, training steps.
hipError_t frameLaunches(rt_batch* b, const rt_view* views, uint32_t n, const std::vector<Group>& groups,
                         const rt_batch_tuning& t, bool anyGuided, bool wavefront, uint32_t pixels, uint32_t maxDepth,
                         hipStream_t stream) {
    const uint8_t* descs = static_cast<const uint8_t*>(b->descs);
    hipError_t e = hipSuccess;
    for (const Group& gr : groups) {
        e = rti::launchBatchGBuffer(descs + b->descBytes * gr.first, gr.count, (uint32_t)gr.tiles, t.xcdSwizzle, stream);
        if (e != hipSuccess) return e;
    }
    for (uint32_t i = 0; i < n; i++)
        if (views[i].settings.guidedProposal) {
            e = rti::launchGuideRefresh(views[i].guide, views[i].cams, views[i].frame, stream);
            if (e != hipSuccess) return e;
        }
    const uint32_t wv = waveVariant(t.waves);
    const uint32_t resident = b->cus * b->perCU[anyGuided ? 1 : 0][wv];
    const uint32_t groupsPT = t.persistentGroups ? t.persistentGroups : resident;
    e = wavefront ? rti::launchBatchWavefront(b->descs, n, pixels, maxDepth, b->wf, groupsPT, stream)
                  : rti::launchBatchPathTrace(b->descs, n, groupsPT, anyGuided, wv, stream);
    if (e != hipSuccess) return e;
    for (uint32_t i = 0; i < n; i++)
        if (views[i].settings.targetFunction) {
            e = rti::launchLearnedRefresh(views[i].scene, views[i].target, views[i].cams, views[i].frame, stream);
            if (e != hipSuccess) return e;
        }
    for (const Group& gr : groups) {
        e = rti::launchBatchReuse(descs + b->descBytes * gr.first, gr.count, (uint32_t)gr.tiles, t.xcdSwizzle,
                                  (gr.form & 1u) != 0u, (gr.form & 2u) != 0u, (gr.form & 4u) != 0u, stream);
        if (e != hipSuccess) return e;
    }
    // one training step per network (rt_mlp_train_frame, rt_gris_guide_train)
    for (uint32_t i = 0; i < n; i++) {
        const rt_view& v = views[i];
        if (v.settings.targetFunction && v.target->batchSize) {
            e = rti::mlpGradients(v.target, v.target->batchPos, v.target->batchTarget, v.target->batchSize, nullptr,
                                  stream);
            if (e == hipSuccess) e = rti::mlpAdam(v.target, stream);
        }
        if (e == hipSuccess && v.settings.guidedProposal) {
            e = rti::launchGuideGather(v.guide, v.cams, v.frame, stream);
            if (e == hipSuccess && v.guide->batchSize) {
                e = rti::mlpGradients(v.guide, v.guide->batchPos, v.guide->batchTarget, v.guide->batchSize, nullptr,
                                      stream);
                if (e == hipSuccess) e = rti::mlpAdam(v.guide, stream);
            }
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

rt_status frame(rt_batch* b, const rt_view* views, uint32_t n, const rt_batch_tuning& t, hipStream_t stream) {
    if (n == 0) return RT_OK;
    if (n > b->maxViews) return invalid("more views than the batch was created for");
    std::vector<const rt_mlp*> nets;
    for (uint32_t i = 0; i < n; i++) {
        rt_status st = checkView(views[i]);
        if (st != RT_OK) return st;
        if (views[i].settings.targetFunction) nets.push_back(views[i].target);
        if (views[i].settings.guidedProposal) nets.push_back(views[i].guide);
    }
    // a network holds one training batch at a time: one view per network
    std::sort(nets.begin(), nets.end());
    if (std::adjacent_find(nets.begin(), nets.end()) != nets.end())
        return invalid("a learned network is bound to more than one view of the frame");
    std::vector<uint32_t>& order = b->order;
    order.resize(n);
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [views](uint32_t a, uint32_t c) {
        return formOf(views[a].settings) < formOf(views[c].settings);
    });
    const int k = (int)(b->frames & 1u);
    RT_HIP(hipEventSynchronize(b->copied[k]));  // staging[k] is free once its last copy ran
    uint8_t* stage = static_cast<uint8_t*>(b->staging[k]);
    // descriptors in form order; tile bases count from the first view of each form
    std::vector<Group> groups;
    bool anyGuided = false;
    const uint32_t qShift = kQueueShift[t.queueTile];
    uint64_t pixels = 0;
    uint32_t maxDepth = 0;
    for (uint32_t j = 0; j < n; j++) {
        const rt_view& v = views[order[j]];
        const uint32_t form = formOf(v.settings);
        if (groups.empty() || groups.back().form != form) groups.push_back(Group{j, 0u, form, 0ull});
        Group& gr = groups.back();
        rti::batchFillDesc(stage + b->descBytes * j, v.scene, v.cams, v.frame, v.settings, (uint32_t)gr.tiles,
                           b->queues + j, qShift, (uint32_t)pixels);
        pixels += (uint64_t)v.frame.width * v.frame.height;
        maxDepth = std::max(maxDepth, v.settings.maxDepth);
        gr.tiles += rti::batchViewTiles(v.frame);
        gr.count++;
        // a launch's grid (tiles rounded up to 8 chunks) stays within 2^26 64-lane workgroups
        if (gr.tiles > kMaxTiles) return invalid("too many tiles in one batch");
        anyGuided = anyGuided || v.settings.guidedProposal;
    }
    // the wavefront variant: plain kernels only (a counting build runs the persistent kernel),
    // views without the learned proposal, one path slot per pixel of the batch
    const bool wavefront = t.shading == 1u && !rti::gCountersEnabled;
    if (wavefront && anyGuided) return invalid("the wavefront shading variant does not run the learned proposal");
    if (wavefront && pixels >= (1ull << 31)) return invalid("the wavefront shading variant holds at most 2^31 pixels");
    if (wavefront && b->wf.capacity < pixels) {
        RT_HIP(hipStreamSynchronize(stream));  // the old state may still be in use
        RT_HIP(hipFree(b->wfMem));
        b->wfMem = nullptr;
        b->wf = rti::WavefrontBuffers{};
        RT_HIP(hipMalloc(&b->wfMem, rti::wavefrontBytes((uint32_t)pixels)));
        rti::wavefrontCarve(b->wfMem, (uint32_t)pixels, b->wf);
    }
    RT_HIP(hipMemcpyAsync(b->descs, stage, b->descBytes * n, hipMemcpyHostToDevice, stream));
    RT_HIP(hipEventRecord(b->copied[k], stream));
    if (t.graph) {
        b->recorded.clear();
        rti::tlsRecorder = &b->recorded;
        const hipError_t e = frameLaunches(b, views, n, groups, t, anyGuided, wavefront, (uint32_t)pixels, maxDepth,
                                           nullptr);
        rti::tlsRecorder = nullptr;
        RT_HIP(e);
        RT_HIP(b->chain.update(b->recorded));
        b->recorded.clear();
        RT_HIP(b->chain.launch(stream));
    } else {
        RT_HIP(frameLaunches(b, views, n, groups, t, anyGuided, wavefront, (uint32_t)pixels, maxDepth, stream));
    }
    b->frames++;
    return RT_OK;
}

// The view with its parity flipped: this frame's G-buffer, reservoirs and
// proposal lobes become the previous frame's and the other way round.
rt_view flipped(const rt_view& v) {
    rt_view w = v;
    std::swap(w.frame.depthNormal, w.frame.depthNormalPrev);
    std::swap(w.frame.albedoMatId, w.frame.albedoMatIdPrev);
    std::swap(w.frame.grisReservoir, w.frame.grisReservoirPrev);
    std::swap(w.frame.guideLobes, w.frame.guideLobesPrev);
    return w;
}

}  // namespace

extern "C" {

rt_status rt_batch_create(uint32_t maxViews, rt_batch** out) {
    if (!out) return invalid("out is null");
    *out = nullptr;
    if (maxViews == 0 || maxViews > 4096) return invalid("maxViews must be in [1, 4096]");
    rt_batch* b = new rt_batch();
    b->maxViews = maxViews;
    // the kernels index the descriptor array as ViewDesc[]: stage at exactly that stride
    b->descBytes = rti::batchDescBytes();
    const size_t bytes = b->descBytes * maxViews;
    bool ok = rti::hipCheck(hipHostMalloc(&b->staging[0], bytes), "hipHostMalloc") &&
              rti::hipCheck(hipHostMalloc(&b->staging[1], bytes), "hipHostMalloc") &&
              rti::hipCheck(hipEventCreateWithFlags(&b->copied[0], hipEventDisableTiming), "hipEventCreate") &&
              rti::hipCheck(hipEventCreateWithFlags(&b->copied[1], hipEventDisableTiming), "hipEventCreate") &&
              rti::hipCheck(hipMalloc(&b->descs, bytes), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&b->queues, sizeof(rtd::WorkQueue) * maxViews), "hipMalloc") &&
              rti::hipCheck(hipMemset(b->queues, 0, sizeof(rtd::WorkQueue) * maxViews), "hipMemset") &&
              rti::hipCheck(rti::batchOccupancy(b->perCU, b->cus), "occupancy");
    if (!ok) { freeBatch(b); return RT_ERR_HIP; }
    *out = b;
    return RT_OK;
}

rt_status rt_batch_destroy(rt_batch* b) {
    if (!b) return invalid("batch is null");
    (void)hipDeviceSynchronize();
    freeBatch(b);
    return RT_OK;
}

rt_status rt_batch_frame(rt_batch* b, const rt_view* views, uint32_t numViews, hipStream_t stream) {
    if (!b || (numViews && !views)) return invalid("null argument");
    return frame(b, views, numViews, b->tuning, stream);
}

rt_status rt_batch_set_tuning(rt_batch* b, const rt_batch_tuning* t) {
    if (!b || !t) return invalid("null argument");
    rt_status st = checkTuning(*t);
    if (st != RT_OK) return st;
    b->tuning = *t;
    return RT_OK;
}

rt_status rt_batch_get_tuning(const rt_batch* b, rt_batch_tuning* t) {
    if (!b || !t) return invalid("null argument");
    *t = b->tuning;
    return RT_OK;
}

rt_status rt_batch_autotune(rt_batch* b, const rt_view* views, uint32_t numViews, uint32_t frames, hipStream_t stream,
                            rt_batch_tuning* best, float* msPerFrame) {
    if (!b || !views || numViews == 0 || frames == 0) return invalid("invalid argument");
    if (numViews > b->maxViews) return invalid("more views than the batch was created for");
    bool guided = false;
    for (uint32_t i = 0; i < numViews; i++) guided = guided || views[i].settings.guidedProposal;
    // frames alternate the views' parity, as a caller's frames do, so that
    // temporal reuse reads the previous frame's reservoirs and G-buffer
    std::vector<rt_view> other(numViews);
    for (uint32_t i = 0; i < numViews; i++) other[i] = flipped(views[i]);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    RT_HIP(hipEventCreate(&e0));
    RT_HIP(hipEventCreate(&e1));
    struct Tried { rt_batch_tuning t; float ms; };
    std::vector<Tried> tried;
    auto same = [](const rt_batch_tuning& a, const rt_batch_tuning& c) {
        return a.persistentGroups == c.persistentGroups && a.xcdSwizzle == c.xcdSwizzle && a.waves == c.waves &&
               a.queueTile == c.queueTile && a.graph == c.graph && a.shading == c.shading;
    };
    uint32_t parity = 0;
    // mean frame time of a tuning over `frames` frames after one warm-up frame
    auto timeOf = [&](const rt_batch_tuning& t, float& ms) -> rt_status {
        for (const Tried& x : tried)
            if (same(x.t, t)) { ms = x.ms; return RT_OK; }
        rt_status st = RT_OK;
        for (uint32_t f = 0; f <= frames && st == RT_OK; f++) {
            if (f == 1 && !rti::hipCheck(hipEventRecord(e0, stream), "hipEventRecord")) return RT_ERR_HIP;
            st = frame(b, parity ? other.data() : views, numViews, t, stream);
            parity ^= 1u;
        }
        if (st != RT_OK) return st;
        if (!rti::hipCheck(hipEventRecord(e1, stream), "hipEventRecord") ||
            !rti::hipCheck(hipEventSynchronize(e1), "hipEventSynchronize") ||
            !rti::hipCheck(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime"))
            return RT_ERR_HIP;
        ms /= (float)frames;
        tried.push_back(Tried{t, ms});
        return RT_OK;
    };
    // coordinate descent from the defaults: register budget, queue tile shape,
    // persistent workgroups, XCD chunk, graph replay -- each searched with the
    // others held at their best so far
    rt_batch_tuning cur{0u, 16u, 4u, 0u, 0u, 0u};
    float curMs = 0.0f;
    rt_status st = timeOf(cur, curMs);
    auto consider = [&](const rt_batch_tuning& t) {
        float ms = 0.0f;
        if (st == RT_OK) st = timeOf(t, ms);
        if (st == RT_OK && ms < curMs) { cur = t; curMs = ms; }
    };
    for (uint32_t w = 0; w < rti::kBatchWaveVariants; w++) {
        rt_batch_tuning t = cur;
        t.waves = (uint32_t)rti::kBatchWaves[w];
        consider(t);
    }
    for (uint32_t q = 0; q < 3u; q++) {
        rt_batch_tuning t = cur;
        t.queueTile = q;
        consider(t);
    }
    const uint32_t full = b->cus * b->perCU[guided ? 1 : 0][waveVariant(cur.waves)];
    for (uint32_t q : {2u, 3u, 4u}) {
        rt_batch_tuning t = cur;
        t.persistentGroups = q == 4u ? 0u : std::max(1u, full * q / 4u);
        consider(t);
    }
    for (uint32_t chunk : {0u, 16u, 128u}) {
        rt_batch_tuning t = cur;
        t.xcdSwizzle = chunk;
        consider(t);
    }
    for (uint32_t g = 0; g < 2u; g++) {
        rt_batch_tuning t = cur;
        t.graph = g;
        consider(t);
    }
    if (!guided) {  // the wavefront variant, at its own grid sizes
        for (uint32_t q : {2u, 4u, 8u}) {
            rt_batch_tuning t = cur;
            t.shading = 1u;
            t.persistentGroups = b->cus * q;
            consider(t);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != RT_OK) return st;
    b->tuning = cur;
    if (best) *best = cur;
    if (msPerFrame) *msPerFrame = curMs;
    return RT_OK;
}

}  // extern "C"
