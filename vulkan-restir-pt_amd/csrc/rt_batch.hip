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
#include "rt_internal.h"
#include <algorithm>
#include <cstring>
#include <vector>

namespace rti {
size_t batchDescBytes();
uint32_t batchViewTiles(const rt_frame& f);
void batchFillDesc(void* dst, const rt_device_scene* scene, const rt_camera* cams, const rt_frame& f,
                   const rt_gris_settings& set, uint32_t tileBase, rtd::WorkQueue* queue);
hipError_t batchOccupancy(uint32_t& perCU, uint32_t& perCUGuided, uint32_t& cus);
hipError_t launchBatchGBuffer(const void* descs, uint32_t numViews, uint32_t totalTiles, uint32_t xcdSwizzle,
                              hipStream_t stream);
hipError_t launchBatchPathTrace(const void* descs, uint32_t numViews, uint32_t persistentGroups, bool guided,
                                hipStream_t stream);
hipError_t launchBatchReuse(const void* descs, uint32_t numViews, uint32_t totalTiles, uint32_t xcdSwizzle,
                            bool learned, bool unbiased, bool guided, hipStream_t stream);
}  // namespace rti

struct rt_batch {
    uint32_t maxViews = 0;
    size_t descBytes = 0;
    void* staging[2] = {nullptr, nullptr};   // pinned host descriptors
    hipEvent_t copied[2] = {nullptr, nullptr};
    void* descs = nullptr;                   // device descriptors
    rtd::WorkQueue* queues = nullptr;        // one per view
    uint64_t frames = 0;
    rt_batch_tuning tuning{0u, 16u};
    uint32_t perCU = 1, perCUGuided = 1, cus = 1;
    std::vector<uint32_t> order;             // views sorted by form (stable)
};

namespace {

rt_status invalid(const char* m) { rti::setError(m); return RT_ERR_INVALID_ARGUMENT; }

constexpr uint64_t kMaxTiles = 1ull << 25;   // tiles of one form per frame (2^31 pixels)
constexpr uint32_t kMaxChunk = 4096;          // tiles per XCD chunk

void freeBatch(rt_batch* b) {
    for (int k = 0; k < 2; k++) {
        if (b->copied[k]) (void)hipEventDestroy(b->copied[k]);
        if (b->staging[k]) (void)hipHostFree(b->staging[k]);
    }
    (void)hipFree(b->descs);
    (void)hipFree(b->queues);
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
    struct Group { uint32_t first, count, form; uint64_t tiles; };
    std::vector<Group> groups;
    bool anyGuided = false;
    for (uint32_t j = 0; j < n; j++) {
        const rt_view& v = views[order[j]];
        const uint32_t form = formOf(v.settings);
        if (groups.empty() || groups.back().form != form) groups.push_back(Group{j, 0u, form, 0ull});
        Group& gr = groups.back();
        rti::batchFillDesc(stage + b->descBytes * j, v.scene, v.cams, v.frame, v.settings, (uint32_t)gr.tiles,
                           b->queues + j);
        gr.tiles += rti::batchViewTiles(v.frame);
        gr.count++;
        // a launch's grid (tiles rounded up to 8 chunks) stays within 2^26 64-lane workgroups
        if (gr.tiles > kMaxTiles) return invalid("too many tiles in one batch");
        anyGuided = anyGuided || v.settings.guidedProposal;
    }
    RT_HIP(hipMemcpyAsync(b->descs, stage, b->descBytes * n, hipMemcpyHostToDevice, stream));
    RT_HIP(hipEventRecord(b->copied[k], stream));
    const uint8_t* descs = static_cast<const uint8_t*>(b->descs);
    for (const Group& gr : groups)
        RT_HIP(rti::launchBatchGBuffer(descs + b->descBytes * gr.first, gr.count, (uint32_t)gr.tiles, t.xcdSwizzle,
                                       stream));
    rt_status st = RT_OK;
    for (uint32_t i = 0; i < n && st == RT_OK; i++)
        if (views[i].settings.guidedProposal)
            st = rt_gris_guide_refresh(views[i].scene, views[i].guide, views[i].cams, &views[i].frame, stream);
    if (st != RT_OK) return st;
    const uint32_t resident = b->cus * (anyGuided ? b->perCUGuided : b->perCU);
    RT_HIP(rti::launchBatchPathTrace(b->descs, n, t.persistentGroups ? t.persistentGroups : resident, anyGuided,
                                     stream));
    for (uint32_t i = 0; i < n && st == RT_OK; i++)
        if (views[i].settings.targetFunction)
            st = rt_gris_learned_refresh(views[i].scene, views[i].target, views[i].cams, &views[i].frame, stream);
    if (st != RT_OK) return st;
    for (const Group& gr : groups)
        RT_HIP(rti::launchBatchReuse(descs + b->descBytes * gr.first, gr.count, (uint32_t)gr.tiles, t.xcdSwizzle,
                                     (gr.form & 1u) != 0u, (gr.form & 2u) != 0u, (gr.form & 4u) != 0u, stream));
    for (uint32_t i = 0; i < n && st == RT_OK; i++) {
        const rt_view& v = views[i];
        if (v.settings.targetFunction && v.target->batchSize) st = rt_mlp_train_frame(v.target, nullptr, stream);
        if (st == RT_OK && v.settings.guidedProposal)
            st = rt_gris_guide_train(v.scene, v.guide, v.cams, &v.frame, nullptr, stream);
    }
    if (st != RT_OK) return st;
    b->frames++;
    return RT_OK;
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
              rti::hipCheck(rti::batchOccupancy(b->perCU, b->perCUGuided, b->cus), "occupancy");
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
    if (t->xcdSwizzle > kMaxChunk) return invalid("xcdSwizzle (tiles per XCD chunk) must be <= 4096");
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
    bool guided = false;
    for (uint32_t i = 0; i < numViews; i++) guided = guided || views[i].settings.guidedProposal;
    const uint32_t full = b->cus * (guided ? b->perCUGuided : b->perCU);
    std::vector<rt_batch_tuning> cand;
    for (uint32_t chunk : {0u, 16u, 128u})
        for (uint32_t q : {2u, 3u, 4u}) cand.push_back({std::max(1u, full * q / 4u), chunk});
    hipEvent_t e0 = nullptr, e1 = nullptr;
    RT_HIP(hipEventCreate(&e0));
    RT_HIP(hipEventCreate(&e1));
    float bestMs = 1e30f;
    rt_batch_tuning bestT = b->tuning;
    rt_status st = RT_OK;
    for (const rt_batch_tuning& t : cand) {
        if ((st = frame(b, views, numViews, t, stream)) != RT_OK) break;  // warm-up
        if (!rti::hipCheck(hipEventRecord(e0, stream), "hipEventRecord")) { st = RT_ERR_HIP; break; }
        for (uint32_t f = 0; f < frames && st == RT_OK; f++) st = frame(b, views, numViews, t, stream);
        if (st != RT_OK) break;
        float ms = 0.0f;
        if (!rti::hipCheck(hipEventRecord(e1, stream), "hipEventRecord") ||
            !rti::hipCheck(hipEventSynchronize(e1), "hipEventSynchronize") ||
            !rti::hipCheck(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) {
            st = RT_ERR_HIP;
            break;
        }
        ms /= (float)frames;
        if (ms < bestMs) { bestMs = ms; bestT = t; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != RT_OK) return st;
    b->tuning = bestT;
    if (best) *best = bestT;
    if (msPerFrame) *msPerFrame = bestMs;
    return RT_OK;
}

}  // extern "C"
