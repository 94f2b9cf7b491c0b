// Batched views behind the rt_batch_* entry points of include/restirpt.h:
// view descriptors are staged in pinned host memory (two buffers, reused
// only after the copy of two frames ago has completed), copied to the device
// on the caller's stream, and the four batch kernels of rt_passes.hip run
// over them. Each view owns a pixel queue of the batch.
#include "rt_internal.h"
#include <cstring>
#include <vector>

namespace rti {
size_t batchDescBytes();
uint32_t batchViewTiles(const rt_frame& f);
void batchFillDesc(void* dst, const rt_device_scene* scene, const rt_camera* cams, const rt_frame& f,
                   const rt_gris_settings& set, uint32_t tileBase, rtd::WorkQueue* queue);
hipError_t batchOccupancy(uint32_t& perCU, uint32_t& cus);
hipError_t launchBatchFrame(const void* descs, uint32_t numViews, uint32_t totalTiles, uint32_t persistentGroups,
                            uint32_t xcdSwizzle, hipStream_t stream);
}  // namespace rti

struct rt_batch {
    uint32_t maxViews = 0;
    size_t descBytes = 0;
    void* staging[2] = {nullptr, nullptr};   // pinned host descriptors
    hipEvent_t copied[2] = {nullptr, nullptr};
    void* descs = nullptr;                   // device descriptors
    rtd::WorkQueue* queues = nullptr;        // one per view
    uint64_t frames = 0;
    rt_batch_tuning tuning{0u, 1u};
    uint32_t perCU = 1, cus = 1;
};

namespace {

rt_status invalid(const char* m) { rti::setError(m); return RT_ERR_INVALID_ARGUMENT; }

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
    if (s.maxDepth == 0 || s.maxDepth > 255 || s.shiftType > 2 || s.spatialNeighbors > 64)
        return invalid("invalid view settings");
    if (s.targetFunction != 0u) return invalid("batched views use the luminance target");
    if (s.unbiasedReuse != 0u) return invalid("batched views use the reference's resampling weights");
    if (s.guidedProposal != 0u) return invalid("batched views sample with the BSDFs (no learned proposal)");
    return RT_OK;
}

rt_status frame(rt_batch* b, const rt_view* views, uint32_t n, const rt_batch_tuning& t, hipStream_t stream) {
    if (n == 0) return RT_OK;
    if (n > b->maxViews) return invalid("more views than the batch was created for");
    for (uint32_t i = 0; i < n; i++) {
        rt_status st = checkView(views[i]);
        if (st != RT_OK) return st;
    }
    const int k = (int)(b->frames & 1u);
    RT_HIP(hipEventSynchronize(b->copied[k]));  // staging[k] is free once its last copy ran
    uint8_t* stage = static_cast<uint8_t*>(b->staging[k]);
    uint64_t tiles = 0;
    for (uint32_t i = 0; i < n; i++) {
        rti::batchFillDesc(stage + b->descBytes * i, views[i].scene, views[i].cams, views[i].frame, views[i].settings,
                           (uint32_t)tiles, b->queues + i);
        tiles += rti::batchViewTiles(views[i].frame);
    }
    if (tiles >= 0xffffffffull) return invalid("too many tiles in one batch");
    RT_HIP(hipMemcpyAsync(b->descs, stage, b->descBytes * n, hipMemcpyHostToDevice, stream));
    RT_HIP(hipEventRecord(b->copied[k], stream));
    const uint32_t groups = t.persistentGroups ? t.persistentGroups : b->cus * b->perCU;
    RT_HIP(rti::launchBatchFrame(b->descs, n, (uint32_t)tiles, groups, t.xcdSwizzle, stream));
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
    const uint32_t full = b->cus * b->perCU;
    std::vector<rt_batch_tuning> cand;
    for (uint32_t swz = 0; swz < 2; swz++)
        for (uint32_t q : {2u, 3u, 4u}) cand.push_back({std::max(1u, full * q / 4u), swz});
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
