// Frame graphs: one GRIS frame of a view as a HIP graph per frame parity.
//
// The frame's launches are recorded through rti::launch (same launchers as
// the eager entry points, so the kernels and arguments are identical), turned
// into a chain of kernel nodes on the first launch of a parity, and on later
// launches re-recorded with the new cameras and written into the instantiated
// graph with hipGraphExecKernelNodeSetParams before it is replayed
// (rti::KernelChain, also used by batched frames). The graph holds no memset or
// copy nodes: counters that a frame needs cleared are cleared by kernels (the
// pixel queue re-arms itself, the gradient tables are zeroed by fill kernels).
#include "rt_internal.h"

namespace rti {
hipError_t launchGBuffer(const rt_device_scene*, const rt_camera&, float*, uint32_t*, uint16_t*, hipStream_t);
hipError_t launchGrisPathTrace(const rt_device_scene*, const rt_camera*, const rt_frame&, const rt_gris_settings&,
                               hipStream_t, rtd::WorkQueue* queue, uint32_t maxGroups);
hipError_t launchGrisTemporal(const rt_device_scene*, const rt_camera*, const rt_frame&, const rt_gris_settings&,
                              hipStream_t);
hipError_t launchGrisSpatial(const rt_device_scene*, const rt_camera*, const rt_frame&, const rt_gris_settings&,
                             hipStream_t);
hipError_t launchLearnedRefresh(const rt_device_scene*, rt_mlp*, const rt_camera*, const rt_frame&, hipStream_t);
hipError_t mlpGradients(rt_mlp* m, const float* pos, const float* target, uint32_t n, float* loss, hipStream_t st);
hipError_t mlpAdam(rt_mlp* m, hipStream_t st);
hipError_t launchGuideRefresh(const rt_mlp* m, const rt_camera* cams, const rt_frame& f, hipStream_t stream);
hipError_t launchGuideGather(rt_mlp* m, const rt_camera* cams, const rt_frame& f, hipStream_t stream);
}  // namespace rti

struct rt_frame_graph {
    const rt_device_scene* scene = nullptr;
    rt_frame frames[2]{};
    rt_gris_settings settings{};
    rt_mlp* mlp = nullptr;
    rt_mlp* guide = nullptr;          // the learned proposal's network (settings.guidedProposal)
    rtd::WorkQueue* queue = nullptr;  // this graph's own pixel queue
    rti::KernelChain chain[2];
};

namespace {

rt_status invalid(const char* m) { rti::setError(m); return RT_ERR_INVALID_ARGUMENT; }

// The frame's launches for one parity, recorded (nothing is launched).
hipError_t recordFrame(rt_frame_graph* g, uint32_t parity, const rt_camera* cams, std::vector<rti::KernelLaunch>& out) {
    out.clear();
    rti::tlsRecorder = &out;
    const rt_frame& f = g->frames[parity];
    const bool learned = g->settings.targetFunction == 1u;
    const bool guided = g->settings.guidedProposal == 1u;
    hipError_t e = rti::launchGBuffer(g->scene, cams[0], const_cast<float*>(f.depthNormal),
                                      const_cast<uint32_t*>(f.albedoMatId), const_cast<uint16_t*>(f.motionVector),
                                      nullptr);
    if (e == hipSuccess && guided) e = rti::launchGuideRefresh(g->guide, cams, f, nullptr);
    if (e == hipSuccess) e = rti::launchGrisPathTrace(g->scene, cams, f, g->settings, nullptr, g->queue, 0u);
    if (e == hipSuccess && learned) e = rti::launchLearnedRefresh(g->scene, g->mlp, cams, f, nullptr);
    if (e == hipSuccess) e = rti::launchGrisTemporal(g->scene, cams, f, g->settings, nullptr);
    if (e == hipSuccess) e = rti::launchGrisSpatial(g->scene, cams, f, g->settings, nullptr);
    if (e == hipSuccess && learned) {
        e = rti::mlpGradients(g->mlp, g->mlp->batchPos, g->mlp->batchTarget, g->mlp->batchSize, nullptr, nullptr);
        if (e == hipSuccess) e = rti::mlpAdam(g->mlp, nullptr);
    }
    if (e == hipSuccess && guided) {  // rt_gris_guide_train: gather, then one step when the batch is not empty
        e = rti::launchGuideGather(g->guide, cams, f, nullptr);
        if (e == hipSuccess && g->guide->batchSize) {
            e = rti::mlpGradients(g->guide, g->guide->batchPos, g->guide->batchTarget, g->guide->batchSize, nullptr,
                                  nullptr);
            if (e == hipSuccess) e = rti::mlpAdam(g->guide, nullptr);
        }
    }
    rti::tlsRecorder = nullptr;
    return e;
}

}  // namespace

extern "C" {

rt_status rt_frame_graph_create(const rt_device_scene* scene, const rt_frame frames[2], const rt_gris_settings* settings,
                                rt_mlp* mlp, rt_mlp* guide, rt_frame_graph** out) {
    if (!out) return invalid("out is null");
    *out = nullptr;
    if (!scene || !frames || !settings) return invalid("null argument");
    if (settings->maxDepth == 0 || settings->maxDepth > 255 || settings->shiftType > 2 ||
        settings->spatialNeighbors > 64 || settings->targetFunction > 1 || settings->unbiasedReuse > 1 ||
        settings->guidedProposal > 1)
        return invalid("invalid settings");
    if (settings->unbiasedReuse && settings->spatialNeighbors > 8)
        return invalid("unbiasedReuse supports at most 8 spatialNeighbors");
    if (settings->targetFunction == 1u && (!mlp || mlp->outputs != (uint32_t)rtm::kOut))
        return invalid("the learned target needs its radiance network (rt_mlp_create)");
    if (settings->guidedProposal == 1u) {
        if (!guide || guide->outputs != (uint32_t)rtm::kGuideOut)
            return invalid("the learned proposal needs its proposal network (rt_guide_create)");
        if (mlp && mlp == guide) return invalid("the target and proposal networks must differ");
        for (int j = 0; j < 2; j++)
            if (!frames[j].guideLobes || !frames[j].guideLobesPrev || !frames[j].guideSamples)
                return invalid("guidedProposal needs guideLobes, guideLobesPrev and guideSamples");
    }
    for (int j = 0; j < 2; j++) {
        const rt_frame& f = frames[j];
        if (f.width == 0 || f.height == 0 || f.width >= 32768 || f.height >= 32768 || f.width != frames[0].width ||
            f.height != frames[0].height)
            return invalid("both parities need the same frame size in [1, 32767]");
        if (!f.depthNormal || !f.albedoMatId || !f.motionVector || !f.grisReservoir || !f.grisReservoirPrev ||
            !f.grisReservoirTemp || !f.indirectOutput)
            return invalid("frame bindings are incomplete");
        // the graph always records temporal reuse, which reads the previous G-buffer
        if (!f.depthNormalPrev || !f.albedoMatIdPrev) return invalid("previous G-buffer is null");
    }
    if (scene->numLights == 0) return invalid("the scene has no lights");
    rt_frame_graph* g = new rt_frame_graph();
    g->scene = scene;
    g->frames[0] = frames[0];
    g->frames[1] = frames[1];
    g->settings = *settings;
    g->mlp = mlp;
    g->guide = settings->guidedProposal == 1u ? guide : nullptr;
    if (!rti::hipCheck(hipMalloc(&g->queue, sizeof(rtd::WorkQueue)), "hipMalloc") ||
        !rti::hipCheck(hipMemset(g->queue, 0, sizeof(rtd::WorkQueue)), "hipMemset")) {
        (void)hipFree(g->queue);
        delete g;
        return RT_ERR_HIP;
    }
    *out = g;
    return RT_OK;
}

rt_status rt_frame_graph_launch(rt_frame_graph* g, uint32_t parity, const rt_camera cams[2], hipStream_t stream) {
    if (!g || !cams) return invalid("null argument");
    if (parity > 1) return invalid("parity must be 0 or 1");
    if (cams[0].filmSize[0] != g->frames[0].width || cams[0].filmSize[1] != g->frames[0].height)
        return invalid("camera filmSize differs from the frame size");
    std::vector<rti::KernelLaunch> L;
    RT_HIP(recordFrame(g, parity, cams, L));
    RT_HIP(g->chain[parity].update(L));
    RT_HIP(g->chain[parity].launch(stream));
    return RT_OK;
}

rt_status rt_frame_graph_destroy(rt_frame_graph* g) {
    if (!g) return invalid("graph is null");
    g->chain[0].reset();
    g->chain[1].reset();
    (void)hipFree(g->queue);
    delete g;
    return RT_OK;
}

}  // extern "C"

namespace rti {

namespace {
hipKernelNodeParams nodeParams(KernelLaunch& k) {
    hipKernelNodeParams p{};
    p.func = const_cast<void*>(k.func);
    p.gridDim = k.grid;
    p.blockDim = k.block;
    p.sharedMemBytes = 0;
    p.kernelParams = k.args.data();
    p.extra = nullptr;
    return p;
}
bool sameDim(const dim3& a, const dim3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
}  // namespace

void KernelChain::reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    nodes.clear();
    funcs.clear();
    grids.clear();
    blocks.clear();
}

hipError_t KernelChain::update(std::vector<KernelLaunch>& L) {
    bool same = exec && L.size() == nodes.size();
    for (size_t i = 0; same && i < L.size(); i++)
        same = L[i].func == funcs[i] && sameDim(L[i].grid, grids[i]) && sameDim(L[i].block, blocks[i]);
    if (same) {
        for (size_t i = 0; i < L.size(); i++) {
            hipKernelNodeParams p = nodeParams(L[i]);
            hipError_t e = hipGraphExecKernelNodeSetParams(exec, nodes[i], &p);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    reset();
    hipError_t e = hipGraphCreate(&graph, 0);
    hipGraphNode_t prev = nullptr;
    for (size_t i = 0; e == hipSuccess && i < L.size(); i++) {
        hipKernelNodeParams p = nodeParams(L[i]);
        hipGraphNode_t node = nullptr;
        e = hipGraphAddKernelNode(&node, graph, prev ? &prev : nullptr, prev ? 1 : 0, &p);
        nodes.push_back(node);
        funcs.push_back(L[i].func);
        grids.push_back(L[i].grid);
        blocks.push_back(L[i].block);
        prev = node;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) reset();
    return e;
}

hipError_t KernelChain::launch(hipStream_t stream) const {
    if (!exec) return hipErrorInvalidValue;
    return hipGraphLaunch(exec, stream);
}

}  // namespace rti
