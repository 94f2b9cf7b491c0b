// C ABI of the gfx950 ReSTIR PT core (declarations and reference citations in
// include/restirpt.h). Validates arguments on the host, owns the device scene
// and launches the kernels of rt_bvh.hip / rt_passes.hip.
#include "rt_internal.h"
#include <cmath>
#include "rt_guide.h"
#include <cstring>
#include <mutex>

namespace rti {

thread_local std::string tlsError;
thread_local std::vector<KernelLaunch>* tlsRecorder = nullptr;
bool gCountersEnabled = false;
static rtd::TraceCounters* gCounters = nullptr;

void setError(const std::string& msg) { tlsError = msg; }
bool hipCheck(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    setError(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}
rtd::TraceCounters* counterBuffer() { return gCounters; }

hipError_t launchTraceRays(const rt_device_scene*, const float*, uint32_t, int, int, float*, float*, uint32_t*,
                           uint32_t*, hipStream_t);
hipError_t launchGBuffer(const rt_device_scene*, const rt_camera&, float*, uint32_t*, uint16_t*, hipStream_t);
hipError_t launchPtPlain(const rt_device_scene*, const rt_camera*, const rt_frame&, uint32_t, uint32_t, hipStream_t);
hipError_t launchGrisPathTrace(const rt_device_scene*, const rt_camera*, const rt_frame&, const rt_gris_settings&,
                               hipStream_t, rtd::WorkQueue* queue = nullptr, uint32_t maxGroups = 0);
hipError_t launchGrisTemporal(const rt_device_scene*, const rt_camera*, const rt_frame&, const rt_gris_settings&,
                              hipStream_t);
hipError_t launchGrisSpatial(const rt_device_scene*, const rt_camera*, const rt_frame&, const rt_gris_settings&,
                             hipStream_t);
hipError_t launchResvReset(void*, uint32_t, uint32_t, hipStream_t);
hipError_t launchLearnedRefresh(const rt_device_scene*, rt_mlp*, const rt_camera*, const rt_frame&, hipStream_t);
hipError_t launchGuideRefresh(const rt_mlp*, const rt_camera*, const rt_frame&, hipStream_t);
hipError_t launchGuideGather(rt_mlp*, const rt_camera*, const rt_frame&, hipStream_t);
hipError_t mlpGradients(rt_mlp* m, const float* pos, const float* target, uint32_t n, float* loss, hipStream_t st);
hipError_t mlpAdam(rt_mlp* m, hipStream_t st);
hipError_t launchDiPathGen(const rt_device_scene*, const rt_camera*, const rt_frame&, const rt_di_reservoirs&,
                           const rt_di_settings&, hipStream_t);
hipError_t launchDiTemporal(const rt_device_scene*, const rt_camera*, const rt_frame&, const rt_di_reservoirs&,
                            const rt_di_settings&, hipStream_t);
hipError_t launchDiSpatial(const rt_device_scene*, const rt_camera*, const rt_frame&, const rt_di_reservoirs&,
                           const rt_di_settings&, hipStream_t);
hipError_t launchDiReset(void*, uint32_t, uint32_t, hipStream_t);
hipError_t launchGiTemporal(const rt_device_scene*, const rt_camera*, const rt_frame&, const rt_gi_reservoirs&,
                            hipStream_t);
hipError_t launchGiReset(void*, uint32_t, uint32_t, hipStream_t);

}  // namespace rti

using rti::setError;

namespace {

rt_status invalid(const std::string& m) { setError(m); return RT_ERR_INVALID_ARGUMENT; }

template <class T>
bool upload(rt_device_scene* s, T** dst, const T* src, size_t count, hipStream_t stream) {
    size_t bytes = sizeof(T) * (count ? count : 1);
    if (!rti::hipCheck(hipMalloc((void**)dst, bytes), "hipMalloc")) return false;
    s->deviceBytes += bytes;
    if (count && !rti::hipCheck(hipMemcpyAsync(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice, stream), "hipMemcpyAsync"))
        return false;
    return true;
}

void freeScene(rt_device_scene* s) {
    (void)hipFree(s->nodes); (void)hipFree(s->nodes4); (void)hipFree(s->leafTris); (void)hipFree(s->primTris); (void)hipFree(s->codes); (void)hipFree(s->order);
    (void)hipFree(s->materials); (void)hipFree(s->materialIndices); (void)hipFree(s->vertices); (void)hipFree(s->indices);
    (void)hipFree(s->instances); (void)hipFree(s->lights); (void)hipFree(s->lightTable); (void)hipFree(s->queues);
    (void)hipFree(s->texDesc); (void)hipFree(s->texels); (void)hipFree(s->srgbLut);
    for (hipEvent_t& e : s->queueEvents)
        if (e) (void)hipEventDestroy(e);
    delete s;
}

rt_status checkScene(const rt_scene_arrays* a) {
    if (!a) return invalid("arrays is null");
    if (a->numMaterials == 0 || !a->materials) return invalid("a scene needs at least one material");
    if (a->numLightTable != a->numLights + 1 || !a->lightTable) return invalid("light table must hold numLights + 1 entries");
    if (a->numLights && !a->lights) return invalid("lights is null");
    if (a->numInstances && !a->instances) return invalid("instances is null");
    if (a->numTextures && !a->textures) return invalid("textures is null");
    for (uint32_t t = 0; t < a->numTextures; t++) {
        const rt_texture& T = a->textures[t];
        if (T.width == 0 || T.height == 0 || T.width > 16384 || T.height > 16384 || !T.rgba8)
            return invalid("texture " + std::to_string(t) + ": size must be in [1, 16384] with texels");
        if (T.filter > 1) return invalid("texture filter must be 0 (nearest) or 1 (linear)");
    }
    for (uint32_t m = 0; m < a->numMaterials; m++)
        if (a->materials[m].textureIdx != 0xffffffffu && a->materials[m].textureIdx >= a->numTextures)
            return invalid("material " + std::to_string(m) + ": textureIdx out of range");
    if (a->lightTable[0].failId != a->numLights) return invalid("light table entry 0 must hold numLights");
    for (uint32_t i = 1; i < a->numLightTable; i++)
        if (a->lightTable[i].failId < 1 || a->lightTable[i].failId > a->numLights)
            return invalid("light table failId out of range");
    uint64_t prims = a->numLights;
    for (uint32_t k = 0; k < a->numInstances; k++) {
        const rt_object_instance& I = a->instances[k];
        if (I.indexCount % 3) return invalid("instance indexCount must be a multiple of 3");
        if (I.indexOffset % 3) return invalid("instance indexOffset must be a multiple of 3");
        if ((uint64_t)I.indexOffset + I.indexCount > a->numIndices) return invalid("instance indices out of range");
        if ((uint64_t)I.indexOffset / 3 + I.indexCount / 3 > a->numMaterialIndices)
            return invalid("instance material indices out of range");
        prims += I.indexCount / 3;
    }
    if (prims >= 0x7fffffffull) return invalid("too many primitives");
    for (uint32_t i = 0; i < a->numIndices; i++)
        if (a->indices[i] >= a->numVertices) return invalid("vertex index out of range");
    for (uint32_t i = 0; i < a->numMaterialIndices; i++)
        if (a->materialIndices[i] < 0 || (uint32_t)a->materialIndices[i] >= a->numMaterials)
            return invalid("material index out of range");
    return RT_OK;
}

rt_status checkFrame(const rt_device_scene* s, const rt_camera* cams, const rt_frame* f) {
    if (!s) return invalid("scene is null");
    if (!cams) return invalid("cams is null");
    if (!f) return invalid("frame is null");
    if (f->width == 0 || f->height == 0 || f->width >= 32768 || f->height >= 32768)
        return invalid("frame size must be in [1, 32767]");
    if (cams[0].filmSize[0] != f->width || cams[0].filmSize[1] != f->height)
        return invalid("camera filmSize differs from the frame size");
    if (!f->depthNormal || !f->albedoMatId) return invalid("frame G-buffer is null");
    return RT_OK;
}

rt_status checkDI(const rt_device_scene* s, const rt_camera* cams, const rt_frame* f, const rt_di_reservoirs* r,
                  const rt_di_settings* set) {
    rt_status st = checkFrame(s, cams, f);
    if (st != RT_OK) return st;
    if (!r) return invalid("reservoirs is null");
    if (!set) return invalid("settings is null");
    if (set->shiftType > 2) return invalid("shiftType must be 0, 1 or 2");
    if (set->sampleType > 2) return invalid("sampleType must be 0, 1 or 2");
    if (s->numLights == 0) return invalid("the scene has no lights");
    return RT_OK;
}

rt_status checkSettings(const rt_gris_settings* st) {
    if (!st) return invalid("settings is null");
    if (st->maxDepth == 0 || st->maxDepth > 255) return invalid("maxDepth must be in [1, 255]");
    if (st->shiftType > 2) return invalid("shiftType must be 0, 1 or 2");
    if (st->spatialNeighbors > 64) return invalid("spatialNeighbors must be <= 64");
    if (st->unbiasedReuse > 1) return invalid("unbiasedReuse must be 0 or 1");
    if (st->unbiasedReuse && st->spatialNeighbors > 8)
        return invalid("unbiasedReuse supports at most 8 spatialNeighbors");
    if (st->targetFunction > 1) return invalid("targetFunction must be 0 (luminance) or 1 (learned)");
    if (st->guidedProposal > 1) return invalid("guidedProposal must be 0 or 1");
    return RT_OK;
}

// The learned proposal's frame bindings, when the settings turn it on.
rt_status checkGuide(const rt_frame* f, const rt_gris_settings* st) {
    if (st->guidedProposal && (!f->guideLobes || !f->guideLobesPrev || !f->guideSamples))
        return invalid("guidedProposal needs guideLobes, guideLobesPrev and guideSamples");
    return RT_OK;
}

#define RT_LAUNCH(expr) do { if (!rti::hipCheck((expr), #expr)) return RT_ERR_HIP; } while (0)

}  // namespace

extern "C" {

const char* rt_last_error(void) { return rti::tlsError.c_str(); }

rt_status rt_device_scene_create(const rt_scene_arrays* a, hipStream_t stream, rt_device_scene** out) {
    if (!out) return invalid("out is null");
    *out = nullptr;
    rt_status st = checkScene(a);
    if (st != RT_OK) return st;
    rt_device_scene* s = new rt_device_scene();
    s->numLights = a->numLights;
    s->numInstances = a->numInstances;
    std::vector<rtd::DevInstance> inst(a->numInstances);
    std::vector<uint32_t> offs(a->numInstances);
    uint32_t prims = a->numLights;
    for (uint32_t k = 0; k < a->numInstances; k++) {
        const rt_object_instance& I = a->instances[k];
        rtd::DevInstance& D = inst[k];
        std::memset(&D, 0, sizeof(D));
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < 3; r++) { D.m[c * 3 + r] = I.transform[c * 4 + r]; D.n[c * 3 + r] = I.transformInvT[c * 4 + r]; }
        D.indexOffset = I.indexOffset;
        D.triCount = I.indexCount / 3;
        offs[k] = prims;
        prims += I.indexCount / 3;
    }
    s->numPrims = prims;
    bool ok = upload(s, &s->materials, a->materials, a->numMaterials, stream) &&
              upload(s, &s->materialIndices, a->materialIndices, a->numMaterialIndices, stream) &&
              upload(s, &s->vertices, a->vertices, a->numVertices, stream) &&
              upload(s, &s->indices, a->indices, a->numIndices, stream) &&
              upload(s, &s->instances, inst.data(), inst.size(), stream) &&
              upload(s, &s->lights, a->lights, a->numLights, stream) &&
              upload(s, &s->lightTable, a->lightTable, a->numLightTable, stream);
    // textures: one texel array (RGBA8 words), a descriptor per texture {first texel, width,
    // height, filter} and the sRGB-to-linear table of the 256 byte values (the transfer function
    // evaluated in double, rounded to float)
    std::vector<uint4> texDesc(a->numTextures);
    std::vector<uint32_t> texels;
    for (uint32_t t = 0; t < a->numTextures; t++) {
        const rt_texture& T = a->textures[t];
        const size_t n = (size_t)T.width * T.height;
        texDesc[t] = make_uint4((uint32_t)texels.size(), T.width, T.height, T.filter);
        const size_t at = texels.size();
        texels.resize(at + n);
        std::memcpy(texels.data() + at, T.rgba8, 4 * n);
    }
    float srgb[256];
    for (int c = 0; c < 256; c++) {
        const double v = c / 255.0;
        srgb[c] = (float)(v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4));
    }
    if (ok && texels.size() >= 0xffffffffull) ok = (invalid("textures hold too many texels"), false);
    ok = ok && upload(s, &s->texDesc, texDesc.data(), texDesc.size(), stream) &&
         upload(s, &s->texels, texels.data(), texels.size(), stream) &&
         upload(s, &s->srgbLut, srgb, 256, stream);
    if (ok) {
        const size_t qb = sizeof(rtd::WorkQueue) * rt_device_scene::kQueueSlots;
        ok = rti::hipCheck(hipMalloc(&s->queues, qb), "hipMalloc") &&
             rti::hipCheck(hipMemsetAsync(s->queues, 0, qb, stream), "hipMemsetAsync");
        s->deviceBytes += qb;
        for (uint32_t k = 0; ok && k < rt_device_scene::kQueueSlots; k++)
            ok = rti::hipCheck(hipEventCreateWithFlags(&s->queueEvents[k], hipEventDisableTiming), "hipEventCreate");
    }
    if (ok) {
        rtd::DevScene& d = s->dev;
        d.materials = s->materials; d.materialIndices = s->materialIndices; d.vertices = s->vertices;
        d.indices = s->indices; d.instances = s->instances; d.lights = s->lights; d.lightTable = s->lightTable;
        d.numPrims = s->numPrims; d.numLights = s->numLights; d.numLightTable = a->numLightTable;
        d.numMaterials = a->numMaterials;
        d.numInstances = a->numInstances;
        d.textures = s->texDesc; d.texels = s->texels; d.srgb = s->srgbLut; d.numTextures = a->numTextures;
        ok = rti::buildLBVH(s, offs, stream);
        d.nodes = s->nodes4; d.leafTris = s->leafTris;
    }
    if (ok) ok = rti::hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (!ok) { freeScene(s); return RT_ERR_HIP; }
    *out = s;
    return RT_OK;
}

rt_status rt_device_scene_destroy(rt_device_scene* s) {
    if (!s) return invalid("scene is null");
    freeScene(s);
    return RT_OK;
}

rt_status rt_device_scene_info(const rt_device_scene* s, rt_scene_info* out) {
    if (!s || !out) return invalid("null argument");
    out->numPrimitives = s->numPrims;
    out->numLights = s->numLights;
    out->bvhNodes = s->numPrims > 1 ? s->numPrims - 1 : (s->numPrims ? 1 : 0);
    out->bvhDepth = s->bvhDepth;
    out->bvh4Nodes = s->numNodes4;
    out->bvh4Depth = s->bvh4Depth;
    for (int a = 0; a < 3; a++) { out->sceneLo[a] = s->sceneLo[a]; out->sceneHi[a] = s->sceneHi[a]; }
    out->deviceBytes = s->deviceBytes;
    out->bvhBuildMs = s->bvhBuildMs;
    out->bvhBinaryMs = s->bvhBinaryMs;
    return RT_OK;
}

rt_status rt_bvh_export(const rt_device_scene* s, uint64_t* codes, uint32_t* order, uint32_t* nodes14) {
    if (!s) return invalid("scene is null");
    const uint32_t n = s->numPrims;
    if (n == 0) return RT_OK;
    if (codes) RT_HIP(hipMemcpy(codes, s->codes, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    if (order) RT_HIP(hipMemcpy(order, s->order, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    if (nodes14) {
        uint32_t nn = n > 1 ? n - 1 : 1;
        std::vector<float4> h((size_t)nn * 4);
        RT_HIP(hipMemcpy(h.data(), s->nodes, sizeof(float4) * 4 * nn, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < nn; i++) std::memcpy(nodes14 + 14 * (size_t)i, &h[4 * (size_t)i], 56);
    }
    return RT_OK;
}

rt_status rt_bvh4_export(const rt_device_scene* s, uint32_t* nodes28) {
    if (!s || !nodes28) return invalid("null argument");
    if (s->numNodes4 == 0) return RT_OK;
    std::vector<float4> h((size_t)s->numNodes4 * 8);
    RT_HIP(hipMemcpy(h.data(), s->nodes4, sizeof(float4) * h.size(), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < s->numNodes4; i++) std::memcpy(nodes28 + 28 * (size_t)i, &h[8 * (size_t)i], 112);
    return RT_OK;
}

rt_status rt_trace_rays(const rt_device_scene* s, const float* rays8, uint32_t n, int mode, int skipLights, float* t,
                        float* bary2, uint32_t* inst, uint32_t* prim, hipStream_t stream) {
    if (!s) return invalid("scene is null");
    if (n && (!rays8 || !t || !bary2 || !inst || !prim)) return invalid("null array");
    if (mode != 0 && mode != 1) return invalid("mode must be 0 (closest) or 1 (any)");
    RT_LAUNCH(rti::launchTraceRays(s, rays8, n, mode, skipLights, t, bary2, inst, prim, stream));
    return RT_OK;
}

rt_status rt_gbuffer_trace(const rt_device_scene* s, const rt_camera* cam, float* depthNormal, uint32_t* albedoMatId,
                           uint16_t* motionVector, hipStream_t stream) {
    if (!s || !cam) return invalid("null argument");
    if (!depthNormal || !albedoMatId || !motionVector) return invalid("null G-buffer image");
    if (cam->filmSize[0] == 0 || cam->filmSize[1] == 0 || cam->filmSize[0] >= 32768 || cam->filmSize[1] >= 32768)
        return invalid("filmSize must be in [1, 32767]");
    RT_LAUNCH(rti::launchGBuffer(s, *cam, depthNormal, albedoMatId, motionVector, stream));
    return RT_OK;
}

rt_status rt_pt_plain(const rt_device_scene* s, const rt_camera cams[2], const rt_frame* f, uint32_t maxDepth,
                      uint32_t which, hipStream_t stream) {
    rt_status st = checkFrame(s, cams, f);
    if (st != RT_OK) return st;
    if (maxDepth == 0 || maxDepth > 255) return invalid("maxDepth must be in [1, 255]");
    if ((which & 1u) && !f->directOutput) return invalid("directOutput is null");
    if ((which & 2u) && !f->indirectOutput) return invalid("indirectOutput is null");
    if (s->numLights == 0) return invalid("the scene has no lights");
    RT_LAUNCH(rti::launchPtPlain(s, cams, *f, maxDepth, which, stream));
    return RT_OK;
}

rt_status rt_gris_path_trace_groups(const rt_device_scene* s, const rt_camera cams[2], const rt_frame* f,
                                    const rt_gris_settings* set, uint32_t workgroups, hipStream_t stream) {
    rt_status st = checkFrame(s, cams, f);
    if (st == RT_OK) st = checkSettings(set);
    if (st == RT_OK) st = checkGuide(f, set);
    if (st != RT_OK) return st;
    if (!f->grisReservoir) return invalid("grisReservoir is null");
    if (s->numLights == 0) return invalid("the scene has no lights");
    RT_LAUNCH(rti::launchGrisPathTrace(s, cams, *f, *set, stream, nullptr, workgroups));
    return RT_OK;
}

rt_status rt_gris_path_trace(const rt_device_scene* s, const rt_camera cams[2], const rt_frame* f,
                             const rt_gris_settings* set, hipStream_t stream) {
    return rt_gris_path_trace_groups(s, cams, f, set, 0u, stream);
}

rt_status rt_gris_temporal(const rt_device_scene* s, const rt_camera cams[2], const rt_frame* f,
                           const rt_gris_settings* set, hipStream_t stream) {
    rt_status st = checkFrame(s, cams, f);
    if (st == RT_OK) st = checkSettings(set);
    if (st == RT_OK) st = checkGuide(f, set);
    if (st != RT_OK) return st;
    if (!f->grisReservoir || !f->grisReservoirPrev || !f->grisReservoirTemp) return invalid("reservoir buffer is null");
    if (!f->depthNormalPrev || !f->albedoMatIdPrev || !f->motionVector) return invalid("previous G-buffer is null");
    if (s->numLights == 0) return invalid("the scene has no lights");
    RT_LAUNCH(rti::launchGrisTemporal(s, cams, *f, *set, stream));
    return RT_OK;
}

rt_status rt_gris_spatial(const rt_device_scene* s, const rt_camera cams[2], const rt_frame* f,
                          const rt_gris_settings* set, hipStream_t stream) {
    rt_status st = checkFrame(s, cams, f);
    if (st == RT_OK) st = checkSettings(set);
    if (st == RT_OK) st = checkGuide(f, set);
    if (st != RT_OK) return st;
    if (!f->grisReservoir || !f->grisReservoirTemp) return invalid("reservoir buffer is null");
    if (!f->indirectOutput) return invalid("indirectOutput is null");
    if (s->numLights == 0) return invalid("the scene has no lights");
    RT_LAUNCH(rti::launchGrisSpatial(s, cams, *f, *set, stream));
    return RT_OK;
}

rt_status rt_gris_render(const rt_device_scene* s, const rt_camera cams[2], const rt_frame* f,
                         const rt_gris_settings* set, hipStream_t stream) {
    if (set && set->targetFunction == 1u)
        return invalid("rt_gris_render runs the luminance target only; the learned target needs "
                       "rt_gris_learned_refresh between the path-trace and temporal passes");
    if (set && set->guidedProposal == 1u)
        return invalid("rt_gris_render runs the BSDF proposal only; the learned proposal needs "
                       "rt_gris_guide_refresh before the path trace and rt_gris_guide_train after the frame");
    rt_status st = rt_gris_path_trace(s, cams, f, set, stream);
    if (st == RT_OK) st = rt_gris_temporal(s, cams, f, set, stream);
    if (st == RT_OK) st = rt_gris_spatial(s, cams, f, set, stream);
    return st;
}

size_t rt_gris_reservoir_bytes(uint32_t w, uint32_t h) { return (size_t)w * h * 96; }

rt_status rt_gris_reservoir_reset(void* r, uint32_t w, uint32_t h, hipStream_t stream) {
    if (!r) return invalid("reservoir buffer is null");
    RT_LAUNCH(rti::launchResvReset(r, w, h, stream));
    return RT_OK;
}

rt_status rt_gris_reservoir_download(const void* r, uint32_t w, uint32_t h, rt_gris_reservoir* host, hipStream_t stream) {
    if (!r || !host) return invalid("null argument");
    size_t n = (size_t)w * h;
    std::vector<float4> pl(n * 5);
    RT_HIP(hipMemcpyAsync(pl.data(), r, sizeof(float4) * 5 * n, hipMemcpyDeviceToHost, stream));
    RT_HIP(hipStreamSynchronize(stream));
    for (size_t i = 0; i < n; i++) {
        rt_gris_reservoir& o = host[i];
        std::memset(&o, 0, sizeof(o));
        const float4 &a = pl[i], &b = pl[n + i], &c = pl[2 * n + i], &d = pl[3 * n + i], &e = pl[4 * n + i];
        o.rcIsec.bary[0] = a.x; o.rcIsec.bary[1] = a.y;
        std::memcpy(&o.rcIsec.instanceIdx, &a.z, 4); std::memcpy(&o.rcIsec.triangleIdx, &a.w, 4);
        o.rcLi[0] = b.x; o.rcLi[1] = b.y; o.rcLi[2] = b.z; std::memcpy(&o.rcRng, &b.w, 4);
        o.rcWi[0] = c.x; o.rcWi[1] = c.y; o.rcWi[2] = c.z; std::memcpy(&o.flags, &c.w, 4);
        o.F[0] = d.x; o.F[1] = d.y; o.F[2] = d.z; std::memcpy(&o.primaryRng, &d.w, 4);
        o.rcPrevSamplePdf = e.x; o.rcJacobian = e.y; o.sampleCount = e.z; o.resampleWeight = e.w;
    }
    return RT_OK;
}

rt_status rt_gris_reservoir_upload(const rt_gris_reservoir* host, uint32_t w, uint32_t h, void* r, hipStream_t stream) {
    if (!r || !host) return invalid("null argument");
    size_t n = (size_t)w * h;
    std::vector<float4> pl(n * 5);
    for (size_t i = 0; i < n; i++) {
        const rt_gris_reservoir& o = host[i];
        float a[4], b[4], c[4], d[4];
        a[0] = o.rcIsec.bary[0]; a[1] = o.rcIsec.bary[1];
        std::memcpy(&a[2], &o.rcIsec.instanceIdx, 4); std::memcpy(&a[3], &o.rcIsec.triangleIdx, 4);
        b[0] = o.rcLi[0]; b[1] = o.rcLi[1]; b[2] = o.rcLi[2]; std::memcpy(&b[3], &o.rcRng, 4);
        c[0] = o.rcWi[0]; c[1] = o.rcWi[1]; c[2] = o.rcWi[2]; std::memcpy(&c[3], &o.flags, 4);
        d[0] = o.F[0]; d[1] = o.F[1]; d[2] = o.F[2]; std::memcpy(&d[3], &o.primaryRng, 4);
        pl[i] = make_float4(a[0], a[1], a[2], a[3]);
        pl[n + i] = make_float4(b[0], b[1], b[2], b[3]);
        pl[2 * n + i] = make_float4(c[0], c[1], c[2], c[3]);
        pl[3 * n + i] = make_float4(d[0], d[1], d[2], d[3]);
        pl[4 * n + i] = make_float4(o.rcPrevSamplePdf, o.rcJacobian, o.sampleCount, o.resampleWeight);
    }
    RT_HIP(hipMemcpyAsync(r, pl.data(), sizeof(float4) * 5 * n, hipMemcpyHostToDevice, stream));
    RT_HIP(hipStreamSynchronize(stream));
    return RT_OK;
}

rt_status rt_gris_learned_download(const void* r, uint32_t w, uint32_t h, float* host4, hipStream_t stream) {
    if (!r || !host4) return invalid("null argument");
    const size_t n = (size_t)w * h;
    RT_HIP(hipMemcpyAsync(host4, static_cast<const float4*>(r) + 5 * n, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
    RT_HIP(hipStreamSynchronize(stream));
    return RT_OK;
}

rt_status rt_gris_learned_upload(const float* host4, uint32_t w, uint32_t h, void* r, hipStream_t stream) {
    if (!r || !host4) return invalid("null argument");
    const size_t n = (size_t)w * h;
    RT_HIP(hipMemcpyAsync(static_cast<float4*>(r) + 5 * n, host4, sizeof(float4) * n, hipMemcpyHostToDevice, stream));
    RT_HIP(hipStreamSynchronize(stream));
    return RT_OK;
}

// ---- ReSTIR DI --------------------------------------------------------------------------------
size_t rt_di_reservoir_bytes(uint32_t w, uint32_t h) { return (size_t)w * h * 64; }

rt_status rt_di_reservoir_reset(void* r, uint32_t w, uint32_t h, hipStream_t stream) {
    if (!r) return invalid("reservoir buffer is null");
    RT_LAUNCH(rti::launchDiReset(r, w, h, stream));
    return RT_OK;
}

// Device plane k holds the k-th 16 bytes of the reference struct.
rt_status rt_di_reservoir_download(const void* r, uint32_t w, uint32_t h, rt_di_reservoir* host, hipStream_t stream) {
    if (!r || !host) return invalid("null argument");
    static_assert(sizeof(rt_di_reservoir) == 64, "DIReservoir is 64 B");
    const size_t n = (size_t)w * h;
    std::vector<float4> pl(n * 4);
    RT_HIP(hipMemcpyAsync(pl.data(), r, sizeof(float4) * 4 * n, hipMemcpyDeviceToHost, stream));
    RT_HIP(hipStreamSynchronize(stream));
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) std::memcpy(reinterpret_cast<char*>(&host[i]) + 16 * k, &pl[k * n + i], 16);
    return RT_OK;
}

rt_status rt_di_reservoir_upload(const rt_di_reservoir* host, uint32_t w, uint32_t h, void* r, hipStream_t stream) {
    if (!r || !host) return invalid("null argument");
    const size_t n = (size_t)w * h;
    std::vector<float4> pl(n * 4);
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) std::memcpy(&pl[k * n + i], reinterpret_cast<const char*>(&host[i]) + 16 * k, 16);
    RT_HIP(hipMemcpyAsync(r, pl.data(), sizeof(float4) * 4 * n, hipMemcpyHostToDevice, stream));
    RT_HIP(hipStreamSynchronize(stream));
    return RT_OK;
}

rt_status rt_di_path_gen(const rt_device_scene* s, const rt_camera cams[2], const rt_frame* f,
                         const rt_di_reservoirs* r, const rt_di_settings* set, hipStream_t stream) {
    rt_status st = checkDI(s, cams, f, r, set);
    if (st != RT_OK) return st;
    if (!r->reservoir) return invalid("reservoir buffer is null");
    RT_LAUNCH(rti::launchDiPathGen(s, cams, *f, *r, *set, stream));
    return RT_OK;
}

rt_status rt_di_temporal(const rt_device_scene* s, const rt_camera cams[2], const rt_frame* f,
                         const rt_di_reservoirs* r, const rt_di_settings* set, hipStream_t stream) {
    rt_status st = checkDI(s, cams, f, r, set);
    if (st != RT_OK) return st;
    if (!r->reservoir || !r->reservoirPrev || !r->reservoirTemp) return invalid("reservoir buffer is null");
    if (!f->depthNormalPrev || !f->albedoMatIdPrev || !f->motionVector) return invalid("previous G-buffer is null");
    RT_LAUNCH(rti::launchDiTemporal(s, cams, *f, *r, *set, stream));
    return RT_OK;
}

rt_status rt_di_spatial(const rt_device_scene* s, const rt_camera cams[2], const rt_frame* f,
                        const rt_di_reservoirs* r, const rt_di_settings* set, hipStream_t stream) {
    rt_status st = checkDI(s, cams, f, r, set);
    if (st != RT_OK) return st;
    if (!r->reservoir || !r->reservoirTemp) return invalid("reservoir buffer is null");
    if (!f->directOutput) return invalid("directOutput is null");
    RT_LAUNCH(rti::launchDiSpatial(s, cams, *f, *r, *set, stream));
    return RT_OK;
}

rt_status rt_di_render(const rt_device_scene* s, const rt_camera cams[2], const rt_frame* f,
                       const rt_di_reservoirs* r, const rt_di_settings* set, hipStream_t stream) {
    rt_status st = rt_di_path_gen(s, cams, f, r, set, stream);
    if (st == RT_OK) st = rt_di_temporal(s, cams, f, r, set, stream);
    if (st == RT_OK) st = rt_di_spatial(s, cams, f, r, set, stream);
    return st;
}

// ---- ReSTIR GI ---------------------------------------------------------------------------------
size_t rt_gi_reservoir_bytes(uint32_t w, uint32_t h) { return (size_t)w * h * 48; }

rt_status rt_gi_reservoir_reset(void* r, uint32_t w, uint32_t h, hipStream_t stream) {
    if (!r) return invalid("reservoir buffer is null");
    RT_LAUNCH(rti::launchGiReset(r, w, h, stream));
    return RT_OK;
}

// Device plane k holds the k-th 16 bytes of the reference struct.
rt_status rt_gi_reservoir_download(const void* r, uint32_t w, uint32_t h, rt_gi_reservoir* host, hipStream_t stream) {
    if (!r || !host) return invalid("null argument");
    static_assert(sizeof(rt_gi_reservoir) == 48, "GIReservoir is 48 B");
    const size_t n = (size_t)w * h;
    std::vector<float4> pl(n * 3);
    RT_HIP(hipMemcpyAsync(pl.data(), r, sizeof(float4) * 3 * n, hipMemcpyDeviceToHost, stream));
    RT_HIP(hipStreamSynchronize(stream));
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) std::memcpy(reinterpret_cast<char*>(&host[i]) + 16 * k, &pl[k * n + i], 16);
    return RT_OK;
}

rt_status rt_gi_reservoir_upload(const rt_gi_reservoir* host, uint32_t w, uint32_t h, void* r, hipStream_t stream) {
    if (!r || !host) return invalid("null argument");
    const size_t n = (size_t)w * h;
    std::vector<float4> pl(n * 3);
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) std::memcpy(&pl[k * n + i], reinterpret_cast<const char*>(&host[i]) + 16 * k, 16);
    RT_HIP(hipMemcpyAsync(r, pl.data(), sizeof(float4) * 3 * n, hipMemcpyHostToDevice, stream));
    RT_HIP(hipStreamSynchronize(stream));
    return RT_OK;
}

rt_status rt_gi_temporal(const rt_device_scene* s, const rt_camera cams[2], const rt_frame* f,
                         const rt_gi_reservoirs* r, hipStream_t stream) {
    rt_status st = checkFrame(s, cams, f);
    if (st != RT_OK) return st;
    if (!r || !r->reservoir || !r->reservoirPrev) return invalid("reservoir buffer is null");
    if (!f->depthNormalPrev || !f->albedoMatIdPrev || !f->motionVector) return invalid("previous G-buffer is null");
    if (!f->indirectOutput) return invalid("indirectOutput is null");
    if (s->numLights == 0) return invalid("the scene has no lights");
    RT_LAUNCH(rti::launchGiTemporal(s, cams, *f, *r, stream));
    return RT_OK;
}

rt_status rt_gris_learned_refresh(const rt_device_scene* s, rt_mlp* mlp, const rt_camera cams[2], const rt_frame* f,
                                  hipStream_t stream) {
    rt_status st = checkFrame(s, cams, f);
    if (st != RT_OK) return st;
    if (!mlp) return invalid("mlp is null");
    if (mlp->outputs != (uint32_t)rtm::kOut) return invalid("mlp is not a radiance network (rt_mlp_create)");
    if (!f->grisReservoir) return invalid("frame reservoir is null");
    RT_LAUNCH(rti::launchLearnedRefresh(s, mlp, cams, *f, stream));
    return RT_OK;
}

size_t rt_guide_lobe_bytes(uint32_t w, uint32_t h) {
    return (size_t)((w + 1u) >> 1) * ((h + 1u) >> 1) * rtg::kPlanes * sizeof(float4);
}

rt_status rt_gris_guide_refresh(const rt_device_scene* s, const rt_mlp* guide, const rt_camera cams[2], const rt_frame* f,
                                hipStream_t stream) {
    rt_status st = checkFrame(s, cams, f);
    if (st != RT_OK) return st;
    if (!guide) return invalid("guide is null");
    if (guide->outputs != (uint32_t)rtm::kGuideOut) return invalid("guide is not a proposal network (rt_guide_create)");
    if (!f->guideLobes) return invalid("frame guideLobes is null");
    RT_LAUNCH(rti::launchGuideRefresh(guide, cams, *f, stream));
    return RT_OK;
}

rt_status rt_gris_guide_train(const rt_device_scene* s, rt_mlp* guide, const rt_camera cams[2], const rt_frame* f,
                              float* loss, hipStream_t stream) {
    rt_status st = checkFrame(s, cams, f);
    if (st != RT_OK) return st;
    if (!guide) return invalid("guide is null");
    if (guide->outputs != (uint32_t)rtm::kGuideOut) return invalid("guide is not a proposal network (rt_guide_create)");
    if (!f->guideSamples) return invalid("frame guideSamples is null");
    RT_LAUNCH(rti::launchGuideGather(guide, cams, *f, stream));
    if (guide->batchSize == 0) return RT_OK;
    RT_LAUNCH(rti::mlpGradients(guide, guide->batchPos, guide->batchTarget, guide->batchSize, loss, stream));
    RT_LAUNCH(rti::mlpAdam(guide, stream));
    return RT_OK;
}

rt_status rt_counters_enable(int enable) {
    if (enable && !rti::gCounters) {
        RT_HIP(hipMalloc(&rti::gCounters, sizeof(rtd::TraceCounters)));
        RT_HIP(hipMemset(rti::gCounters, 0, sizeof(rtd::TraceCounters)));
    }
    rti::gCountersEnabled = enable != 0;
    return RT_OK;
}

rt_status rt_counters_read(uint64_t out3[3], int reset) {
    if (!out3) return invalid("out is null");
    out3[0] = out3[1] = out3[2] = 0;
    if (!rti::gCounters) return RT_OK;
    rtd::TraceCounters c;
    RT_HIP(hipDeviceSynchronize());
    RT_HIP(hipMemcpy(&c, rti::gCounters, sizeof(c), hipMemcpyDeviceToHost));
    out3[0] = c.nodes; out3[1] = c.tris; out3[2] = c.rays;
    if (reset) RT_HIP(hipMemset(rti::gCounters, 0, sizeof(rtd::TraceCounters)));
    return RT_OK;
}

}  // extern "C"
