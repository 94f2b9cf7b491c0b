/*
 * restirpt.h — C ABI of the MI355X (gfx950) ReSTIR PT core.
 *
 * This is the boundary the reference's C++ pass surface calls into. Each entry
 * point names the reference interface it replaces (paths relative to the
 * HummaWhite/Vulkan-ReSTIR-PT source tree). All device pointers are caller
 * owned HIP device allocations; all calls are stream ordered on the given
 * hipStream_t (0 = the null stream) and return an rt_status. On failure
 * rt_last_error() returns a message for the calling thread.
 *
 * Data layouts follow the reference's std430 GLSL structs
 * (src/shader/layouts.glsl:6-164) and G-buffer formats
 * (src/GBufferPass.h:23-26):
 *   depthNormal  RGBA32F  {euclidean depth, normal.xyz}, depth 0 = background
 *   albedoMatId  RG32UI   {packUnorm4x8(albedo), matIndex << 16 | instance}
 *   motionVector RG16F    {prevUV - curUV} as IEEE half
 *   outputs      RGBA32F  (the reference's uDirectOutput / uIndirectOutput)
 * Images are row-major, pixel (x, y) at index y * width + x, row 0 at the top.
 * GRIS reservoir buffers are opaque device buffers of
 * rt_gris_reservoir_bytes(width, height) bytes (planar, 96 B per pixel).
 */
#ifndef RESTIRPT_H
#define RESTIRPT_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rt_status;
enum {
    RT_OK = 0,
    RT_ERR_INVALID_ARGUMENT = 1,
    RT_ERR_HIP = 2,
    RT_ERR_UNSUPPORTED = 3,
    RT_ERR_CAPACITY = 4
};

/* ---- scene data model: src/shader/layouts.glsl:6-94, src/Scene.h:15-38 ---- */
typedef struct rt_material {          /* layouts.glsl:6-25, 32 B */
    float baseColor[3];
    uint32_t type;                    /* 1 Lambert, 2 MetallicWorkflow, 3 Metal, 4 Dielectric, 6 Fake */
    uint32_t textureIdx;              /* 0xffffffff: none; else an index into rt_scene_arrays.textures */
    float metallic, roughness, ior;
} rt_material;

typedef struct rt_mesh_vertex {       /* layouts.glsl:53-58, 32 B */
    float pos[3]; float uvx;
    float norm[3]; float uvy;
} rt_mesh_vertex;

typedef struct rt_object_instance {   /* layouts.glsl:60-72, 224 B; matrices column-major */
    float transform[16];
    float transformInv[16];
    float transformInvT[16];
    float radiance[3]; float pad0;
    uint32_t indexOffset, indexCount, matIndex; float pad2;
} rt_object_instance;

typedef struct rt_triangle_light {    /* layouts.glsl:74-83, 64 B */
    float v0[3]; float nx;
    float v1[3]; float ny;
    float v2[3]; float nz;
    float radiance[3]; float area;
} rt_triangle_light;

typedef struct rt_light_table_entry { /* layouts.glsl:85-88; entry 0 = {sumPower, numLights} */
    float prob; uint32_t failId;
} rt_light_table_entry;

typedef struct rt_intersection {      /* layouts.glsl:90-94 */
    float bary[2]; uint32_t instanceIdx, triangleIdx;
} rt_intersection;

typedef struct rt_camera {            /* layouts.glsl:27-51, 352 B (src/Camera.h:47-69) */
    float view[16], proj[16], projView[16], lastProjView[16];
    float pos[3]; float FOV;
    float angle[3]; float nearZ;
    float front[3]; float farZ;
    float right[3]; float lensRadius;
    float up[3]; float focalDist;
    uint32_t filmSize[2]; uint32_t frameIndex; uint32_t seed;
} rt_camera;

typedef struct rt_gris_reservoir {    /* layouts.glsl:132-156, 96 B (host view only) */
    rt_intersection rcIsec;
    float rcLi[3]; uint32_t rcRng;
    float rcWi[3]; uint32_t flags;
    float pad[2]; float rcPrevSamplePdf; float rcJacobian;
    float F[3]; uint32_t primaryRng;
    float sampleCount, resampleWeight, contribWeight, pad0;
} rt_gris_reservoir;

/* An albedo texture (uTextures, layouts.glsl:178; loaded by Resource::addImage,
 * src/Resource.cpp:20-35, as 4-channel 8-bit images and bound as
 * R8G8B8A8_SRGB with a repeat sampler, src/Scene.cpp:428-445,
 * zvk/core/Memory.cpp:75-92): width x height RGBA8 texels, row 0 first.
 * Sampling converts sRGB to linear per texel, then filters: filter 0 =
 * nearest, 1 = bilinear (weights with 8 fractional bits, rounded to nearest);
 * repeat wrap; level 0 (the reference creates no mip data). */
typedef struct rt_texture {
    uint32_t width, height;
    uint32_t filter;
    uint32_t pad;
    const uint8_t* rgba8;
} rt_texture;

/* Host arrays of a scene, as the reference binds them in descriptor set 1
 * (layouts.glsl:171-178; filled by Scene::load, src/Scene.cpp:104-322). */
typedef struct rt_scene_arrays {
    const rt_material* materials;             uint32_t numMaterials;
    const int32_t* materialIndices;           uint32_t numMaterialIndices;
    const rt_mesh_vertex* vertices;           uint32_t numVertices;
    const uint32_t* indices;                  uint32_t numIndices;
    const rt_object_instance* instances;      uint32_t numInstances;
    const rt_triangle_light* lights;          uint32_t numLights;
    const rt_light_table_entry* lightTable;   uint32_t numLightTable;
    const rt_texture* textures;               uint32_t numTextures;
} rt_scene_arrays;

/* GRISReSTIR::Settings (src/GRISReSTIR.h:11-17, push constant
 * gris_reservoir.glsl:21-31) followed by the kernel constants the reference
 * hard-codes. */
typedef struct rt_gris_settings {
    uint32_t shiftType;        /* 0 Reconnection, 1 Replay, 2 Hybrid */
    float rrScale;
    uint32_t temporalReuse;
    uint32_t spatialReuse;
    uint32_t cap;
    uint32_t maxDepth;         /* MaxTracingDepth, gris_path_trace.glsl:46 (15) */
    uint32_t spatialNeighbors; /* ResampleNum, gris_resample_spatial.glsl:65 (3) */
    float spatialRadius;       /* ResampleRadius in pixels, gris_resample_spatial.glsl:66 (20) */
    /* Target function of the temporal and spatial resampling (GRISToScalar,
     * gris_reservoir.glsl:33-35, at the sites gris_retrace.glsl:204,215 and
     * gris_resample_spatial.glsl:128): 0 = luminance (the reference),
     * 1 = learned (luminance of the contribution with the reconnection
     * radiance replaced by an rt_mlp prediction; needs rt_gris_learned_refresh
     * after each path-trace pass). */
    uint32_t targetFunction;
    /* Resampling weights of the temporal and spatial merges: 0 = the reference's
     * (each candidate counts its M, final weight divided by the total M,
     * gris_resample_spatial.glsl:128; biased where a shift fails), 1 = unbiased
     * (the total M is replaced by the sum of the M of the domains whose inverse
     * shift of the picked sample is valid with a positive target -- the "1/Z"
     * contribution weights; M-cap applied once per pass; at most 8 spatial
     * neighbours), with shifts held to what each domain's path tracer
     * generates: reconnection Jacobians with the triangle's geometric normal,
     * surface reconnection vertices on either side of their shading normal,
     * the path tracer's own cosines and reconnection-vertex classification.
     * No counterpart in the reference. */
    uint32_t unbiasedReuse;
    /* Learned directional proposal at the primary vertex (no counterpart in the
     * reference): 1 = sample the primary vertex's direction from the frame's
     * 8-lobe vMF mixture with probability 1/2, else from the BSDF, with the
     * one-sample MIS density (the frame's guideLobes, guideLobesPrev and
     * guideSamples must be bound; rt_gris_guide_refresh before the path-trace
     * pass). 0 = the reference's BSDF sampling. */
    uint32_t guidedProposal;
} rt_gris_settings;

/* Descriptor set 2 (RayImageDescSet) of the reference: layouts.glsl:180-195,
 * bound per frame parity by Renderer::initDescriptor (src/Renderer.cpp:319-353). */
typedef struct rt_frame {
    uint32_t width, height;
    float* directOutput;            /* binding 0, RGBA32F */
    float* indirectOutput;          /* binding 1, RGBA32F */
    const float* depthNormal;       /* binding 2 */
    const float* depthNormalPrev;   /* binding 3 */
    const uint32_t* albedoMatId;    /* binding 4 */
    const uint32_t* albedoMatIdPrev;/* binding 5 */
    const uint16_t* motionVector;   /* binding 6 */
    void* grisReservoir;            /* binding 12 (this frame) */
    void* grisReservoirPrev;        /* binding 13 */
    void* grisReservoirTemp;        /* binding 14 */
    /* Learned vMF-mixture proposal (no reference counterpart; read only when
     * rt_gris_settings.guidedProposal = 1): the per-pixel lobe planes of this
     * frame (written by rt_gris_guide_refresh) and of the previous frame,
     * rt_guide_lobe_bytes(width, height) each, and the per-pixel training
     * record the path-trace pass writes (16 B/px). Null otherwise. */
    void* guideLobes;
    void* guideLobesPrev;
    void* guideSamples;
} rt_frame;

typedef struct rt_scene_info {
    uint32_t numPrimitives;    /* light triangles + instanced object triangles */
    uint32_t numLights;
    uint32_t bvhNodes;         /* internal nodes */
    uint32_t bvhDepth;         /* max root-to-leaf internal-node count */
    float sceneLo[3], sceneHi[3];
    size_t deviceBytes;
    uint32_t bvh4Nodes;        /* nodes of the collapsed 4-wide tree that is traversed */
    uint32_t bvh4Depth;
    float bvhBuildMs;          /* GPU time of the LBVH build: world triangles to the 4-wide tree */
    float bvhBinaryMs;         /* of which the binary LBVH (bounds, Morton codes, sort, topology, refit) */
} rt_scene_info;

typedef struct rt_device_scene rt_device_scene;

const char* rt_last_error(void);

/* Replaces DeviceScene::DeviceScene (src/Scene.cpp:324-330): uploads the
 * buffers of createBufferAndImages (:368-446) and builds the acceleration
 * structure of createAccelerationStructure (:448-547) as a GPU LBVH over all
 * light and instanced object triangles. Synchronises the stream. */
rt_status rt_device_scene_create(const rt_scene_arrays* arrays, hipStream_t stream, rt_device_scene** out);
/* Replaces DeviceScene::destroy (src/Scene.cpp:332-349). */
rt_status rt_device_scene_destroy(rt_device_scene* scene);
rt_status rt_device_scene_info(const rt_device_scene* scene, rt_scene_info* out);
/* Copies the LBVH to host memory (for parity checks): sorted Morton codes
 * [N], leaf->primitive order [N], internal nodes [N-1] as 14 words each
 * (left box min/max, right box min/max as float bits, left, right child;
 * bit 31 marks a leaf). Any pointer may be null. */
rt_status rt_bvh_export(const rt_device_scene* scene, uint64_t* codes, uint32_t* order, uint32_t* nodes14);
/* Copies the collapsed 4-wide tree that the ray queries traverse: 28 words per
 * node (child boxes as lo[axis][slot], hi[axis][slot] float bits, then 4 child
 * references; 0xffffffff = empty slot, bit 31 = leaf). rt_scene_info.bvh4Nodes
 * gives the count. */
rt_status rt_bvh4_export(const rt_device_scene* scene, uint32_t* nodes28);

/* Allocation-free variant of the binary LBVH build that rt_device_scene_create
 * runs (and rt_bvh_export reads back). No counterpart in the reference: it is
 * for callers that cannot allocate -- code under HIP graph capture, serving
 * loops that rebuild dynamic geometry each frame. primTris: n world-space
 * triangles on the device, 3 float4 each ({v0, instance custom index as float
 * bits}, {v1, primitive index as float bits}, {v2, 0}). Writes the sorted
 * Morton codes (n), the leaf -> primitive order (n) and the internal nodes
 * (max(n - 1, 1) x 16 floats: the 14 words of rt_bvh_export, then 2 pad words)
 * to caller-owned device buffers, bit-identical to the scene build of the
 * same triangles. Stream ordered, no host synchronisation, no hipMalloc:
 * its scratch is a static 8 MiB device array per GPU, so n is limited to
 * [1, 131072] (RT_ERR_CAPACITY beyond) and calls that overlap on one GPU race.
 * The device of the stream must be the current device. */
rt_status rt_bvh_build_noalloc(const float* primTris, uint32_t n, uint64_t* codes, uint32_t* order, float* nodes,
                               hipStream_t stream);

/* Ray queries of src/shader/ray_query.glsl:6-70 (traceClosestHit /
 * traceShadow) over device arrays: rays8 = {ox, oy, oz, dx, dy, dz, tmin, tmax}.
 * mode 0 = closest hit (t, bary, instance custom index, primitive index;
 * instance 0xffffffff on a miss), mode 1 = any hit (inst = 1 if occluded).
 * skipLights excludes the light triangles (instance 0). */
rt_status rt_trace_rays(const rt_device_scene* scene, const float* rays8, uint32_t n, int mode, int skipLights,
                        float* t, float* bary2, uint32_t* inst, uint32_t* prim, hipStream_t stream);

/* Replaces GBufferPass::render (src/GBufferPass.cpp:22-56) with a primary-ray
 * kernel at pixel centres producing the same three images for the frame
 * parity being rendered. cam->filmSize gives the image size. */
rt_status rt_gbuffer_trace(const rt_device_scene* scene, const rt_camera* cam, float* depthNormal,
                           uint32_t* albedoMatId, uint16_t* motionVector, hipStream_t stream);

/* Plain path tracer with NEE+MIS: RayTracing::execute of di_naive.comp and
 * gi_naive.comp (src/Renderer.cpp:453-469; src/shader/di_naive.glsl:79-167,
 * gi_naive.glsl:28-149). which: bit 0 direct, bit 1 indirect. cams[0] is
 * uCamera, cams[1] uPrevCamera. */
rt_status rt_pt_plain(const rt_device_scene* scene, const rt_camera cams[2], const rt_frame* frame,
                      uint32_t maxDepth, uint32_t which, hipStream_t stream);

/* The three passes of GRISReSTIR::render (src/GRISReSTIR.cpp:9-53).
 * rt_gris_path_trace (also inside rt_gris_render) is a persistent kernel that
 * takes its pixel queue from 64 slots per scene, handed out round robin; a
 * launch whose slot's previous launch may still run (on another stream) first
 * waits for it on its own stream, so concurrent launches never share a queue.
 * Under stream capture no such wait is recorded: at most 64 captured launches
 * on one scene may then be in flight at once. */
rt_status rt_gris_path_trace(const rt_device_scene* scene, const rt_camera cams[2], const rt_frame* frame,
                             const rt_gris_settings* settings, hipStream_t stream);
/* rt_gris_path_trace with at most `workgroups` one-wave workgroups in the
 * persistent launch (0: as many as the device holds at once, which is what
 * rt_gris_path_trace launches). No counterpart in the reference. The result
 * does not depend on the count; a smaller launch leaves CUs to kernels running
 * beside it on other streams (rt_pipeline's frame overlap). */
rt_status rt_gris_path_trace_groups(const rt_device_scene* scene, const rt_camera cams[2], const rt_frame* frame,
                                    const rt_gris_settings* settings, uint32_t workgroups, hipStream_t stream);
rt_status rt_gris_temporal(const rt_device_scene* scene, const rt_camera cams[2], const rt_frame* frame,
                           const rt_gris_settings* settings, hipStream_t stream);
rt_status rt_gris_spatial(const rt_device_scene* scene, const rt_camera cams[2], const rt_frame* frame,
                          const rt_gris_settings* settings, hipStream_t stream);
/* GRISReSTIR::render: path trace, temporal reuse, spatial reuse, in order.
 * Luminance target and BSDF proposal only: targetFunction = 1 and
 * guidedProposal = 1 are rejected, since their refresh and training steps run
 * between the passes (rt_gris_learned_refresh, rt_gris_guide_refresh/train,
 * or the frame graph, which records them). */
rt_status rt_gris_render(const rt_device_scene* scene, const rt_camera cams[2], const rt_frame* frame,
                         const rt_gris_settings* settings, hipStream_t stream);

/* Reservoir buffers (the GRISReservoir storage buffers allocated by
 * Renderer::createRayImage, src/Renderer.cpp:236-250): 96 B per pixel, the
 * reference's reservoir size. Five 16-B planes hold the reference fields; the
 * sixth holds the learned target data (predicted reconnection radiance and
 * the target value the resample weight refers to) and is only used when
 * targetFunction = 1. */
size_t rt_gris_reservoir_bytes(uint32_t width, uint32_t height);
/* Sets every reservoir to the reset state (GRISReservoirReset,
 * gris_reservoir.glsl:75-79; other fields zero). */
rt_status rt_gris_reservoir_reset(void* reservoirs, uint32_t width, uint32_t height, hipStream_t stream);
/* Converts between the device layout and the reference's 96-B struct on the host. */
rt_status rt_gris_reservoir_download(const void* reservoirs, uint32_t width, uint32_t height,
                                     rt_gris_reservoir* host, hipStream_t stream);
rt_status rt_gris_reservoir_upload(const rt_gris_reservoir* host, uint32_t width, uint32_t height,
                                   void* reservoirs, hipStream_t stream);
/* Host copy of the learned-target plane: 4 floats per pixel (predicted
 * reconnection radiance RGB, target value of the resample weight). */
rt_status rt_gris_learned_download(const void* reservoirs, uint32_t width, uint32_t height, float* host4,
                                   hipStream_t stream);
rt_status rt_gris_learned_upload(const float* host4, uint32_t width, uint32_t height, void* reservoirs,
                                 hipStream_t stream);

/* ---- ReSTIR DI ------------------------------------------------------------
 * The direct-lighting method "ReSTIR DI" of the reference (TestReSTIR,
 * src/TestReSTIR.h:7-38, selected by Renderer's directMethod = ResampledDI,
 * src/Renderer.cpp:457-459): path generation (one light sample and one BSDF
 * sample per pixel streamed into a reservoir), temporal reuse and spatial
 * reuse (10 neighbours in 20 px, M-cap 40) with the reconnection or random
 * replay shift, the spatial pass shading the pixel into directOutput. */
typedef struct rt_di_settings {   /* TestReSTIR::Settings, src/TestReSTIR.h:12-17 (push constant) */
    uint32_t shiftType;           /* 0 Reconnection, 1 Replay, 2 Hybrid (merges nothing in the reference) */
    uint32_t sampleType;          /* 0 Light, 1 BSDF, 2 Both */
    uint32_t temporalReuse;
    uint32_t spatialReuse;
} rt_di_settings;

typedef struct rt_di_reservoir {  /* DIReservoir, layouts.glsl:97-115, 64 B (host view only) */
    rt_intersection isec;
    float Li[3]; float pad0;
    float jacobian; float samplePdf; uint32_t rng; uint32_t isLightSample;
    uint32_t sampleCount; float resampleWeight; float contribWeight; float weight;
} rt_di_reservoir;

/* Descriptor set 2 bindings 7/8/9 (layouts.glsl:187-189): this frame's,
 * the previous frame's and the temporal reservoirs, each
 * rt_di_reservoir_bytes(width, height) of device memory; swapped per frame
 * parity like the reference's mDIReservoir[.][2] (src/Renderer.cpp:339-341). */
typedef struct rt_di_reservoirs {
    void* reservoir;
    void* reservoirPrev;
    void* reservoirTemp;
} rt_di_reservoirs;

/* 64 B per pixel, the reference's DIReservoirData (src/Renderer.cpp:33). */
size_t rt_di_reservoir_bytes(uint32_t width, uint32_t height);
/* Sets every reservoir to the reset state with an invalid sample. */
rt_status rt_di_reservoir_reset(void* reservoirs, uint32_t width, uint32_t height, hipStream_t stream);
/* Converts between the device layout and the reference's 64-B struct on the host. */
rt_status rt_di_reservoir_download(const void* reservoirs, uint32_t width, uint32_t height, rt_di_reservoir* host,
                                   hipStream_t stream);
rt_status rt_di_reservoir_upload(const rt_di_reservoir* host, uint32_t width, uint32_t height, void* reservoirs,
                                 hipStream_t stream);
/* The three passes of TestReSTIR::render (src/TestReSTIR.cpp:9-36): di_path_gen,
 * di_temporal and di_spatial (.comp/.glsl). Path generation writes
 * reservoirs->reservoir; temporal reads it and reservoirPrev and writes
 * reservoirTemp; spatial reads reservoirTemp, writes reservoir and accumulates
 * frame->directOutput. The frame's GRIS reservoir bindings are not used. */
rt_status rt_di_path_gen(const rt_device_scene* scene, const rt_camera cams[2], const rt_frame* frame,
                         const rt_di_reservoirs* reservoirs, const rt_di_settings* settings, hipStream_t stream);
rt_status rt_di_temporal(const rt_device_scene* scene, const rt_camera cams[2], const rt_frame* frame,
                         const rt_di_reservoirs* reservoirs, const rt_di_settings* settings, hipStream_t stream);
rt_status rt_di_spatial(const rt_device_scene* scene, const rt_camera cams[2], const rt_frame* frame,
                        const rt_di_reservoirs* reservoirs, const rt_di_settings* settings, hipStream_t stream);
/* TestReSTIR::render: path generation, temporal reuse, spatial reuse, in order. */
rt_status rt_di_render(const rt_device_scene* scene, const rt_camera cams[2], const rt_frame* frame,
                       const rt_di_reservoirs* reservoirs, const rt_di_settings* settings, hipStream_t stream);

/* ---- ReSTIR GI (temporal) -------------------------------------------------
 * The indirect method "ReSTIR GI" of the reference (gi_resample_temporal.comp /
 * .glsl:34-213, run by Renderer::recordRenderCommand for indirectMethod =
 * ResampledGI, src/Renderer.cpp:470-471): one pass that traces a path from the
 * primary vertex (reconnection vertex = the first secondary hit, its outgoing
 * radiance gathered with NEE + MIS, at most 15 bounces), merges it into the
 * reprojected previous reservoir (M-cap 40), shades the reservoir's sample
 * through a visibility ray and accumulates frame->indirectOutput. */
typedef struct rt_gi_reservoir {  /* GIReservoir, layouts.glsl:117-129, 48 B (host view only) */
    rt_intersection rcIsec;
    float rcLo[3]; uint32_t rcPrevCoord;
    uint32_t sampleCount; float resampleWeight; float contribWeight; float pad0;
} rt_gi_reservoir;
/* Descriptor set 2 bindings 10/11 (layouts.glsl:190-191), swapped per frame
 * parity like mGIReservoir[.][2] (src/Renderer.cpp:342-343). */
typedef struct rt_gi_reservoirs {
    void* reservoir;
    void* reservoirPrev;
} rt_gi_reservoirs;
/* 48 B per pixel, the reference's GIReservoirData (src/Renderer.cpp:34). */
size_t rt_gi_reservoir_bytes(uint32_t width, uint32_t height);
rt_status rt_gi_reservoir_reset(void* reservoirs, uint32_t width, uint32_t height, hipStream_t stream);
rt_status rt_gi_reservoir_download(const void* reservoirs, uint32_t width, uint32_t height, rt_gi_reservoir* host,
                                   hipStream_t stream);
rt_status rt_gi_reservoir_upload(const rt_gi_reservoir* host, uint32_t width, uint32_t height, void* reservoirs,
                                 hipStream_t stream);
rt_status rt_gi_temporal(const rt_device_scene* scene, const rt_camera cams[2], const rt_frame* frame,
                         const rt_gi_reservoirs* reservoirs, hipStream_t stream);

/* ---- Post-processing -------------------------------------------------------
 * PostProcessFrag (src/PostProcessFrag.h:5-45, post_proc.frag:16-43): the
 * direct and indirect outputs summed (each can be switched off), tone mapped
 * (0 none, 1 Filmic, 2 ACES), gamma corrected (1/2.2) and stored as the
 * 8-bit image the reference presents; bytes R, G, B, A per pixel, row 0 at the
 * top (the screenshot layout of Renderer::writeScreenshot, src/Renderer.cpp:758-793).
 * rgba8 is a device buffer of width * height * 4 bytes. */
typedef struct rt_post_settings {  /* PostProcessFrag::PushConstant; Renderer defaults {1, 1, 0, 0} */
    uint32_t toneMapping;
    uint32_t correctGamma;
    uint32_t noDirect;
    uint32_t noIndirect;
} rt_post_settings;
rt_status rt_post_process(const rt_frame* frame, const rt_post_settings* settings, uint8_t* rgba8, hipStream_t stream);

/* ---- Learned target function (no counterpart in the reference) ----------
 * A multiresolution hash-grid encoding of world position (levels x 2 FP16
 * features, 2^log2TableSize entries per level) feeding an MLP with two hidden
 * layers of 64 (FP16 weights on the matrix cores, FP32 accumulation, ReLU,
 * exp output) that predicts RGB radiance. Trained online with Adam on a
 * relative L2 loss mean((y^ - y)^2 / (sg(y^)^2 + 0.01)). */
typedef struct rt_mlp rt_mlp;
typedef struct rt_mlp_config {
    uint32_t levels;            /* 16 (only value supported) */
    uint32_t featuresPerLevel;  /* 2 (only value supported) */
    uint32_t log2TableSize;     /* 19; 10..24 */
    uint32_t baseResolution;    /* 16 */
    uint32_t maxResolution;     /* 2048 */
    uint32_t hiddenWidth;       /* 64 (only value supported) */
    uint32_t hiddenLayers;      /* 2 (only value supported) */
    uint32_t seed;              /* parameter initialisation */
    float learningRate;         /* 1e-2 */
    float beta1, beta2, epsilon;/* 0.9, 0.99, 1e-15 */
    float boundsLo[3], boundsHi[3];  /* world box mapped to the unit cube */
} rt_mlp_config;

rt_status rt_mlp_create(const rt_mlp_config* config, hipStream_t stream, rt_mlp** out);
rt_status rt_mlp_destroy(rt_mlp* mlp);
/* Parameter counts: grid = levels * 2^log2TableSize * 2, weights = 32*64 + 64*64 + 64*3. */
rt_status rt_mlp_param_counts(const rt_mlp* mlp, uint64_t* gridParams, uint64_t* weightParams);
/* FP32 master parameters on the host: grid [level][entry][feature], weights
 * W1 [32][64], W2 [64][64], W3 [64][3] (input-major). Setting also refreshes
 * the FP16 copies the kernels read and clears the Adam state. */
rt_status rt_mlp_get_params(const rt_mlp* mlp, float* grid, float* weights);
rt_status rt_mlp_set_params(rt_mlp* mlp, const float* grid, const float* weights);
/* Inference on device arrays: pos [n][3] world positions -> rgb [n][3]. */
rt_status rt_mlp_infer(const rt_mlp* mlp, const float* pos, uint32_t n, float* rgb, hipStream_t stream);
/* Gradients of the loss over a batch (device arrays; a sample whose target has
 * a negative first component is masked out). loss (device, may be null)
 * receives the mean loss. The gradients stay on the device for the next
 * rt_mlp_adam_step; rt_mlp_read_gradients copies them to the host. */
rt_status rt_mlp_gradients(rt_mlp* mlp, const float* pos, const float* target, uint32_t n, float* loss,
                           hipStream_t stream);
rt_status rt_mlp_read_gradients(const rt_mlp* mlp, float* gridGrad, float* weightGrad);
rt_status rt_mlp_adam_step(rt_mlp* mlp, hipStream_t stream);
/* rt_mlp_gradients + rt_mlp_adam_step. */
rt_status rt_mlp_train_step(rt_mlp* mlp, const float* pos, const float* target, uint32_t n, float* loss,
                            hipStream_t stream);
/* One training step on the samples gathered by the last rt_gris_learned_refresh. */
rt_status rt_mlp_train_frame(rt_mlp* mlp, float* loss, hipStream_t stream);
/* Host copy of that batch: *count samples of position [3] and target [3]
 * (arrays of at least 65536 x 3 floats; either may be null). */
rt_status rt_mlp_batch_download(const rt_mlp* mlp, float* pos, float* target, uint32_t* count);

/* After rt_gris_path_trace: predicts the reconnection radiance of every
 * pixel's reservoir with the network, rescales the resample weight to the
 * learned target and fills the learned-target plane; also gathers the frame's
 * training batch (reconnection vertex position, reconnection radiance) from a
 * fixed stride of pixels. */
rt_status rt_gris_learned_refresh(const rt_device_scene* scene, rt_mlp* mlp, const rt_camera cams[2],
                                  const rt_frame* frame, hipStream_t stream);

/* ---- Learned directional proposal -----------------------------------------
 * No counterpart in the reference. A second network of the same encoding and
 * hidden layers whose 32 outputs are an 8-lobe von Mises-Fisher mixture per
 * position (lobe k: weight logit, log concentration squashed into (-4, 5), polar
 * and azimuth angles through sigmoids; fixed offsets put the untrained lobes at
 * the eight octant directions), trained online on the direction the
 * path tracer sampled at each primary vertex weighted by the path's luminance
 * (negative log-likelihood weighted by the sample weight over the batch mean),
 * and sampled at the primary vertex by settings.guidedProposal = 1. All
 * rt_mlp_* functions apply: rt_mlp_infer returns the 32 raw outputs per
 * sample, training targets are 4 floats (unit direction, weight; weight <= 0
 * masks the sample), weights are W1 [32][64], W2 [64][64], W3 [64][32]. */
rt_status rt_guide_create(const rt_mlp_config* config, hipStream_t stream, rt_mlp** out);
/* Bytes of one lobe-plane buffer: one mixture per 2x2 pixel block, 16 float4
 * planes of ceil(width/2) * ceil(height/2) blocks (256 B per block). */
size_t rt_guide_lobe_bytes(uint32_t width, uint32_t height);
/* After the G-buffer pass: evaluates the proposal network once per 2x2 pixel
 * block, at the primary vertex of the block's first non-background pixel in
 * raster order, and writes the mixture into frame->guideLobes at block index
 * (y/2) * ceil(width/2) + x/2 (lobe k: plane 2k = (mu, kappa), plane 2k+1 =
 * (pi kappa / (2 pi (1 - e^{-2 kappa})), e^{-2 kappa}, cumulative pi, pi));
 * every pixel of the block samples that mixture; all-background blocks are
 * not written. */
rt_status rt_gris_guide_refresh(const rt_device_scene* scene, const rt_mlp* guide, const rt_camera cams[2],
                                const rt_frame* frame, hipStream_t stream);
/* After the frame's passes: gathers the training batch (primary position,
 * sampled direction, path luminance) of about 2^14 pixels from
 * frame->guideSamples at a fixed stride (offset by the camera seed) and takes one training step;
 * loss (device, may be null) receives the batch's mean loss before the step. */
rt_status rt_gris_guide_train(const rt_device_scene* scene, rt_mlp* guide, const rt_camera cams[2],
                              const rt_frame* frame, float* loss, hipStream_t stream);

/* ---- Frame graphs -------------------------------------------------------
 * The per-frame command recording of Renderer::recordRenderCommand
 * (src/Renderer.cpp:400-503, one command buffer per frame) as HIP graphs: one
 * GRIS frame of a view (G-buffer, [proposal refresh], path trace, [learned
 * refresh], temporal, spatial, [training steps]) per frame parity, built on the first launch
 * of that parity; later launches update the kernel arguments (cameras, Adam
 * step) and replay. frames[j] are the bindings of parity j (same size); the
 * graph owns its pixel queue. mlp (the radiance network) is required when
 * settings->targetFunction = 1; guide (the proposal network) when
 * settings->guidedProposal = 1, whose refresh after the G-buffer and training
 * step (gather + step, as rt_gris_guide_train) the graph records too. */
typedef struct rt_frame_graph rt_frame_graph;
rt_status rt_frame_graph_create(const rt_device_scene* scene, const rt_frame frames[2],
                                const rt_gris_settings* settings, rt_mlp* mlp, rt_mlp* guide,
                                rt_frame_graph** out);
rt_status rt_frame_graph_launch(rt_frame_graph* graph, uint32_t parity, const rt_camera cams[2], hipStream_t stream);
rt_status rt_frame_graph_destroy(rt_frame_graph* graph);

/* ---- Batched views --------------------------------------------------------
 * Many independent views (each its own scene, cameras, frame bindings and
 * settings) rendered together, the batched form of Renderer::drawFrame
 * (src/Renderer.cpp:539-569) for many views: one persistent GRIS path-trace
 * launch whose waves drain all views' pixel queues in turn, and per form of
 * the views (target function, resampling weights, learned proposal) one
 * G-buffer, one temporal and one spatial launch over those views' 8x8 tiles.
 * A view with the learned target or proposal brings its own network (one
 * view per network); the batch runs that network's steps where a single
 * view's frame runs them (proposal refresh before the path trace, target
 * refresh after it, one training step each after spatial reuse), so every
 * view renders exactly as it would alone. The caller flips each view's frame
 * parity between frames; calls of one batch are ordered on one stream. */
typedef struct rt_view {
    const rt_device_scene* scene;
    rt_camera cams[2];            /* uCamera, uPrevCamera */
    rt_frame frame;               /* the bindings of the parity being rendered */
    rt_gris_settings settings;
    rt_mlp* target;               /* radiance network when settings.targetFunction = 1, else ignored */
    rt_mlp* guide;                /* proposal network when settings.guidedProposal = 1, else ignored */
} rt_view;

typedef struct rt_batch_tuning {
    uint32_t persistentGroups;    /* workgroups of the persistent path trace (0: CUs x resident) */
    uint32_t xcdSwizzle;          /* per-pixel passes: each XCD takes runs of this many consecutive
                                     8x8 tiles, every 8th run (0: dispatch order; at most 4096) */
    uint32_t waves;               /* register budget of the persistent path trace, in waves per
                                     SIMD: 3, 4 or 5 (0: 4) */
    uint32_t queueTile;           /* pixels of one path-trace queue item (64 lanes refilled together):
                                     0 = an 8x8 tile, 1 = 16x4, 2 = 64x1 */
    uint32_t graph;               /* 1: the frame's launches replay as one HIP graph whose kernel
                                     nodes are updated per frame; 0: launched one by one */
    uint32_t shading;             /* 0: the persistent path trace (one lane per path, lanes refilled
                                     from the pixel queues); 1: the wavefront variant -- path state
                                     in HBM, per bounce a closest-hit launch, a counting sort of
                                     the live paths by the hit's material type and a shading
                                     launch in that order (views without the learned proposal;
                                     allocates 256 B per pixel of the batch on first use) */
} rt_batch_tuning;

typedef struct rt_batch rt_batch;
rt_status rt_batch_create(uint32_t maxViews, rt_batch** out);
rt_status rt_batch_destroy(rt_batch* batch);
rt_status rt_batch_frame(rt_batch* batch, const rt_view* views, uint32_t numViews, hipStream_t stream);
rt_status rt_batch_set_tuning(rt_batch* batch, const rt_batch_tuning* tuning);
rt_status rt_batch_get_tuning(const rt_batch* batch, rt_batch_tuning* tuning);
/* Autotuner: coordinate descent over the tuning from {persistentGroups 0,
 * xcdSwizzle 16, waves 4, queueTile 0, graph 0, shading 0}: the register
 * budget (3, 4, 5), then the queue tile shape, the persistent workgroups (1/2,
 * 3/4 and 1 x the resident limit of the chosen budget), the XCD chunk (0, 16,
 * 128 tiles), graph replay and, when no view uses the learned proposal, the
 * wavefront variant (with 2, 4 and 8 workgroups per CU), each searched with
 * the others at their best so far. Each
 * candidate is timed with HIP events over `frames` frames of these views after
 * one warm-up frame; the frames alternate the views' parity (this frame's
 * G-buffer, reservoirs and proposal lobes swapped with the previous frame's),
 * as a caller's frames do, so temporal reuse works on the previous frame's
 * output (the cameras are used as given: a camera flagged as a sequence's
 * first frame switches temporal reuse off in every tuning frame, so tune on
 * a later frame). Keeps the fastest and returns it with its mean frame time. The
 * views' reservoirs are left in an arbitrary state (and their networks trained
 * on those frames): reset them before rendering for real. */
rt_status rt_batch_autotune(rt_batch* batch, const rt_view* views, uint32_t numViews, uint32_t frames,
                            hipStream_t stream, rt_batch_tuning* best, float* msPerFrame);

/* Traversal work counters (internal nodes visited, triangles tested, rays)
 * accumulated by every kernel launched while enabled; off by default. The
 * counting build of each kernel is only used while enabled. */
rt_status rt_counters_enable(int enable);
rt_status rt_counters_read(uint64_t out3[3], int reset);

#ifdef __cplusplus
}
#endif
#endif /* RESTIRPT_H */
