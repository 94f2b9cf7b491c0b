/*
 * restirpt_host.h — C entry points of the C++ host layer (librestirpt_host.so).
 *
 * The host layer is the reference's C++ surface rebuilt over the C ABI of
 * restirpt.h: Scene + light table (src/Scene.cpp:192-322,
 * src/util/AliasTable.h), Camera (src/Camera.cpp), GBufferPass, RayTracing,
 * GRISReSTIR and a headless frame driver standing in for Renderer
 * (src/Renderer.cpp:193-257, 319-368, 400-503, 539-577). These functions expose
 * it to callers that are not C++ (tests, benchmark drivers).
 */
#ifndef RESTIRPT_HOST_H
#define RESTIRPT_HOST_H

#include "restirpt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rt_host_scene rt_host_scene;
typedef struct rt_renderer rt_renderer;
typedef struct rt_pipeline rt_pipeline;

/* Procedural scenes (the workloads this build is measured on):
 *   0  closed box room of 12 quads + 64 icospheres (subdivision 2), Lambert,
 *      one ceiling quad light of radiance 10 (~20.5k triangles)
 *   1  512x512-cell fBm heightfield (6 octaves, 4 triangles per cell) + 200
 *      instanced icospheres, MetallicWorkflow GGX, 16 emissive triangles of
 *      log-uniform radiance in [1, 100] (~1.1M triangles)
 *   2  a small test scene: one floor quad under one quad light
 *   3  as 1 with a 200x200-cell heightfield (~225k triangles)
 *   4  as 1 with an 860x860-cell heightfield (~3.0M triangles)
 *   5  1024x1024-cell heightfield + 2000 instanced icospheres, 64 emissive
 *      triangles, GGX roughness in [0.02, 1] (~4.8M triangles)
 * Everything is drawn from PCG32 streams seeded by `seed`. */
rt_status rt_scenegen(uint32_t config, uint32_t seed, rt_host_scene** out);
/* Scene::load (src/Scene.cpp:104-303): a scene file in the reference's XML
 * format (integrator film size, camera, model instances of Wavefront OBJ
 * meshes with transforms, material overrides and emitters of given power).
 * Textures are not supported. The camera is the file's, still, with a seeded
 * path of 1024 frames for rt_host_scene_camera. Errors (unreadable or
 * malformed files, textures, zero-power lights) return
 * RT_ERR_INVALID_ARGUMENT with the message in rt_last_error(). */
rt_status rt_scene_load(const char* path, rt_host_scene** out);
/* The film size of a loaded scene file (<integrator><size>); 0 x 0 for procedural scenes. */
rt_status rt_host_scene_film_size(const rt_host_scene* scene, uint32_t* width, uint32_t* height);
rt_status rt_host_scene_destroy(rt_host_scene* scene);
/* Borrowed views of the scene arrays (valid until the scene is destroyed). */
rt_status rt_host_scene_arrays(const rt_host_scene* scene, rt_scene_arrays* out);
/* Camera of frame `frame` of the scene's seeded camera path at width x height:
 * out[0] = uCamera (frame's seed, lastProjView of the previous frame,
 * frameIndex 0, clear flag on frame 0), out[1] = uPrevCamera. */
rt_status rt_host_scene_camera(const rt_host_scene* scene, uint32_t frame, uint32_t width, uint32_t height,
                               rt_camera out[2]);

/* DiscreteSampler1D<float>::build (src/util/AliasTable.h:26-71): n + 1 entries. */
rt_status rt_light_table_build(const float* distribution, uint32_t n, rt_light_table_entry* out);

/* Headless frame driver: owns the frame images (two G-buffer parities,
 * motion vectors, three reservoir buffers, direct/indirect outputs) and runs
 * GBufferPass + GRISReSTIR (method 0) or GBufferPass + naive DI + naive GI
 * (method 1) per frame, flipping the frame parity afterwards
 * (Renderer::drawFrame, src/Renderer.cpp:539-569). */
rt_status rt_renderer_create(rt_device_scene* scene, uint32_t width, uint32_t height, rt_renderer** out);
rt_status rt_renderer_destroy(rt_renderer* r);
rt_status rt_renderer_set_settings(rt_renderer* r, const rt_gris_settings* settings);
rt_status rt_renderer_frame(rt_renderer* r, const rt_camera cams[2], uint32_t method, hipStream_t stream);
/* One frame with the reference's separate method choices (Renderer::Settings
 * directMethod / indirectMethod, src/Renderer.h:25-31, dispatched as in
 * Renderer::recordRenderCommand, src/Renderer.cpp:440-470): G-buffer, then the
 * direct method (0 None, 1 Naive, 2 ReSTIR DI), then the indirect method
 * (0 None, 1 Naive, 2 ReSTIR GI, 3 ReSTIR PT); flips the frame parity. */
rt_status rt_renderer_frame_methods(rt_renderer* r, const rt_camera cams[2], uint32_t directMethod,
                                    uint32_t indirectMethod, hipStream_t stream);
/* TestReSTIR::settings of the renderer's ReSTIR DI pass. */
rt_status rt_renderer_set_di_settings(rt_renderer* r, const rt_di_settings* settings);
/* Current frame's DI reservoir bindings (descriptor bindings 7/8/9). */
rt_status rt_renderer_di_bindings(const rt_renderer* r, rt_di_reservoirs* out);
/* Current frame's bindings (device pointers) for callers that launch passes themselves. */
rt_status rt_renderer_bindings(const rt_renderer* r, rt_frame* out);
/* Flips the frame parity (mCurFrame ^= 1). rt_renderer_frame does this itself. */
rt_status rt_renderer_flip(rt_renderer* r);
/* Copies an RGBA32F output (0 direct, 1 indirect) to host memory. */
rt_status rt_renderer_read_output(const rt_renderer* r, uint32_t which, float* host, hipStream_t stream);
/* Resets the GRIS and DI reservoir buffers (the GUI's clearReservoir, src/Renderer.cpp:567). */
rt_status rt_renderer_reset_reservoirs(rt_renderer* r, hipStream_t stream);
/* Creates the renderer's learned target function (hash grid 16 x 2 x
 * 2^log2TableSize, MLP 2 x 64, bounds = scene box + 10%) and binds it to the
 * GRIS passes; frames with settings.targetFunction = 1 then run the learned
 * refresh after path tracing and one training step after spatial reuse.
 * *mlp (may be null) receives the network, owned by the renderer. */
rt_status rt_renderer_enable_learned(rt_renderer* r, uint32_t seed, uint32_t log2TableSize, hipStream_t stream,
                                     rt_mlp** mlp);
/* Creates the renderer's learned vMF-mixture proposal (rt_guide_create, same
 * network shape and bounds as above), allocates its frame buffers (two lobe
 * planes of 256 B per 2x2 pixel block, 16 B/px of training records) and binds it to the GRIS
 * passes; frames with settings.guidedProposal = 1 then run the proposal's
 * refresh before the path trace and one training step after spatial reuse.
 * No counterpart in the reference. *mlp (may be null) receives the network,
 * owned by the renderer. */
rt_status rt_renderer_enable_proposal(rt_renderer* r, uint32_t seed, uint32_t log2TableSize, hipStream_t stream,
                                      rt_mlp** mlp);

/* GRIS frames overlapped on two streams: no counterpart in the reference,
 * whose Renderer::drawFrame (src/Renderer.cpp:539-569) records each frame
 * after the last. Frame k + 1's G-buffer and path trace need only its camera,
 * so they run on traceStream beside frame k's temporal and spatial reuse on
 * reuseStream (events order the two). Each frame computes exactly what
 * rt_renderer_frame (method 0) computes for it with the same settings and
 * cameras: the same reservoirs and output, bit for bit. The pipeline owns
 * three G-buffers, two motion-vector images, two path-trace and two final
 * reservoir buffers, one Temp buffer and the indirect output.
 * Settings: the luminance target with BSDF sampling (targetFunction 0,
 * guidedProposal 0), either resampling weights; rt_pipeline_set_settings
 * rejects the others with RT_ERR_INVALID_ARGUMENT. The two streams may be the
 * same stream (then the frames run one after the other); the gain needs two
 * streams on different hardware queues. */
rt_status rt_pipeline_create(rt_device_scene* scene, uint32_t width, uint32_t height, rt_pipeline** out);
rt_status rt_pipeline_destroy(rt_pipeline* p);
rt_status rt_pipeline_set_settings(rt_pipeline* p, const rt_gris_settings* settings);
/* Workgroups of each frame's persistent path trace (rt_gris_path_trace_groups;
 * 0, the default: all the device holds). Fewer leave CUs to the reuse passes
 * of the frame before, which run beside it; the results do not change. */
rt_status rt_pipeline_set_trace_groups(rt_pipeline* p, uint32_t workgroups);
/* Enqueues one frame; cams[0].filmSize must equal the pipeline's size. The
 * frame's output and final reservoirs are complete once reuseStream has run
 * up to this call. */
rt_status rt_pipeline_frame(rt_pipeline* p, const rt_camera cams[2], hipStream_t traceStream, hipStream_t reuseStream);
/* The bindings after the last enqueued frame, as rt_renderer_bindings gives
 * them before the flip: that frame's G-buffer and motion vectors, its final
 * reservoirs (grisReservoir), the frame before's (grisReservoirPrev), the
 * indirect output. */
rt_status rt_pipeline_bindings(const rt_pipeline* p, rt_frame* out);
/* The direct light of each frame (Renderer::Settings directMethod), as
 * rt_renderer_frame_methods(r, cams, directMethod, 3) computes it: 0 None (the
 * default); 1 Naive (rt_pt_plain's direct part, depth settings.maxDepth, on
 * traceStream after the G-buffer); 2 ReSTIR DI with rt_pipeline_set_di_settings
 * (path generation on traceStream, temporal and spatial reuse on reuseStream
 * beside the path trace; the first call with 2 allocates five DI reservoir
 * buffers). A change of method waits for the device first. */
rt_status rt_pipeline_set_direct_method(rt_pipeline* p, uint32_t directMethod);
/* The stream of ReSTIR DI's temporal and spatial reuse (null, the default: the
 * reuse stream). A third stream lets them run beside both the path trace and
 * the GRIS reuse; the ReSTIR DI output is then complete once this stream has
 * run up to the frame. The stream is used from the next rt_pipeline_frame on. */
rt_status rt_pipeline_set_direct_stream(rt_pipeline* p, hipStream_t stream);
/* TestReSTIR::Settings of the pipeline's ReSTIR DI (default {0, 0, 0, 1}). */
rt_status rt_pipeline_set_di_settings(rt_pipeline* p, const rt_di_settings* settings);
/* ReSTIR DI's final reservoirs after the last enqueued frame (reservoir), the
 * frame before's (reservoirPrev) and Temp, as rt_renderer_di_bindings gives
 * them before the flip. */
rt_status rt_pipeline_di_bindings(const rt_pipeline* p, rt_di_reservoirs* out);
/* Copies an RGBA32F output (0 direct, 1 indirect) to host memory on stream:
 * pass the trace stream for the direct output and the reuse stream for the
 * indirect one (the copy then follows the last frame's pass that wrote it). */
rt_status rt_pipeline_read_output(const rt_pipeline* p, uint32_t which, float* host, hipStream_t stream);
/* Waits for the device, resets every reservoir buffer on stream, waits for
 * the reset, and restarts the frame count. */
rt_status rt_pipeline_reset_reservoirs(rt_pipeline* p, hipStream_t stream);

/* Writes an 8-bit RGBA image (host memory, row 0 at the top) as a PNG file:
 * the stbi_write_png call of Renderer::writeScreenshot (src/Renderer.cpp:787-788).
 * Image data is stored in uncompressed deflate blocks. */
rt_status rt_write_png(const char* path, const uint8_t* rgba8, uint32_t width, uint32_t height);
/* Reads an 8-bit PNG file (grey, grey + alpha, RGB, RGBA or palette; not
 * interlaced) as RGBA8, row 0 at the top: the albedo-texture load of
 * zvk::HostImage::createFromFile (zvk/core/HostImage.cpp:65-96, 4 channels).
 * With rgba8 null only the size is returned. */
rt_status rt_read_png(const char* path, uint32_t* width, uint32_t* height, uint8_t* rgba8);
/* Renderer::writeScreenshot after PostProcessFrag: post-processes the current
 * outputs (the last frame drawn) and writes them to path as a PNG. */
rt_status rt_renderer_screenshot(rt_renderer* r, const rt_post_settings* settings, const char* path, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* RESTIRPT_HOST_H */
