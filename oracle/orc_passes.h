// ORACLE — test infrastructure only. Nothing in the product links or calls this.
//
// Per-pixel passes restated from the reference GLSL:
//   gbuffer      : src/shader/GBuffer.vert:19-29 + GBuffer.frag:20-53, evaluated by a
//                  primary ray at the pixel centre instead of rasterisation
//   di / gi      : src/shader/di_naive.glsl:79-167, gi_naive.glsl:28-149
//   GRIS         : src/shader/gris_path_trace.glsl:45-306, gris_retrace.glsl:42-236,
//                  gris_resample_temporal.glsl:11-83, gris_resample_spatial.glsl:11-135
//   ReSTIR GI    : src/shader/gi_resample_temporal.glsl:9-213 (orc_gi.cpp)
//   ReSTIR DI    : src/shader/di_path_gen.glsl:9-36, di_temporal.glsl:9-88,
//                  di_spatial.glsl:9-119, di_reservoir.glsl:28-290 (orc_di.cpp)
//   wrappers     : *.comp main() accumulation (frameIndex & mask)
// Uninitialised GLSL locals are taken as zero.
#pragma once
#include "orc_bvh.h"

namespace orc {

struct GRISSettings {
    uint32_t shiftMode;     // gris_reservoir.glsl:21-31 push constant
    float rrScale;
    uint32_t temporalReuse;
    uint32_t spatialReuse;
    uint32_t cap;
    uint32_t maxDepth;          // MaxTracingDepth (gris_path_trace.glsl:46, 15)
    uint32_t spatialNeighbors;  // ResampleNum (gris_resample_spatial.glsl:65, 3)
    float spatialRadius;        // ResampleRadius (gris_resample_spatial.glsl:66, 20)
    uint32_t targetFunction;    // 0 luminance (GRISToScalar), 1 learned
    uint32_t unbiasedReuse;     // 0 the reference's 1/M weights, 1 the "1/Z" contribution weights
    uint32_t guidedProposal;    // 1: learned vMF-mixture proposal at the primary vertex (orc_guide.h)
};
static const uint32_t UnbiasedMaxNeighbors = 8;
// Learned-target data of a reservoir: predicted reconnection radiance and the
// target value of the sample that resampleWeight refers to.
struct LearnedData { float L[3]; float pRef; };
static const uint32_t ShiftReconnection = 0, ShiftReplay = 1, ShiftHybrid = 2;
static const uint32_t RcVertexTypeLightSampled = 0, RcVertexTypeLightScattered = 1, RcVertexTypeSurface = 2;
static const float GRISDistanceThreshold = 0.01f;
static const uint32_t CameraClearFlag = 0x80000000u;
static const uint32_t CameraFrameIndexMask = 0x7fffffffu;

struct GBufView {
    const float* depthNormal;    // RGBA32F
    const uint32_t* albedoMatId; // RG32UI
    const uint16_t* motion;      // RG16F (may be null)
    uint32_t W, H;
};

struct Ctx {
    const SceneArrays* S;
    const BVH* B;
    TraceStats* st;
    // learned proposal (settings.guidedProposal): lobe planes of this and the previous
    // frame (float4 planes of W*H pixels) and the per-pixel training record written by
    // tracePath (direction sampled at the primary vertex, stream-RIS weight sum)
    const float* guideLobes = nullptr;
    const float* guideLobesPrev = nullptr;
    float* guideSamples = nullptr;
};

static inline uint32_t index1D(uint32_t W, uint32_t x, uint32_t y) { return W * y + x; }

// ---- camera.glsl:28-42 -----------------------------------------------------------
struct Ray { vec3 ori, dir; };
static inline Ray pinholeCameraSampleRay(const Camera& cam, vec2 uv, vec2 r) {
    vec2 texelSize = V2(1.0f / (float)cam.filmSize[0], 1.0f / (float)cam.filmSize[1]);
    vec2 biased = uv + texelSize * r;
    vec2 ndc = V2(biased.x * 2.0f - 1.0f, biased.y * 2.0f - 1.0f);
    float aspect = (float)cam.filmSize[0] / (float)cam.filmSize[1];
    float tanFOV = rt_tan(cam.FOV * 0.5f * 0.017453292519943295f);
    vec3 pFocus = V3(ndc.x * aspect * tanFOV, ndc.y * 1.0f * tanFOV, 1.0f);
    vec3 dir = normalize(pFocus);
    dir = normalize(L3(cam.right) * dir.x + L3(cam.up) * dir.y + L3(cam.front) * dir.z);
    return {L3(cam.pos), dir};
}

// ---- G-buffer access -------------------------------------------------------------
// texture() with a linear filter and Repeat addressing (GBufferPass.cpp:108,
// zvk/core/Memory.cpp:79-83): 8-bit subtexel weights, rounded to nearest.
static inline void bilinearSetup(float u, uint32_t W, int& i0, int& i1, float& a) {
    float x = u * (float)W - 0.5f;
    float fl = std::floor(x);
    float fr = x - fl;
    int q = (int)(fr * 256.0f + 0.5f);
    int ii = (int)fl;
    if (q >= 256) { ii += 1; q = 0; }
    int w = (int)W;
    i0 = ((ii % w) + w) % w;
    i1 = (((ii + 1) % w) + w) % w;
    a = (float)q * (1.0f / 256.0f);
}
static inline vec4 texDepthNormal(const GBufView& g, vec2 uv) {
    int x0, x1, y0, y1; float a, b;
    bilinearSetup(uv.x, g.W, x0, x1, a);
    bilinearSetup(uv.y, g.H, y0, y1, b);
    const float* t00 = g.depthNormal + 4 * (size_t)index1D(g.W, x0, y0);
    const float* t10 = g.depthNormal + 4 * (size_t)index1D(g.W, x1, y0);
    const float* t01 = g.depthNormal + 4 * (size_t)index1D(g.W, x0, y1);
    const float* t11 = g.depthNormal + 4 * (size_t)index1D(g.W, x1, y1);
    float r[4];
    for (int k = 0; k < 4; k++) {
        float top = t00[k] * (1.0f - a) + t10[k] * a;
        float bot = t01[k] * (1.0f - a) + t11[k] * a;
        r[k] = top * (1.0f - b) + bot * b;
    }
    return {r[0], r[1], r[2], r[3]};
}
// texelFetch: out-of-range texels read as zero.
static inline void fetchAlbedoMatId(const GBufView& g, int x, int y, uint32_t out[2]) {
    if (x < 0 || y < 0 || x >= (int)g.W || y >= (int)g.H) { out[0] = out[1] = 0; return; }
    const uint32_t* p = g.albedoMatId + 2 * (size_t)index1D(g.W, x, y);
    out[0] = p[0]; out[1] = p[1];
}
static inline vec2 fetchMotion(const GBufView& g, int x, int y) {
    if (!g.motion || x < 0 || y < 0 || x >= (int)g.W || y >= (int)g.H) return V2(0, 0);
    const uint16_t* p = g.motion + 2 * (size_t)index1D(g.W, x, y);
    return V2(f16_to_f32(p[0]), f16_to_f32(p[1]));
}
// gbuffer_util.glsl:4-10 (packUnorm4x8 / unpackUnorm4x8)
static inline uint32_t packUnorm8(float c) {
    float v = gclamp(c, 0.0f, 1.0f) * 255.0f;
    return (uint32_t)(v + 0.5f);
}
static inline uint32_t packAlbedo(vec3 a) {
    return packUnorm8(a.x) | (packUnorm8(a.y) << 8) | (packUnorm8(a.z) << 16) | (packUnorm8(1.0f) << 24);
}
static inline vec3 unpackAlbedo(uint32_t p) {
    return V3((float)(p & 0xffu) / 255.0f, (float)((p >> 8) & 0xffu) / 255.0f, (float)((p >> 16) & 0xffu) / 255.0f);
}
// gbuffer_util.glsl:24-35
static inline bool unpackGBuffer(vec4 dn, const uint32_t am[2], float& depth, vec3& normal, vec3& albedo, int& matMeshId) {
    depth = dn.x;
    if (depth == 0.0f) return false;
    normal = V3(dn.y, dn.z, dn.w);
    albedo = unpackAlbedo(am[0]);
    matMeshId = (int)am[1];
    return true;
}

// ---- traversal wrappers (ray_query.glsl:6-70) ----------------------------------------
static inline Intersection traceClosestHit(const Ctx& c, vec3 o, float tmin, vec3 d, float tmax) {
    return traceClosest4(*c.B, makeRayIn(o, d, tmin, tmax), false, c.st).isec;
}
static inline bool traceShadow(const Ctx& c, vec3 o, float tmin, vec3 d, float tmax) {
    return traceAny4(*c.B, makeRayIn(o, d, tmin, tmax), c.st);
}
static inline bool traceVisibility(const Ctx& c, vec3 from, vec3 to) {
    return !traceShadow(c, from, MinRayDistance, normalize(to - from), distance(to, from) - MinRayDistance);
}

// ---- G-buffer producer --------------------------------------------------------------
void gbufferPixel(const Ctx& c, const Camera& cam, uint32_t x, uint32_t y, float* dn, uint32_t* am, uint16_t* mv);

// ---- passes -----------------------------------------------------------------------
vec3 directIllumination2(const Ctx& c, const Camera& cam, const GBufView& g, uint32_t x, uint32_t y);
vec3 indirectIllumination(const Ctx& c, const Camera& cam, const GBufView& g, uint32_t x, uint32_t y, uint32_t maxDepth);
vec3 tracePath(const Ctx& c, const Camera& cam, const GBufView& g, const GRISSettings& set, uint32_t x, uint32_t y,
               GRISReservoir& out, bool& written);
bool temporalReuse(const Ctx& c, const Camera& cam, const Camera& prevCam, const GBufView& g, const GBufView& gPrev,
                   const GRISSettings& set, const GRISReservoir* resvThis, const GRISReservoir* resvPrev,
                   uint32_t x, uint32_t y, GRISReservoir& outTemp, const LearnedData* learnThis = nullptr,
                   const LearnedData* learnPrev = nullptr, LearnedData* outLearn = nullptr);
vec3 spatialReuse(const Ctx& c, const Camera& cam, const GBufView& g, const GRISSettings& set,
                  const GRISReservoir* resvTemp, uint32_t x, uint32_t y, GRISReservoir& outThis, bool& written,
                  const LearnedData* learnTemp = nullptr, LearnedData* outLearn = nullptr);
// Learned-target refresh of one pixel's reservoir after the path-trace pass,
// given the network's prediction Lhat at its rc vertex. Returns false for
// background pixels (nothing written). need: the sample has a surface rc
// vertex (pos = its position).
bool learnedRefresh(const Ctx& c, const GBufView& g, uint32_t x, uint32_t y, GRISReservoir& resv, const float Lhat[3],
                    LearnedData& out, bool& need, vec3& pos);

// ---- ReSTIR DI (TestReSTIR) ---------------------------------------------------------
struct DIPathSample {        // layouts.glsl:97-106, 48 B
    Intersection isec;
    float Li[3]; float pad0;
    float jacobian; float samplePdf; uint32_t rng; uint32_t isLightSample;
};
struct DIReservoir {         // layouts.glsl:108-115, 64 B
    DIPathSample pathSample;
    uint32_t sampleCount; float resampleWeight; float contribWeight; float weight;
};
static_assert(sizeof(DIReservoir) == 64, "DI reservoir layout");
struct DISettings {          // TestReSTIR::Settings, src/TestReSTIR.h:12-17
    uint32_t shiftType;      // 0 Reconnection, 1 Replay, 2 Hybrid (merges nothing)
    uint32_t sampleType;     // 0 Light, 1 BSDF, 2 Both
    uint32_t temporalReuse;
    uint32_t spatialReuse;
};
static const uint32_t DISampleLight = 0, DISampleBSDF = 1, DISampleBoth = 2;

// ---- ReSTIR GI (temporal) -------------------------------------------------------------
struct GIPathSample {        // layouts.glsl:117-121, 32 B
    Intersection rcIsec;
    float rcLo[3]; uint32_t rcPrevCoord;
};
struct GIReservoir {         // layouts.glsl:123-129, 48 B
    GIPathSample pathSample;
    uint32_t sampleCount; float resampleWeight; float contribWeight; float pad0;
};
static_assert(sizeof(GIReservoir) == 48, "GI reservoir layout");
// Writes resvThis[pixel] (a background pixel's reservoir is reset in place);
// returns the radiance before clampColor.
vec3 giTemporal(const Ctx& c, const Camera& cam, const Camera& prevCam, const GBufView& g, const GBufView& gPrev,
                GIReservoir* resvThis, const GIReservoir* resvPrev, uint32_t x, uint32_t y);

// Each returns false (nothing written) for a background pixel.
bool diPathGen(const Ctx& c, const Camera& cam, const GBufView& g, const DISettings& set, uint32_t x, uint32_t y,
               DIReservoir& out);
bool diTemporal(const Ctx& c, const Camera& cam, const Camera& prevCam, const GBufView& g, const GBufView& gPrev,
                const DISettings& set, const DIReservoir* resvThis, const DIReservoir* resvPrev, uint32_t x,
                uint32_t y, DIReservoir& out);
vec3 diSpatial(const Ctx& c, const Camera& cam, const GBufView& g, const DISettings& set,
               const DIReservoir* resvTemp, uint32_t x, uint32_t y, DIReservoir& outThis, bool& written);

}  // namespace orc
