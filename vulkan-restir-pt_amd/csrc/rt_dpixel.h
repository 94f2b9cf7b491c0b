// Per-pixel helpers shared by the pass kernels: the 8x8 tile mapping, the
// pinhole camera ray, G-buffer access (texture() and texelFetch semantics),
// the traversal wrappers of ray_query.glsl and output accumulation.
#pragma once
#include "rt_internal.h"
#include <hip/hip_fp16.h>

namespace rtp {

using namespace rtd;

constexpr int kTileX = 8, kTileY = 8, kTileThreads = 64;

struct GB {   // one G-buffer parity (descriptor bindings 2/4/6 or 3/5)
    const float4* dn;
    const uint2* am;
    const uint32_t* mv;
    uint32_t W, H;
};

// ---- camera.glsl:28-42 ------------------------------------------------------------------
struct CamRay { f3 ori, dir; };
RT_DEV CamRay pinholeRay(const rt_camera& cam, f2 uv) {
    f2 texel = mk2(1.0f / (float)cam.filmSize[0], 1.0f / (float)cam.filmSize[1]);
    f2 biased = uv + texel * mk2(0.0f, 0.0f);
    f2 ndc = mk2(biased.x * 2.0f - 1.0f, biased.y * 2.0f - 1.0f);
    float aspect = (float)cam.filmSize[0] / (float)cam.filmSize[1];
    float tanFOV = rt_tan(cam.FOV * 0.5f * 0.017453292519943295f);
    f3 pf = mk3(ndc.x * aspect * tanFOV, ndc.y * 1.0f * tanFOV, 1.0f);
    f3 dir = normalize(pf);
    dir = normalize(ld3(cam.right) * dir.x + ld3(cam.up) * dir.y + ld3(cam.front) * dir.z);
    CamRay r; r.ori = ld3(cam.pos); r.dir = dir;
    return r;
}
RT_DEV f2 pixelUV(uint32_t x, uint32_t y, uint32_t W, uint32_t H) {
    return mk2(((float)x + 0.5f) / (float)W, ((float)y + 0.5f) / (float)H);
}

// ---- G-buffer access ----------------------------------------------------------------------
// texture() with linear filtering and Repeat addressing, 8-bit subtexel weights.
RT_DEV void bilinearSetup(float u, uint32_t W, int& i0, int& i1, float& a) {
    float x = u * (float)W - 0.5f;
    float fl = floorf(x);
    float fr = x - fl;
    int q = (int)(fr * 256.0f + 0.5f);
    int ii = (int)fl;
    if (q >= 256) { ii += 1; q = 0; }
    int w = (int)W;
    i0 = ((ii % w) + w) % w;
    i1 = (((ii + 1) % w) + w) % w;
    a = (float)q * (1.0f / 256.0f);
}
RT_DEV f4 texDN(const GB& g, f2 uv) {
    int x0, x1, y0, y1; float a, b;
    bilinearSetup(uv.x, g.W, x0, x1, a);
    bilinearSetup(uv.y, g.H, y0, y1, b);
    float4 t00 = g.dn[(size_t)g.W * y0 + x0], t10 = g.dn[(size_t)g.W * y0 + x1];
    float4 t01 = g.dn[(size_t)g.W * y1 + x0], t11 = g.dn[(size_t)g.W * y1 + x1];
    float oa = 1.0f - a, ob = 1.0f - b;
    f4 r;
    r.x = (t00.x * oa + t10.x * a) * ob + (t01.x * oa + t11.x * a) * b;
    r.y = (t00.y * oa + t10.y * a) * ob + (t01.y * oa + t11.y * a) * b;
    r.z = (t00.z * oa + t10.z * a) * ob + (t01.z * oa + t11.z * a) * b;
    r.w = (t00.w * oa + t10.w * a) * ob + (t01.w * oa + t11.w * a) * b;
    return r;
}
// At a pixel centre the 8-bit weights round to (0, 0) or (1, 1), so texture()
// returns exactly the texel (valid for images narrower than 32768 pixels).
RT_DEV f4 texDNCentre(const GB& g, uint32_t x, uint32_t y) {
    float4 t = g.dn[(size_t)g.W * y + x];
    f4 r; r.x = t.x; r.y = t.y; r.z = t.z; r.w = t.w;
    return r;
}
RT_DEV uint2 fetchAM(const GB& g, int x, int y) {
    if (x < 0 || y < 0 || x >= (int)g.W || y >= (int)g.H) return make_uint2(0u, 0u);
    return g.am[(size_t)g.W * y + x];
}
RT_DEV f2 fetchMotion(const GB& g, int x, int y) {
    uint32_t v = g.mv[(size_t)g.W * y + x];
    return mk2(__half2float(__ushort_as_half((unsigned short)(v & 0xffffu))),
               __half2float(__ushort_as_half((unsigned short)(v >> 16))));
}
RT_DEV f3 unpackAlbedo(uint32_t p) {
    return mk3((float)(p & 0xffu) / 255.0f, (float)((p >> 8) & 0xffu) / 255.0f, (float)((p >> 16) & 0xffu) / 255.0f);
}
RT_DEV uint32_t packUnorm8(float c) { return (uint32_t)(gclamp(c, 0.0f, 1.0f) * 255.0f + 0.5f); }

struct GPix { float depth; f3 norm; f3 albedo; int matMeshId; };
// gbuffer_util.glsl:24-35
RT_DEV bool unpackGB(f4 dn, uint2 am, GPix& p) {
    p.depth = dn.x;
    if (p.depth == 0.0f) return false;
    p.norm = mk3(dn.y, dn.z, dn.w);
    p.albedo = unpackAlbedo(am.x);
    p.matMeshId = (int)am.y;
    return true;
}

// ---- traversal wrappers (ray_query.glsl:6-70) ------------------------------------------------
template <bool kCount, bool kCull = true>
RT_DEV Isec traceClosestHit(const DevScene& S, TStack st, f3 o, float tmin, f3 d, float tmax, TraceCounters* ctr,
                            bool skipLights = false, float* tOut = nullptr) {
    RayT r = makeRay(o, d, tmin, tmax);
    Isec h; h.b1 = 0.0f; h.b2 = 0.0f; h.inst = RT_INVALID_HIT; h.prim = 0;
    float tBest = tmax;
    bool hit = traverse<false, kCount, kCull>(S, r, skipLights, st, tBest, h, ctr);
    if (!hit) { h.inst = RT_INVALID_HIT; h.b1 = 0.0f; h.b2 = 0.0f; h.prim = 0; }
    if (tOut) *tOut = hit ? tBest : 0.0f;
    return h;
}
template <bool kCount>
RT_DEV bool traceShadow(const DevScene& S, TStack st, f3 o, float tmin, f3 d, float tmax, TraceCounters* ctr) {
    RayT r = makeRay(o, d, tmin, tmax);
    Isec h; float tBest = tmax;
    return traverse<true, kCount>(S, r, false, st, tBest, h, ctr);
}
template <bool kCount>
RT_DEV bool traceVisibility(const DevScene& S, TStack st, f3 from, f3 to, TraceCounters* ctr) {
    return !traceShadow<kCount>(S, st, from, RT_MIN_RAY, normalize(to - from), distance(to, from) - RT_MIN_RAY, ctr);
}

RT_DEV void accumulate(float* out, size_t p, f3 c, uint32_t frameIndex) {
    float k = (float)(frameIndex & 0x7fffffffu);
    float4 o = reinterpret_cast<float4*>(out)[p];
    f3 acc = mk3(o.x, o.y, o.z);
    acc = (acc * k + c) / (k + 1.0f);
    reinterpret_cast<float4*>(out)[p] = make_float4(acc.x, acc.y, acc.z, 1.0f);
}

#define PIXEL_INDEX                                                              \
    const uint32_t x = blockIdx.x * kTileX + (threadIdx.x & (kTileX - 1));      \
    const uint32_t y = blockIdx.y * kTileY + (threadIdx.x / kTileX);

inline dim3 tileGrid(uint32_t W, uint32_t H) { return dim3((W + kTileX - 1) / kTileX, (H + kTileY - 1) / kTileY, 1); }
inline GB makeGB(const float* dn, const uint32_t* am, const uint16_t* mv, uint32_t W, uint32_t H) {
    GB g; g.dn = reinterpret_cast<const float4*>(dn); g.am = reinterpret_cast<const uint2*>(am);
    g.mv = reinterpret_cast<const uint32_t*>(mv); g.W = W; g.H = H;
    return g;
}

}  // namespace rtp
