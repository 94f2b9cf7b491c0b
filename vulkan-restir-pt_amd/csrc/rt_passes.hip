// Per-pixel kernels for gfx950: G-buffer producer, plain NEE+MIS path tracer
// and the three GRIS (ReSTIR PT) passes.
//
// Dispatch mirrors the reference's 8x8 workgroups (src/shader/HostDevice.h:36-37,
// zvk/core/ComputePipeline.cpp:11-25): one 64-lane wave per 8x8 pixel tile.
// Each lane keeps its path state in registers and traverses the LBVH with an
// LDS short stack. GRIS reservoirs live in HBM as five float4 planes
// (80 B/pixel) so that every plane access of a wave is one contiguous 1-KiB
// transaction:
//   P0 rcIsec (bary.xy, instance, primitive)   P1 rcLi.xyz, rcRng
//   P2 rcWi.xyz, flags                         P3 F.xyz, primaryRng
//   P4 rcPrevSamplePdf, rcJacobian, sampleCount, resampleWeight
// Arithmetic follows the reference GLSL expression by expression (see the
// citations at each function) under the FP contract of rt_dmath.h.
#include "rt_dpixel.h"
#include "rt_guide.h"
#include <algorithm>
#include <cstring>

using namespace rtd;
using namespace rtp;

namespace {

// Occupancy that each kernel's register allocation is held to, in waves per
// SIMD (a lane then has at most 512 / waves VGPRs). Chosen by measurement on
// the 1080p terrain frame: the GRIS kernels run fastest at 4 waves even with
// some spilling (path trace 2.63 ms at 4 waves, 2.72 at 5, 2.78 at 3; spatial
// 1.93 ms at 4, 2.19 at 5, 2.35 at 6), faster than at the 2-3 waves their
// unconstrained allocation (150-170 VGPRs) gives. Each can be overridden at
// build time.
#ifndef RT_WAVES_TRACE
#define RT_WAVES_TRACE 4
#endif
#ifndef RT_WAVES_GBUFFER
#define RT_WAVES_GBUFFER 4
#endif
#ifndef RT_WAVES_PT_PLAIN
#define RT_WAVES_PT_PLAIN 4
#endif
#ifndef RT_WAVES_PATH_TRACE
#define RT_WAVES_PATH_TRACE 4
#endif
#ifndef RT_WAVES_TEMPORAL
#define RT_WAVES_TEMPORAL 4
#endif
#ifndef RT_WAVES_SPATIAL
#define RT_WAVES_SPATIAL 4
#endif
// The reuse passes with the unbiased weights (a shift check per merged pixel
// on top of the merges) hold more state; their variants have their own count.
#ifndef RT_WAVES_TEMPORAL_UB
#define RT_WAVES_TEMPORAL_UB RT_WAVES_TEMPORAL
#endif
#ifndef RT_WAVES_SPATIAL_UB
#define RT_WAVES_SPATIAL_UB RT_WAVES_SPATIAL
#endif

// ================================================================================================
// Ray query test entry
template <bool kCount>
__global__ __launch_bounds__(kTileThreads, RT_WAVES_TRACE) void k_trace_rays(DevScene S, const float* __restrict__ rays8, uint32_t n,
                                                             int mode, int skipLights, float* t, float* bary2,
                                                             uint32_t* inst, uint32_t* prim, TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    uint32_t i = blockIdx.x * kTileThreads + threadIdx.x;
    if (i >= n) return;
    TStack st = laneStack(stk);
    const float* q = rays8 + 8 * (size_t)i;
    f3 o = mk3(q[0], q[1], q[2]), d = mk3(q[3], q[4], q[5]);
    if (mode == 1) {
        bool occ = traceShadow<kCount>(S, st, o, q[6], d, q[7], ctr);
        inst[i] = occ ? 1u : 0u; prim[i] = 0; t[i] = 0.0f; bary2[2 * i] = 0.0f; bary2[2 * i + 1] = 0.0f;
        return;
    }
    float th;
    Isec h = traceClosestHit<kCount>(S, st, o, q[6], d, q[7], ctr, skipLights != 0, &th);
    t[i] = th; bary2[2 * i] = h.b1; bary2[2 * i + 1] = h.b2; inst[i] = h.inst; prim[i] = h.prim;
}

// ================================================================================================
// G-buffer: GBuffer.vert:19-29 + GBuffer.frag:20-53 evaluated at pixel centres.
// Light triangles are not drawn by the reference's raster pass; the clip
// planes act as ray distances [near, far].
template <bool kCount>
RT_DEV void gbufferPixel(const DevScene& S, const rt_camera& cam, float4* __restrict__ dnOut, uint2* __restrict__ amOut,
                         uint32_t* __restrict__ mvOut, uint32_t x, uint32_t y, TStack st, TraceCounters* ctr) {
    const uint32_t W = cam.filmSize[0], H = cam.filmSize[1];
    if (x >= W || y >= H) return;
    size_t p = (size_t)y * W + x;
    f2 uv = pixelUV(x, y, W, H);
    CamRay ray = pinholeRay(cam, mk2(uv.x, 1.0f - uv.y));
    Isec h = traceClosestHit<kCount>(S, st, ray.ori, cam.nearZ, ray.dir, cam.farZ, ctr, true);
    if (h.inst == RT_INVALID_HIT) {
        dnOut[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        amOut[p] = make_uint2(0u, 0u);
        mvOut[p] = 0u;
        return;
    }
    uint32_t meshIdx = h.inst - 1;
    const DevInstance& I = S.instances[meshIdx];
    uint32_t tri = h.prim;
    float b1 = h.b1, b2 = h.b2, b0 = 1.0f - b1 - b2;
    f3 P[3], N[3];
    float U[3], V[3];
    for (int k = 0; k < 3; k++) {
        const rt_mesh_vertex& v = S.vertices[S.indices[I.indexOffset + tri * 3 + k]];
        U[k] = v.uvx; V[k] = v.uvy;
        P[k] = colMul4(I.m, ld3(v.pos), 1.0f);
        f3 nn = mk3(I.n[0], I.n[1], I.n[2]) * v.norm[0] + mk3(I.n[3], I.n[4], I.n[5]) * v.norm[1] +
                mk3(I.n[6], I.n[7], I.n[8]) * v.norm[2];
        N[k] = normalize(nn);
    }
    f3 pos = P[0] * b0 + P[1] * b1 + P[2] * b2;
    f3 nrm = normalize(N[0] * b0 + N[1] * b1 + N[2] * b2);
    uint32_t matIndex = (uint32_t)S.materialIndices[I.indexOffset / 3 + tri];
    const rt_material& m = S.materials[matIndex];
    const float* L = cam.lastProjView;
    float lx = L[0] * pos.x + L[4] * pos.y + L[8] * pos.z + L[12] * 1.0f;
    float ly = L[1] * pos.x + L[5] * pos.y + L[9] * pos.z + L[13] * 1.0f;
    float lw = L[3] * pos.x + L[7] * pos.y + L[11] * pos.z + L[15] * 1.0f;
    lx = lx / lw; ly = ly / lw;
    f2 lastC = mk2(lx * 0.5f + 0.5f, ly * 0.5f + 0.5f);
    dnOut[p] = make_float4(length(ld3(cam.pos) - pos), nrm.x, nrm.y, nrm.z);
    // GBuffer.frag:33-37: the texture at the interpolated uv, or baseColor
    const f3 albedo = materialAlbedo(S, m, U[0] * b0 + U[1] * b1 + U[2] * b2, V[0] * b0 + V[1] * b1 + V[2] * b2);
    uint32_t alb = packUnorm8(albedo.x) | (packUnorm8(albedo.y) << 8) | (packUnorm8(albedo.z) << 16) |
                   (packUnorm8(1.0f) << 24);
    amOut[p] = make_uint2(alb, (matIndex << 16) | meshIdx);
    uint32_t mx = __half_as_ushort(__float2half_rn(lastC.x - uv.x));
    uint32_t my = __half_as_ushort(__float2half_rn(lastC.y - uv.y));
    mvOut[p] = mx | (my << 16);
}
template <bool kCount>
__global__ __launch_bounds__(kTileThreads, RT_WAVES_GBUFFER) void k_gbuffer(DevScene S, rt_camera cam, float4* __restrict__ dnOut,
                                                          uint2* __restrict__ amOut, uint32_t* __restrict__ mvOut,
                                                          TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    PIXEL_INDEX
    gbufferPixel<kCount>(S, cam, dnOut, amOut, mvOut, x, y, laneStack(stk), ctr);
}

// ================================================================================================
// Plain PT: di_naive.glsl:79-167 (bit 0) and gi_naive.glsl:28-149 (bit 1) with
// the accumulation of di_naive.comp / gi_naive.comp.
template <bool kCount>
RT_DEV f3 directIllumination2(const DevScene& S, TStack st, const rt_camera& cam, const GB& g, uint32_t x, uint32_t y,
                              TraceCounters* ctr) {
    f2 uv = pixelUV(x, y, g.W, g.H);
    GPix gp;
    if (!unpackGB(texDNCentre(g, x, y), fetchAM(g, (int)x, (int)y), gp)) return s3(0.0f);
    int matId = gp.matMeshId >> 16;
    CamRay ray = pinholeRay(cam, mk2(uv.x, 1.0f - uv.y));
    uint32_t rng = makeSeedPix(cam.seed, x, y);
    f3 radiance = s3(0.0f);
    f3 pos = ray.ori + ray.dir * (gp.depth - 1e-4f);
    f3 wo = -ray.dir;
    Mat mat = loadMat(S, (uint32_t)matId);
    f3 sLi = s3(0.0f); float sW = 0.0f, sSum = 0.0f;
    if (!isBSDFDelta(mat)) {
        f3 lightDir; float lightDist, lightPdf, jac; f2 bary; uint32_t id;
        f3 Lr = sampleLightByPower(S, pos, lightDir, lightDist, lightPdf, jac, bary, id, sample4f(rng));
        bool shadowed = traceShadow<kCount>(S, st, pos, RT_MIN_RAY, lightDir, lightDist - 1e-4f, ctr);
        if (!shadowed && lightPdf > 1e-6f) {
            float bsdfPdf = evalPdf(mat, gp.norm, -ray.dir, lightDir);
            float w = MISWeight(lightPdf, bsdfPdf);
            f3 contrib = Lr * evalBSDF(mat, gp.albedo, gp.norm, wo, lightDir) * satDot(gp.norm, lightDir) / lightPdf * w;
            float r = sample1f(rng);
            sSum += 100.0f;
            if (r * sSum < 100.0f) { sW = 100.0f; sLi = contrib; }
            radiance = radiance + contrib;
        }
    }
    BSDFSample s; s.type = 0; s.pdf = 0.0f;
    if (sampleBSDF(mat, gp.albedo, gp.norm, wo, sample3f(rng), s) && s.pdf > 1e-6f) {
        Isec is = traceClosestHit<kCount>(S, st, pos, RT_MIN_RAY, s.wi, RT_MAX_RAY, ctr);
        if (is.inst != RT_INVALID_HIT) {
            Surf surf; loadSurfaceInfo(S, is, surf);
            float cosTheta = -dot(s.wi, surf.norm);
            if (surf.isLight && cosTheta > 0.0f) {
                float dist = length(surf.pos - pos);
                float sumPower = S.lightTable[0].prob;
                float lightPdf = luminance(surf.albedo) / sumPower * dist * dist / fabsf(cosTheta);
                float w = isSampleTypeDelta(s.type) ? 1.0f : MISWeight(s.pdf, lightPdf);
                float cosTerm = isSampleTypeDelta(s.type) ? 1.0f : satDot(gp.norm, s.wi);
                f3 contrib = surf.albedo * s.bsdf * cosTerm / s.pdf * w;
                float r = sample1f(rng);
                sSum += 1.0f;
                if (r * sSum < 1.0f) { sW = 1.0f; sLi = contrib; }
                radiance = radiance + contrib;
            }
        }
    }
    if (sW > 0.0f && sSum > 0.0f) radiance = sLi * sSum / sW;
    else radiance = s3(0.0f);
    return clampColor(radiance);
}

template <bool kCount>
RT_DEV f3 indirectIllumination(const DevScene& S, TStack st, const rt_camera& cam, const GB& g, uint32_t x, uint32_t y,
                               uint32_t maxDepth, TraceCounters* ctr) {
    f2 uv = pixelUV(x, y, g.W, g.H);
    GPix gp;
    if (!unpackGB(texDNCentre(g, x, y), fetchAM(g, (int)x, (int)y), gp)) return s3(0.0f);
    int matId = gp.matMeshId >> 16;
    CamRay ray = pinholeRay(cam, mk2(uv.x, 1.0f - uv.y));
    uint32_t rng = makeSeedPix(cam.seed, x, y);
    f3 radiance = s3(0.0f), throughput = s3(1.0f), lastPos = s3(0.0f);
    f3 wo = -ray.dir;
    Surf surf;
    surf.pos = ray.ori + ray.dir * (gp.depth - 1e-4f);
    surf.norm = gp.norm; surf.albedo = gp.albedo; surf.isLight = false; surf.matIndex = (uint32_t)matId;
    Mat mat = loadMat(S, (uint32_t)matId);
    BSDFSample s; s.type = 0; s.pdf = 0.0f; s.wi = s3(0.0f); s.bsdf = s3(0.0f);
    f3 rdir = ray.dir, rori = ray.ori;
    for (uint32_t bounce = 0; bounce < maxDepth; bounce++) {
        if (bounce > 0) {
            Isec is = traceClosestHit<kCount>(S, st, rori, RT_MIN_RAY, rdir, RT_MAX_RAY, ctr);
            if (is.inst == RT_INVALID_HIT) break;
            loadSurfaceInfo(S, is, surf);
            mat = loadMat(S, surf.matIndex);
        }
        if (surf.isLight) {
            float cosTheta = -dot(rdir, surf.norm);
            if (bounce > 1 && cosTheta > 0.0f) {
                float w = 1.0f;
                if (!isSampleTypeDelta(s.type)) {
                    float dist = length(surf.pos - lastPos);
                    float sumPower = S.lightTable[0].prob;
                    float lightPdf = luminance(surf.albedo) / sumPower * dist * dist / fabsf(cosTheta);
                    w = MISWeight(s.pdf, lightPdf);
                }
                radiance = radiance + surf.albedo * w * throughput;
            }
            break;
        }
        if (bounce > 0 && !isBSDFDelta(mat)) {
            f3 lightDir; float lightDist, lightPdf, jac; f2 bary; uint32_t id;
            f3 Lr = sampleLightByPower(S, surf.pos, lightDir, lightDist, lightPdf, jac, bary, id, sample4f(rng));
            bool shadowed = traceShadow<kCount>(S, st, surf.pos, RT_MIN_RAY, lightDir, lightDist - RT_MIN_RAY, ctr);
            if (!shadowed && lightPdf > 1e-6f) {
                // gi_naive.glsl:116-118 forces the NEE MIS weight to 1
                f3 contrib = Lr * evalBSDF(mat, surf.albedo, surf.norm, wo, lightDir) * satDot(surf.norm, lightDir) /
                             lightPdf * 1.0f * throughput;
                radiance = radiance + contrib;
            }
        }
        if (bounce > 4) {
            float pdfTerminate = gmax(1.0f - luminance(throughput), 0.0f);
            if (sample1f(rng) < pdfTerminate) break;
            throughput = throughput / (1.0f - pdfTerminate);
        }
        if (!sampleBSDF(mat, surf.albedo, surf.norm, wo, sample3f(rng), s) || s.pdf < 1e-6f) break;
        float cosTheta = isSampleTypeDelta(s.type) ? 1.0f : absDot(surf.norm, s.wi);
        throughput = throughput * (s.bsdf * cosTheta / s.pdf);
        lastPos = surf.pos;
        wo = -s.wi;
        rdir = s.wi;
        rori = surf.pos + rdir * 1e-4f;
    }
    return clampColor(radiance);
}

template <bool kCount>
__global__ __launch_bounds__(kTileThreads, RT_WAVES_PT_PLAIN) void k_pt_plain(DevScene S, rt_camera cam, GB g, uint32_t maxDepth,
                                                           uint32_t which, float* directOut, float* indirectOut,
                                                           TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    PIXEL_INDEX
    if (x >= g.W || y >= g.H) return;
    TStack st = laneStack(stk);
    size_t p = (size_t)y * g.W + x;
    // both estimators: the grid's two z layers take one each (indirect first), so a pixel's two
    // independent estimates run on two lanes instead of one after the other
    const uint32_t part = which == 3u ? (blockIdx.z == 0u ? 2u : 1u) : which;
    if (part & 1u) accumulate(directOut, p, directIllumination2<kCount>(S, st, cam, g, x, y, ctr), cam.frameIndex);
    if (part & 2u) accumulate(indirectOut, p, indirectIllumination<kCount>(S, st, cam, g, x, y, maxDepth, ctr), cam.frameIndex);
}

// ================================================================================================
// GRIS reservoir (layouts.glsl:132-156) in registers, planar in HBM.
struct Resv {
    Isec rcIsec;
    f3 rcLi; uint32_t rcRng;
    f3 rcWi; uint32_t flags;
    float rcPrevSamplePdf, rcJacobian;
    f3 F; uint32_t primaryRng;
    float sampleCount, resampleWeight;
    // learned target only (plane P5): predicted reconnection radiance and the
    // target value of the sample that resampleWeight refers to
    f3 Lhat; float pRef;
};
struct ResvPlanes { float4* p; size_t n; };

template <bool kLearned = false>
RT_DEV Resv loadResv(const ResvPlanes& R, size_t i) {
    float4 a = R.p[i], b = R.p[R.n + i], c = R.p[2 * R.n + i], d = R.p[3 * R.n + i], e = R.p[4 * R.n + i];
    Resv r;
    r.rcIsec.b1 = a.x; r.rcIsec.b2 = a.y; r.rcIsec.inst = __float_as_uint(a.z); r.rcIsec.prim = __float_as_uint(a.w);
    r.rcLi = mk3(b.x, b.y, b.z); r.rcRng = __float_as_uint(b.w);
    r.rcWi = mk3(c.x, c.y, c.z); r.flags = __float_as_uint(c.w);
    r.F = mk3(d.x, d.y, d.z); r.primaryRng = __float_as_uint(d.w);
    r.rcPrevSamplePdf = e.x; r.rcJacobian = e.y; r.sampleCount = e.z; r.resampleWeight = e.w;
    if (kLearned) {
        float4 l = R.p[5 * R.n + i];
        r.Lhat = mk3(l.x, l.y, l.z); r.pRef = l.w;
    } else {
        r.Lhat = s3(0.0f); r.pRef = 0.0f;
    }
    return r;
}
template <bool kLearned = false>
RT_DEV void storeResv(const ResvPlanes& R, size_t i, const Resv& r) {
    R.p[i] = make_float4(r.rcIsec.b1, r.rcIsec.b2, __uint_as_float(r.rcIsec.inst), __uint_as_float(r.rcIsec.prim));
    R.p[R.n + i] = make_float4(r.rcLi.x, r.rcLi.y, r.rcLi.z, __uint_as_float(r.rcRng));
    R.p[2 * R.n + i] = make_float4(r.rcWi.x, r.rcWi.y, r.rcWi.z, __uint_as_float(r.flags));
    R.p[3 * R.n + i] = make_float4(r.F.x, r.F.y, r.F.z, __uint_as_float(r.primaryRng));
    R.p[4 * R.n + i] = make_float4(r.rcPrevSamplePdf, r.rcJacobian, r.sampleCount, r.resampleWeight);
    if (kLearned) R.p[5 * R.n + i] = make_float4(r.Lhat.x, r.Lhat.y, r.Lhat.z, r.pRef);
}
RT_DEV Resv zeroResv() {
    Resv r;
    r.rcIsec.b1 = 0.0f; r.rcIsec.b2 = 0.0f; r.rcIsec.inst = 0u; r.rcIsec.prim = 0u;
    r.rcLi = s3(0.0f); r.rcRng = 0u; r.rcWi = s3(0.0f); r.flags = 0u;
    r.rcPrevSamplePdf = 0.0f; r.rcJacobian = 0.0f; r.F = s3(0.0f); r.primaryRng = 0u;
    r.sampleCount = 0.0f; r.resampleWeight = 0.0f;
    r.Lhat = s3(0.0f); r.pRef = 0.0f;
    return r;
}
// gris_reservoir.glsl:37-136
RT_DEV uint32_t flagsRcVertexId(uint32_t f) { return f & 0xffu; }
RT_DEV uint32_t flagsRcVertexType(uint32_t f) { return (f >> 16) & 0xffu; }
RT_DEV void flagsSetRcVertexId(uint32_t& f, uint32_t id) { f = (f & 0xffffff00u) | (id & 0xffu); }
RT_DEV void flagsSetPathLength(uint32_t& f, uint32_t v) { f = (f & 0xffff00ffu) | ((v & 0xffu) << 8); }
RT_DEV void flagsSetRcVertexType(uint32_t& f, uint32_t t) { f = (f & 0xff00ffffu) | ((t & 0xffu) << 16); }
RT_DEV void resvReset(Resv& r) { r.rcIsec.inst = RT_INVALID_HIT; r.sampleCount = 0.0f; r.resampleWeight = 0.0f; }
RT_DEV bool resvIsValid(const Resv& r) { return !isnan(r.resampleWeight) && r.resampleWeight >= 0.0f; }
RT_DEV void resvMerge(Resv& r, const Resv& rhs, float rnd) {
    r.sampleCount += rhs.sampleCount;
    r.resampleWeight += rhs.resampleWeight;
    if (rnd * r.resampleWeight < rhs.resampleWeight) {
        float sc = r.sampleCount, rw = r.resampleWeight;
        r = rhs;
        r.sampleCount = sc; r.resampleWeight = rw;
    }
}
RT_DEV void resvCap(Resv& r, float cap) {
    if (r.sampleCount > cap) { r.resampleWeight *= cap / r.sampleCount; r.sampleCount = cap; }
}

#define RC_LIGHT_SAMPLED 0u
#define RC_LIGHT_SCATTERED 1u
#define RC_SURFACE 2u
#define GRIS_DIST_THRESHOLD 0.01f

// gris_path_trace.glsl:45-306, run as a persistent kernel with path
// regeneration: a lane whose path ends takes the next pixel of a tile-ordered
// work queue at once, so a wave keeps tracing while any pixel is left. Each
// pixel's arithmetic and random-number stream are those of the reference's
// one-invocation-per-pixel loop; only the assignment of pixels to lanes moves.

// Locals of the reference's bounce loop that live from one vertex to the next.
struct PTPath {
    uint32_t bounce, rng, sampleState, lastSampleState;
    bool isLastVertexConnectible;
    bool picked;  // a stream-RIS candidate has been picked (it is kept in LDS)
    f3 throughput, rcThroughput, lastPos, rdir, rori;  // wo = -rdir
    BSDFSample s;
    Isec isec;
    Resv ps;  // the path sample
    float streamWeight, streamSum;
    f3 dir0;  // direction sampled at the primary vertex (training record of the learned proposal)
};

// The stream-RIS pick of a path: 18 words at p[0], p[stride], ... -- in LDS for
// the persistent kernels (stride = lanes per workgroup), in the path-state
// planes for the wavefront variant. It is written only when a candidate is
// picked and read once at the end of the path, so keeping it out of the
// registers buys occupancy.
constexpr int kStreamWords = 18;
struct StreamSlot { float* p; uint32_t stride; };
RT_DEV void storeStream(StreamSlot slot, const Resv& r) {
    const float v[kStreamWords] = {r.rcIsec.b1, r.rcIsec.b2, __uint_as_float(r.rcIsec.inst), __uint_as_float(r.rcIsec.prim),
                                   r.rcLi.x, r.rcLi.y, r.rcLi.z, __uint_as_float(r.rcRng),
                                   r.rcWi.x, r.rcWi.y, r.rcWi.z, __uint_as_float(r.flags),
                                   r.rcPrevSamplePdf, r.rcJacobian, r.F.x, r.F.y, r.F.z, __uint_as_float(r.primaryRng)};
#pragma unroll
    for (int i = 0; i < kStreamWords; i++) slot.p[i * slot.stride] = v[i];
}
RT_DEV Resv loadStream(StreamSlot slot) {
    float v[kStreamWords];
#pragma unroll
    for (int i = 0; i < kStreamWords; i++) v[i] = slot.p[i * slot.stride];
    Resv r = zeroResv();
    r.rcIsec.b1 = v[0]; r.rcIsec.b2 = v[1]; r.rcIsec.inst = __float_as_uint(v[2]); r.rcIsec.prim = __float_as_uint(v[3]);
    r.rcLi = mk3(v[4], v[5], v[6]); r.rcRng = __float_as_uint(v[7]);
    r.rcWi = mk3(v[8], v[9], v[10]); r.flags = __float_as_uint(v[11]);
    r.rcPrevSamplePdf = v[12]; r.rcJacobian = v[13];
    r.F = mk3(v[14], v[15], v[16]); r.primaryRng = __float_as_uint(v[17]);
    return r;
}
// A destination reservoir whose sample lives in an LDS stream slot (the count
// and weight stay in registers): the same merge and cap as on a Resv.
struct LdsResv { StreamSlot slot; float sampleCount, resampleWeight; };
RT_DEV void resvMerge(LdsResv& r, const Resv& rhs, float rnd) {
    r.sampleCount += rhs.sampleCount;
    r.resampleWeight += rhs.resampleWeight;
    if (rnd * r.resampleWeight < rhs.resampleWeight) storeStream(r.slot, rhs);
}
RT_DEV void resvCap(LdsResv& r, float cap) {
    if (r.sampleCount > cap) { r.resampleWeight *= cap / r.sampleCount; r.sampleCount = cap; }
}

// Start of the invocation (gris_path_trace.glsl:45-100): primary vertex from
// the G-buffer. Returns false for background pixels, whose reservoir is left
// untouched as in the reference.
RT_DEV bool ptBegin(const DevScene& S, const rt_camera& cam, const GB& g, uint32_t x, uint32_t y, PTPath& P, Surf& surf,
                    Mat& mat) {
    f2 uv = pixelUV(x, y, g.W, g.H);
    GPix gp;
    if (!unpackGB(texDNCentre(g, x, y), fetchAM(g, (int)x, (int)y), gp)) return false;
    int matId = gp.matMeshId >> 16;
    CamRay ray = pinholeRay(cam, mk2(uv.x, 1.0f - uv.y));
    P.rng = makeSeedPix(cam.seed, x, y);
    P.bounce = 0;
    P.throughput = s3(1.0f); P.rcThroughput = s3(0.0f);
    P.lastPos = s3(0.0f);
    P.isLastVertexConnectible = false;
    P.sampleState = 0; P.lastSampleState = 0;
    surf.pos = ray.ori + ray.dir * (gp.depth - 1e-4f);
    surf.norm = gp.norm; surf.albedo = gp.albedo; surf.isLight = false; surf.matIndex = (uint32_t)matId;
    mat = loadMat(S, (uint32_t)matId);
    P.s.type = 0; P.s.pdf = 0.0f; P.s.wi = s3(0.0f); P.s.bsdf = s3(0.0f);
    P.isec.b1 = 0.0f; P.isec.b2 = 0.0f; P.isec.inst = 0u; P.isec.prim = 0u;
    P.rdir = ray.dir; P.rori = ray.ori;
    P.ps = zeroResv();  // the path sample (sampleCount/resampleWeight unused here)
    P.ps.rcIsec.inst = RT_INVALID_HIT;
    P.ps.primaryRng = P.rng;
    P.picked = false;
    P.streamWeight = 0.0f; P.streamSum = 0.0f;
    P.dir0 = s3(0.0f);
    return true;
}

// One vertex of the bounce loop, after its closest-hit query. Returns false
// where the reference leaves the loop; otherwise the next ray is in rori/rdir.
template <bool kCount>
RT_DEV bool ptVertex(const DevScene& S, TStack st, StreamSlot slot, const rt_gris_settings& set, PTPath& P, const Surf& surf,
                     const Mat& mat, const uint32_t bounce, TraceCounters* ctr,
                     const rtg::GuidePix& gpx = rtg::noGuide()) {
    uint32_t& rng = P.rng;
    uint32_t& sampleState = P.sampleState;
    uint32_t& lastSampleState = P.lastSampleState;
    bool& isLastVertexConnectible = P.isLastVertexConnectible;
    f3& throughput = P.throughput;
    f3& rcThroughput = P.rcThroughput;
    const f3 wo = -P.rdir;
    f3& lastPos = P.lastPos;
    f3& rdir = P.rdir;
    f3& rori = P.rori;
    BSDFSample& s = P.s;
    const Isec& isec = P.isec;
    Resv& ps = P.ps;
    float& streamWeight = P.streamWeight;
    float& streamSum = P.streamSum;
    {
        flagsSetPathLength(ps.flags, bounce + 1);
        float cosPrevWi = dot(rdir, surf.norm);
        float distToPrev = distance(lastPos, surf.pos);
        float geometryJacobian = fabsf(cosPrevWi) / square(distToPrev);
        bool isThisVertexConnectible = surf.isLight || isBSDFConnectible(mat);
        lastSampleState = sampleState;
        sampleState = (sampleState == 2) ? 2u : (!isThisVertexConnectible ? 0u : sampleState + 1u);
        if (set.shiftType == 0u && bounce == 1 && !surf.isLight) { sampleState = 2; lastSampleState = 1; }
        float resvRandSample = sample1f(rng);

        if (surf.isLight) {
            if (bounce > 1 && cosPrevWi < 0.0f) {
                float w = 1.0f;
                float sumPower = S.lightTable[0].prob;
                float lightPdf = luminance(surf.albedo) / sumPower / geometryJacobian;
                if (!isSampleTypeDelta(s.type)) w = MISWeight(s.pdf, lightPdf);
                f3 weightedLi = surf.albedo * w;
                if (sampleState == 2 && lastSampleState == 2) {
                    ps.rcLi = ps.rcLi + weightedLi * rcThroughput;
                    ps.F = ps.F + weightedLi * throughput;
                } else if ((sampleState == 2 && lastSampleState == 1) && isLastVertexConnectible &&
                           distToPrev > GRIS_DIST_THRESHOLD) {
                    ps.rcIsec = isec;
                    ps.rcRng = rng;
                    ps.rcPrevSamplePdf = s.pdf;
                    ps.rcJacobian = geometryJacobian;
                    // unbiased reuse: the MIS weight depends on the predecessor and is re-evaluated
                    // wherever the sample is reconnected, so the vertex keeps the raw radiance
                    ps.rcLi = set.unbiasedReuse ? surf.albedo : weightedLi;
                    ps.rcWi = s3(0.0f);
                    ps.F = weightedLi * throughput;
                    flagsSetRcVertexId(ps.flags, bounce);
                    flagsSetRcVertexType(ps.flags, RC_LIGHT_SCATTERED);
                    float wgt = luminance(ps.F);
                    streamSum += wgt;
                    if (resvRandSample * streamSum < wgt) { streamWeight = wgt; storeStream(slot, ps); P.picked = true; }
                }
            }
            return false;
        }

        bool connectible = isThisVertexConnectible && isLastVertexConnectible && distToPrev > GRIS_DIST_THRESHOLD;
        if ((sampleState == 2 && lastSampleState == 1) && (connectible || set.shiftType == 0u)) {
            ps.rcIsec = isec;
            ps.rcRng = rng;
            ps.rcPrevSamplePdf = s.pdf;
            // unbiased reuse: the Jacobian with the triangle's geometric normal
            ps.rcJacobian = set.unbiasedReuse ? fabsf(dot(rdir, geomNormal(S, isec))) / square(distToPrev)
                                              : geometryJacobian;
            flagsSetRcVertexId(ps.flags, bounce);
            flagsSetRcVertexType(ps.flags, RC_SURFACE);
            rcThroughput = s3(1.0f);
        }

        f4 lightRand = sample4f(rng);
        resvRandSample = sample1f(rng);

        if (bounce > 0 && !isBSDFDelta(mat)) {
            f3 lightDir; float lightDist, lightPdf, lightJac; f2 lightBary; uint32_t lightId;
            f3 Lr = sampleLightByPower(S, surf.pos, lightDir, lightDist, lightPdf, lightJac, lightBary, lightId, lightRand);
            bool shadowed = traceShadow<kCount>(S, st, surf.pos, RT_MIN_RAY, lightDir, lightDist - RT_MIN_RAY, ctr);
            if (!shadowed && lightPdf > 1e-6f) {
                float bsdfPdf = absDot(surf.norm, lightDir) * RT_PI_INV;
                float w = MISWeight(lightPdf, bsdfPdf);
                f3 scatterTerm = evalBSDF(mat, surf.albedo, surf.norm, wo, lightDir) * satDot(surf.norm, lightDir);
                f3 weightedLi = Lr / lightPdf * w;
                if (sampleState == 2 && lastSampleState == 2) {
                    ps.rcLi = ps.rcLi + weightedLi * scatterTerm * rcThroughput;
                    ps.F = ps.F + weightedLi * scatterTerm * throughput;
                } else if (sampleState == 2 && lastSampleState == 1) {
                    ps.rcLi = weightedLi;
                    ps.rcWi = lightDir;
                    ps.F = weightedLi * scatterTerm * throughput;
                    float wgt = luminance(ps.F);
                    streamSum += wgt;
                    if (resvRandSample * streamSum < wgt) { streamWeight = wgt; storeStream(slot, ps); P.picked = true; }
                } else if (sampleState == 1 && isThisVertexConnectible && lightDist > GRIS_DIST_THRESHOLD) {
                    ps.rcIsec.b1 = lightBary.x; ps.rcIsec.b2 = lightBary.y;
                    ps.rcIsec.inst = 0u; ps.rcIsec.prim = lightId;
                    ps.rcRng = rng;
                    ps.rcPrevSamplePdf = lightPdf;
                    ps.rcJacobian = lightJac;
                    ps.rcLi = set.unbiasedReuse ? Lr : Lr * w;
                    ps.rcWi = s3(0.0f);
                    ps.F = weightedLi * scatterTerm * throughput;
                    flagsSetRcVertexId(ps.flags, bounce + 1);
                    flagsSetRcVertexType(ps.flags, RC_LIGHT_SAMPLED);
                    float wgt = luminance(ps.F);
                    streamSum += wgt;
                    if (resvRandSample * streamSum < wgt) { streamWeight = wgt; storeStream(slot, ps); P.picked = true; }
                }
            }
        }

        if (bounce > 4) {
            float pdfTerminate = gmax(1.0f - luminance(throughput) * set.rrScale, 0.0f);
            if (sample1f(rng) < pdfTerminate) return false;
            throughput = throughput / (1.0f - pdfTerminate);
            rcThroughput = rcThroughput / (1.0f - pdfTerminate);
        }

        bool sampled;
        if (bounce == 0u && gpx.L) {  // learned proposal at the primary vertex
            const f4 gu = sample4f(rng);
            const f3 rb = sample3f(rng);
            sampled = rtg::guidedSampleBSDF(gpx, mat, surf.albedo, surf.norm, wo, gu, rb, s);
            P.dir0 = s.wi;
        } else {
            sampled = sampleBSDF(mat, surf.albedo, surf.norm, wo, sample3f(rng), s);
        }
        if (!sampled || s.pdf < 1e-6f) return false;
        float cosTheta = isSampleTypeDelta(s.type) ? 1.0f : absDot(surf.norm, s.wi);
        f3 scatterTerm = s.bsdf * cosTheta / s.pdf;
        throughput = throughput * scatterTerm;
        if (sampleState == 2 && lastSampleState == 2) {
            rcThroughput = rcThroughput * scatterTerm;
        } else if (sampleState == 2 && lastSampleState == 1) {
            ps.rcLi = s3(0.0f);
            ps.rcWi = s.wi;
            ps.F = s3(0.0f);
            rcThroughput = rcThroughput / s.pdf;
        }
        lastPos = surf.pos;
        rdir = s.wi;
        rori = surf.pos + rdir * 1e-4f;
        isLastVertexConnectible = isThisVertexConnectible;
    }
    return true;
}

// After the bounce loop (gris_path_trace.glsl:280-306): last stream-RIS
// candidate and the reservoir written for the pixel.
RT_DEV Resv ptFinish(PTPath& P, StreamSlot slot) {
    Resv out;
    bool lastPicked = false;
    if (P.sampleState == 2 && P.lastSampleState == 2) {
        float wgt = luminance(P.ps.F);
        float r = sample1f(P.rng);
        P.streamSum += wgt;
        if (r * P.streamSum < wgt) {
            P.streamWeight = wgt;
            out = P.ps;
            lastPicked = true;
        }
    }
    if (!lastPicked) out = P.picked ? loadStream(slot) : zeroResv();
    out.resampleWeight = 0.0f;
    if (P.streamSum > 0.0f && P.streamWeight > 0.0f) {
        float k = P.streamSum / P.streamWeight;
        out.F = out.F * k;
        out.rcLi = out.rcLi * k;
        out.resampleWeight = luminance(out.F);
    } else {
        out.rcIsec.inst = RT_INVALID_HIT;
        out.F = s3(0.0f);
    }
    out.sampleCount = 1.0f;
    return out;
}

// Pixels are handed out in tile order (64-pixel tiles of 2^qShift x 2^(6-qShift)
// pixels, 8x8 for qShift 3), so the lanes refilled together by one wave take
// neighbouring pixels. Every lane leaves the loop once the queue
// is drained; the last wave to finish re-arms the queue for the next launch.
// Drains one view's pixel queue (the body of the persistent kernels): returns
// when every lane of the wave is out of work.
template <bool kCount>
RT_DEV void ptRun(const DevScene& S, const rt_camera& cam, const GB& g, const rt_gris_settings& set, const ResvPlanes& outR,
                  WorkQueue* q, uint32_t qShift, TStack st, float* lds, TraceCounters* ctr,
                  const rtg::GuideFrame& gf) {
    const StreamSlot slot{lds, (uint32_t)kTileThreads};
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tw = 1u << qShift, th = 64u >> qShift;
    const uint32_t tilesX = (g.W + tw - 1) / tw;
    const uint32_t total = tilesX * ((g.H + th - 1) / th) * kTileThreads;
    PTPath P;
    uint32_t pix = 0;
    bool live = false, drained = false;
    // the queue's band this wave claims from (workgroup b starts at band b mod kQueueParts, so
    // the waves of one XCD start in one band) and the bands it has found exhausted
    const uint32_t tiles = total / kTileThreads;
    uint32_t part = blockIdx.x % kQueueParts, exhausted = 0;
    while (true) {
        // lanes without a path take the next pixels; a background pixel or a
        // path that ends at its primary vertex sends the lane back for more
        while (true) {
            const bool want = !live && !drained;
            const unsigned long long mask = __ballot(want);
            if (mask == 0ull) break;
            if (exhausted == kQueueParts) {  // every band is claimed: the lanes are done
                drained = true;
                continue;
            }
            const uint32_t lo = (uint32_t)((uint64_t)tiles * part / kQueueParts) * kTileThreads;
            const uint32_t hi = (uint32_t)((uint64_t)tiles * (part + 1u) / kQueueParts) * kTileThreads;
            const int leader = __ffsll((long long)mask) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(&q->part[part].next, (uint32_t)__popcll(mask));
            base = (uint32_t)__shfl((int)base, leader);
            if (lo + base >= hi) {  // this band is claimed: move to the next one
                part = part + 1u == kQueueParts ? 0u : part + 1u;
                exhausted++;
                continue;
            }
            if (!want) continue;
            const uint32_t w = lo + base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
            if (w >= hi) continue;  // past the band's end: the lane claims again
            const uint32_t tile = w / kTileThreads, t = w % kTileThreads;
            const uint32_t x = (tile % tilesX) * tw + (t & (tw - 1u)), y = (tile / tilesX) * th + (t >> qShift);
            Surf surf; Mat mat;
            if (x >= g.W || y >= g.H || !ptBegin(S, cam, g, x, y, P, surf, mat)) continue;
            pix = y * g.W + x;
            const rtg::GuidePix gpx = rtg::guidePixel(gf.lobes, g.W, g.H, x, y);
            live = ptVertex<kCount>(S, st, slot, set, P, surf, mat, 0u, ctr, gpx) && ++P.bounce < set.maxDepth;
            if (!live) {
                storeResv(outR, pix, ptFinish(P, slot));
                if (gf.samples) gf.samples[pix] = make_float4(P.dir0.x, P.dir0.y, P.dir0.z, P.streamSum);
            }
        }
        if (__ballot(live) == 0ull) break;
        if (live) {
            P.isec = traceClosestHit<kCount>(S, st, P.rori, RT_MIN_RAY, P.rdir, RT_MAX_RAY, ctr);
            live = P.isec.inst != RT_INVALID_HIT;
            if (live) {
                Surf surf;
                loadSurfaceInfo(S, P.isec, surf);
                Mat mat = loadMat(S, surf.matIndex);
                live = ptVertex<kCount>(S, st, slot, set, P, surf, mat, P.bounce, ctr) && ++P.bounce < set.maxDepth;
            }
            if (!live) {
                storeResv(outR, pix, ptFinish(P, slot));
                if (gf.samples) gf.samples[pix] = make_float4(P.dir0.x, P.dir0.y, P.dir0.z, P.streamSum);
            }
        }
    }
}

// Counts this workgroup as finished with queue q; the last one re-arms it.
RT_DEV void queueDone(WorkQueue* q) {
    if ((threadIdx.x & 63u) == 0u) {
        __threadfence();
        if (atomicAdd(&q->done, 1u) == gridDim.x - 1u) {
            for (uint32_t k = 0; k < kQueueParts; k++) atomicExch(&q->part[k].next, 0u);
            atomicExch(&q->done, 0u);
        }
    }
}

// kGuided: the learned proposal's variant (the other variants compile without its code).
template <bool kCount, bool kGuided>
__global__ __launch_bounds__(kTileThreads, RT_WAVES_PATH_TRACE) void k_gris_path_trace(
    DevScene S, rt_camera cam, GB g, rt_gris_settings set, ResvPlanes outR, WorkQueue* q, rtg::GuideFrame gf,
    TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    __shared__ float picks[kStreamWords * kTileThreads];
    ptRun<kCount>(S, cam, g, set, outR, q, 3u, laneStack(stk), picks + threadIdx.x, ctr,
                  kGuided ? gf : rtg::GuideFrame{nullptr, nullptr, nullptr});
    queueDone(q);
}

// ---- gris_retrace.glsl:42-136 --------------------------------------------------------------------
struct RcData { Isec rcPrevIsec; f3 rcPrevWo; f3 rcPrevThroughput; };

// kCanon (unbiased reuse): the shifted path must also be one the destination's
// path tracer classifies the same way -- its reconnection vertex is the first
// vertex whose predecessor is connectible after a run of one connectible vertex
// (the path tracer's sampleState machine) -- so the replayed prefix may not
// hold an earlier pair of consecutive connectible vertices. Without this the
// hybrid shift is not a bijection between the domains.
template <bool kCount, bool kCanon = false>
RT_DEV void replayHybrid(const DevScene& S, TStack st, const rt_gris_settings& set, Isec isec, Surf surf, f3 rdir, f3 rori,
                         uint32_t targetFlags, uint32_t rng, RcData& rc, TraceCounters* ctr, const rtg::GuidePix& gpx) {
    f3 throughput = s3(1.0f);
    f3 wo = -rdir;
    Mat mat = loadMat(S, surf.matIndex);
    BSDFSample s;
    rc.rcPrevThroughput = s3(0.0f);
    rc.rcPrevWo = s3(0.0f);
    rc.rcPrevIsec.b1 = 0.0f; rc.rcPrevIsec.b2 = 0.0f; rc.rcPrevIsec.prim = 0u;
    rc.rcPrevIsec.inst = RT_INVALID_HIT;
    uint32_t targetId = flagsRcVertexId(targetFlags);
    if (targetId == 1) { rc.rcPrevIsec = isec; rc.rcPrevWo = wo; rc.rcPrevThroughput = throughput; return; }
    uint32_t state = 0;  // the path tracer's sampleState over the replayed vertices
    for (uint32_t bounce = 0; bounce < set.maxDepth; bounce++) {
        if (bounce > 0) {
            isec = traceClosestHit<kCount, false>(S, st, rori, RT_MIN_RAY, rdir, RT_MAX_RAY, ctr);
            if (isec.inst == RT_INVALID_HIT) break;
            loadSurfaceInfo(S, isec, surf);
            mat = loadMat(S, surf.matIndex);
        }
        bool conn = isBSDFConnectible(mat);
        if (kCanon) state = state == 2u ? 2u : (conn ? state + 1u : 0u);
        (void)sample1f(rng);
        if (surf.isLight) break;
        if (bounce == targetId - 1u) {
            if (conn && (!kCanon || state == 1u)) {
                rc.rcPrevIsec = isec; rc.rcPrevWo = wo; rc.rcPrevThroughput = throughput;
            }
            break;
        }
        (void)sample4f(rng);
        (void)sample1f(rng);
        if (bounce > 4) {
            float pdfTerminate = gmax(1.0f - luminance(throughput) * set.rrScale, 0.0f);
            if (sample1f(rng) < pdfTerminate) break;
            throughput = throughput / (1.0f - pdfTerminate);
        }
        bool sampled;
        if (bounce == 0u && gpx.L) {  // the destination's learned proposal, same random numbers
            const f4 gu = sample4f(rng);
            const f3 rb = sample3f(rng);
            sampled = rtg::guidedSampleBSDF(gpx, mat, surf.albedo, surf.norm, wo, gu, rb, s);
        } else {
            sampled = sampleBSDF(mat, surf.albedo, surf.norm, wo, sample3f(rng), s);
        }
        if (!sampled || s.pdf < 1e-6f) break;
        float cosTheta = isSampleTypeDelta(s.type) ? 1.0f : absDot(surf.norm, s.wi);
        throughput = throughput * (s.bsdf * cosTheta / s.pdf);
        wo = -s.wi;
        rdir = s.wi;
        rori = surf.pos + rdir * 1e-4f;
    }
}

// Reconnection through the rc vertex: the shading shared by
// GRISReservoirReuseAndMerge (gris_retrace.glsl:188-213) and the final
// shading of spatialReuse (gris_resample_spatial.glsl:114-125). rcLi is the
// sample's reconnection radiance (or, for the learned target, its prediction).
// generated (unbiased reuse): the cosines the path tracer applies when it
// generates the sample -- |cos| for a direction sampled from a metallic BSDF
// (nonzero with both directions below the shading normal), the clamped cosine
// for a light sample and for a Lambert vertex (which samples only its upper
// hemisphere) -- instead of the reference's clamped cosines, so that a shift
// is valid exactly where the domain generates.
RT_DEV f3 reconnectLi(const DevScene& S, const Resv& src, f3 rcLi, const Surf& rcPrevSurf, const Surf& rcSurf, f3 wi,
                      const RcData& rc, bool generated = false) {
    Mat rcMat = loadMat(S, rcSurf.matIndex);
    Mat rcPrevMat = loadMat(S, rcPrevSurf.matIndex);
    f3 Li = rcLi;
    uint32_t rcType = flagsRcVertexType(src.flags);
    if (rcType == RC_SURFACE && length(src.rcWi) > 0.5f) {
        const float c = (generated && rcMat.type != kLambert) ? absDot(rcSurf.norm, src.rcWi)
                                                              : satDot(rcSurf.norm, src.rcWi);
        Li = Li * (evalBSDF(rcMat, rcSurf.albedo, rcSurf.norm, -wi, src.rcWi) * c);
    }
    const float cp = (generated && rcType != RC_LIGHT_SAMPLED && rcPrevMat.type != kLambert) ? absDot(rcPrevSurf.norm, wi)
                                                                                             : satDot(rcPrevSurf.norm, wi);
    Li = Li * (evalBSDF(rcPrevMat, rcPrevSurf.albedo, rcPrevSurf.norm, rc.rcPrevWo, wi) * cp);
    Li = Li * rc.rcPrevThroughput;
    Li = Li / src.rcPrevSamplePdf;
    return Li;
}

// Learned target: the reconnection radiance of a surface rc vertex is replaced
// by the network's prediction at that vertex; light rc vertices keep their
// exact emitted radiance.
RT_DEV f3 targetRcLi(const Resv& r) { return flagsRcVertexType(r.flags) == RC_SURFACE ? r.Lhat : r.rcLi; }

// Unbiased reuse: MIS weight of a light reconnection vertex at the current
// predecessor (the path tracer then stores the vertex's raw radiance): light
// sampling against the cosine pdf (gris_path_trace.glsl:203-204) or the BSDF hit
// against light sampling (:128-133). 1 for a surface vertex.
RT_DEV float rcLightMIS(const DevScene& S, uint32_t rcType, const Surf& rcPrevSurf, f3 rcPrevWo, const Surf& rcSurf,
                        f3 wi, float geomJac) {
    if (rcType == RC_SURFACE) return 1.0f;
    const float lightPdf = luminance(rcSurf.albedo) / S.lightTable[0].prob / geomJac;
    if (rcType == RC_LIGHT_SAMPLED) return MISWeight(lightPdf, absDot(rcPrevSurf.norm, wi) * RT_PI_INV);
    return MISWeight(evalPdf(loadMat(S, rcPrevSurf.matIndex), rcPrevSurf.norm, rcPrevWo, wi), lightPdf);
}

// gris_retrace.glsl:138-236. With kLearned the target function at the sites
// :204 (dstPHat) and :215 (srcPHat) is the learned one; the sample's target
// value in its new pixel travels with it as pRef.
// kUnbiased (the unbiased reuse mode, no reference counterpart): the merged
// sample keeps F consistent with its new reconnection pdf (F is rescaled by
// pdf_src / pdf_dst, so the next shift's target ratio is exact, as it is for a
// fresh sample), a shift to a zero-pdf reconnection is rejected, and the M-cap
// is left to the caller (applied once per pass after the 1/Z normalisation).
template <bool kCount, bool kLearned = false, bool kUnbiased = false, class Dst = Resv>
RT_DEV void reuseAndMerge(const DevScene& S, TStack st, const rt_gris_settings& set, Dst& dst, const Surf& dstPrimarySurf,
                          const Isec& dstPrimaryIsec, f3 primDir, f3 primOri, Resv src, uint32_t& rng, TraceCounters* ctr,
                          const rtg::GuidePix& gpx) {
    RcData rc;
    Surf rcPrevSurf, rcSurf;
    f3 wi = s3(0.0f), Li = s3(0.0f);
    bool shiftable = false;  // every condition of the shift but the visibility of the reconnection
    float dstJacobian = 0.0f, jacobian = 0.0f, dstPHat = 0.0f, dstSamplePdf = 0.0f;
    if (src.rcIsec.inst != RT_INVALID_HIT) {
        replayHybrid<kCount, kUnbiased>(S, st, set, dstPrimaryIsec, dstPrimarySurf, primDir, primOri, src.flags,
                                        src.primaryRng, rc, ctr, gpx);
        if (rc.rcPrevIsec.inst != RT_INVALID_HIT) {
            if (rc.rcPrevIsec.inst == RT_SPECIAL_HIT) rcPrevSurf = dstPrimarySurf;
            else loadSurfaceInfo(S, rc.rcPrevIsec, rcPrevSurf);
            loadSurfaceInfo(S, src.rcIsec, rcSurf);
            Mat rcPrevMat = loadMat(S, rcPrevSurf.matIndex);
            float dist = distance(rcPrevSurf.pos, rcSurf.pos);
            wi = normalize(rcSurf.pos - rcPrevSurf.pos);
            float cosTheta = -dot(rcSurf.norm, wi);
            // unbiased reuse: a surface vertex's Jacobian uses its geometric normal
            const bool geoJ = kUnbiased && flagsRcVertexType(src.flags) == RC_SURFACE;
            dstJacobian = fabsf(geoJ ? dot(geomNormal(S, src.rcIsec), wi) : cosTheta) / square(dist);
            jacobian = dstJacobian / src.rcJacobian;
            // unbiased reuse: a surface rc vertex may face away from its predecessor (interpolated
            // shading normal), as the path tracer generates it; light vertices must face it
            const bool facing = (kUnbiased && flagsRcVertexType(src.flags) == RC_SURFACE) ? cosTheta != 0.0f
                                                                                         : cosTheta > 0.0f;
            shiftable = dist > GRIS_DIST_THRESHOLD && facing && !isnan(jacobian) && src.rcJacobian > 0.0f &&
                        isBSDFConnectible(rcPrevMat);
        }
    }
    // The shifted sample is formed before its reconnection's visibility is traced, so that only
    // it and the ray stay live across the traversal; an occluded or failed shift merges with
    // weight 0 (only its M counts, as a weight-0 candidate is never picked).
    if (shiftable) {
        const f3 visFrom = rcPrevSurf.pos, visTo = rcSurf.pos;
        uint32_t rcType = flagsRcVertexType(src.flags);
        f3 LiL = s3(0.0f);
        if (!isnan(src.rcPrevSamplePdf) && src.rcPrevSamplePdf > 1e-6f) {
            const float misW = kUnbiased ? rcLightMIS(S, rcType, rcPrevSurf, rc.rcPrevWo, rcSurf, wi, dstJacobian) : 1.0f;
            Li = reconnectLi(S, src, kUnbiased ? src.rcLi * misW : src.rcLi, rcPrevSurf, rcSurf, wi, rc, kUnbiased);
            if (kLearned)
                LiL = reconnectLi(S, src, kUnbiased ? targetRcLi(src) * misW : targetRcLi(src), rcPrevSurf, rcSurf, wi, rc,
                                  kUnbiased);
            if (!isBlack(Li) && !hasNan(Li)) dstPHat = luminance((kLearned ? LiL : Li) * jacobian);
            if (rcType == RC_LIGHT_SAMPLED) {
                float sumPower = S.lightTable[0].prob;
                dstSamplePdf = luminance(rcSurf.albedo) / sumPower / dstJacobian;
            } else {
                Mat rcPrevMat = loadMat(S, rcPrevSurf.matIndex);
                // a primary-vertex predecessor samples with the learned proposal when it is on
                dstSamplePdf = rtg::guidedPdf(rc.rcPrevIsec.inst == RT_SPECIAL_HIT ? gpx : rtg::noGuide(), rcPrevMat,
                                              rcPrevSurf.norm, rc.rcPrevWo, wi);
            }
        }
        float srcPHat = kLearned ? src.pRef : luminance(src.F);
        if (kUnbiased) {
            const float r = (dstSamplePdf > 1e-6f) ? src.rcPrevSamplePdf / dstSamplePdf : 0.0f;
            if (r == 0.0f) dstPHat = 0.0f;
            Li = Li * r;
            LiL = LiL * r;
        }
        if (kLearned) src.pRef = luminance(LiL);
        src.rcJacobian = dstJacobian;
        src.rcPrevSamplePdf = dstSamplePdf;
        src.F = Li;
        if (src.rcPrevSamplePdf < 1e-6f || isnan(src.rcPrevSamplePdf)) src.rcPrevSamplePdf = 0.0f;
        src.resampleWeight *= dstPHat / srcPHat;
        if (!traceVisibility<kCount>(S, st, visFrom, visTo, ctr)) src.resampleWeight = 0.0f;
    } else {
        src.resampleWeight = 0.0f;
    }
    if (resvIsValid(src)) resvMerge(dst, src, sample1f(rng));
    if (!kUnbiased) resvCap(dst, (float)set.cap);
}

// Unbiased reuse: can the domain whose primary vertex is (surf, isec, ray) have
// produced the picked sample y? True when the hybrid shift of y onto that
// vertex is valid under the conditions of reuseAndMerge (replay to the
// reconnection predecessor, connectible, distance threshold, front-facing,
// visible) and the shifted sample's own contribution there -- with its pdf at
// that predecessor -- has a positive target value.
template <bool kCount, bool kLearned>
RT_DEV bool shiftSupports(const DevScene& S, TStack st, const rt_gris_settings& set, const Resv& y, const Surf& primSurf,
                          const Isec& primIsec, f3 primDir, f3 primOri, TraceCounters* ctr, const rtg::GuidePix& gpx) {
    if (y.rcIsec.inst == RT_INVALID_HIT) return false;
    RcData rc;
    replayHybrid<kCount, true>(S, st, set, primIsec, primSurf, primDir, primOri, y.flags, y.primaryRng, rc, ctr, gpx);
    if (rc.rcPrevIsec.inst == RT_INVALID_HIT) return false;
    Surf rcPrevSurf, rcSurf;
    if (rc.rcPrevIsec.inst == RT_SPECIAL_HIT) rcPrevSurf = primSurf;
    else loadSurfaceInfo(S, rc.rcPrevIsec, rcPrevSurf);
    loadSurfaceInfo(S, y.rcIsec, rcSurf);
    const Mat rcPrevMat = loadMat(S, rcPrevSurf.matIndex);
    const float dist = distance(rcPrevSurf.pos, rcSurf.pos);
    const f3 wi = normalize(rcSurf.pos - rcPrevSurf.pos);
    const float cosTheta = -dot(rcSurf.norm, wi);
    const float jac =
        fabsf(flagsRcVertexType(y.flags) == RC_SURFACE ? dot(geomNormal(S, y.rcIsec), wi) : cosTheta) / square(dist);
    const bool facing = flagsRcVertexType(y.flags) == RC_SURFACE ? cosTheta != 0.0f : cosTheta > 0.0f;
    if (!(dist > GRIS_DIST_THRESHOLD && facing && jac > 0.0f && !isnan(jac) && isBSDFConnectible(rcPrevMat)))
        return false;
    Resv x = y;
    x.rcPrevSamplePdf = flagsRcVertexType(y.flags) == RC_LIGHT_SAMPLED
                            ? luminance(rcSurf.albedo) / S.lightTable[0].prob / jac
                            : rtg::guidedPdf(rc.rcPrevIsec.inst == RT_SPECIAL_HIT ? gpx : rtg::noGuide(), rcPrevMat,
                                             rcPrevSurf.norm, rc.rcPrevWo, wi);
    if (isnan(x.rcPrevSamplePdf) || !(x.rcPrevSamplePdf > 1e-6f)) return false;
    const float misW = rcLightMIS(S, flagsRcVertexType(y.flags), rcPrevSurf, rc.rcPrevWo, rcSurf, wi, jac);
    const f3 Li = reconnectLi(S, x, x.rcLi * misW, rcPrevSurf, rcSurf, wi, rc, true);
    // x is a candidate of that domain iff its resampling weight luminance(F) is positive
    if (hasNan(Li) || !(luminance(Li) > 0.0f)) return false;
    if (kLearned && !(luminance(reconnectLi(S, x, targetRcLi(x) * misW, rcPrevSurf, rcSurf, wi, rc, true)) > 0.0f))
        return false;
    // the visibility of the reconnection last: a shift that fails the other conditions traces no ray
    return traceVisibility<kCount>(S, st, rcPrevSurf.pos, rcSurf.pos, ctr);
}

// Primary vertex of pixel (x, y) as its own passes rebuilt it (texel-centre
// G-buffer, camera ray through the pixel centre); false for background.
RT_DEV bool primaryOf(const rt_camera& cam, const GB& g, int x, int y, Surf& surf, Isec& isec, CamRay& ray) {
    if (x < 0 || y < 0 || x >= (int)g.W || y >= (int)g.H) return false;
    GPix gp;
    if (!unpackGB(texDNCentre(g, (uint32_t)x, (uint32_t)y), fetchAM(g, x, y), gp)) return false;
    const f2 uv = pixelUV((uint32_t)x, (uint32_t)y, g.W, g.H);
    ray = pinholeRay(cam, mk2(uv.x, 1.0f - uv.y));
    surf.pos = ray.ori + ray.dir * (gp.depth - 1e-4f);
    surf.norm = gp.norm; surf.albedo = gp.albedo; surf.matIndex = (uint32_t)(gp.matMeshId >> 16); surf.isLight = false;
    isec.inst = RT_SPECIAL_HIT; isec.b1 = uv.x; isec.b2 = uv.y; isec.prim = 0u;
    return true;
}

// 1/Z normalisation of a merged reservoir: resampleWeight * M / Z, so that the
// final weight resampleWeight / M equals the unnormalised sum over Z.
RT_DEV void normaliseZ(Resv& r, float z) {
    if (z > 0.0f) r.resampleWeight = r.resampleWeight * (r.sampleCount / z);
}

template <bool kLearned>
RT_DEV Resv resvAt(const ResvPlanes& R, uint32_t W, uint32_t H, int x, int y) {
    unsigned long long idx = (unsigned long long)W * (uint32_t)y + (uint32_t)x;
    if (x < 0 || y < 0 || idx >= (unsigned long long)W * H) return zeroResv();
    return loadResv<kLearned>(R, (size_t)idx);
}

// gris_resample_temporal.glsl:11-83
template <bool kCount, bool kLearned, bool kUnbiased = false>
RT_DEV void temporalPixel(const DevScene& S, const rt_camera& cam, const rt_camera& prevCam, const GB& g, const GB& gPrev,
                          const rt_gris_settings& set, const ResvPlanes& thisR, const ResvPlanes& prevR,
                          const ResvPlanes& tempR, uint32_t x, uint32_t y, TStack st, TraceCounters* ctr,
                          const rtg::GuideFrame& gf) {
    if (x >= g.W || y >= g.H) return;
    f2 uv = pixelUV(x, y, g.W, g.H);
    GPix gp;
    if (!unpackGB(texDNCentre(g, x, y), fetchAM(g, (int)x, (int)y), gp)) return;  // Temp not written
    f2 motion = fetchMotion(g, (int)x, (int)y);
    int matId = gp.matMeshId >> 16;
    CamRay ray = pinholeRay(cam, mk2(uv.x, 1.0f - uv.y));
    uint32_t rng = makeSeedPix(cam.seed, x, y) ^ 1u;
    uint32_t resvRng = ~rng;
    f3 pos = ray.ori + ray.dir * (gp.depth - 1e-4f);
    size_t p = (size_t)y * g.W + x;
    Resv resv = loadResv<kLearned>(thisR, p);
    Isec dstPrimaryIsec; dstPrimaryIsec.inst = RT_SPECIAL_HIT; dstPrimaryIsec.b1 = uv.x; dstPrimaryIsec.b2 = uv.y;
    dstPrimaryIsec.prim = 0u;
    const float mCanon = resv.sampleCount;
    float mPrev = 0.0f;
    int prevX = -1, prevY = -1;
    if (set.temporalReuse) {
        Surf dstPrimarySurf; dstPrimarySurf.pos = pos; dstPrimarySurf.norm = gp.norm; dstPrimarySurf.albedo = gp.albedo;
        dstPrimarySurf.matIndex = (uint32_t)matId; dstPrimarySurf.isLight = false;
        if ((cam.frameIndex & 0x80000000u) == 0u) {
            // findPreviousReservoir (gris_resample_temporal.glsl:11-34)
            f2 puv = uv + motion;
            // a non-finite reprojection (half overflow of the motion vector) is rejected
            if (!(puv.x < 0.0f || puv.y < 0.0f || puv.x > 1.0f || puv.y > 1.0f || isnan(puv.x) || isnan(puv.y))) {
                int px = (int)(puv.x * (float)cam.filmSize[0]);
                int py = (int)(puv.y * (float)cam.filmSize[1]);
                GPix pp;
                if (unpackGB(texDN(gPrev, puv), fetchAM(gPrev, px, py), pp)) {
                    CamRay pr = pinholeRay(prevCam, mk2(puv.x, 1.0f - puv.y));
                    f3 posPrev = pr.ori + pr.dir * (pp.depth - 1e-4f);
                    if (!(pp.matMeshId != gp.matMeshId || dot(pp.norm, gp.norm) < 0.95f || distance(pos, posPrev) > 0.5f)) {
                        Resv tr = resvAt<kLearned>(prevR, cam.filmSize[0], cam.filmSize[1], px, py);
                        if (resvIsValid(tr)) {
                            reuseAndMerge<kCount, kLearned, kUnbiased>(
                                S, st, set, resv, dstPrimarySurf, dstPrimaryIsec, ray.dir, ray.ori, tr, resvRng, ctr,
                                rtg::guidePixel(gf.lobes, g.W, g.H, x, y));
                            mPrev = tr.sampleCount; prevX = px; prevY = py;
                        }
                    }
                }
            }
        }
    }
    if (kUnbiased) {
        if (mPrev > 0.0f && resvIsValid(resv) && resv.rcIsec.inst != RT_INVALID_HIT) {
            float z = mCanon;
            Surf ps; Isec pi; CamRay pr;
            const rtg::GuidePix prevGp = rtg::guidePixel(gf.lobesPrev, g.W, g.H, (uint32_t)prevX, (uint32_t)prevY);
            if (primaryOf(prevCam, gPrev, prevX, prevY, ps, pi, pr) &&
                shiftSupports<kCount, kLearned>(S, st, set, resv, ps, pi, pr.dir, pr.ori, ctr, prevGp))
                z += mPrev;
            normaliseZ(resv, z);
        }
        resvCap(resv, (float)set.cap);
    }
    if (!resvIsValid(resv)) resvReset(resv);
    storeResv<kLearned>(tempR, p, resv);
}
template <bool kCount, bool kLearned, bool kUnbiased, bool kGuided>
__global__ __launch_bounds__(kTileThreads, kUnbiased ? RT_WAVES_TEMPORAL_UB : RT_WAVES_TEMPORAL) void k_gris_temporal(DevScene S, rt_camera cam, rt_camera prevCam, GB g,
                                                                GB gPrev, rt_gris_settings set, ResvPlanes thisR,
                                                                ResvPlanes prevR, ResvPlanes tempR, rtg::GuideFrame gf,
                                                                TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    PIXEL_INDEX
    temporalPixel<kCount, kLearned, kUnbiased>(S, cam, prevCam, g, gPrev, set, thisR, prevR, tempR, x, y,
                                    laneStack(stk), ctr,
                                    kGuided ? gf : rtg::GuideFrame{nullptr, nullptr, nullptr});
}

// gris_resample_spatial.glsl:11-135 and gris_resample_spatial.comp accumulation
template <bool kCount, bool kLearned, bool kUnbiased = false>
RT_DEV void spatialPixel(const DevScene& S, const rt_camera& cam, const GB& g, const rt_gris_settings& set,
                         const ResvPlanes& tempR, const ResvPlanes& thisR, float* indirectOut, uint32_t x, uint32_t y,
                         TStack st, TraceCounters* ctr, const rtg::GuideFrame& gf, StreamSlot dstSlot) {
    // the luminance target with the reference's weights keeps the merged sample in LDS between
    // the neighbours (the forms with more per-sample state keep it in registers)
    constexpr bool kDstInLds = !kLearned && !kUnbiased;
    if (x >= g.W || y >= g.H) return;
    size_t p = (size_t)y * g.W + x;
    const rtg::GuidePix gpx = rtg::guidePixel(gf.lobes, g.W, g.H, x, y);
    f2 uv = pixelUV(x, y, g.W, g.H);
    GPix gp;
    f3 radiance = s3(0.0f);
    if (unpackGB(texDNCentre(g, x, y), fetchAM(g, (int)x, (int)y), gp)) {
        uint32_t rng = makeSeedPix(cam.seed, x, y) ^ 2u;
        CamRay ray = pinholeRay(cam, mk2(uv.x, 1.0f - uv.y));
        f3 pos = ray.ori + ray.dir * (gp.depth - 1e-4f);
        int matId = gp.matMeshId >> 16;
        Resv resv = loadResv<kLearned>(tempR, p);
        Isec dstPrimaryIsec; dstPrimaryIsec.inst = RT_SPECIAL_HIT; dstPrimaryIsec.b1 = uv.x; dstPrimaryIsec.b2 = uv.y;
        dstPrimaryIsec.prim = 0u;
        Surf dstPrimarySurf; dstPrimarySurf.pos = pos; dstPrimarySurf.norm = gp.norm; dstPrimarySurf.albedo = gp.albedo;
        dstPrimarySurf.matIndex = (uint32_t)matId; dstPrimarySurf.isLight = false;
        const float mCanon = resv.sampleCount;
        // unbiased mode: the merged neighbours (pixel, M), at most 8 (the settings check)
        uint32_t nbPix[kUnbiased ? 8 : 1];
        float nbM[kUnbiased ? 8 : 1];
        uint32_t nbCount = 0;
        LdsResv dstL{dstSlot, resv.sampleCount, resv.resampleWeight};
        if constexpr (kDstInLds) storeStream(dstSlot, resv);
        if (set.spatialReuse) {
            f2 texel = mk2(1.0f / (float)g.W, 1.0f / (float)g.H);
            for (uint32_t i = 0; i < set.spatialNeighbors; i++) {
                f2 dsk = toConcentricDisk(sample2f(rng));
                f2 nb = uv + (dsk * set.spatialRadius) * texel;
                // findNeighborReservoir (gris_resample_spatial.glsl:11-34)
                if (nb.x < 0.0f || nb.y < 0.0f || nb.x > 1.0f || nb.y > 1.0f) continue;
                int px = (int)(nb.x * (float)cam.filmSize[0]);
                int py = (int)(nb.y * (float)cam.filmSize[1]);
                GPix np;
                if (!unpackGB(texDN(g, nb), fetchAM(g, px, py), np)) continue;
                CamRay nr = pinholeRay(cam, mk2(nb.x, 1.0f - nb.y));
                f3 posN = nr.ori + nr.dir * (np.depth - 1e-4f);
                if (dot(np.norm, gp.norm) < 0.9f || distance(pos, posN) > 0.4f) continue;
                Resv nres = resvAt<kLearned>(tempR, cam.filmSize[0], cam.filmSize[1], px, py);
                if (resvIsValid(nres)) {
                    if constexpr (kDstInLds)
                        reuseAndMerge<kCount, kLearned, kUnbiased>(S, st, set, dstL, dstPrimarySurf, dstPrimaryIsec,
                                                                    ray.dir, ray.ori, nres, rng, ctr, gpx);
                    else
                        reuseAndMerge<kCount, kLearned, kUnbiased>(S, st, set, resv, dstPrimarySurf, dstPrimaryIsec,
                                                                    ray.dir, ray.ori, nres, rng, ctr, gpx);
                    if (kUnbiased && nbCount < 8u) {
                        nbPix[nbCount] = ((uint32_t)px << 16) | (uint32_t)py;
                        nbM[nbCount] = nres.sampleCount;
                        nbCount++;
                    }
                }
            }
        }
        if constexpr (kDstInLds) {
            resv = loadStream(dstSlot);
            resv.sampleCount = dstL.sampleCount;
            resv.resampleWeight = dstL.resampleWeight;
        }
        if (kUnbiased) {
            if (nbCount > 0u && resvIsValid(resv) && resv.rcIsec.inst != RT_INVALID_HIT) {
                float z = mCanon;
                for (uint32_t j = 0; j < nbCount; j++) {
                    if (!(nbM[j] > 0.0f)) continue;
                    Surf ns; Isec ni; CamRay nr;
                    const uint32_t nx = nbPix[j] >> 16, ny = nbPix[j] & 0xffffu;
                    const rtg::GuidePix ngp = rtg::guidePixel(gf.lobes, g.W, g.H, nx, ny);
                    if (primaryOf(cam, g, (int)nx, (int)ny, ns, ni, nr) &&
                        shiftSupports<kCount, kLearned>(S, st, set, resv, ns, ni, nr.dir, nr.ori, ctr, ngp))
                        z += nbM[j];
                }
                normaliseZ(resv, z);
            }
            resvCap(resv, (float)set.cap);
        }
        if (!resvIsValid(resv)) resvReset(resv);
        storeResv<kLearned>(thisR, p, resv);
        if (resvIsValid(resv) && resv.sampleCount > 0.0f && resv.rcIsec.inst != RT_INVALID_HIT) {
            RcData rc;
            replayHybrid<kCount>(S, st, set, dstPrimaryIsec, dstPrimarySurf, ray.dir, ray.ori, resv.flags, resv.primaryRng,
                                 rc, ctr, gpx);
            if (rc.rcPrevIsec.inst != RT_INVALID_HIT) {
                Surf rcPrevSurf, rcSurf;
                if (rc.rcPrevIsec.inst == RT_SPECIAL_HIT) rcPrevSurf = dstPrimarySurf;
                else loadSurfaceInfo(S, rc.rcPrevIsec, rcPrevSurf);
                loadSurfaceInfo(S, resv.rcIsec, rcSurf);
                f3 wi = normalize(rcSurf.pos - rcPrevSurf.pos);
                float misW = 1.0f;
                if (kUnbiased) {
                    const float gj = fabsf(dot(rcSurf.norm, wi)) / square(distance(rcPrevSurf.pos, rcSurf.pos));
                    misW = rcLightMIS(S, flagsRcVertexType(resv.flags), rcPrevSurf, rc.rcPrevWo, rcSurf, wi, gj);
                }
                f3 Li = reconnectLi(S, resv, kUnbiased ? resv.rcLi * misW : resv.rcLi, rcPrevSurf, rcSurf, wi, rc,
                                    kUnbiased);
                float pHat = kLearned ? luminance(reconnectLi(S, resv, kUnbiased ? targetRcLi(resv) * misW : targetRcLi(resv),
                                                              rcPrevSurf, rcSurf, wi, rc, kUnbiased))
                                      : luminance(Li);
                if (!isBlack(Li) && !hasNan(Li)) radiance = Li / pHat * resv.resampleWeight / resv.sampleCount;
            }
        }
        radiance = clampColor(radiance);
    }
    accumulate(indirectOut, p, radiance, cam.frameIndex);
}
template <bool kCount, bool kLearned, bool kUnbiased, bool kGuided>
__global__ __launch_bounds__(kTileThreads, kUnbiased ? RT_WAVES_SPATIAL_UB : RT_WAVES_SPATIAL) void k_gris_spatial(DevScene S, rt_camera cam, GB g, rt_gris_settings set,
                                                               ResvPlanes tempR, ResvPlanes thisR, float* indirectOut,
                                                               rtg::GuideFrame gf, TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    __shared__ float dstWords[(!kLearned && !kUnbiased) ? kStreamWords * kTileThreads : 1];
    PIXEL_INDEX
    spatialPixel<kCount, kLearned, kUnbiased>(S, cam, g, set, tempR, thisR, indirectOut, x, y,
                                              laneStack(stk), ctr,
                                              kGuided ? gf : rtg::GuideFrame{nullptr, nullptr, nullptr},
                                              StreamSlot{dstWords + threadIdx.x, (uint32_t)kTileThreads});
}

// ---- Batched views ---------------------------------------------------------------------------
// Everything one view's passes read, in one descriptor; the batch kernels take
// a device array of them. The view of a workgroup is made wave-uniform, so its
// descriptor fields are read with scalar loads.
struct ViewDesc {
    DevScene S;
    rt_camera cam, prevCam;
    GB g, gPrev;
    ResvPlanes cur, prev, temp;
    float* indirectOut;
    float4* dnOut;
    uint2* amOut;
    uint32_t* mvOut;
    rt_gris_settings set;
    uint32_t tileBase, tilesX, tiles;  // tileBase: first tile of the view in its launch
    uint32_t queueShift;               // path-trace queue tiles of 2^queueShift x 2^(6-queueShift) pixels
    uint32_t pixBase, pad[3];          // wavefront variant: first path (pixel) of the view in the batch
    WorkQueue* queue;
    rtg::GuideFrame gf;  // the learned proposal's bindings (null unless the view samples with it)
};

// View of a linear tile index: the last view whose tileBase <= tile.
RT_DEV uint32_t viewOfTile(const ViewDesc* __restrict__ V, uint32_t numViews, uint32_t tile) {
    uint32_t lo = 0, hi = numViews;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (V[mid].tileBase <= tile) lo = mid;
        else hi = mid;
    }
    return lo;
}
// Workgroup -> tile. Consecutive workgroups are dealt round robin to the 8
// XCDs; with a chunk of C tiles, XCD k takes runs of C consecutive tiles
// (chunks k, k + 8, k + 16, ...), so neighbouring tiles share that XCD's L2
// while every XCD still gets work from all over the views. C = 0: dispatch order.
RT_DEV uint32_t batchTile(uint32_t b, uint32_t totalTiles, uint32_t chunk) {
    if (!chunk) return b;
    const uint32_t xcd = b & 7u, i = b >> 3;
    return ((i / chunk) * 8u + xcd) * chunk + i % chunk;
}
RT_DEV bool batchPixel(const ViewDesc* __restrict__ V, uint32_t numViews, uint32_t totalTiles, uint32_t swz,
                       uint32_t& v, uint32_t& x, uint32_t& y) {
    const uint32_t tile = batchTile(blockIdx.x, totalTiles, swz);  // swz: tiles per XCD chunk
    if (tile >= totalTiles) return false;
    v = __builtin_amdgcn_readfirstlane(viewOfTile(V, numViews, tile));
    const uint32_t t = tile - V[v].tileBase, tx = V[v].tilesX;
    x = (t % tx) * kTileX + (threadIdx.x & (kTileX - 1));
    y = (t / tx) * kTileY + threadIdx.x / kTileX;
    return true;
}

template <bool kCount>
__global__ __launch_bounds__(kTileThreads, RT_WAVES_GBUFFER) void k_gbuffer_batch(const ViewDesc* __restrict__ V,
                                                                                  uint32_t numViews, uint32_t totalTiles,
                                                                                  uint32_t swz, TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    uint32_t v, x, y;
    if (!batchPixel(V, numViews, totalTiles, swz, v, x, y)) return;
    const ViewDesc& D = V[v];
    gbufferPixel<kCount>(D.S, D.cam, D.dnOut, D.amOut, D.mvOut, x, y, laneStack(stk), ctr);
}

// All views' GRIS path tracing in one persistent launch: workgroup b starts at
// view b % numViews and drains the views' pixel queues in turn, so waves move
// to the views that still have pixels. kGuided: some views sample with the
// learned proposal (a view whose gf is null samples with its BSDFs). kWaves:
// the register budget, in waves per SIMD (a tuning choice of the batch).
template <bool kCount, bool kGuided, int kWaves = RT_WAVES_PATH_TRACE>
__global__ __launch_bounds__(kTileThreads, kWaves) void k_gris_path_trace_batch(const ViewDesc* __restrict__ V,
                                                                               uint32_t numViews, TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    __shared__ float picks[kStreamWords * kTileThreads];
    for (uint32_t i = 0; i < numViews; i++) {
        const uint32_t v = __builtin_amdgcn_readfirstlane((blockIdx.x + i) % numViews);
        const ViewDesc& D = V[v];
        ptRun<kCount>(D.S, D.cam, D.g, D.set, D.cur, D.queue, D.queueShift, laneStack(stk),
                      picks + threadIdx.x, ctr, kGuided ? D.gf : rtg::GuideFrame{nullptr, nullptr, nullptr});
        queueDone(D.queue);
    }
}

// The reuse passes of one launch run one form (target function, resampling
// weights, proposal) for all of its views.
template <bool kCount, bool kLearned, bool kUnbiased, bool kGuided>
__global__ __launch_bounds__(kTileThreads, kUnbiased ? RT_WAVES_TEMPORAL_UB : RT_WAVES_TEMPORAL) void k_gris_temporal_batch(const ViewDesc* __restrict__ V,
                                                                                        uint32_t numViews,
                                                                                        uint32_t totalTiles, uint32_t swz,
                                                                                        TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    uint32_t v, x, y;
    if (!batchPixel(V, numViews, totalTiles, swz, v, x, y)) return;
    const ViewDesc& D = V[v];
    temporalPixel<kCount, kLearned, kUnbiased>(D.S, D.cam, D.prevCam, D.g, D.gPrev, D.set, D.cur, D.prev, D.temp, x, y,
                                               laneStack(stk), ctr,
                                               kGuided ? D.gf : rtg::GuideFrame{nullptr, nullptr, nullptr});
}

template <bool kCount, bool kLearned, bool kUnbiased, bool kGuided>
__global__ __launch_bounds__(kTileThreads, kUnbiased ? RT_WAVES_SPATIAL_UB : RT_WAVES_SPATIAL) void k_gris_spatial_batch(const ViewDesc* __restrict__ V,
                                                                                      uint32_t numViews, uint32_t totalTiles,
                                                                                      uint32_t swz, TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    __shared__ float dstWords[(!kLearned && !kUnbiased) ? kStreamWords * kTileThreads : 1];
    uint32_t v, x, y;
    if (!batchPixel(V, numViews, totalTiles, swz, v, x, y)) return;
    const ViewDesc& D = V[v];
    spatialPixel<kCount, kLearned, kUnbiased>(D.S, D.cam, D.g, D.set, D.temp, D.cur, D.indirectOut, x, y,
                                              laneStack(stk), ctr,
                                              kGuided ? D.gf : rtg::GuideFrame{nullptr, nullptr, nullptr},
                                              StreamSlot{dstWords + threadIdx.x, (uint32_t)kTileThreads});
}

// ---- Wavefront variant of the batched path trace: material-sorted shading ----
// The persistent kernels keep each path in one lane from its primary vertex to
// its end. This variant instead keeps every path's state in HBM (kWfWords
// planes of one word per path; path = pixel of the batch) and runs each bounce
// as four launches over a device-side list of live paths:
//   extend  -- closest hit of each live path's next ray; its shading bucket
//              (miss, light, or the hit material's type) and per-bucket counts,
//              aggregated per wave with ballots;
//   scan    -- bucket offsets (one wave), and the next list's length cleared;
//   scatter -- the live paths reordered by bucket (wave-aggregated cursors);
//   shade   -- the vertex in bucket order, so the lanes of a wave run the same
//              BSDF code, with its light-sample shadow ray; a path that goes
//              on is appended to the next list (ballot + one atomic per wave),
//              a finished one writes its reservoir.
// Every path computes exactly what the persistent kernel computes for it (the
// same ptBegin / ptVertex / ptFinish on its own random numbers), so the
// reservoirs are identical; only the order of work differs.
constexpr uint32_t kWfWords = 60;
constexpr uint32_t kWfBuckets = 8;  // 0 miss, 1 light, 2..5 material types 1..4, 6 other types
constexpr uint32_t kWfSlot = 42;    // first plane of the stream-RIS pick

struct WfState {
    float* w;            // kWfWords x n words
    uint32_t* lists;     // 2 x n: live paths of this bounce and of the next
    uint32_t* sorted;    // n: this bounce's live paths in bucket order
    uint32_t* keys;      // n: bucket of lists[cur][j]
    uint32_t* ctr;       // [0], [1] list lengths; [2, 2 + K) bucket counts; [2 + K, 2 + 2K) cursors
    uint32_t n;
};

RT_DEV StreamSlot wfSlot(const WfState& B, uint32_t i) { return StreamSlot{B.w + (size_t)kWfSlot * B.n + i, B.n}; }

RT_DEV void wfStore(const WfState& B, uint32_t i, uint32_t view, uint32_t pix, const PTPath& P) {
    float* w = B.w + i;
    const size_t n = B.n;
    auto put = [&](uint32_t k, float v) { w[k * n] = v; };
    auto putU = [&](uint32_t k, uint32_t v) { w[k * n] = __uint_as_float(v); };
    putU(0, view); putU(1, pix);
    putU(2, P.bounce | (P.sampleState << 8) | (P.lastSampleState << 10) | ((P.isLastVertexConnectible ? 1u : 0u) << 12) |
                ((P.picked ? 1u : 0u) << 13));
    putU(3, P.rng);
    put(4, P.throughput.x); put(5, P.throughput.y); put(6, P.throughput.z);
    put(7, P.rcThroughput.x); put(8, P.rcThroughput.y); put(9, P.rcThroughput.z);
    put(10, P.lastPos.x); put(11, P.lastPos.y); put(12, P.lastPos.z);
    put(13, P.rdir.x); put(14, P.rdir.y); put(15, P.rdir.z);
    put(16, P.s.pdf); putU(17, P.s.type);
    storeStream(StreamSlot{w + 22 * n, (uint32_t)n}, P.ps);
    put(40, P.streamWeight); put(41, P.streamSum);
}

// The path after its last closest-hit query: the state and that query's hit.
RT_DEV void wfLoad(const WfState& B, uint32_t i, uint32_t& view, uint32_t& pix, PTPath& P) {
    const float* w = B.w + i;
    const size_t n = B.n;
    auto get = [&](uint32_t k) { return w[k * n]; };
    auto getU = [&](uint32_t k) { return __float_as_uint(w[k * n]); };
    view = getU(0); pix = getU(1);
    const uint32_t f = getU(2);
    P.bounce = f & 0xffu; P.sampleState = (f >> 8) & 3u; P.lastSampleState = (f >> 10) & 3u;
    P.isLastVertexConnectible = ((f >> 12) & 1u) != 0u; P.picked = ((f >> 13) & 1u) != 0u;
    P.rng = getU(3);
    P.throughput = mk3(get(4), get(5), get(6));
    P.rcThroughput = mk3(get(7), get(8), get(9));
    P.lastPos = mk3(get(10), get(11), get(12));
    P.rdir = mk3(get(13), get(14), get(15));
    P.rori = P.lastPos + P.rdir * 1e-4f;  // as ptVertex set it
    P.s.pdf = get(16); P.s.type = getU(17); P.s.wi = s3(0.0f); P.s.bsdf = s3(0.0f);
    P.isec.b1 = get(18); P.isec.b2 = get(19); P.isec.inst = getU(20); P.isec.prim = getU(21);
    P.ps = loadStream(StreamSlot{const_cast<float*>(w) + 22 * n, (uint32_t)n});
    P.streamWeight = get(40); P.streamSum = get(41);
    P.dir0 = s3(0.0f);
}

// View of a batch pixel: the last view whose pixBase <= i.
RT_DEV uint32_t viewOfPixel(const ViewDesc* __restrict__ V, uint32_t numViews, uint32_t i) {
    uint32_t lo = 0, hi = numViews;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (V[mid].pixBase <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Appends the lanes with `take` to list L (length at *len): one atomic per wave.
RT_DEV void wfAppend(uint32_t* L, uint32_t* len, bool take, uint32_t value) {
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long m = __ballot(take);
    if (m == 0ull) return;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(len, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, leader);
    if (take) L[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = value;
}

__global__ void k_wf_clear(WfState B) {
    if (threadIdx.x < 2u + 2u * kWfBuckets) B.ctr[threadIdx.x] = 0u;
}

// Primary vertex of every pixel of the batch (bounce 0: no closest-hit query).
__global__ __launch_bounds__(kTileThreads, RT_WAVES_PATH_TRACE) void k_wf_begin(const ViewDesc* __restrict__ V,
                                                                               uint32_t numViews, uint32_t total,
                                                                               WfState B) {
    __shared__ StackLDS<kTileThreads> stk;
    const TStack st = laneStack(stk);
    for (uint32_t base = blockIdx.x * kTileThreads; base < total; base += gridDim.x * kTileThreads) {
        const uint32_t i = base + threadIdx.x;
        bool live = false;
        if (i < total) {
            const uint32_t v = viewOfPixel(V, numViews, i);
            const ViewDesc& D = V[v];
            const uint32_t p = i - D.pixBase, x = p % D.g.W, y = p / D.g.W;
            PTPath P;
            Surf surf; Mat mat;
            if (ptBegin(D.S, D.cam, D.g, x, y, P, surf, mat)) {
                live = ptVertex<false>(D.S, st, wfSlot(B, i), D.set, P, surf, mat, 0u, nullptr) &&
                       ++P.bounce < D.set.maxDepth;
                if (live) wfStore(B, i, v, p, P);
                else storeResv(D.cur, p, ptFinish(P, wfSlot(B, i)));
            }
        }
        wfAppend(B.lists, B.ctr, live, i);
    }
}

// Shading bucket of a hit: 0 miss, 1 light, 2 + (material type - 1) for types
// 1..4, 6 for any other type.
RT_DEV uint32_t wfBucket(const DevScene& S, const Isec& h) {
    if (h.inst == RT_INVALID_HIT) return 0u;
    if (h.inst == 0u) return 1u;
    const DevInstance& I = S.instances[h.inst - 1];
    const uint32_t type = S.materials[S.materialIndices[I.indexOffset / 3 + h.prim]].type;
    return (type >= 1u && type <= 4u) ? type + 1u : 6u;
}

// Adds each lane's key to counts[key], one atomic per distinct key of the wave;
// with cursors, also returns the lane's position (cursor base + rank in its key).
RT_DEV uint32_t wfKeyAtomic(uint32_t* counts, bool valid, uint32_t key) {
    const uint32_t lane = threadIdx.x & 63u;
    unsigned long long todo = __ballot(valid);
    uint32_t pos = 0;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const uint32_t k = (uint32_t)__shfl((int)key, leader);
        const unsigned long long same = __ballot(valid && key == k);
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(counts + k, (uint32_t)__popcll(same));
        base = (uint32_t)__shfl((int)base, leader);
        if (valid && key == k) pos = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        todo &= ~same;
    }
    return pos;
}

__global__ __launch_bounds__(kTileThreads, RT_WAVES_GBUFFER) void k_wf_extend(const ViewDesc* __restrict__ V,
                                                                             WfState B, uint32_t cur) {
    __shared__ StackLDS<kTileThreads> stk;
    const TStack st = laneStack(stk);
    const uint32_t n = B.ctr[cur];
    const uint32_t* L = B.lists + (size_t)cur * B.n;
    for (uint32_t base = blockIdx.x * kTileThreads; base < n; base += gridDim.x * kTileThreads) {
        const uint32_t j = base + threadIdx.x;
        const bool valid = j < n;
        uint32_t key = 0;
        if (valid) {
            const uint32_t i = L[j];
            float* w = B.w + i;
            const size_t N = B.n;
            const ViewDesc& D = V[__float_as_uint(w[0])];
            const f3 lastPos = mk3(w[10 * N], w[11 * N], w[12 * N]);
            const f3 rdir = mk3(w[13 * N], w[14 * N], w[15 * N]);
            const Isec h = traceClosestHit<false>(D.S, st, lastPos + rdir * 1e-4f, RT_MIN_RAY, rdir, RT_MAX_RAY, nullptr);
            w[18 * N] = h.b1; w[19 * N] = h.b2; w[20 * N] = __uint_as_float(h.inst); w[21 * N] = __uint_as_float(h.prim);
            key = wfBucket(D.S, h);
            B.keys[j] = key;
        }
        (void)wfKeyAtomic(B.ctr + 2, valid, key);
    }
}

// Bucket offsets into the cursors, counts cleared for the next bounce, and the
// next list emptied.
__global__ void k_wf_scan(WfState B, uint32_t cur) {
    if (threadIdx.x != 0u) return;
    uint32_t off = 0;
    for (uint32_t k = 0; k < kWfBuckets; k++) {
        B.ctr[2 + kWfBuckets + k] = off;
        off += B.ctr[2 + k];
        B.ctr[2 + k] = 0u;
    }
    B.ctr[cur ^ 1u] = 0u;
}

__global__ __launch_bounds__(kTileThreads) void k_wf_scatter(WfState B, uint32_t cur) {
    const uint32_t n = B.ctr[cur];
    const uint32_t* L = B.lists + (size_t)cur * B.n;
    for (uint32_t base = blockIdx.x * kTileThreads; base < n; base += gridDim.x * kTileThreads) {
        const uint32_t j = base + threadIdx.x;
        const bool valid = j < n;
        const uint32_t key = valid ? B.keys[j] : 0u;
        const uint32_t pos = wfKeyAtomic(B.ctr + 2 + kWfBuckets, valid, key);
        if (valid) B.sorted[pos] = L[j];
    }
}

__global__ __launch_bounds__(kTileThreads, RT_WAVES_PATH_TRACE) void k_wf_shade(const ViewDesc* __restrict__ V,
                                                                               WfState B, uint32_t cur) {
    __shared__ StackLDS<kTileThreads> stk;
    const TStack st = laneStack(stk);
    const uint32_t n = B.ctr[cur];
    uint32_t* next = B.lists + (size_t)(cur ^ 1u) * B.n;
    for (uint32_t base = blockIdx.x * kTileThreads; base < n; base += gridDim.x * kTileThreads) {
        const uint32_t j = base + threadIdx.x;
        bool live = false;
        uint32_t i = 0;
        if (j < n) {
            i = B.sorted[j];
            uint32_t v, p;
            PTPath P;
            wfLoad(B, i, v, p, P);
            const ViewDesc& D = V[v];
            live = P.isec.inst != RT_INVALID_HIT;
            if (live) {
                Surf surf;
                loadSurfaceInfo(D.S, P.isec, surf);
                const Mat mat = loadMat(D.S, surf.matIndex);
                live = ptVertex<false>(D.S, st, wfSlot(B, i), D.set, P, surf, mat, P.bounce, nullptr) &&
                       ++P.bounce < D.set.maxDepth;
            }
            if (live) wfStore(B, i, v, p, P);
            else storeResv(D.cur, p, ptFinish(P, wfSlot(B, i)));
        }
        wfAppend(next, B.ctr + (cur ^ 1u), live, i);
    }
}

__global__ void k_resv_reset(ResvPlanes R) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R.n) return;
    Resv r = zeroResv();
    r.rcIsec.inst = RT_INVALID_HIT;
    storeResv<true>(R, i, r);
}

// Learned-target refresh after the path-trace pass. Each wave (four per
// workgroup, sharing one LDS copy of the network's weights) walks 64-pixel
// chunks; each lane reads its pixel's reservoir and writes the
// training slot at once, while the pixels whose sample has a surface rc vertex
// (the only ones the network serves, often ~10 %) are compacted into an LDS
// queue; whenever 32 are queued the wave encodes them (lanes 0-31 and 32-63
// each gathering half the grid levels of the same sample) and runs the network
// on the matrix cores. For such a sample F_L = F * Lhat / rcLi per channel
// gives the learned target value and the resample weight, which refers to
// luminance(F) after stream RIS, is rescaled to refer to it; other pixels keep
// their weight and record luminance(F). Every pixel with p % stride == offset
// also writes training slot p / stride (position, reconnection radiance;
// target -1 if unusable).
struct RefreshItem { uint32_t p; float y[3], F[3], rcLi[3], rw; };
constexpr int kRefreshQueue = 96;  // 32 in flight + up to 63 more from one chunk
constexpr uint32_t kRefreshWaves = 4;

// Orders this wave's LDS accesses before and after it (a wave's LDS operations
// complete in issue order; this keeps the compiler from moving them across).
RT_DEV void waveLdsSync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(64 * kRefreshWaves, 3) void k_gris_learned_refresh(DevScene S, rtm::MlpView M, GB g,
                                                                             ResvPlanes R, float* __restrict__ batchPos,
                                                                             float* __restrict__ batchTarget,
                                                                             uint32_t stride, uint32_t offset) {
    __shared__ _Float16 sW[rtm::kWeights];
    __shared__ RefreshItem qs[kRefreshWaves][kRefreshQueue];
    for (int i = threadIdx.x; i < rtm::kWeights; i += blockDim.x) sW[i] = M.w[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    RefreshItem* q = qs[wave];
    const uint32_t n = g.W * g.H;
    const uint32_t chunks = (n + 63u) / 64u;
    uint32_t queued = 0;  // wave-uniform
    // runs the network on queue entries 0-31 (fewer when flushing) and retires them
    auto drain = [&](uint32_t count) {
        const uint32_t j = lane & 31u;
        const bool has = j < count;
        float u = 0.0f, v = 0.0f, w = 0.0f;
        if (has) rtm::normalizePos(M, q[j].y[0], q[j].y[1], q[j].y[2], u, v, w);
        rtm::h4 enc[4];
        rtm::encodeLane(M, lane >> 5, u, v, w, enc);
        rtm::Fwd F;
        rtm::forwardTile(sW, lane, enc, F);
        if (lane < 32u && has) {
            const RefreshItem& it = q[j];
            f3 L = mk3(rtm::outAct(F.z3[0]), rtm::outAct(F.z3[1]), rtm::outAct(F.z3[2]));
            f3 Fv = mk3(it.F[0], it.F[1], it.F[2]);
            f3 FL;
            FL.x = it.rcLi[0] > 0.0f ? Fv.x * (L.x / it.rcLi[0]) : Fv.x;
            FL.y = it.rcLi[1] > 0.0f ? Fv.y * (L.y / it.rcLi[1]) : Fv.y;
            FL.z = it.rcLi[2] > 0.0f ? Fv.z * (L.z / it.rcLi[2]) : Fv.z;
            const float pOld = luminance(Fv), pNew = luminance(FL);
            float rw = it.rw;
            if (pOld > 0.0f) rw = rw * (pNew / pOld);
            R.p[4 * R.n + it.p].w = rw;
            R.p[5 * R.n + it.p] = make_float4(L.x, L.y, L.z, pNew);
        }
        waveLdsSync();
        // move the entries behind the drained ones to the front
        const uint32_t rest = queued - count;
        RefreshItem keep;
        if (lane < rest) keep = q[count + lane];
        waveLdsSync();
        if (lane < rest) q[lane] = keep;
        waveLdsSync();
        queued = rest;
    };
    for (uint32_t c = blockIdx.x * kRefreshWaves + wave; c < chunks; c += gridDim.x * kRefreshWaves) {
        const uint32_t p = c * 64u + lane;
        bool need = false;
        Resv r = zeroResv();
        f3 y = s3(0.0f);
        if (p < n && g.dn[p].x != 0.0f) {
            r = loadResv(R, p);
            need = resvIsValid(r) && r.rcIsec.inst != RT_INVALID_HIT && flagsRcVertexType(r.flags) == RC_SURFACE;
            if (need) {
                Surf sf;
                loadSurfaceInfo(S, r.rcIsec, sf);
                y = sf.pos;
            } else {
                R.p[5 * R.n + p] = make_float4(0.0f, 0.0f, 0.0f, luminance(r.F));
            }
        }
        if (p < n && p % stride == offset) {
            const uint32_t slot = p / stride;
            batchPos[3 * (size_t)slot] = y.x;
            batchPos[3 * (size_t)slot + 1] = y.y;
            batchPos[3 * (size_t)slot + 2] = y.z;
            batchTarget[3 * (size_t)slot] = need ? r.rcLi.x : -1.0f;
            batchTarget[3 * (size_t)slot + 1] = need ? r.rcLi.y : 0.0f;
            batchTarget[3 * (size_t)slot + 2] = need ? r.rcLi.z : 0.0f;
        }
        const unsigned long long mask = __ballot(need);
        if (need) {
            RefreshItem& it = q[queued + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))];
            it.p = p;
            it.y[0] = y.x; it.y[1] = y.y; it.y[2] = y.z;
            it.F[0] = r.F.x; it.F[1] = r.F.y; it.F[2] = r.F.z;
            it.rcLi[0] = r.rcLi.x; it.rcLi[1] = r.rcLi.y; it.rcLi[2] = r.rcLi.z;
            it.rw = r.resampleWeight;
        }
        queued += (uint32_t)__popcll(mask);
        waveLdsSync();
        while (queued >= 32u) drain(32u);
    }
    if (queued > 0u) drain(queued);
}

inline ResvPlanes planes(void* p, uint32_t W, uint32_t H) {
    ResvPlanes r; r.p = reinterpret_cast<float4*>(p); r.n = (size_t)W * H; return r;
}

}  // namespace

// ===================================================================================================
// Launchers (called by the C ABI after argument checks)
namespace rti {

// Counting or plain build of a kernel, launched (or recorded) through rti::launch.
#define RT_DISPATCH(KERNEL, GRID, BLOCK, STREAM, ...)                                                  \
    do {                                                                                                 \
        if (gCountersEnabled) (void)launch(KERNEL<true>, GRID, BLOCK, STREAM, __VA_ARGS__, counterBuffer()); \
        else (void)launch(KERNEL<false>, GRID, BLOCK, STREAM, __VA_ARGS__, nullptr);                        \
    } while (0)

// Same, for kernels that also take the target function, the resampling-weight mode and the
// learned proposal (template arguments <counting, learned, unbiased, guided>).
#define RT_LAUNCH_LUG_(KERNEL, C, G, LEARNED, UNBIASED, GRID, BLOCK, STREAM, ...)                           \
    do {                                                                                                  \
        if (LEARNED) {                                                                                    \
            if (UNBIASED) (void)launch(KERNEL<C, true, true, G>, GRID, BLOCK, STREAM, __VA_ARGS__);       \
            else (void)launch(KERNEL<C, true, false, G>, GRID, BLOCK, STREAM, __VA_ARGS__);               \
        } else {                                                                                          \
            if (UNBIASED) (void)launch(KERNEL<C, false, true, G>, GRID, BLOCK, STREAM, __VA_ARGS__);      \
            else (void)launch(KERNEL<C, false, false, G>, GRID, BLOCK, STREAM, __VA_ARGS__);              \
        }                                                                                                 \
    } while (0)
#define RT_DISPATCH_LUG(KERNEL, LEARNED, UNBIASED, GUIDED, GRID, BLOCK, STREAM, ...)                           \
    do {                                                                                                      \
        TraceCounters* ctr_ = gCountersEnabled ? counterBuffer() : nullptr;                                     \
        if (gCountersEnabled) {                                                                               \
            if (GUIDED) RT_LAUNCH_LUG_(KERNEL, true, true, LEARNED, UNBIASED, GRID, BLOCK, STREAM, __VA_ARGS__, ctr_);  \
            else RT_LAUNCH_LUG_(KERNEL, true, false, LEARNED, UNBIASED, GRID, BLOCK, STREAM, __VA_ARGS__, ctr_);        \
        } else {                                                                                              \
            if (GUIDED) RT_LAUNCH_LUG_(KERNEL, false, true, LEARNED, UNBIASED, GRID, BLOCK, STREAM, __VA_ARGS__, ctr_); \
            else RT_LAUNCH_LUG_(KERNEL, false, false, LEARNED, UNBIASED, GRID, BLOCK, STREAM, __VA_ARGS__, ctr_);       \
        }                                                                                                     \
    } while (0)


hipError_t launchTraceRays(const rt_device_scene* s, const float* rays8, uint32_t n, int mode, int skipLights, float* t,
                           float* bary2, uint32_t* inst, uint32_t* prim, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    dim3 grid((n + kTileThreads - 1) / kTileThreads);
    RT_DISPATCH(k_trace_rays, grid, kTileThreads, stream, s->dev, rays8, n, mode, skipLights, t, bary2, inst, prim);
    return hipGetLastError();
}

hipError_t launchGBuffer(const rt_device_scene* s, const rt_camera& cam, float* dn, uint32_t* am, uint16_t* mv,
                         hipStream_t stream) {
    RT_DISPATCH(k_gbuffer, tileGrid(cam.filmSize[0], cam.filmSize[1]), kTileThreads, stream, s->dev, cam,
                reinterpret_cast<float4*>(dn), reinterpret_cast<uint2*>(am), reinterpret_cast<uint32_t*>(mv));
    return hipGetLastError();
}

hipError_t launchPtPlain(const rt_device_scene* s, const rt_camera* cams, const rt_frame& f, uint32_t maxDepth,
                         uint32_t which, hipStream_t stream) {
    GB g = makeGB(f.depthNormal, f.albedoMatId, f.motionVector, f.width, f.height);
    dim3 grid = tileGrid(f.width, f.height);
    if (which == 3u) grid.z = 2;
    RT_DISPATCH(k_pt_plain, grid, kTileThreads, stream, s->dev, cams[0], g, maxDepth, which, f.directOutput,
                f.indirectOutput);
    return hipGetLastError();
}

// Workgroups of a persistent pass: as many as the device holds at once, never
// more than there are 8x8 tiles.
template <class Kernel>
static hipError_t persistentGrid(Kernel kernel, uint32_t W, uint32_t H, uint32_t& grid) {
    int dev = 0, cus = 0, perCU = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, reinterpret_cast<const void*>(kernel), kTileThreads, 0);
    if (e != hipSuccess) return e;
    const uint64_t tiles = (uint64_t)((W + kTileX - 1) / kTileX) * ((H + kTileY - 1) / kTileY);
    grid = (uint32_t)std::min<uint64_t>(tiles, (uint64_t)std::max(1, cus) * (uint64_t)std::max(1, perCU));
    return hipSuccess;
}

// The learned proposal's bindings of a frame (all null unless settings.guidedProposal).
rtg::GuideFrame guideFrame(const rt_frame& f, const rt_gris_settings& set) {
    if (!set.guidedProposal) return rtg::GuideFrame{nullptr, nullptr, nullptr};
    return rtg::GuideFrame{reinterpret_cast<const float4*>(f.guideLobes), reinterpret_cast<const float4*>(f.guideLobesPrev),
                           reinterpret_cast<float4*>(f.guideSamples)};
}

// queue: the pixel queue of this launch (null: the scene's next round-robin slot).
// maxGroups: at most this many workgroups (0: as many as the device holds at once).
hipError_t launchGrisPathTrace(const rt_device_scene* s, const rt_camera* cams, const rt_frame& f,
                               const rt_gris_settings& set, hipStream_t stream, WorkQueue* queue, uint32_t maxGroups) {
    GB g = makeGB(f.depthNormal, f.albedoMatId, f.motionVector, f.width, f.height);
    uint32_t grid = 0;
    const bool guided = set.guidedProposal == 1u;
    hipError_t e = gCountersEnabled
                       ? (guided ? persistentGrid(k_gris_path_trace<true, true>, f.width, f.height, grid)
                                 : persistentGrid(k_gris_path_trace<true, false>, f.width, f.height, grid))
                       : (guided ? persistentGrid(k_gris_path_trace<false, true>, f.width, f.height, grid)
                                 : persistentGrid(k_gris_path_trace<false, false>, f.width, f.height, grid));
    if (e != hipSuccess) return e;
    if (maxGroups) grid = std::min(grid, maxGroups);
    TraceCounters* ctr = gCountersEnabled ? counterBuffer() : nullptr;
    const ResvPlanes R = planes(f.grisReservoir, f.width, f.height);
    // the scene's next queue slot: wait (on this stream) for the slot's last launch if it may
    // still run; a stream under capture is left alone (its recorded graph orders its own work)
    hipEvent_t slotEvent = nullptr;
    WorkQueue* q = queue;
    if (!q) {
        const uint32_t slot = s->nextQueueSlot();
        q = s->queues + slot;
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        if (!tlsRecorder) e = hipStreamIsCapturing(stream, &cap);
        if (e != hipSuccess) return e;
        if (!tlsRecorder && cap == hipStreamCaptureStatusNone) {
            slotEvent = s->queueEvents[slot];
            const hipError_t qe = hipEventQuery(slotEvent);
            if (qe == hipErrorNotReady) e = hipStreamWaitEvent(stream, slotEvent, 0);
            else if (qe != hipSuccess) e = qe;
            if (e != hipSuccess) return e;
        }
    }
    const rtg::GuideFrame gf = guideFrame(f, set);
    if (gCountersEnabled) {
        if (guided) (void)launch(k_gris_path_trace<true, true>, dim3(grid), kTileThreads, stream, s->dev, cams[0], g, set, R, q, gf, ctr);
        else (void)launch(k_gris_path_trace<true, false>, dim3(grid), kTileThreads, stream, s->dev, cams[0], g, set, R, q, gf, ctr);
    } else {
        if (guided) (void)launch(k_gris_path_trace<false, true>, dim3(grid), kTileThreads, stream, s->dev, cams[0], g, set, R, q, gf, ctr);
        else (void)launch(k_gris_path_trace<false, false>, dim3(grid), kTileThreads, stream, s->dev, cams[0], g, set, R, q, gf, ctr);
    }
    e = hipGetLastError();
    if (e == hipSuccess && slotEvent) e = hipEventRecord(slotEvent, stream);
    return e;
}

hipError_t launchGrisTemporal(const rt_device_scene* s, const rt_camera* cams, const rt_frame& f,
                              const rt_gris_settings& set, hipStream_t stream) {
    GB g = makeGB(f.depthNormal, f.albedoMatId, f.motionVector, f.width, f.height);
    GB gp = makeGB(f.depthNormalPrev, f.albedoMatIdPrev, nullptr, f.width, f.height);
    RT_DISPATCH_LUG(k_gris_temporal, set.targetFunction == 1u, set.unbiasedReuse == 1u, set.guidedProposal == 1u,
                    tileGrid(f.width, f.height),
                   kTileThreads, stream, s->dev,
                  cams[0], cams[1], g, gp, set, planes(f.grisReservoir, f.width, f.height),
                  planes(f.grisReservoirPrev, f.width, f.height), planes(f.grisReservoirTemp, f.width, f.height),
                  guideFrame(f, set));
    return hipGetLastError();
}

hipError_t launchGrisSpatial(const rt_device_scene* s, const rt_camera* cams, const rt_frame& f,
                             const rt_gris_settings& set, hipStream_t stream) {
    GB g = makeGB(f.depthNormal, f.albedoMatId, f.motionVector, f.width, f.height);
    RT_DISPATCH_LUG(k_gris_spatial, set.targetFunction == 1u, set.unbiasedReuse == 1u, set.guidedProposal == 1u,
                    tileGrid(f.width, f.height),
                   kTileThreads, stream, s->dev,
                  cams[0], g, set, planes(f.grisReservoirTemp, f.width, f.height),
                  planes(f.grisReservoir, f.width, f.height), f.indirectOutput, guideFrame(f, set));
    return hipGetLastError();
}

hipError_t launchLearnedRefresh(const rt_device_scene* s, rt_mlp* m, const rt_camera* cams, const rt_frame& f,
                                hipStream_t stream) {
    const uint32_t n = f.width * f.height;
    const uint32_t stride = std::max(1u, (n + rtm::kBatchCap - 1) / rtm::kBatchCap);
    const uint32_t offset = (cams[0].frameIndex & 0x7fffffffu) % stride;
    m->batchSize = offset < n ? (n - offset + stride - 1) / stride : 0u;
    GB g = makeGB(f.depthNormal, f.albedoMatId, f.motionVector, f.width, f.height);
    const uint32_t chunks = (n + 63u) / 64u;
    const uint32_t blocks = std::min<uint32_t>((chunks + kRefreshWaves - 1u) / kRefreshWaves, 1280u);
    (void)launch(k_gris_learned_refresh, dim3(blocks), dim3(64 * kRefreshWaves), stream, s->dev, m->view, g,
                 planes(f.grisReservoir, f.width, f.height), m->batchPos, m->batchTarget, stride, offset);
    return hipGetLastError();
}

// ---- batched views ----
size_t batchDescBytes() { return sizeof(ViewDesc); }

uint32_t batchViewTiles(const rt_frame& f) {
    return ((f.width + kTileX - 1) / kTileX) * ((f.height + kTileY - 1) / kTileY);
}

// Writes the descriptor of one view into host memory at dst.
void batchFillDesc(void* dst, const rt_device_scene* scene, const rt_camera* cams, const rt_frame& f,
                   const rt_gris_settings& set, uint32_t tileBase, WorkQueue* queue, uint32_t queueShift,
                   uint32_t pixBase) {
    ViewDesc d;
    std::memset(&d, 0, sizeof(d));
    d.S = scene->dev;
    d.cam = cams[0];
    d.prevCam = cams[1];
    d.g = makeGB(f.depthNormal, f.albedoMatId, f.motionVector, f.width, f.height);
    d.gPrev = makeGB(f.depthNormalPrev, f.albedoMatIdPrev, nullptr, f.width, f.height);
    d.cur = planes(f.grisReservoir, f.width, f.height);
    d.prev = planes(f.grisReservoirPrev, f.width, f.height);
    d.temp = planes(f.grisReservoirTemp, f.width, f.height);
    d.indirectOut = f.indirectOutput;
    d.dnOut = reinterpret_cast<float4*>(const_cast<float*>(f.depthNormal));
    d.amOut = reinterpret_cast<uint2*>(const_cast<uint32_t*>(f.albedoMatId));
    d.mvOut = reinterpret_cast<uint32_t*>(const_cast<uint16_t*>(f.motionVector));
    d.set = set;
    d.tileBase = tileBase;
    d.tilesX = (f.width + kTileX - 1) / kTileX;
    d.tiles = batchViewTiles(f);
    d.queue = queue;
    d.queueShift = queueShift;
    d.pixBase = pixBase;
    d.gf = guideFrame(f, set);
    std::memcpy(dst, &d, sizeof(d));
}

// The batched persistent path trace's kernel for a counting build, the
// learned proposal and a register budget of kBatchWaves[w] waves per SIMD.
static const void* batchPathTraceKernel(bool counting, bool guided, uint32_t w) {
    if (counting) return guided ? reinterpret_cast<const void*>(k_gris_path_trace_batch<true, true>)
                                : reinterpret_cast<const void*>(k_gris_path_trace_batch<true, false>);
    static const void* const k[2][kBatchWaveVariants] = {
        {reinterpret_cast<const void*>(k_gris_path_trace_batch<false, false, kBatchWaves[0]>),
         reinterpret_cast<const void*>(k_gris_path_trace_batch<false, false, kBatchWaves[1]>),
         reinterpret_cast<const void*>(k_gris_path_trace_batch<false, false, kBatchWaves[2]>)},
        {reinterpret_cast<const void*>(k_gris_path_trace_batch<false, true, kBatchWaves[0]>),
         reinterpret_cast<const void*>(k_gris_path_trace_batch<false, true, kBatchWaves[1]>),
         reinterpret_cast<const void*>(k_gris_path_trace_batch<false, true, kBatchWaves[2]>)}};
    return k[guided ? 1 : 0][w];
}

// Resident workgroups per CU of each batched persistent path-trace variant
// (perCU[guided][w]), and the CU count.
hipError_t batchOccupancy(uint32_t perCU[2][kBatchWaveVariants], uint32_t& cus) {
    int dev = 0, c = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
    for (int gd = 0; gd < 2; gd++)
        for (uint32_t w = 0; w < kBatchWaveVariants; w++) {
            int o = 0;
            if (e == hipSuccess)
                e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, batchPathTraceKernel(false, gd == 1, w),
                                                                 kTileThreads, 0);
            perCU[gd][w] = (uint32_t)std::max(1, o);
        }
    cus = (uint32_t)std::max(1, c);
    return e;
}

// The passes of one frame over device arrays of view descriptors. The
// G-buffer and reuse launches cover the numViews descriptors at descs, whose
// tile bases count from 0 over totalTiles tiles; the persistent path trace
// covers every view of the frame.
// Grid of a tile launch: with chunks of C tiles, every tile below totalTiles is
// reached when the grid is a multiple of 8 C (the caller bounds C and totalTiles).
static dim3 batchTileGrid(uint32_t totalTiles, uint32_t chunk) {
    if (!chunk) return dim3(totalTiles);
    const uint64_t span = 8ull * chunk;
    return dim3((uint32_t)((totalTiles + span - 1) / span * span));
}

hipError_t launchBatchGBuffer(const void* descs, uint32_t numViews, uint32_t totalTiles, uint32_t xcdSwizzle,
                              hipStream_t stream) {
    const ViewDesc* V = reinterpret_cast<const ViewDesc*>(descs);
    const uint32_t swz = xcdSwizzle;
    RT_DISPATCH(k_gbuffer_batch, batchTileGrid(totalTiles, swz), kTileThreads, stream, V, numViews, totalTiles, swz);
    return hipGetLastError();
}

hipError_t launchBatchPathTrace(const void* descs, uint32_t numViews, uint32_t persistentGroups, bool guided,
                                uint32_t waveVariant, hipStream_t stream) {
    const ViewDesc* V = reinterpret_cast<const ViewDesc*>(descs);
    const dim3 grid(std::max(1u, persistentGroups));
    TraceCounters* ctr = gCountersEnabled ? counterBuffer() : nullptr;
    // every variant has the parameters (const ViewDesc*, uint32_t, TraceCounters*)
    auto* kernel = reinterpret_cast<void (*)(const ViewDesc*, uint32_t, TraceCounters*)>(
        const_cast<void*>(batchPathTraceKernel(gCountersEnabled, guided, waveVariant)));
    (void)launch(kernel, grid, kTileThreads, stream, V, numViews, ctr);
    return hipGetLastError();
}

// Path-state bytes of the wavefront variant for n paths, and their layout.
size_t wavefrontBytes(uint32_t n) {
    return ((size_t)kWfWords + 4u) * 4u * n + 4u * (2u + 2u * kWfBuckets);
}
void wavefrontCarve(void* base, uint32_t n, WavefrontBuffers& wb) {
    char* p = static_cast<char*>(base);
    wb.words = reinterpret_cast<float*>(p);
    p += (size_t)kWfWords * 4u * n;
    wb.lists = reinterpret_cast<uint32_t*>(p);
    p += 8u * (size_t)n;
    wb.sorted = reinterpret_cast<uint32_t*>(p);
    p += 4u * (size_t)n;
    wb.keys = reinterpret_cast<uint32_t*>(p);
    p += 4u * (size_t)n;
    wb.counters = reinterpret_cast<uint32_t*>(p);
    wb.capacity = n;
}

// The wavefront variant of launchBatchPathTrace (non-guided views; the plain
// kernels, not the counting builds): clear, primary vertices, then per bounce
// extend, scan, scatter and shade, each a grid-stride launch of `groups`
// one-wave workgroups over the device-side lists.
hipError_t launchBatchWavefront(const void* descs, uint32_t numViews, uint32_t totalPixels, uint32_t maxDepth,
                                const WavefrontBuffers& wb, uint32_t groups, hipStream_t stream) {
    const ViewDesc* V = reinterpret_cast<const ViewDesc*>(descs);
    const WfState B{wb.words, wb.lists, wb.sorted, wb.keys, wb.counters, wb.capacity};
    const dim3 grid(std::max(1u, groups));
    (void)launch(k_wf_clear, dim3(1), dim3(64), stream, B);
    (void)launch(k_wf_begin, grid, dim3(kTileThreads), stream, V, numViews, totalPixels, B);
    uint32_t cur = 0;
    for (uint32_t b = 1; b < maxDepth; b++) {
        (void)launch(k_wf_extend, grid, dim3(kTileThreads), stream, V, B, cur);
        (void)launch(k_wf_scan, dim3(1), dim3(64), stream, B, cur);
        (void)launch(k_wf_scatter, grid, dim3(kTileThreads), stream, B, cur);
        (void)launch(k_wf_shade, grid, dim3(kTileThreads), stream, V, B, cur);
        cur ^= 1u;
    }
    return hipGetLastError();
}

hipError_t launchBatchReuse(const void* descs, uint32_t numViews, uint32_t totalTiles, uint32_t xcdSwizzle,
                            bool learned, bool unbiased, bool guided, hipStream_t stream) {
    const ViewDesc* V = reinterpret_cast<const ViewDesc*>(descs);
    const uint32_t swz = xcdSwizzle;
    const dim3 grid = batchTileGrid(totalTiles, swz);
    RT_DISPATCH_LUG(k_gris_temporal_batch, learned, unbiased, guided, grid, kTileThreads, stream, V, numViews,
                    totalTiles, swz);
    RT_DISPATCH_LUG(k_gris_spatial_batch, learned, unbiased, guided, grid, kTileThreads, stream, V, numViews,
                    totalTiles, swz);
    return hipGetLastError();
}

hipError_t launchResvReset(void* resv, uint32_t W, uint32_t H, hipStream_t stream) {
    ResvPlanes R = planes(resv, W, H);
    if (R.n == 0) return hipSuccess;
    k_resv_reset<<<(unsigned)((R.n + 255) / 256), 256, 0, stream>>>(R);
    return hipGetLastError();
}

}  // namespace rti
