// ReSTIR DI for gfx950: the three passes of TestReSTIR::render
// (src/TestReSTIR.cpp:9-36) -- path generation (di_path_gen.glsl:9-63),
// temporal reuse (di_temporal.glsl:36-88) and spatial reuse with shading
// (di_spatial.glsl:36-119) -- over the reservoir operations and shift mappings
// of di_reservoir.glsl.
//
// One 64-lane wave per 8x8 pixel tile, as the GRIS passes. A DI reservoir
// (layouts.glsl:97-115, 64 B) lives in HBM as four float4 planes, plane k
// holding the k-th 16 bytes of the reference struct, so a wave's access to one
// plane is a contiguous 1-KiB transaction:
//   P0 isec (bary.xy, instance, primitive)   P1 Li.xyz, pad0
//   P2 jacobian, samplePdf, rng, isLightSample
//   P3 sampleCount (uint), resampleWeight, contribWeight, weight
// Locals the GLSL leaves uninitialised (the path sample and weight of a fresh
// reservoir) are zero. A sample reference read from reservoir memory is
// bounds-checked against the scene before its surface is loaded (an
// out-of-range one loads a zero surface), so stale or uncleared buffers cannot
// make a kernel read outside the scene arrays.
#include "rt_dpixel.h"

using namespace rtd;
using namespace rtp;

namespace {

#ifndef RT_WAVES_DI
#define RT_WAVES_DI 4
#endif

constexpr uint32_t kSampleLight = 0, kSampleBSDF = 1;   // TestReSTIR::SampleType (Both = 2)
constexpr uint32_t kShiftReconnection = 0, kShiftReplay = 1;
constexpr uint32_t kDICap = 40;                        // DIReservoirCapSample(resv, 40)
constexpr uint32_t kDINeighbors = 10;                  // ResampleNum, di_spatial.glsl:72
constexpr float kDIRadius = 20.0f;                     // ResampleRadius, di_spatial.glsl:73

struct DISample { Isec isec; f3 Li; float pad0; float jacobian, samplePdf; uint32_t rng, isLight; };
struct DIResv { DISample s; uint32_t sampleCount; float resampleWeight, contribWeight, weight; };
struct DIPlanes { float4* p; size_t n; };

RT_DEV DIResv diZero() {
    DIResv r;
    r.s.isec.b1 = 0.0f; r.s.isec.b2 = 0.0f; r.s.isec.inst = 0u; r.s.isec.prim = 0u;
    r.s.Li = s3(0.0f); r.s.pad0 = 0.0f; r.s.jacobian = 0.0f; r.s.samplePdf = 0.0f; r.s.rng = 0u; r.s.isLight = 0u;
    r.sampleCount = 0u; r.resampleWeight = 0.0f; r.contribWeight = 0.0f; r.weight = 0.0f;
    return r;
}
RT_DEV DIResv diLoad(const DIPlanes& R, size_t i) {
    const float4 a = R.p[i], b = R.p[R.n + i], c = R.p[2 * R.n + i], d = R.p[3 * R.n + i];
    DIResv r;
    r.s.isec.b1 = a.x; r.s.isec.b2 = a.y; r.s.isec.inst = __float_as_uint(a.z); r.s.isec.prim = __float_as_uint(a.w);
    r.s.Li = mk3(b.x, b.y, b.z); r.s.pad0 = b.w;
    r.s.jacobian = c.x; r.s.samplePdf = c.y; r.s.rng = __float_as_uint(c.z); r.s.isLight = __float_as_uint(c.w);
    r.sampleCount = __float_as_uint(d.x); r.resampleWeight = d.y; r.contribWeight = d.z; r.weight = d.w;
    return r;
}
RT_DEV void diStore(const DIPlanes& R, size_t i, const DIResv& r) {
    R.p[i] = make_float4(r.s.isec.b1, r.s.isec.b2, __uint_as_float(r.s.isec.inst), __uint_as_float(r.s.isec.prim));
    R.p[R.n + i] = make_float4(r.s.Li.x, r.s.Li.y, r.s.Li.z, r.s.pad0);
    R.p[2 * R.n + i] = make_float4(r.s.jacobian, r.s.samplePdf, __uint_as_float(r.s.rng), __uint_as_float(r.s.isLight));
    R.p[3 * R.n + i] = make_float4(__uint_as_float(r.sampleCount), r.resampleWeight, r.contribWeight, r.weight);
}
// uDIReservoir*[index1D(pixel)]: reads outside the buffer are zero
RT_DEV DIResv diAt(const DIPlanes& R, uint32_t W, uint32_t H, int x, int y) {
    const unsigned long long idx = (unsigned long long)W * (uint32_t)y + (uint32_t)x;
    if (x < 0 || y < 0 || idx >= (unsigned long long)W * H) return diZero();
    return diLoad(R, (size_t)idx);
}

// ---- di_reservoir.glsl:28-77 -----------------------------------------------------------------
RT_DEV bool diSampleValid(const DISample& s) { return s.isec.inst != RT_INVALID_HIT; }
RT_DEV void diReset(DIResv& r) { r.sampleCount = 0u; r.resampleWeight = 0.0f; r.contribWeight = 0.0f; }
RT_DEV bool diValid(const DIResv& r) { return !isnan(r.resampleWeight); }
RT_DEV void diAddSample(DIResv& r, const DISample& s, float w, float u) {
    r.resampleWeight += w;
    r.sampleCount++;
    if (u * r.resampleWeight < w) { r.s = s; r.weight = w; }
}
RT_DEV void diMerge(DIResv& r, const DIResv& rhs, float u) {
    r.resampleWeight += rhs.resampleWeight;
    r.sampleCount += rhs.sampleCount;
    if (u * r.resampleWeight < rhs.resampleWeight) { r.s = rhs.s; r.weight = rhs.weight; }
}
RT_DEV void diCap(DIResv& r, uint32_t cap) {
    if (r.sampleCount > cap) {
        r.resampleWeight *= (float)cap / (float)r.sampleCount;
        r.sampleCount = cap;
    }
}

// loadSurfaceInfo of a sample reference, bounds-checked against the scene.
RT_DEV void diSurface(const DevScene& S, const Isec& is, Surf& sf) {
    const bool ok = is.inst == 0u ? is.prim < S.numLights
                                  : (is.inst - 1u < S.numInstances && is.prim < S.instances[is.inst - 1u].triCount);
    if (ok) { loadSurfaceInfo(S, is, sf); return; }
    sf.pos = s3(0.0f); sf.norm = s3(0.0f); sf.albedo = s3(0.0f); sf.matIndex = 0u; sf.isLight = is.inst == 0u;
}

// di_reservoir.glsl:79-185
template <bool kCount>
RT_DEV f3 sampleLi(const DevScene& S, TStack st, uint32_t sampleMode, const Surf& surf, const Mat& mat, f3 wo,
                   uint32_t rng, uint32_t& resvRng, DIResv& resv, TraceCounters* ctr) {
    f3 radiance = s3(0.0f);
    DISample ps = diZero().s;
    ps.rng = rng;
    const f4 lightRand = sample4f(rng);
    const f3 scatterRand = sample3f(rng);
    if (sampleMode != kSampleBSDF && !isBSDFDelta(mat)) {
        f3 lightDir; f2 lightBary; float lightDist, lightPdf, lightJac; uint32_t lightId;
        const f3 Lr = sampleLightByPower(S, surf.pos, lightDir, lightDist, lightPdf, lightJac, lightBary, lightId, lightRand);
        const bool shadowed = traceShadow<kCount>(S, st, surf.pos, RT_MIN_RAY, lightDir, lightDist - RT_MIN_RAY, ctr);
        if (!shadowed && lightPdf > 1e-6f) {
            const float bsdfPdf = evalPdf(mat, surf.norm, wo, lightDir);
            float weight = MISWeight(lightPdf, bsdfPdf);
            if (sampleMode == kSampleLight) weight = 1.0f;
            const f3 contrib =
                Lr * evalBSDF(mat, surf.albedo, surf.norm, wo, lightDir) * satDot(surf.norm, lightDir) / lightPdf * weight;
            float sw = luminance(contrib);
            if (isnan(sw) || sw < 0.0f) sw = 0.0f;
            ps.isec.b1 = lightBary.x; ps.isec.b2 = lightBary.y; ps.isec.inst = 0u; ps.isec.prim = lightId;
            ps.Li = Lr * weight;
            ps.jacobian = lightJac;
            ps.samplePdf = lightPdf;
            ps.isLight = 1u;
            diAddSample(resv, ps, sw, sample1f(resvRng));
            radiance = radiance + contrib;
        }
    }
    ps.isLight = 0u;
    BSDFSample s; s.type = 0u; s.pdf = 0.0f; s.wi = s3(0.0f); s.bsdf = s3(0.0f);
    if (sampleMode != kSampleLight && sampleBSDF(mat, surf.albedo, surf.norm, wo, scatterRand, s) && s.pdf > 1e-6f) {
        const Isec isec = traceClosestHit<kCount>(S, st, surf.pos, RT_MIN_RAY, s.wi, RT_MAX_RAY, ctr);
        if (isec.inst != RT_INVALID_HIT) {
            Surf hit; loadSurfaceInfo(S, isec, hit);
            const float cosTheta = -dot(s.wi, hit.norm);
            if (hit.isLight && cosTheta > 0.0f) {
                const float dist = length(hit.pos - surf.pos);
                const float sumPower = S.lightTable[0].prob;
                const float lightPdf = luminance(hit.albedo) / sumPower * dist * dist / fabsf(cosTheta);
                float weight = MISWeight(s.pdf, lightPdf);
                if (sampleMode == kSampleBSDF || isSampleTypeDelta(s.type)) weight = 1.0f;
                const float cosTerm = isSampleTypeDelta(s.type) ? 1.0f : satDot(surf.norm, s.wi);
                const f3 contrib = hit.albedo * s.bsdf * cosTerm / s.pdf * weight;
                ps.isec = isec;
                ps.Li = hit.albedo * weight;
                ps.jacobian = fabsf(cosTheta) / square(dist);
                ps.samplePdf = s.pdf;
                ps.isLight = 0u;
                // the BSDF sample's resampling weight is luminance(contrib) unsanitised (di_reservoir.glsl:167)
                diAddSample(resv, ps, luminance(contrib), sample1f(resvRng));
                radiance = radiance + contrib;
            }
        }
    }
    if (!diValid(resv)) diReset(resv);
    // importance sampling, not resampling (di_reservoir.glsl:173-184)
    if (resv.sampleCount > 0u && diSampleValid(resv.s) && resv.weight > 0.0f) {
        resv.s.Li = resv.s.Li * (resv.resampleWeight / resv.weight);
        resv.weight = resv.resampleWeight;
    } else {
        resv.s.isec.inst = RT_INVALID_HIT;
        resv.weight = 0.0f;
        resv.resampleWeight = 0.0f;
    }
    resv.sampleCount = 1u;
    return radiance;
}

// di_reservoir.glsl:187-221
template <bool kCount>
RT_DEV void randomReplay(const DevScene& S, TStack st, uint32_t sampleMode, DIResv& dst, const Surf& dstSurf,
                         const DIResv& src, f3 wo, uint32_t& rng, TraceCounters* ctr) {
    const Mat dstMat = loadMat(S, dstSurf.matIndex);
    DIResv rr = diZero();
    (void)sampleLi<kCount>(S, st, sampleMode, dstSurf, dstMat, wo, src.s.rng, rng, rr, ctr);
    const float jacobian = 1.0f;
    if (diSampleValid(rr.s)) {
        Surf rs; diSurface(S, rr.s.isec, rs);
        const f3 wi = normalize(rs.pos - dstSurf.pos);
        const f3 Li = rr.s.Li * evalBSDF(dstMat, dstSurf.albedo, dstSurf.norm, wo, wi) * satDot(dstSurf.norm, wi) /
                      rr.s.samplePdf;
        const float dstPHat = luminance(Li * jacobian);
        rr.resampleWeight = src.resampleWeight * dstPHat / src.weight;
        rr.sampleCount = src.sampleCount;
    } else {
        rr.resampleWeight = 0.0f;
    }
    if (diValid(rr)) diMerge(dst, rr, sample1f(rng));
}

// di_reservoir.glsl:223-281
template <bool kCount>
RT_DEV void reconnection(const DevScene& S, TStack st, DIResv& dst, const Surf& dstSurf, DIResv src, f3 wo,
                         uint32_t& rng, TraceCounters* ctr) {
    DISample ss = src.s;
    Surf rc; diSurface(S, ss.isec, rc);
    const float dist = distance(rc.pos, dstSurf.pos);
    const f3 wi = normalize(rc.pos - dstSurf.pos);
    const float cosTheta = -dot(rc.norm, wi);
    const float dstJacobian = fabsf(cosTheta) / square(dist);
    const float jacobian = dstJacobian / ss.jacobian;
    float dstPHat = 0.0f, dstSamplePdf = 0.0f;
    // The shifted sample is formed before the visibility of its reconnection is traced, so that
    // only it and the ray stay live across the traversal; an occluded shift merges with weight 0
    // (only its M counts: a weight-0 candidate is never picked).
    if (diSampleValid(ss) && dist > 1e-4f && cosTheta > 0.0f && !isnan(jacobian) && ss.jacobian > 0.0f) {
        const Mat dstMat = loadMat(S, dstSurf.matIndex);
        if (!isnan(ss.samplePdf) && ss.samplePdf > 1e-6f) {
            const f3 Li = src.s.Li * evalBSDF(dstMat, dstSurf.albedo, dstSurf.norm, wo, wi) * satDot(dstSurf.norm, wi) /
                          ss.samplePdf;
            dstPHat = luminance(Li * jacobian);
        }
        if (ss.isLight) dstSamplePdf = luminance(rc.albedo) / S.lightTable[0].prob / dstJacobian;
        else dstSamplePdf = evalPdf(dstMat, dstSurf.norm, wo, wi);
        ss.jacobian = dstJacobian;
        ss.samplePdf = dstSamplePdf;
        if (ss.samplePdf < 1e-6f || isnan(ss.samplePdf)) ss.samplePdf = 0.0f;
        src.s = ss;
        src.resampleWeight *= dstPHat / src.weight;
        if (isnan(src.resampleWeight)) src.resampleWeight = 0.0f;
        if (!traceVisibility<kCount>(S, st, dstSurf.pos, rc.pos, ctr)) src.resampleWeight = 0.0f;
    } else {
        src.resampleWeight = 0.0f;
    }
    if (diValid(src)) diMerge(dst, src, sample1f(rng));
}

// DIReservoirReuseAndMerge (di_reservoir.glsl:283-290); the Hybrid shift merges nothing.
template <bool kCount, uint32_t kShift>
RT_DEV void reuseAndMerge(const DevScene& S, TStack st, uint32_t sampleMode, DIResv& dst, const Surf& dstSurf,
                          const DIResv& src, f3 wo, uint32_t& rng, TraceCounters* ctr) {
    if (kShift == kShiftReconnection) reconnection<kCount>(S, st, dst, dstSurf, src, wo, rng, ctr);
    else if (kShift == kShiftReplay) randomReplay<kCount>(S, st, sampleMode, dst, dstSurf, src, wo, rng, ctr);
}

// Visibility of the final sample (di_temporal.glsl:72-81, di_spatial.glsl:99-108).
template <bool kCount>
RT_DEV void visibilityCheck(const DevScene& S, TStack st, DIResv& resv, f3 pos, TraceCounters* ctr) {
    if (diValid(resv) && diSampleValid(resv.s)) {
        Surf sf; diSurface(S, resv.s.isec, sf);
        if (!traceVisibility<kCount>(S, st, pos, sf.pos, ctr)) resv.resampleWeight = 0.0f;
    }
}

struct DIPixel { f2 uv; CamRay ray; f3 pos; Surf surf; GPix gp; };
RT_DEV bool diPixel(const rt_camera& cam, const GB& g, uint32_t x, uint32_t y, DIPixel& px) {
    px.uv = pixelUV(x, y, g.W, g.H);
    if (!unpackGB(texDNCentre(g, x, y), fetchAM(g, (int)x, (int)y), px.gp)) return false;
    px.ray = pinholeRay(cam, mk2(px.uv.x, 1.0f - px.uv.y));
    px.pos = px.ray.ori + px.ray.dir * (px.gp.depth - 1e-4f);
    px.surf.pos = px.pos; px.surf.norm = px.gp.norm; px.surf.albedo = px.gp.albedo;
    px.surf.matIndex = (uint32_t)(px.gp.matMeshId >> 16); px.surf.isLight = false;
    return true;
}

// ================================================================================================
// di_path_gen.glsl:9-36 (the pass stores the reservoir; its radiance output is not written,
// di_path_gen.comp:28)
template <bool kCount>
__global__ __launch_bounds__(kTileThreads, RT_WAVES_DI) void k_di_path_gen(DevScene S, rt_camera cam, GB g,
                                                                          uint32_t sampleMode, DIPlanes thisR,
                                                                          TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    PIXEL_INDEX
    if (x >= g.W || y >= g.H) return;
    DIPixel px;
    if (!diPixel(cam, g, x, y, px)) return;
    uint32_t rng = makeSeedPix(cam.seed, x, y);
    uint32_t resvRng = ~rng;
    DIResv resv = diZero();
    (void)sampleLi<kCount>(S, laneStack(stk), sampleMode, px.surf, loadMat(S, px.surf.matIndex),
                           -px.ray.dir, rng, resvRng, resv, ctr);
    diStore(thisR, (size_t)y * g.W + x, resv);
}

// di_temporal.glsl:9-88 (the pass stores the temporal reservoir; its radiance output is not
// written, di_temporal.comp:28)
template <bool kCount, uint32_t kShift>
__global__ __launch_bounds__(kTileThreads, RT_WAVES_DI) void k_di_temporal(DevScene S, rt_camera cam, rt_camera prevCam,
                                                                          GB g, GB gPrev, rt_di_settings set,
                                                                          DIPlanes thisR, DIPlanes prevR, DIPlanes tempR,
                                                                          TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    PIXEL_INDEX
    if (x >= g.W || y >= g.H) return;
    TStack st = laneStack(stk);
    DIPixel px;
    if (!diPixel(cam, g, x, y, px)) return;
    const f2 motion = fetchMotion(g, (int)x, (int)y);
    const uint32_t rng = makeSeedPix(cam.seed, x, y) ^ 1u;
    uint32_t resvRng = ~rng;
    const size_t p = (size_t)y * g.W + x;
    DIResv resv = diLoad(thisR, p);
    if (set.temporalReuse) {
        if ((cam.frameIndex & 0x80000000u) == 0u) {
            // findPreviousReservoir (di_temporal.glsl:9-34)
            const f2 puv = px.uv + motion;
            // a non-finite reprojection (half overflow of the motion vector) is rejected
            if (!(puv.x < 0.0f || puv.y < 0.0f || puv.x > 1.0f || puv.y > 1.0f || isnan(puv.x) || isnan(puv.y))) {
                const int qx = (int)(puv.x * (float)cam.filmSize[0]);
                const int qy = (int)(puv.y * (float)cam.filmSize[1]);
                GPix pp;
                if (unpackGB(texDN(gPrev, puv), fetchAM(gPrev, qx, qy), pp)) {
                    const CamRay pr = pinholeRay(prevCam, mk2(puv.x, 1.0f - puv.y));
                    const f3 posPrev = pr.ori + pr.dir * (pp.depth - 1e-4f);
                    if (!(pp.matMeshId != px.gp.matMeshId || dot(pp.norm, px.gp.norm) < 0.95f ||
                          distance(px.pos, posPrev) > 0.5f)) {
                        const DIResv tr = diAt(prevR, cam.filmSize[0], cam.filmSize[1], qx, qy);
                        if (diValid(tr))
                            reuseAndMerge<kCount, kShift>(S, st, set.sampleType, resv, px.surf, tr, -px.ray.dir, resvRng,
                                                          ctr);
                    }
                }
            }
        }
        visibilityCheck<kCount>(S, st, resv, px.pos, ctr);
    }
    diCap(resv, kDICap);
    if (!diValid(resv)) diReset(resv);
    diStore(tempR, p, resv);
}

// di_spatial.glsl:9-119 and the accumulation of di_spatial.comp:21-27
template <bool kCount, uint32_t kShift>
__global__ __launch_bounds__(kTileThreads, RT_WAVES_DI) void k_di_spatial(DevScene S, rt_camera cam, GB g,
                                                                         rt_di_settings set, DIPlanes tempR,
                                                                         DIPlanes thisR, float* directOut,
                                                                         TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    PIXEL_INDEX
    if (x >= g.W || y >= g.H) return;
    TStack st = laneStack(stk);
    const size_t p = (size_t)y * g.W + x;
    f3 radiance = s3(0.0f);
    DIPixel px;
    if (diPixel(cam, g, x, y, px)) {
        uint32_t rng = makeSeedPix(cam.seed, x, y) ^ 2u;
        const f3 wo = -px.ray.dir;
        DIResv resv = diLoad(tempR, p);
        if (set.spatialReuse) {
            const f2 texel = mk2(1.0f / (float)g.W, 1.0f / (float)g.H);
            for (uint32_t i = 0; i < kDINeighbors; i++) {
                const f2 nb = px.uv + (toConcentricDisk(sample2f(rng)) * kDIRadius) * texel;
                // findNeighborReservoir (di_spatial.glsl:9-34)
                if (nb.x < 0.0f || nb.y < 0.0f || nb.x > 1.0f || nb.y > 1.0f) continue;
                const int qx = (int)(nb.x * (float)cam.filmSize[0]);
                const int qy = (int)(nb.y * (float)cam.filmSize[1]);
                GPix np;
                if (!unpackGB(texDN(g, nb), fetchAM(g, qx, qy), np)) continue;
                const CamRay nr = pinholeRay(cam, mk2(nb.x, 1.0f - nb.y));
                const f3 posN = nr.ori + nr.dir * (np.depth - 1e-4f);
                if (dot(np.norm, px.gp.norm) < 0.9f || distance(px.pos, posN) > 0.4f) continue;
                const DIResv nres = diAt(tempR, cam.filmSize[0], cam.filmSize[1], qx, qy);
                if (diValid(nres))
                    reuseAndMerge<kCount, kShift>(S, st, set.sampleType, resv, px.surf, nres, wo, rng, ctr);
            }
            visibilityCheck<kCount>(S, st, resv, px.pos, ctr);
        }
        diCap(resv, kDICap);
        if (!diValid(resv)) diReset(resv);
        diStore(thisR, p, resv);
        if (diValid(resv) && diSampleValid(resv.s)) {
            Surf sf; diSurface(S, resv.s.isec, sf);
            const f3 wi = normalize(sf.pos - px.pos);
            if (resv.sampleCount > 0u) {
                const Mat mat = loadMat(S, px.surf.matIndex);
                const f3 Li = resv.s.Li * evalBSDF(mat, px.gp.albedo, px.gp.norm, wo, wi) * satDot(px.gp.norm, wi) /
                              resv.s.samplePdf;
                if (!isBlack(Li)) radiance = Li / luminance(Li) * resv.resampleWeight / (float)resv.sampleCount;
            }
        }
        radiance = clampColor(radiance);
    }
    accumulate(directOut, p, radiance, cam.frameIndex);
}

__global__ void k_di_reset(DIPlanes R) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R.n) return;
    DIResv r = diZero();
    r.s.isec.inst = RT_INVALID_HIT;
    diStore(R, i, r);
}

inline DIPlanes diPlanes(void* p, uint32_t W, uint32_t H) {
    DIPlanes r; r.p = reinterpret_cast<float4*>(p); r.n = (size_t)W * H; return r;
}

}  // namespace

// ===================================================================================================
// Launchers (called by the C ABI after argument checks)
namespace rti {

// Counting or plain build of a kernel templated on the shift mapping.
#define RT_DI_SHIFT_CASE(KERNEL, SHIFT, GRID, STREAM, ...)                                                      \
    (gCountersEnabled ? launch(KERNEL<true, SHIFT>, GRID, kTileThreads, STREAM, __VA_ARGS__, counterBuffer())     \
                      : launch(KERNEL<false, SHIFT>, GRID, kTileThreads, STREAM, __VA_ARGS__, nullptr))
#define RT_DI_DISPATCH(KERNEL, SHIFT, GRID, STREAM, ...)                                                        \
    do {                                                                                                           \
        if ((SHIFT) == 0u) (void)RT_DI_SHIFT_CASE(KERNEL, 0u, GRID, STREAM, __VA_ARGS__);                          \
        else if ((SHIFT) == 1u) (void)RT_DI_SHIFT_CASE(KERNEL, 1u, GRID, STREAM, __VA_ARGS__);                     \
        else (void)RT_DI_SHIFT_CASE(KERNEL, 2u, GRID, STREAM, __VA_ARGS__);                                        \
    } while (0)

hipError_t launchDiPathGen(const rt_device_scene* s, const rt_camera* cams, const rt_frame& f,
                           const rt_di_reservoirs& r, const rt_di_settings& set, hipStream_t stream) {
    GB g = makeGB(f.depthNormal, f.albedoMatId, f.motionVector, f.width, f.height);
    if (gCountersEnabled)
        (void)launch(k_di_path_gen<true>, tileGrid(f.width, f.height), kTileThreads, stream, s->dev, cams[0], g,
                     set.sampleType, diPlanes(r.reservoir, f.width, f.height), counterBuffer());
    else
        (void)launch(k_di_path_gen<false>, tileGrid(f.width, f.height), kTileThreads, stream, s->dev, cams[0], g,
                     set.sampleType, diPlanes(r.reservoir, f.width, f.height), nullptr);
    return hipGetLastError();
}

hipError_t launchDiTemporal(const rt_device_scene* s, const rt_camera* cams, const rt_frame& f,
                            const rt_di_reservoirs& r, const rt_di_settings& set, hipStream_t stream) {
    GB g = makeGB(f.depthNormal, f.albedoMatId, f.motionVector, f.width, f.height);
    GB gp = makeGB(f.depthNormalPrev, f.albedoMatIdPrev, nullptr, f.width, f.height);
    RT_DI_DISPATCH(k_di_temporal, set.shiftType, tileGrid(f.width, f.height), stream, s->dev, cams[0], cams[1], g, gp,
                   set, diPlanes(r.reservoir, f.width, f.height), diPlanes(r.reservoirPrev, f.width, f.height),
                   diPlanes(r.reservoirTemp, f.width, f.height));
    return hipGetLastError();
}

hipError_t launchDiSpatial(const rt_device_scene* s, const rt_camera* cams, const rt_frame& f,
                           const rt_di_reservoirs& r, const rt_di_settings& set, hipStream_t stream) {
    GB g = makeGB(f.depthNormal, f.albedoMatId, f.motionVector, f.width, f.height);
    RT_DI_DISPATCH(k_di_spatial, set.shiftType, tileGrid(f.width, f.height), stream, s->dev, cams[0], g, set,
                   diPlanes(r.reservoirTemp, f.width, f.height), diPlanes(r.reservoir, f.width, f.height),
                   f.directOutput);
    return hipGetLastError();
}

hipError_t launchDiReset(void* resv, uint32_t W, uint32_t H, hipStream_t stream) {
    const DIPlanes R = diPlanes(resv, W, H);
    if (R.n == 0) return hipSuccess;
    (void)launch(k_di_reset, dim3((unsigned)((R.n + 255) / 256)), dim3(256), stream, R);
    return hipGetLastError();
}

}  // namespace rti
