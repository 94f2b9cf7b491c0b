// ReSTIR GI temporal reuse for gfx950: gi_resample_temporal.glsl:34-213 with
// the accumulation of gi_resample_temporal.comp:14-27 (the reference's
// indirect method "ReSTIR GI", src/Renderer.cpp:470-471). One pass per frame:
// trace a path from the primary vertex (reconnection vertex = the first
// secondary hit; radiance leaving it gathered with NEE + MIS), merge it into
// the reprojected previous reservoir, cap M at 40 and shade the reservoir's
// sample through a visibility ray.
//
// GI reservoirs (layouts.glsl:117-129, 48 B) live in HBM as three float4
// planes, plane k holding the k-th 16 bytes of the reference struct:
//   P0 rcIsec (bary.xy, instance, primitive)   P1 rcLo.xyz, rcPrevCoord
//   P2 sampleCount (uint), resampleWeight, contribWeight, pad0
// Uninitialised GLSL locals (the primary scatter term and pdf when the path
// ends at the primary vertex) are zero. A sample reference read from reservoir
// memory is bounds-checked against the scene before its surface is loaded.
#include "rt_dpixel.h"

using namespace rtd;
using namespace rtp;

namespace {

#ifndef RT_WAVES_GI
#define RT_WAVES_GI 4
#endif

constexpr int kGIMaxDepth = 15;      // MaxTracingDepth, gi_resample_temporal.glsl:37
constexpr uint32_t kGICap = 40;      // GIReservoirCapSample(resv, 40), :188

struct GIResv { Isec isec; f3 rcLo; uint32_t rcPrevCoord; uint32_t sampleCount; float resampleWeight, contribWeight, pad0; };
struct GIPlanes { float4* p; size_t n; };

RT_DEV GIResv giZero() {
    GIResv r;
    r.isec.b1 = 0.0f; r.isec.b2 = 0.0f; r.isec.inst = 0u; r.isec.prim = 0u;
    r.rcLo = s3(0.0f); r.rcPrevCoord = 0u; r.sampleCount = 0u; r.resampleWeight = 0.0f; r.contribWeight = 0.0f;
    r.pad0 = 0.0f;
    return r;
}
RT_DEV GIResv giLoad(const GIPlanes& R, size_t i) {
    const float4 a = R.p[i], b = R.p[R.n + i], c = R.p[2 * R.n + i];
    GIResv r;
    r.isec.b1 = a.x; r.isec.b2 = a.y; r.isec.inst = __float_as_uint(a.z); r.isec.prim = __float_as_uint(a.w);
    r.rcLo = mk3(b.x, b.y, b.z); r.rcPrevCoord = __float_as_uint(b.w);
    r.sampleCount = __float_as_uint(c.x); r.resampleWeight = c.y; r.contribWeight = c.z; r.pad0 = c.w;
    return r;
}
RT_DEV void giStore(const GIPlanes& R, size_t i, const GIResv& r) {
    R.p[i] = make_float4(r.isec.b1, r.isec.b2, __uint_as_float(r.isec.inst), __uint_as_float(r.isec.prim));
    R.p[R.n + i] = make_float4(r.rcLo.x, r.rcLo.y, r.rcLo.z, __uint_as_float(r.rcPrevCoord));
    R.p[2 * R.n + i] = make_float4(__uint_as_float(r.sampleCount), r.resampleWeight, r.contribWeight, r.pad0);
}

// gi_reservoir.glsl:34-56
RT_DEV void giReset(GIResv& r) { r.sampleCount = 0u; r.resampleWeight = 0.0f; r.contribWeight = 0.0f; }
RT_DEV bool giValid(const GIResv& r) { return !isnan(r.resampleWeight) && r.resampleWeight >= 0.0f; }

// loadSurfaceInfo of a stored sample reference, bounds-checked against the scene.
RT_DEV void giSurface(const DevScene& S, const Isec& is, Surf& sf) {
    const bool ok = is.inst == 0u ? is.prim < S.numLights
                                  : (is.inst - 1u < S.numInstances && is.prim < S.instances[is.inst - 1u].triCount);
    if (ok) { loadSurfaceInfo(S, is, sf); return; }
    sf.pos = s3(0.0f); sf.norm = s3(0.0f); sf.albedo = s3(0.0f); sf.matIndex = 0u; sf.isLight = is.inst == 0u;
}

template <bool kCount>
__global__ __launch_bounds__(kTileThreads, RT_WAVES_GI) void k_gi_temporal(DevScene S, rt_camera cam, rt_camera prevCam,
                                                                          GB g, GB gPrev, GIPlanes thisR, GIPlanes prevR,
                                                                          float* indirectOut, TraceCounters* ctr) {
    __shared__ StackLDS<kTileThreads> stk;
    PIXEL_INDEX
    if (x >= g.W || y >= g.H) return;
    TStack st = laneStack(stk);
    const size_t p = (size_t)y * g.W + x;
    const f2 uv = pixelUV(x, y, g.W, g.H);
    GPix gp;
    f3 radiance = s3(0.0f);
    if (!unpackGB(texDNCentre(g, x, y), fetchAM(g, (int)x, (int)y), gp)) {
        // GIReservoirReset(uGIReservoir[index]) in place: the path sample is kept
        GIResv r = giLoad(thisR, p);
        giReset(r);
        giStore(thisR, p, r);
        accumulate(indirectOut, p, radiance, cam.frameIndex);
        return;
    }
    const f2 motion = fetchMotion(g, (int)x, (int)y);
    const int matId = gp.matMeshId >> 16;
    CamRay ray = pinholeRay(cam, mk2(uv.x, 1.0f - uv.y));
    uint32_t rng = makeSeedPix(cam.seed, x, y);
    f3 throughputAfter = s3(1.0f), lastPos = s3(0.0f);
    f3 wo = -ray.dir;
    Surf surf;
    surf.pos = ray.ori + ray.dir * (gp.depth - 1e-4f);
    surf.norm = gp.norm; surf.albedo = gp.albedo; surf.isLight = false; surf.matIndex = (uint32_t)matId;
    GIResv ps = giZero();   // the path sample (rcIsec, rcLo, rcPrevCoord)
    ps.isec.inst = RT_INVALID_HIT;
    ps.rcPrevCoord = (y << 16) | x;
    const f3 primaryPos = surf.pos, primaryWo = -ray.dir;
    f3 primaryScatter = s3(0.0f);
    float primaryPdf = 0.0f;
    Mat mat = loadMat(S, (uint32_t)matId);
    BSDFSample s; s.type = 0u; s.pdf = 0.0f; s.wi = s3(0.0f); s.bsdf = s3(0.0f);
    f3 rdir = ray.dir, rori = ray.ori;
    for (int bounce = 0; bounce < kGIMaxDepth; bounce++) {
        if (bounce > 0) {
            const Isec is = traceClosestHit<kCount>(S, st, rori, RT_MIN_RAY, rdir, RT_MAX_RAY, ctr);
            if (is.inst == RT_INVALID_HIT) break;
            loadSurfaceInfo(S, is, surf);
            mat = loadMat(S, surf.matIndex);
            if (bounce == 1 && !surf.isLight) ps.isec = is;
        }
        if (surf.isLight) {
            const float cosTheta = -dot(rdir, surf.norm);
            if (bounce > 1 && cosTheta > 0.0f) {
                float w = 1.0f;
                if (!isSampleTypeDelta(s.type)) {
                    const float dist = length(surf.pos - lastPos);
                    const float lightPdf = luminance(surf.albedo) / S.lightTable[0].prob * dist * dist / fabsf(cosTheta);
                    w = MISWeight(s.pdf, lightPdf);
                }
                ps.rcLo = ps.rcLo + surf.albedo * w * throughputAfter;
            }
            break;
        }
        if (bounce > 0 && !isBSDFDelta(mat)) {
            f3 lightDir; float lightDist, lightPdf, jac; f2 bary; uint32_t id;
            const f3 Lr = sampleLightByPower(S, surf.pos, lightDir, lightDist, lightPdf, jac, bary, id, sample4f(rng));
            const bool shadowed = traceShadow<kCount>(S, st, surf.pos, RT_MIN_RAY, lightDir, lightDist - RT_MIN_RAY, ctr);
            if (!shadowed && lightPdf > 1e-6f) {
                const float bsdfPdf = absDot(surf.norm, lightDir) * RT_PI_INV;
                const float w = MISWeight(lightPdf, bsdfPdf);
                ps.rcLo = ps.rcLo + Lr * evalBSDF(mat, surf.albedo, surf.norm, wo, lightDir) * satDot(surf.norm, lightDir) /
                                        lightPdf * w * throughputAfter;
            }
        }
        if (bounce > 4) {
            const float pdfTerminate = gmax(1.0f - luminance(throughputAfter), 0.0f);
            if (sample1f(rng) < pdfTerminate) break;
            throughputAfter = throughputAfter / (1.0f - pdfTerminate);
        }
        if (!sampleBSDF(mat, surf.albedo, surf.norm, wo, sample3f(rng), s) || s.pdf < 1e-6f) break;
        if (bounce == 0) primaryPdf = s.pdf;
        const float cosTheta = isSampleTypeDelta(s.type) ? 1.0f : absDot(surf.norm, s.wi);
        const f3 scatterTerms = s.bsdf * cosTheta / s.pdf;
        if (bounce == 0) {
            primaryScatter = s.bsdf * cosTheta;
            primaryPdf = s.pdf;
        } else {
            throughputAfter = throughputAfter * scatterTerms;
        }
        lastPos = surf.pos;
        wo = -s.wi;
        rdir = s.wi;
        rori = surf.pos + rdir * 1e-4f;
    }
    radiance = ps.rcLo * primaryScatter / primaryPdf;

    GIResv resv = giZero();
    giReset(resv);
    if ((cam.frameIndex & 0x80000000u) == 0u) {
        // findPreviousReservoir (gi_resample_temporal.glsl:9-32)
        const f2 puv = uv + motion;
        // a non-finite reprojection (half overflow of the motion vector) is rejected
        if (!(puv.x < 0.0f || puv.y < 0.0f || puv.x > 1.0f || puv.y > 1.0f || isnan(puv.x) || isnan(puv.y))) {
            const int qx = (int)(puv.x * (float)cam.filmSize[0]);
            const int qy = (int)(puv.y * (float)cam.filmSize[1]);
            GPix pp;
            if (unpackGB(texDN(gPrev, puv), fetchAM(gPrev, qx, qy), pp)) {
                if (!(pp.matMeshId != gp.matMeshId || dot(pp.norm, gp.norm) < 0.9f || fabsf(pp.depth - gp.depth) > 5.0f)) {
                    const unsigned long long idx = (unsigned long long)cam.filmSize[0] * (uint32_t)qy + (uint32_t)qx;
                    if (qx >= 0 && qy >= 0 && idx < (unsigned long long)cam.filmSize[0] * cam.filmSize[1])
                        resv = giLoad(prevR, (size_t)idx);
                    else
                        resv = giZero();
                }
            }
        }
    }
    if (ps.isec.inst != RT_INVALID_HIT) {
        float sampleWeight = luminance(radiance);
        if (isnan(sampleWeight) || sampleWeight < 0.0f || primaryPdf < 1e-6f) sampleWeight = 0.0f;
        // GIReservoirAddSample (gi_reservoir.glsl:49-56)
        const float u = sample1f(rng);
        resv.resampleWeight += sampleWeight;
        resv.sampleCount++;
        if (u * resv.resampleWeight < sampleWeight) { resv.isec = ps.isec; resv.rcLo = ps.rcLo; resv.rcPrevCoord = ps.rcPrevCoord; }
    }
    if (!giValid(resv)) giReset(resv);
    if (resv.sampleCount > kGICap) {
        resv.resampleWeight *= (float)kGICap / (float)resv.sampleCount;
        resv.sampleCount = kGICap;
    }
    const Mat primaryMat = loadMat(S, (uint32_t)matId);
    if (giValid(resv) && resv.sampleCount > 0u && !isBSDFDelta(primaryMat)) {
        Surf rs; giSurface(S, resv.isec, rs);
        const f3 primaryWi = normalize(rs.pos - primaryPos);
        const float weight = resv.resampleWeight / (float)resv.sampleCount;
        const f3 Li = resv.rcLo * evalBSDF(primaryMat, gp.albedo, gp.norm, primaryWo, primaryWi) * satDot(gp.norm, primaryWi);
        if (!isBlack(Li) && traceVisibility<kCount>(S, st, primaryPos, rs.pos, ctr)) radiance = Li / luminance(Li) * weight;
    }
    giStore(thisR, p, resv);
    accumulate(indirectOut, p, clampColor(radiance), cam.frameIndex);
}

__global__ void k_gi_reset(GIPlanes R) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R.n) return;
    GIResv r = giZero();
    r.isec.inst = RT_INVALID_HIT;
    giStore(R, i, r);
}

inline GIPlanes giPlanes(void* p, uint32_t W, uint32_t H) {
    GIPlanes r; r.p = reinterpret_cast<float4*>(p); r.n = (size_t)W * H; return r;
}

}  // namespace

namespace rti {

hipError_t launchGiTemporal(const rt_device_scene* s, const rt_camera* cams, const rt_frame& f,
                            const rt_gi_reservoirs& r, hipStream_t stream) {
    GB g = makeGB(f.depthNormal, f.albedoMatId, f.motionVector, f.width, f.height);
    GB gp = makeGB(f.depthNormalPrev, f.albedoMatIdPrev, nullptr, f.width, f.height);
    if (gCountersEnabled)
        (void)launch(k_gi_temporal<true>, tileGrid(f.width, f.height), dim3(kTileThreads), stream, s->dev, cams[0], cams[1],
                     g, gp, giPlanes(r.reservoir, f.width, f.height), giPlanes(r.reservoirPrev, f.width, f.height),
                     f.indirectOutput, counterBuffer());
    else
        (void)launch(k_gi_temporal<false>, tileGrid(f.width, f.height), dim3(kTileThreads), stream, s->dev, cams[0],
                     cams[1], g, gp, giPlanes(r.reservoir, f.width, f.height),
                     giPlanes(r.reservoirPrev, f.width, f.height), f.indirectOutput, nullptr);
    return hipGetLastError();
}

hipError_t launchGiReset(void* resv, uint32_t W, uint32_t H, hipStream_t stream) {
    const GIPlanes R = giPlanes(resv, W, H);
    if (R.n == 0) return hipSuccess;
    (void)launch(k_gi_reset, dim3((unsigned)((R.n + 255) / 256)), dim3(256), stream, R);
    return hipGetLastError();
}

}  // namespace rti
