// ORACLE — test infrastructure only. Nothing in the product links or calls this.
//
// ReSTIR DI ("TestReSTIR", src/TestReSTIR.cpp:9-36) restated from the
// reference GLSL: path generation (di_path_gen.glsl:9-36), temporal reuse
// (di_temporal.glsl:9-88) and spatial reuse with shading (di_spatial.glsl:9-119)
// over the reservoir operations and shift mappings of di_reservoir.glsl.
// Uninitialised GLSL locals are taken as zero. A sample reference read from a
// reservoir that lies outside the scene loads a zero surface (the reference
// reads such entries through robust buffer access).
#include "orc_passes.h"
#include <cstring>

namespace orc {

namespace {

const uint32_t DICap = 40;             // DIReservoirCapSample(resv, 40)
const uint32_t DIResampleNum = 10;     // di_spatial.glsl:72
const float DIResampleRadius = 20.0f;  // di_spatial.glsl:73

void setV(float* d, vec3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }
DIReservoir zeroResv() { DIReservoir r; std::memset(&r, 0, sizeof(r)); return r; }

// ---- di_reservoir.glsl:28-77 -----------------------------------------------------
void pathSampleInit(DIPathSample& p) { p.isec.instanceIdx = InvalidHitIndex; }
bool pathSampleIsValid(const DIPathSample& p) { return p.isec.instanceIdx != InvalidHitIndex; }
void reservoirReset(DIReservoir& r) { r.sampleCount = 0; r.resampleWeight = 0.0f; r.contribWeight = 0.0f; }
bool reservoirIsValid(const DIReservoir& r) { return !std::isnan(r.resampleWeight); }
void reservoirResetIfInvalid(DIReservoir& r) { if (!reservoirIsValid(r)) reservoirReset(r); }
void reservoirAddSample(DIReservoir& r, const DIPathSample& s, float w, float u) {
    r.resampleWeight += w;
    r.sampleCount++;
    if (u * r.resampleWeight < w) { r.pathSample = s; r.weight = w; }
}
void reservoirMerge(DIReservoir& r, const DIReservoir& rhs, float u) {
    r.resampleWeight += rhs.resampleWeight;
    r.sampleCount += rhs.sampleCount;
    if (u * r.resampleWeight < rhs.resampleWeight) { r.pathSample = rhs.pathSample; r.weight = rhs.weight; }
}
void reservoirCapSample(DIReservoir& r, uint32_t cap) {
    if (r.sampleCount > cap) {
        r.resampleWeight *= (float)cap / (float)r.sampleCount;
        r.sampleCount = cap;
    }
}

// loadSurfaceInfo of a stored sample reference, range-checked against the scene.
void sampleSurface(const SceneArrays& S, const Intersection& isec, SurfaceInfo& sf) {
    bool ok;
    if (isec.instanceIdx == 0) ok = isec.triangleIdx < S.numLights;
    else ok = isec.instanceIdx - 1 < S.numInstances &&
              isec.triangleIdx < S.instances[isec.instanceIdx - 1].indexCount / 3;
    if (ok) { loadSurfaceInfo(S, isec, sf); return; }
    sf = SurfaceInfo{V3s(0.0f), V3s(0.0f), V3s(0.0f), 0u, isec.instanceIdx == 0};
}

// sampleLi, di_reservoir.glsl:79-185
vec3 sampleLi(const Ctx& c, uint32_t sampleMode, const SurfaceInfo& surf, const Material& mat, vec3 wo, uint32_t rng,
              uint32_t& resvRng, DIReservoir& resv) {
    const SceneArrays& S = *c.S;
    vec3 radiance = V3s(0.0f);
    DIPathSample pathSample; std::memset(&pathSample, 0, sizeof(pathSample));
    pathSample.rng = rng;
    vec4 lightRandSample = sample4f(rng);
    vec3 scatterRandSample = sample3f(rng);
    if (sampleMode != DISampleBSDF && !isBSDFDelta(mat)) {
        vec3 lightDir; vec2 lightBary; float lightDist, lightPdf, lightJacobian; uint32_t lightId;
        vec3 lightRadiance = sampleLightByPower(S, surf.pos, lightDir, lightDist, lightPdf, lightJacobian, lightBary,
                                                lightId, lightRandSample);
        bool shadowed = traceShadow(c, surf.pos, MinRayDistance, lightDir, lightDist - MinRayDistance);
        if (!shadowed && lightPdf > 1e-6f) {
            float bsdfPdf = evalPdf(mat, surf.norm, wo, lightDir);
            float weight = MISWeight(lightPdf, bsdfPdf);
            if (sampleMode == DISampleLight) weight = 1.0f;
            vec3 contrib = lightRadiance * evalBSDF(mat, surf.albedo, surf.norm, wo, lightDir) *
                           satDot(surf.norm, lightDir) / lightPdf * weight;
            float sampleWeight = luminance(contrib);
            if (std::isnan(sampleWeight) || sampleWeight < 0.0f) sampleWeight = 0.0f;
            pathSample.isec.bary[0] = lightBary.x; pathSample.isec.bary[1] = lightBary.y;
            pathSample.isec.instanceIdx = 0; pathSample.isec.triangleIdx = lightId;
            setV(pathSample.Li, lightRadiance * weight);
            pathSample.jacobian = lightJacobian;
            pathSample.samplePdf = lightPdf;
            pathSample.isLightSample = 1;
            reservoirAddSample(resv, pathSample, sampleWeight, sample1f(resvRng));
            radiance = radiance + contrib;
        }
    }
    BSDFSample s{};
    pathSample.isLightSample = 0;
    if (sampleMode != DISampleLight && sampleBSDF(mat, surf.albedo, surf.norm, wo, scatterRandSample, s) &&
        s.pdf > 1e-6f) {
        Intersection isec = traceClosestHit(c, surf.pos, MinRayDistance, s.wi, MaxRayDistance);
        if (isec.instanceIdx != InvalidHitIndex) {
            SurfaceInfo hit; loadSurfaceInfo(S, isec, hit);
            float cosTheta = -dot(s.wi, hit.norm);
            if (hit.isLight && cosTheta > 0.0f) {
                float dist = length(hit.pos - surf.pos);
                float sumPower = S.lightTable[0].prob;
                float lightPdf = luminance(hit.albedo) / sumPower * dist * dist / std::fabs(cosTheta);
                float weight = MISWeight(s.pdf, lightPdf);
                if (sampleMode == DISampleBSDF || isSampleTypeDelta(s.type)) weight = 1.0f;
                float cosTerm = isSampleTypeDelta(s.type) ? 1.0f : satDot(surf.norm, s.wi);
                vec3 contrib = hit.albedo * s.bsdf * cosTerm / s.pdf * weight;
                pathSample.isec = isec;
                setV(pathSample.Li, hit.albedo * weight);
                pathSample.jacobian = std::fabs(cosTheta) / square(dist);
                pathSample.samplePdf = s.pdf;
                pathSample.isLightSample = 0;
                // resampled with luminance(contrib) itself, not the sanitised weight (di_reservoir.glsl:167)
                reservoirAddSample(resv, pathSample, luminance(contrib), sample1f(resvRng));
                radiance = radiance + contrib;
            }
        }
    }
    reservoirResetIfInvalid(resv);
    if (resv.sampleCount > 0 && pathSampleIsValid(resv.pathSample) && resv.weight > 0.0f) {
        setV(resv.pathSample.Li, L3(resv.pathSample.Li) * (resv.resampleWeight / resv.weight));
        resv.weight = resv.resampleWeight;
    } else {
        pathSampleInit(resv.pathSample);
        resv.weight = 0.0f;
        resv.resampleWeight = 0.0f;
    }
    resv.sampleCount = 1;
    return radiance;
}

// randomReplay, di_reservoir.glsl:187-221
void randomReplay(const Ctx& c, const DISettings& set, DIReservoir& dstResv, const SurfaceInfo& dstSurf,
                  const DIReservoir& srcResv, vec3 wo, uint32_t& rng) {
    const SceneArrays& S = *c.S;
    const Material& dstMat = S.materials[dstSurf.matIndex];
    DIReservoir replayResv = zeroResv();
    reservoirReset(replayResv);
    vec3 Li = sampleLi(c, set.sampleType, dstSurf, dstMat, wo, srcResv.pathSample.rng, rng, replayResv);
    float jacobian = 1.0f;
    DIPathSample replaySample = replayResv.pathSample;
    if (pathSampleIsValid(replaySample)) {
        SurfaceInfo replaySurf; sampleSurface(S, replaySample.isec, replaySurf);
        vec3 wi = normalize(replaySurf.pos - dstSurf.pos);
        Li = L3(replaySample.Li) * evalBSDF(dstMat, dstSurf.albedo, dstSurf.norm, wo, wi) * satDot(dstSurf.norm, wi) /
             replayResv.pathSample.samplePdf;
        float dstPHat = luminance(Li * jacobian);
        replayResv.resampleWeight = srcResv.resampleWeight * dstPHat / srcResv.weight;
        replayResv.sampleCount = srcResv.sampleCount;
    } else {
        replayResv.resampleWeight = 0.0f;
    }
    if (reservoirIsValid(replayResv)) reservoirMerge(dstResv, replayResv, sample1f(rng));
}

// reconnection, di_reservoir.glsl:223-281
void reconnection(const Ctx& c, DIReservoir& dstResv, const SurfaceInfo& dstSurf, DIReservoir srcResv, vec3 wo,
                  uint32_t& rng) {
    const SceneArrays& S = *c.S;
    const Material& dstMat = S.materials[dstSurf.matIndex];
    DIPathSample srcSample = srcResv.pathSample;
    SurfaceInfo rcSurf; sampleSurface(S, srcSample.isec, rcSurf);
    float dist = distance(rcSurf.pos, dstSurf.pos);
    vec3 wi = normalize(rcSurf.pos - dstSurf.pos);
    float cosTheta = -dot(rcSurf.norm, wi);
    float dstJacobian = std::fabs(cosTheta) / square(dist);
    float jacobian = dstJacobian / srcSample.jacobian;
    bool srcSampleValid = false;
    float dstPHat = 0.0f, dstSamplePdf = 0.0f;
    if (pathSampleIsValid(srcSample) && dist > 1e-4f) {
        if (cosTheta > 0.0f && !std::isnan(jacobian) && srcSample.jacobian > 0.0f) {
            if (traceVisibility(c, dstSurf.pos, rcSurf.pos)) {
                srcSampleValid = true;
                if (!std::isnan(srcSample.samplePdf) && srcSample.samplePdf > 1e-6f) {
                    vec3 Li = L3(srcResv.pathSample.Li) * evalBSDF(dstMat, dstSurf.albedo, dstSurf.norm, wo, wi) *
                              satDot(dstSurf.norm, wi) / srcSample.samplePdf;
                    dstPHat = luminance(Li * jacobian);
                }
                if (srcSample.isLightSample) {
                    float sumPower = S.lightTable[0].prob;
                    dstSamplePdf = luminance(rcSurf.albedo) / sumPower / dstJacobian;
                } else {
                    dstSamplePdf = evalPdf(dstMat, dstSurf.norm, wo, wi);
                }
            }
        }
    }
    if (srcSampleValid) {
        srcSample.jacobian = dstJacobian;
        srcSample.samplePdf = dstSamplePdf;
        if (srcSample.samplePdf < 1e-6f || std::isnan(srcSample.samplePdf)) srcSample.samplePdf = 0.0f;
        srcResv.pathSample = srcSample;
        srcResv.resampleWeight *= dstPHat / srcResv.weight;
        if (std::isnan(srcResv.resampleWeight)) srcResv.resampleWeight = 0.0f;
    } else {
        srcResv.resampleWeight = 0.0f;
    }
    if (reservoirIsValid(srcResv)) reservoirMerge(dstResv, srcResv, sample1f(rng));
}

// DIReservoirReuseAndMerge, di_reservoir.glsl:283-290 (Hybrid merges nothing)
void reuseAndMerge(const Ctx& c, const DISettings& set, DIReservoir& dstResv, const SurfaceInfo& dstSurf,
                   const DIReservoir& srcResv, vec3 wo, uint32_t& rng) {
    if (set.shiftType == ShiftReconnection) reconnection(c, dstResv, dstSurf, srcResv, wo, rng);
    else if (set.shiftType == ShiftReplay) randomReplay(c, set, dstResv, dstSurf, srcResv, wo, rng);
}

// di_temporal.glsl:72-81 / di_spatial.glsl:99-108
void visibilityCheck(const Ctx& c, DIReservoir& resv, vec3 pos) {
    if (reservoirIsValid(resv) && pathSampleIsValid(resv.pathSample)) {
        SurfaceInfo surf; sampleSurface(*c.S, resv.pathSample.isec, surf);
        if (!traceVisibility(c, pos, surf.pos)) resv.resampleWeight = 0.0f;
    }
}

DIReservoir reservoirAt(const DIReservoir* buf, uint32_t W, uint32_t H, int x, int y) {
    uint64_t idx = (uint64_t)W * (uint32_t)y + (uint32_t)x;
    if (x < 0 || y < 0 || idx >= (uint64_t)W * H) return zeroResv();
    return buf[idx];
}

}  // namespace

// di_path_gen.glsl:9-36
bool diPathGen(const Ctx& c, const Camera& cam, const GBufView& g, const DISettings& set, uint32_t x, uint32_t y,
               DIReservoir& out) {
    const SceneArrays& S = *c.S;
    vec2 uv = V2(((float)x + 0.5f) / (float)g.W, ((float)y + 0.5f) / (float)g.H);
    float depth; vec3 norm, albedo; int matMeshId;
    uint32_t am[2]; fetchAlbedoMatId(g, (int)x, (int)y, am);
    if (!unpackGBuffer(texDepthNormal(g, uv), am, depth, norm, albedo, matMeshId)) return false;
    int matId = matMeshId >> 16;
    Ray ray = pinholeCameraSampleRay(cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    uint32_t rng = makeSeedPix(cam.seed, x, y);
    uint32_t resvRng = ~rng;
    vec3 pos = ray.ori + ray.dir * (depth - 1e-4f);
    vec3 wo = -ray.dir;
    const Material& mat = S.materials[matId];
    SurfaceInfo surf{pos, norm, albedo, (uint32_t)matId, false};
    DIReservoir resv = zeroResv();
    reservoirReset(resv);
    sampleLi(c, set.sampleType, surf, mat, wo, rng, resvRng, resv);
    out = resv;
    return true;
}

// di_temporal.glsl:9-88
bool diTemporal(const Ctx& c, const Camera& cam, const Camera& prevCam, const GBufView& g, const GBufView& gPrev,
                const DISettings& set, const DIReservoir* resvThis, const DIReservoir* resvPrev, uint32_t x,
                uint32_t y, DIReservoir& out) {
    vec2 uv = V2(((float)x + 0.5f) / (float)g.W, ((float)y + 0.5f) / (float)g.H);
    float depth; vec3 norm, albedo; int matMeshId;
    uint32_t am[2]; fetchAlbedoMatId(g, (int)x, (int)y, am);
    if (!unpackGBuffer(texDepthNormal(g, uv), am, depth, norm, albedo, matMeshId)) return false;
    vec2 motion = fetchMotion(g, (int)x, (int)y);
    int matId = matMeshId >> 16;
    Ray ray = pinholeCameraSampleRay(cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    uint32_t rng = makeSeedPix(cam.seed, x, y) ^ 1u;
    uint32_t resvRng = ~rng;
    vec3 pos = ray.ori + ray.dir * (depth - 1e-4f);
    vec3 wo = -ray.dir;
    DIReservoir resv = resvThis[index1D(g.W, x, y)];
    if (set.temporalReuse) {
        DIReservoir temporalResv = zeroResv();
        reservoirReset(temporalResv);
        SurfaceInfo dstSurf{pos, norm, albedo, (uint32_t)matId, false};
        bool found = false;
        if ((cam.frameIndex & CameraClearFlag) == 0) {
            // findPreviousReservoir (di_temporal.glsl:9-34)
            vec2 puv = uv + motion;
            // GLSL leaves ivec2(NaN) undefined: a non-finite reprojection is rejected
            if (!(puv.x < 0.0f || puv.y < 0.0f || puv.x > 1.0f || puv.y > 1.0f) && !std::isnan(puv.x) &&
                !std::isnan(puv.y)) {
                int px = (int)(puv.x * (float)cam.filmSize[0]);
                int py = (int)(puv.y * (float)cam.filmSize[1]);
                float depthPrev; vec3 normalPrev, albedoPrev; int matMeshIdPrev;
                uint32_t amp[2]; fetchAlbedoMatId(gPrev, px, py, amp);
                if (unpackGBuffer(texDepthNormal(gPrev, puv), amp, depthPrev, normalPrev, albedoPrev, matMeshIdPrev)) {
                    Ray pr = pinholeCameraSampleRay(prevCam, V2(puv.x, 1.0f - puv.y), V2(0, 0));
                    vec3 posPrev = pr.ori + pr.dir * (depthPrev - 1e-4f);
                    if (!(matMeshIdPrev != matMeshId || dot(normalPrev, norm) < 0.95f || distance(pos, posPrev) > 0.5f)) {
                        temporalResv = reservoirAt(resvPrev, cam.filmSize[0], cam.filmSize[1], px, py);
                        found = true;
                    }
                }
            }
        }
        if (found && reservoirIsValid(temporalResv)) reuseAndMerge(c, set, resv, dstSurf, temporalResv, wo, resvRng);
        visibilityCheck(c, resv, pos);
    }
    reservoirCapSample(resv, DICap);
    reservoirResetIfInvalid(resv);
    out = resv;
    return true;
}

// di_spatial.glsl:9-119
vec3 diSpatial(const Ctx& c, const Camera& cam, const GBufView& g, const DISettings& set,
               const DIReservoir* resvTemp, uint32_t x, uint32_t y, DIReservoir& outThis, bool& written) {
    written = false;
    const SceneArrays& S = *c.S;
    vec2 uv = V2(((float)x + 0.5f) / (float)g.W, ((float)y + 0.5f) / (float)g.H);
    float depth; vec3 norm, albedo; int matMeshId;
    uint32_t am[2]; fetchAlbedoMatId(g, (int)x, (int)y, am);
    if (!unpackGBuffer(texDepthNormal(g, uv), am, depth, norm, albedo, matMeshId)) return V3s(0.0f);
    int matId = matMeshId >> 16;
    Ray ray = pinholeCameraSampleRay(cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    uint32_t rng = makeSeedPix(cam.seed, x, y) ^ 2u;
    vec3 pos = ray.ori + ray.dir * (depth - 1e-4f);
    vec3 wo = -ray.dir;
    const Material& mat = S.materials[matId];
    vec3 radiance = V3s(0.0f);
    DIReservoir resv = resvTemp[index1D(g.W, x, y)];
    if (set.spatialReuse) {
        vec2 texelSize = V2(1.0f / (float)g.W, 1.0f / (float)g.H);
        SurfaceInfo dstSurf{pos, norm, albedo, (uint32_t)matId, false};
        for (uint32_t i = 0; i < DIResampleNum; i++) {
            vec2 neighbor = uv + (toConcentricDisk(sample2f(rng)) * DIResampleRadius) * texelSize;
            // findNeighborReservoir (di_spatial.glsl:9-34)
            if (neighbor.x < 0.0f || neighbor.y < 0.0f || neighbor.x > 1.0f || neighbor.y > 1.0f) continue;
            int px = (int)(neighbor.x * (float)cam.filmSize[0]);
            int py = (int)(neighbor.y * (float)cam.filmSize[1]);
            float depthN; vec3 normalN, albedoN; int matMeshIdN;
            uint32_t amn[2]; fetchAlbedoMatId(g, px, py, amn);
            if (!unpackGBuffer(texDepthNormal(g, neighbor), amn, depthN, normalN, albedoN, matMeshIdN)) continue;
            Ray nr = pinholeCameraSampleRay(cam, V2(neighbor.x, 1.0f - neighbor.y), V2(0, 0));
            vec3 posN = nr.ori + nr.dir * (depthN - 1e-4f);
            if (dot(normalN, norm) < 0.9f || distance(pos, posN) > 0.4f) continue;
            DIReservoir neighborResv = reservoirAt(resvTemp, cam.filmSize[0], cam.filmSize[1], px, py);
            if (reservoirIsValid(neighborResv)) reuseAndMerge(c, set, resv, dstSurf, neighborResv, wo, rng);
        }
        visibilityCheck(c, resv, pos);
    }
    reservoirCapSample(resv, DICap);
    reservoirResetIfInvalid(resv);
    outThis = resv;
    written = true;
    if (reservoirIsValid(resv) && pathSampleIsValid(resv.pathSample)) {
        const DIPathSample& ps = resv.pathSample;
        SurfaceInfo surf; sampleSurface(S, ps.isec, surf);
        vec3 wi = normalize(surf.pos - pos);
        if (resv.sampleCount > 0) {
            vec3 Li = L3(ps.Li) * evalBSDF(mat, albedo, norm, wo, wi) * satDot(norm, wi) / ps.samplePdf;
            if (!isBlack(Li)) radiance = Li / luminance(Li) * resv.resampleWeight / (float)resv.sampleCount;
        }
    }
    return clampColor(radiance);
}

}  // namespace orc
