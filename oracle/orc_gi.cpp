// ORACLE — test infrastructure only. Nothing in the product links or calls this.
//
// ReSTIR GI temporal reuse restated from gi_resample_temporal.glsl:9-213 and the
// reservoir operations of gi_reservoir.glsl:9-63, with the accumulation of
// gi_resample_temporal.comp:14-27. Uninitialised GLSL locals are taken as zero;
// a sample reference read from a reservoir that lies outside the scene loads a
// zero surface (robust buffer access).
#include "orc_passes.h"
#include <cstring>

namespace orc {

namespace {

void giReset(GIReservoir& r) { r.sampleCount = 0; r.resampleWeight = 0.0f; r.contribWeight = 0.0f; }
bool giValid(const GIReservoir& r) { return !std::isnan(r.resampleWeight) && r.resampleWeight >= 0.0f; }
void setV(float* d, vec3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

void sampleSurface(const SceneArrays& S, const Intersection& isec, SurfaceInfo& sf) {
    bool ok;
    if (isec.instanceIdx == 0) ok = isec.triangleIdx < S.numLights;
    else ok = isec.instanceIdx - 1 < S.numInstances &&
              isec.triangleIdx < S.instances[isec.instanceIdx - 1].indexCount / 3;
    if (ok) { loadSurfaceInfo(S, isec, sf); return; }
    sf = SurfaceInfo{V3s(0.0f), V3s(0.0f), V3s(0.0f), 0u, isec.instanceIdx == 0};
}

}  // namespace

// gi_resample_temporal.glsl:34-213; returns the radiance before clampColor.
vec3 giTemporal(const Ctx& c, const Camera& cam, const Camera& prevCam, const GBufView& g, const GBufView& gPrev,
                GIReservoir* resvThis, const GIReservoir* resvPrev, uint32_t x, uint32_t y) {
    (void)prevCam;
    const SceneArrays& S = *c.S;
    const size_t p = index1D(g.W, x, y);
    vec2 uv = V2(((float)x + 0.5f) / (float)g.W, ((float)y + 0.5f) / (float)g.H);
    float depth; vec3 norm, albedo; int matMeshId;
    uint32_t am[2]; fetchAlbedoMatId(g, (int)x, (int)y, am);
    if (!unpackGBuffer(texDepthNormal(g, uv), am, depth, norm, albedo, matMeshId)) {
        giReset(resvThis[p]);
        return V3s(0.0f);
    }
    vec2 motion = fetchMotion(g, (int)x, (int)y);
    int matId = matMeshId >> 16;
    Ray ray = pinholeCameraSampleRay(cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    uint32_t rng = makeSeedPix(cam.seed, x, y);
    vec3 throughputAfter = V3s(1.0f), lastPos = V3s(0.0f);
    vec3 wo = -ray.dir;
    SurfaceInfo surf{ray.ori + ray.dir * (depth - 1e-4f), norm, albedo, (uint32_t)matId, false};
    GIPathSample pathSample; std::memset(&pathSample, 0, sizeof(pathSample));
    pathSample.rcIsec.instanceIdx = InvalidHitIndex;
    pathSample.rcPrevCoord = (y << 16) | x;
    vec3 primaryPos = surf.pos, primaryWo = -ray.dir;
    vec3 primaryScatter = V3s(0.0f);
    float primaryPdf = 0.0f;
    Material mat = S.materials[matId];
    BSDFSample s{};
    for (int bounce = 0; bounce < 15; bounce++) {
        if (bounce > 0) {
            Intersection isec = traceClosestHit(c, ray.ori, MinRayDistance, ray.dir, MaxRayDistance);
            if (isec.instanceIdx == InvalidHitIndex) break;
            loadSurfaceInfo(S, isec, surf);
            mat = S.materials[surf.matIndex];
            if (bounce == 1 && !surf.isLight) pathSample.rcIsec = isec;
        }
        if (surf.isLight) {
            float cosTheta = -dot(ray.dir, surf.norm);
            if (bounce > 1 && cosTheta > 0.0f) {
                float weight = 1.0f;
                if (bounce > 0 && !isSampleTypeDelta(s.type)) {
                    float dist = length(surf.pos - lastPos);
                    float sumPower = S.lightTable[0].prob;
                    float lightPdf = luminance(surf.albedo) / sumPower * dist * dist / std::fabs(cosTheta);
                    weight = MISWeight(s.pdf, lightPdf);
                }
                vec3 weightedLi = surf.albedo * weight;
                setV(pathSample.rcLo, L3(pathSample.rcLo) + weightedLi * throughputAfter);
            }
            break;
        }
        if (bounce > 0 && !isBSDFDelta(mat)) {
            vec3 lightDir; float lightDist, lightPdf, jac; vec2 bary; uint32_t id;
            vec3 lightRadiance = sampleLightByPower(S, surf.pos, lightDir, lightDist, lightPdf, jac, bary, id, sample4f(rng));
            bool shadowed = traceShadow(c, surf.pos, MinRayDistance, lightDir, lightDist - MinRayDistance);
            if (!shadowed && lightPdf > 1e-6f) {
                float bsdfPdf = absDot(surf.norm, lightDir) * PiInv;
                float weight = MISWeight(lightPdf, bsdfPdf);
                setV(pathSample.rcLo, L3(pathSample.rcLo) + lightRadiance * evalBSDF(mat, surf.albedo, surf.norm, wo, lightDir) *
                                                                satDot(surf.norm, lightDir) / lightPdf * weight * throughputAfter);
            }
        }
        if (bounce > 4) {
            float pdfTerminate = gmax(1.0f - luminance(throughputAfter), 0.0f);
            if (sample1f(rng) < pdfTerminate) break;
            throughputAfter = throughputAfter / (1.0f - pdfTerminate);
        }
        if (!sampleBSDF(mat, surf.albedo, surf.norm, wo, sample3f(rng), s) || s.pdf < 1e-6f) break;
        if (bounce == 0) primaryPdf = s.pdf;
        float cosTheta = isSampleTypeDelta(s.type) ? 1.0f : absDot(surf.norm, s.wi);
        vec3 scatterTerms = s.bsdf * cosTheta / s.pdf;
        if (bounce == 0) {
            primaryScatter = s.bsdf * cosTheta;
            primaryPdf = s.pdf;
        } else {
            throughputAfter = throughputAfter * scatterTerms;
        }
        lastPos = surf.pos;
        wo = -s.wi;
        ray.dir = s.wi;
        ray.ori = surf.pos + ray.dir * 1e-4f;
    }
    vec3 radiance = L3(pathSample.rcLo) * primaryScatter / primaryPdf;

    GIReservoir resv; std::memset(&resv, 0, sizeof(resv));
    giReset(resv);
    if ((cam.frameIndex & CameraClearFlag) == 0) {
        // findPreviousReservoir (gi_resample_temporal.glsl:9-32)
        vec2 puv = uv + motion;
        if (!(puv.x < 0.0f || puv.y < 0.0f || puv.x > 1.0f || puv.y > 1.0f) && !std::isnan(puv.x) && !std::isnan(puv.y)) {
            int px = (int)(puv.x * (float)cam.filmSize[0]);
            int py = (int)(puv.y * (float)cam.filmSize[1]);
            float depthPrev; vec3 normalPrev, albedoPrev; int matMeshIdPrev;
            uint32_t amp[2]; fetchAlbedoMatId(gPrev, px, py, amp);
            if (unpackGBuffer(texDepthNormal(gPrev, puv), amp, depthPrev, normalPrev, albedoPrev, matMeshIdPrev) &&
                !(matMeshIdPrev != matMeshId || dot(normalPrev, norm) < 0.9f || std::fabs(depthPrev - depth) > 5.0f)) {
                uint64_t idx = (uint64_t)cam.filmSize[0] * (uint32_t)py + (uint32_t)px;
                if (px >= 0 && py >= 0 && idx < (uint64_t)cam.filmSize[0] * cam.filmSize[1]) resv = resvPrev[idx];
                else std::memset(&resv, 0, sizeof(resv));
            }
        }
    }
    if (pathSample.rcIsec.instanceIdx != InvalidHitIndex) {
        float sampleWeight = luminance(radiance);
        if (std::isnan(sampleWeight) || sampleWeight < 0.0f || primaryPdf < 1e-6f) sampleWeight = 0.0f;
        // GIReservoirAddSample (gi_reservoir.glsl:49-56)
        float r = sample1f(rng);
        resv.resampleWeight += sampleWeight;
        resv.sampleCount++;
        if (r * resv.resampleWeight < sampleWeight) resv.pathSample = pathSample;
    }
    if (!giValid(resv)) giReset(resv);
    if (resv.sampleCount > 40u) {
        resv.resampleWeight *= (float)40u / (float)resv.sampleCount;
        resv.sampleCount = 40u;
    }
    const Material& primaryMat = S.materials[matId];
    if (giValid(resv) && resv.sampleCount > 0 && !isBSDFDelta(primaryMat)) {
        SurfaceInfo rs; sampleSurface(S, resv.pathSample.rcIsec, rs);
        vec3 primaryWi = normalize(rs.pos - primaryPos);
        float weight = resv.resampleWeight / (float)resv.sampleCount;
        vec3 Li = L3(resv.pathSample.rcLo) * evalBSDF(primaryMat, albedo, norm, primaryWo, primaryWi) * satDot(norm, primaryWi);
        if (!isBlack(Li) && traceVisibility(c, primaryPos, rs.pos)) radiance = Li / luminance(Li) * weight;
    }
    resvThis[p] = resv;
    return radiance;
}

}  // namespace orc
