// ORACLE — test infrastructure only. Nothing in the product links or calls this.
// Transliteration of the per-pixel GLSL passes; see orc_passes.h for sources.
#include "orc_passes.h"
#include "orc_guide.h"
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------------
// G-buffer (GBuffer.vert:19-29, GBuffer.frag:20-53) evaluated at the pixel centre.
// Light triangles are not drawn by the reference's raster pass, so the primary
// ray ignores them. Clip planes are applied as ray distances [near, far].
void gbufferPixel(const Ctx& c, const Camera& cam, uint32_t x, uint32_t y, float* dn, uint32_t* am, uint16_t* mv) {
    const uint32_t W = cam.filmSize[0], H = cam.filmSize[1];
    vec2 uv = V2(((float)x + 0.5f) / (float)W, ((float)y + 0.5f) / (float)H);
    Ray ray = pinholeCameraSampleRay(cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    Hit h = traceClosest4(*c.B, makeRayIn(ray.ori, ray.dir, cam.nearZ, cam.farZ), true, c.st);
    if (!h.hit) {
        dn[0] = dn[1] = dn[2] = dn[3] = 0.0f; am[0] = am[1] = 0; mv[0] = mv[1] = 0;
        return;
    }
    const SceneArrays& S = *c.S;
    uint32_t meshIdx = h.isec.instanceIdx - 1;
    const ObjectInstance& I = S.instances[meshIdx];
    uint32_t tri = h.isec.triangleIdx;
    float b1 = h.isec.bary[0], b2 = h.isec.bary[1], b0 = 1.0f - b1 - b2;
    vec3 P[3], N[3];
    float U[3], Vt[3];
    for (int k = 0; k < 3; k++) {
        const MeshVertex& v = S.vertices[S.indices[I.indexOffset + tri * 3 + k]];
        U[k] = v.uvx; Vt[k] = v.uvy;                        // vsOut.uv
        P[k] = mat4MulPoint(I.transform, L3(v.pos), 1.0f);  // vsOut.pos
        const float* m = I.transformInvT;                    // mat3(transformInvT) * aNorm
        vec3 nn = V3(m[0], m[1], m[2]) * v.norm[0] + V3(m[4], m[5], m[6]) * v.norm[1] + V3(m[8], m[9], m[10]) * v.norm[2];
        N[k] = normalize(nn);
    }
    vec3 pos = P[0] * b0 + P[1] * b1 + P[2] * b2;
    vec3 nrm = normalize(N[0] * b0 + N[1] * b1 + N[2] * b2);
    uint32_t matIndex = (uint32_t)S.materialIndices[I.indexOffset / 3 + tri];
    // GBuffer.frag:33-37: the texture at the interpolated uv, or baseColor
    vec3 albedo = materialAlbedo(S, S.materials[matIndex], U[0] * b0 + U[1] * b1 + U[2] * b2,
                                 Vt[0] * b0 + Vt[1] * b1 + Vt[2] * b2);
    vec4 last = mat4MulVec4(cam.lastProjView, {pos.x, pos.y, pos.z, 1.0f});
    float w = last.w;
    last = {last.x / w, last.y / w, last.z / w, last.w / w};
    vec2 lastC = V2(last.x * 0.5f + 0.5f, last.y * 0.5f + 0.5f);
    vec2 thisC = V2(((float)x + 0.5f) / (float)W, ((float)y + 0.5f) / (float)H);
    dn[0] = length(L3(cam.pos) - pos);
    dn[1] = nrm.x; dn[2] = nrm.y; dn[3] = nrm.z;
    am[0] = packAlbedo(albedo);
    am[1] = (matIndex << 16) | meshIdx;
    mv[0] = f32_to_f16(lastC.x - thisC.x);
    mv[1] = f32_to_f16(lastC.y - thisC.y);
}

// ---------------------------------------------------------------------------------
// di_naive.glsl:79-167
vec3 directIllumination2(const Ctx& c, const Camera& cam, const GBufView& g, uint32_t x, uint32_t y) {
    const SceneArrays& S = *c.S;
    vec2 uv = V2(((float)x + 0.5f) / (float)g.W, ((float)y + 0.5f) / (float)g.H);
    float depth; vec3 norm, albedo; int matMeshId;
    uint32_t am[2]; fetchAlbedoMatId(g, (int)x, (int)y, am);
    if (!unpackGBuffer(texDepthNormal(g, uv), am, depth, norm, albedo, matMeshId)) return V3s(0.0f);
    int matId = matMeshId >> 16;
    Ray ray = pinholeCameraSampleRay(cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    uint32_t rng = makeSeedPix(cam.seed, x, y);
    vec3 radiance = V3s(0.0f);
    vec3 pos = ray.ori + ray.dir * (depth - 1e-4f);
    vec3 wo = -ray.dir;
    const Material& mat = S.materials[matId];
    // StreamSampler (di_naive.glsl:54-77)
    vec3 sLi = V3s(0.0f); float sWeight = 0.0f, sSum = 0.0f;
    auto addStream = [&](vec3 Li, float weight, float r) {
        sSum += weight;
        if (r * sSum < weight) { sWeight = weight; sLi = Li; }
    };
    if (!isBSDFDelta(mat)) {
        vec3 lightDir; float lightDist, lightPdf, jac; vec2 bary; uint32_t id;
        vec3 lightRadiance = sampleLightByPower(S, pos, lightDir, lightDist, lightPdf, jac, bary, id, sample4f(rng));
        bool shadowed = traceShadow(c, pos, MinRayDistance, lightDir, lightDist - 1e-4f);
        if (!shadowed && lightPdf > 1e-6f) {
            float bsdfPdf = evalPdf(mat, norm, -ray.dir, lightDir);
            float weight = MISWeight(lightPdf, bsdfPdf);
            vec3 contrib = lightRadiance * evalBSDF(mat, albedo, norm, wo, lightDir) * satDot(norm, lightDir) / lightPdf * weight;
            addStream(contrib, 100.0f, sample1f(rng));
            radiance = radiance + contrib;
        }
    }
    BSDFSample s{};
    if (sampleBSDF(mat, albedo, norm, wo, sample3f(rng), s) && s.pdf > 1e-6f) {
        Intersection isec = traceClosestHit(c, pos, MinRayDistance, s.wi, MaxRayDistance);
        if (isec.instanceIdx != InvalidHitIndex) {
            SurfaceInfo surf; loadSurfaceInfo(S, isec, surf);
            float cosTheta = -dot(s.wi, surf.norm);
            if (surf.isLight && cosTheta > 0.0f) {
                float dist = length(surf.pos - pos);
                float sumPower = S.lightTable[0].prob;
                float lightPdf = luminance(surf.albedo) / sumPower * dist * dist / std::fabs(cosTheta);
                float weight = isSampleTypeDelta(s.type) ? 1.0f : MISWeight(s.pdf, lightPdf);
                float cosTerm = isSampleTypeDelta(s.type) ? 1.0f : satDot(norm, s.wi);
                vec3 contrib = surf.albedo * s.bsdf * cosTerm / s.pdf * weight;
                addStream(contrib, 1.0f, sample1f(rng));
                radiance = radiance + contrib;
            }
        }
    }
    if (sWeight > 0.0f && sSum > 0.0f) radiance = sLi * sSum / sWeight;
    else radiance = V3s(0.0f);
    return clampColor(radiance);
}

// ---------------------------------------------------------------------------------
// gi_naive.glsl:28-149
vec3 indirectIllumination(const Ctx& c, const Camera& cam, const GBufView& g, uint32_t x, uint32_t y, uint32_t maxDepth) {
    const SceneArrays& S = *c.S;
    vec2 uv = V2(((float)x + 0.5f) / (float)g.W, ((float)y + 0.5f) / (float)g.H);
    float depth; vec3 norm, albedo; int matMeshId;
    uint32_t am[2]; fetchAlbedoMatId(g, (int)x, (int)y, am);
    if (!unpackGBuffer(texDepthNormal(g, uv), am, depth, norm, albedo, matMeshId)) return V3s(0.0f);
    int matId = matMeshId >> 16;
    Ray ray = pinholeCameraSampleRay(cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    uint32_t rng = makeSeedPix(cam.seed, x, y);
    vec3 radiance = V3s(0.0f), throughput = V3s(1.0f), lastPos = V3s(0.0f);
    vec3 wo = -ray.dir;
    SurfaceInfo surf;
    surf.pos = ray.ori + ray.dir * (depth - 1e-4f);
    surf.norm = norm; surf.albedo = albedo; surf.isLight = false; surf.matIndex = (uint32_t)matId;
    Material mat = S.materials[matId];
    BSDFSample s{};
    for (uint32_t bounce = 0; bounce < maxDepth; bounce++) {
        if (bounce > 0) {
            Intersection isec = traceClosestHit(c, ray.ori, MinRayDistance, ray.dir, MaxRayDistance);
            if (isec.instanceIdx == InvalidHitIndex) break;
            loadSurfaceInfo(S, isec, surf);
            mat = S.materials[surf.matIndex];
        }
        if (surf.isLight) {
            float cosTheta = -dot(ray.dir, surf.norm);
            if (bounce > 1 && cosTheta > 0.0f) {
                float weight = 1.0f;
                if (!isSampleTypeDelta(s.type)) {
                    float dist = length(surf.pos - lastPos);
                    float sumPower = S.lightTable[0].prob;
                    float lightPdf = luminance(surf.albedo) / sumPower * dist * dist / std::fabs(cosTheta);
                    weight = MISWeight(s.pdf, lightPdf);
                }
                radiance = radiance + surf.albedo * weight * throughput;
            }
            break;
        }
        if (bounce > 0 && !isBSDFDelta(mat)) {
            vec3 lightDir; float lightDist, lightPdf, jac; vec2 bary; uint32_t id;
            vec3 lightRadiance = sampleLightByPower(S, surf.pos, lightDir, lightDist, lightPdf, jac, bary, id, sample4f(rng));
            bool shadowed = traceShadow(c, surf.pos, MinRayDistance, lightDir, lightDist - MinRayDistance);
            if (!shadowed && lightPdf > 1e-6f) {
                // gi_naive.glsl:116-118: the MIS weight is computed and then forced to 1.
                float weight = 1.0f;
                vec3 contrib = lightRadiance * evalBSDF(mat, surf.albedo, surf.norm, wo, lightDir) *
                               satDot(surf.norm, lightDir) / lightPdf * weight * throughput;
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
        ray.dir = s.wi;
        ray.ori = surf.pos + ray.dir * 1e-4f;
    }
    return clampColor(radiance);
}

// ---------------------------------------------------------------------------------
// gris_reservoir.glsl:37-136
static inline uint32_t flagsRcVertexId(uint32_t f) { return f & 0xffu; }
static inline void flagsSetRcVertexId(uint32_t& f, uint32_t id) { f = (f & 0xffffff00u) | (id & 0xffu); }
static inline void flagsSetPathLength(uint32_t& f, uint32_t id) { f = (f & 0xffff00ffu) | ((id & 0xffu) << 8); }
static inline uint32_t flagsRcVertexType(uint32_t f) { return (f >> 16) & 0xffu; }
static inline void flagsSetRcVertexType(uint32_t& f, uint32_t t) { f = (f & 0xff00ffffu) | ((t & 0xffu) << 16); }
static inline void setV(float* d, vec3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }
static inline bool pathSampleIsValid(const GRISPathSample& p) { return p.rcIsec.instanceIdx != InvalidHitIndex; }
static inline void reservoirReset(GRISReservoir& r) {
    r.pathSample.rcIsec.instanceIdx = InvalidHitIndex;
    r.sampleCount = 0.0f;
    r.resampleWeight = 0.0f;
}
static inline bool reservoirIsValid(const GRISReservoir& r) {
    return !std::isnan(r.resampleWeight) && r.resampleWeight >= 0.0f;
}
static inline bool reservoirMerge(GRISReservoir& r, const GRISReservoir& rhs, float rnd) {
    r.sampleCount += rhs.sampleCount;
    r.resampleWeight += rhs.resampleWeight;
    if (rnd * r.resampleWeight < rhs.resampleWeight) { r.pathSample = rhs.pathSample; return true; }
    return false;
}
static inline void reservoirCap(GRISReservoir& r, float cap) {
    if (r.sampleCount > cap) { r.resampleWeight *= cap / r.sampleCount; r.sampleCount = cap; }
}

// gris_path_trace.glsl:35-43
static inline uint32_t nextRcVertexSampleState(uint32_t state, bool connectible) {
    if (state == 2) return 2;
    if (!connectible) return 0;
    return state + (connectible ? 1u : 0u);
}

// gris_path_trace.glsl:45-306
vec3 tracePath(const Ctx& c, const Camera& cam, const GBufView& g, const GRISSettings& set, uint32_t x, uint32_t y,
               GRISReservoir& out, bool& written) {
    written = false;
    const SceneArrays& S = *c.S;
    GRISReservoir resv; std::memset(&resv, 0, sizeof(resv));
    vec2 uv = V2(((float)x + 0.5f) / (float)g.W, ((float)y + 0.5f) / (float)g.H);
    float depth; vec3 norm, albedo; int matMeshId;
    uint32_t am[2]; fetchAlbedoMatId(g, (int)x, (int)y, am);
    if (!unpackGBuffer(texDepthNormal(g, uv), am, depth, norm, albedo, matMeshId)) {
        // the reference returns before writing the reservoir; the buffer keeps
        // its previous contents
        return V3s(0.0f);
    }
    int matId = matMeshId >> 16;
    Ray ray = pinholeCameraSampleRay(cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    uint32_t rng = makeSeedPix(cam.seed, x, y);

    vec3 radiance = V3s(0.0f), throughput = V3s(1.0f), rcThroughput = V3s(0.0f);
    vec3 wo = -ray.dir;
    vec3 lastPos = V3s(0.0f);
    bool isLastVertexConnectible = false;
    uint32_t sampleState = 0, lastSampleState = 0;
    SurfaceInfo surf;
    surf.pos = ray.ori + ray.dir * (depth - 1e-4f);
    surf.norm = norm; surf.albedo = albedo; surf.isLight = false; surf.matIndex = (uint32_t)matId;
    Material mat = S.materials[matId];
    BSDFSample s{};
    Intersection isec{};

    GRISPathSample pathSample; std::memset(&pathSample, 0, sizeof(pathSample));
    pathSample.rcIsec.instanceIdx = InvalidHitIndex;
    pathSample.primaryRng = rng;

    reservoirReset(resv);
    // StreamSampler (gris_path_trace.glsl:10-33)
    GRISPathSample streamSample; std::memset(&streamSample, 0, sizeof(streamSample));
    float streamWeight = 0.0f, streamSum = 0.0f;
    const GuidePix gpx = set.guidedProposal ? guidePixel(c.guideLobes, g.W, g.H, x, y) : noGuide();
    vec3 dir0 = V3s(0.0f);
    auto addStream = [&](const GRISPathSample& ps, float weight, float r) {
        streamSum += weight;
        if (r * streamSum < weight) { streamWeight = weight; streamSample = ps; }
    };

    for (uint32_t bounce = 0; bounce < set.maxDepth; bounce++) {
        if (bounce > 0) {
            isec = traceClosestHit(c, ray.ori, MinRayDistance, ray.dir, MaxRayDistance);
            if (isec.instanceIdx == InvalidHitIndex) break;
            loadSurfaceInfo(S, isec, surf);
            mat = S.materials[surf.matIndex];
        }
        flagsSetPathLength(pathSample.flags, bounce + 1);
        float cosPrevWi = dot(ray.dir, surf.norm);
        float distToPrev = distance(lastPos, surf.pos);
        float geometryJacobian = std::fabs(cosPrevWi) / square(distToPrev);
        bool isThisVertexConnectible = surf.isLight || isBSDFConnectible(mat);
        lastSampleState = sampleState;
        sampleState = nextRcVertexSampleState(sampleState, isThisVertexConnectible);
        if (set.shiftMode == ShiftReconnection && bounce == 1 && !surf.isLight) {
            sampleState = 2; lastSampleState = 1;
        }
        float resvRandSample = sample1f(rng);

        if (surf.isLight) {
            if (bounce > 1 && cosPrevWi < 0.0f) {
                float weight = 1.0f;
                float sumPower = S.lightTable[0].prob;
                float lightPdf = luminance(surf.albedo) / sumPower / geometryJacobian;
                if (!isSampleTypeDelta(s.type)) weight = MISWeight(s.pdf, lightPdf);
                vec3 weightedLi = surf.albedo * weight;
                if (sampleState == 2 && lastSampleState == 2) {
                    setV(pathSample.rcLi, L3(pathSample.rcLi) + weightedLi * rcThroughput);
                    setV(pathSample.F, L3(pathSample.F) + weightedLi * throughput);
                } else if ((sampleState == 2 && lastSampleState == 1) && isLastVertexConnectible &&
                           distToPrev > GRISDistanceThreshold) {
                    pathSample.rcIsec = isec;
                    pathSample.rcRng = rng;
                    pathSample.rcPrevSamplePdf = s.pdf;
                    pathSample.rcJacobian = geometryJacobian;
                    // unbiased reuse keeps the raw radiance (MIS re-evaluated at each reconnection)
                    setV(pathSample.rcLi, set.unbiasedReuse ? surf.albedo : weightedLi);
                    setV(pathSample.rcWi, V3s(0.0f));
                    setV(pathSample.F, weightedLi * throughput);
                    flagsSetRcVertexId(pathSample.flags, bounce);
                    flagsSetRcVertexType(pathSample.flags, RcVertexTypeLightScattered);
                    addStream(pathSample, luminance(L3(pathSample.F)), resvRandSample);
                }
                radiance = radiance + weightedLi * throughput;
            }
            break;
        }

        bool connectible = isThisVertexConnectible && isLastVertexConnectible && distToPrev > GRISDistanceThreshold;
        if ((sampleState == 2 && lastSampleState == 1) && (connectible || set.shiftMode == ShiftReconnection)) {
            pathSample.rcIsec = isec;
            pathSample.rcRng = rng;
            pathSample.rcPrevSamplePdf = s.pdf;
            // unbiased reuse: the Jacobian with the triangle's geometric normal
            pathSample.rcJacobian = set.unbiasedReuse ? std::fabs(dot(ray.dir, geomNormal(S, isec))) / square(distToPrev)
                                                      : geometryJacobian;
            flagsSetRcVertexId(pathSample.flags, bounce);
            flagsSetRcVertexType(pathSample.flags, RcVertexTypeSurface);
            rcThroughput = V3s(1.0f);
        }

        vec4 lightRandSample = sample4f(rng);
        resvRandSample = sample1f(rng);

        if (bounce > 0 && !isBSDFDelta(mat)) {
            vec3 lightDir; float lightDist, lightPdf, lightJacobian; vec2 lightBary; uint32_t lightId;
            vec3 lightRadiance = sampleLightByPower(S, surf.pos, lightDir, lightDist, lightPdf, lightJacobian, lightBary,
                                                    lightId, lightRandSample);
            bool shadowed = traceShadow(c, surf.pos, MinRayDistance, lightDir, lightDist - MinRayDistance);
            if (!shadowed && lightPdf > 1e-6f) {
                float bsdfPdf = absDot(surf.norm, lightDir) * PiInv;
                float weight = MISWeight(lightPdf, bsdfPdf);
                vec3 scatterTerm = evalBSDF(mat, surf.albedo, surf.norm, wo, lightDir) * satDot(surf.norm, lightDir);
                vec3 weightedLi = lightRadiance / lightPdf * weight;
                if (sampleState == 2 && lastSampleState == 2) {
                    setV(pathSample.rcLi, L3(pathSample.rcLi) + weightedLi * scatterTerm * rcThroughput);
                    setV(pathSample.F, L3(pathSample.F) + weightedLi * scatterTerm * throughput);
                } else if (sampleState == 2 && lastSampleState == 1) {
                    setV(pathSample.rcLi, weightedLi);
                    setV(pathSample.rcWi, lightDir);
                    setV(pathSample.F, weightedLi * scatterTerm * throughput);
                    addStream(pathSample, luminance(L3(pathSample.F)), resvRandSample);
                } else if (sampleState == 1 && isThisVertexConnectible && lightDist > GRISDistanceThreshold) {
                    pathSample.rcIsec.bary[0] = lightBary.x;
                    pathSample.rcIsec.bary[1] = lightBary.y;
                    pathSample.rcIsec.instanceIdx = 0;
                    pathSample.rcIsec.triangleIdx = lightId;
                    pathSample.rcRng = rng;
                    pathSample.rcPrevSamplePdf = lightPdf;
                    pathSample.rcJacobian = lightJacobian;
                    setV(pathSample.rcLi, set.unbiasedReuse ? lightRadiance : lightRadiance * weight);
                    setV(pathSample.rcWi, V3s(0.0f));
                    setV(pathSample.F, weightedLi * scatterTerm * throughput);
                    flagsSetRcVertexId(pathSample.flags, bounce + 1);
                    flagsSetRcVertexType(pathSample.flags, RcVertexTypeLightSampled);
                    addStream(pathSample, luminance(L3(pathSample.F)), resvRandSample);
                }
                radiance = radiance + weightedLi * scatterTerm * throughput;
            }
        }

        if (bounce > 4) {
            float pdfTerminate = gmax(1.0f - luminance(throughput) * set.rrScale, 0.0f);
            if (sample1f(rng) < pdfTerminate) break;
            throughput = throughput / (1.0f - pdfTerminate);
            rcThroughput = rcThroughput / (1.0f - pdfTerminate);
        }

        bool sampled;
        if (bounce == 0 && gpx.L) {  // learned proposal at the primary vertex
            const vec4 gu = sample4f(rng);
            const vec3 rb = sample3f(rng);
            sampled = guidedSampleBSDF(gpx, mat, surf.albedo, surf.norm, wo, gu, rb, s);
            dir0 = s.wi;
        } else {
            sampled = sampleBSDF(mat, surf.albedo, surf.norm, wo, sample3f(rng), s);
        }
        if (!sampled || s.pdf < 1e-6f) break;
        float cosTheta = isSampleTypeDelta(s.type) ? 1.0f : absDot(surf.norm, s.wi);
        vec3 scatterTerm = s.bsdf * cosTheta / s.pdf;
        throughput = throughput * scatterTerm;
        if (sampleState == 2 && lastSampleState == 2) {
            rcThroughput = rcThroughput * scatterTerm;
        } else if (sampleState == 2 && lastSampleState == 1) {
            setV(pathSample.rcLi, V3s(0.0f));
            setV(pathSample.rcWi, s.wi);
            setV(pathSample.F, V3s(0.0f));
            rcThroughput = rcThroughput / s.pdf;
        }
        lastPos = surf.pos;
        wo = -s.wi;
        ray.dir = s.wi;
        ray.ori = surf.pos + ray.dir * 1e-4f;
        isLastVertexConnectible = isThisVertexConnectible;
    }

    if (sampleState == 2 && lastSampleState == 2) {
        addStream(pathSample, luminance(L3(pathSample.F)), sample1f(rng));
    }
    pathSample = streamSample;
    if (streamSum > 0.0f && streamWeight > 0.0f) {
        float k = streamSum / streamWeight;
        setV(pathSample.F, L3(pathSample.F) * k);
        setV(pathSample.rcLi, L3(pathSample.rcLi) * k);
        resv.resampleWeight = luminance(L3(pathSample.F));
    } else {
        pathSample.rcIsec.instanceIdx = InvalidHitIndex;
        setV(pathSample.F, V3s(0.0f));
    }
    resv.pathSample = pathSample;
    resv.sampleCount = 1.0f;
    out = resv;
    written = true;
    if (set.guidedProposal && c.guideSamples) {
        float* rec = c.guideSamples + 4 * (size_t)index1D(g.W, x, y);
        rec[0] = dir0.x; rec[1] = dir0.y; rec[2] = dir0.z; rec[3] = streamSum;
    }
    if (!pathSampleIsValid(pathSample)) return V3s(0.0f);
    return radiance;
}

// ---------------------------------------------------------------------------------
// gris_retrace.glsl:42-136
struct RcData { Intersection rcPrevIsec; vec3 rcPrevWo; vec3 rcPrevThroughput; };

// canonical (unbiased reuse, no reference counterpart): the replayed prefix must
// classify the reconnection vertex as the destination's path tracer would --
// its sampleState over the replayed vertices is exactly 1 at the predecessor,
// i.e. no earlier pair of consecutive connectible vertices.
static void traceReplayPathForHybridShift(const Ctx& c, const GRISSettings& set, Intersection isec, SurfaceInfo surf,
                                          Ray ray, uint32_t targetFlags, uint32_t rng, RcData& rc, const GuidePix& gpx,
                                          bool canonical = false) {
    const SceneArrays& S = *c.S;
    vec3 throughput = V3s(1.0f);
    vec3 wo = -ray.dir;
    Material mat = S.materials[surf.matIndex];
    BSDFSample s{};
    rc.rcPrevThroughput = V3s(0.0f);
    rc.rcPrevWo = V3s(0.0f);
    rc.rcPrevIsec = Intersection{};
    rc.rcPrevIsec.instanceIdx = InvalidHitIndex;
    uint32_t targetId = flagsRcVertexId(targetFlags);
    if (targetId == 1) {
        rc.rcPrevIsec = isec; rc.rcPrevWo = wo; rc.rcPrevThroughput = throughput;
        return;
    }
    uint32_t sampleState = 0;
    for (uint32_t bounce = 0; bounce < set.maxDepth; bounce++) {
        if (bounce > 0) {
            isec = traceClosestHit(c, ray.ori, MinRayDistance, ray.dir, MaxRayDistance);
            if (isec.instanceIdx == InvalidHitIndex) break;
            loadSurfaceInfo(S, isec, surf);
            mat = S.materials[surf.matIndex];
        }
        bool isThisVertexConnectible = isBSDFConnectible(mat);
        sampleState = (sampleState == 2) ? 2u : (!isThisVertexConnectible ? 0u : sampleState + 1u);
        (void)sample1f(rng);
        if (surf.isLight) break;
        if (bounce == targetId - 1u) {
            if (isThisVertexConnectible && (!canonical || sampleState == 1u)) {
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
        if (bounce == 0 && gpx.L) {  // the destination's learned proposal, same random numbers
            const vec4 gu = sample4f(rng);
            const vec3 rb = sample3f(rng);
            sampled = guidedSampleBSDF(gpx, mat, surf.albedo, surf.norm, wo, gu, rb, s);
        } else {
            sampled = sampleBSDF(mat, surf.albedo, surf.norm, wo, sample3f(rng), s);
        }
        if (!sampled || s.pdf < 1e-6f) break;
        float cosTheta = isSampleTypeDelta(s.type) ? 1.0f : absDot(surf.norm, s.wi);
        throughput = throughput * (s.bsdf * cosTheta / s.pdf);
        wo = -s.wi;
        ray.dir = s.wi;
        ray.ori = surf.pos + ray.dir * 1e-4f;
    }
}

// gris_retrace.glsl:138-236
// Reconnection shading of gris_retrace.glsl:188-203 with the reconnection
// radiance rcLi given (the sample's own, or the learned prediction).
// generated (unbiased reuse, no reference counterpart): the path tracer's own
// cosines -- |cos| for a direction sampled from a metallic BSDF, clamped for a
// light sample and for a Lambert vertex (upper-hemisphere sampling only).
static vec3 reconnectLi(const GRISPathSample& srcSample, vec3 rcLi, const Material& rcMat, const SurfaceInfo& rcSurf,
                        const Material& rcPrevMat, const SurfaceInfo& rcPrevSurf, vec3 wi, const RcData& rc,
                        bool generated = false) {
    vec3 Li = rcLi;
    uint32_t rcType = flagsRcVertexType(srcSample.flags);
    if (rcType == RcVertexTypeSurface && length(L3(srcSample.rcWi)) > 0.5f) {
        const float cr = (generated && rcMat.type != MatLambert) ? absDot(rcSurf.norm, L3(srcSample.rcWi))
                                                                 : satDot(rcSurf.norm, L3(srcSample.rcWi));
        Li = Li * (evalBSDF(rcMat, rcSurf.albedo, rcSurf.norm, -wi, L3(srcSample.rcWi)) * cr);
    }
    const float cp = (generated && rcType != RcVertexTypeLightSampled && rcPrevMat.type != MatLambert)
                         ? absDot(rcPrevSurf.norm, wi)
                         : satDot(rcPrevSurf.norm, wi);
    Li = Li * (evalBSDF(rcPrevMat, rcPrevSurf.albedo, rcPrevSurf.norm, rc.rcPrevWo, wi) * cp);
    Li = Li * rc.rcPrevThroughput;
    Li = Li / srcSample.rcPrevSamplePdf;
    return Li;
}
// Unbiased reuse: MIS weight of a light reconnection vertex at the current
// predecessor (light sampling vs the cosine pdf, gris_path_trace.glsl:203-204;
// BSDF hit vs light sampling, :128-133); 1 for a surface vertex.
static float rcLightMIS(const SceneArrays& S, uint32_t rcType, const SurfaceInfo& rcPrevSurf, vec3 rcPrevWo,
                        const SurfaceInfo& rcSurf, vec3 wi, float geomJac) {
    if (rcType == RcVertexTypeSurface) return 1.0f;
    float lightPdf = luminance(rcSurf.albedo) / S.lightTable[0].prob / geomJac;
    if (rcType == RcVertexTypeLightSampled) return MISWeight(lightPdf, absDot(rcPrevSurf.norm, wi) * PiInv);
    return MISWeight(evalPdf(S.materials[rcPrevSurf.matIndex], rcPrevSurf.norm, rcPrevWo, wi), lightPdf);
}
// Learned target: a surface rc vertex's radiance is replaced by the prediction.
static vec3 targetRcLi(const GRISPathSample& ps, const LearnedData& L) {
    return flagsRcVertexType(ps.flags) == RcVertexTypeSurface ? V3(L.L[0], L.L[1], L.L[2]) : L3(ps.rcLi);
}

// dstL/srcL are null for the luminance target; with the learned target the
// sites gris_retrace.glsl:204 (dstPHat) and :215 (srcPHat) use the learned
// values and the picked sample's learned data moves with it.
static void reuseAndMerge(const Ctx& c, const GRISSettings& set, GRISReservoir& dstResv, LearnedData* dstL,
                          const SurfaceInfo& dstPrimarySurf, const Intersection& dstPrimaryIsec, const Ray& primaryRay,
                          GRISReservoir srcResv, LearnedData srcL, bool learned, uint32_t& rng, bool unbiased,
                          const GuidePix& gpx) {
    const SceneArrays& S = *c.S;
    GRISPathSample srcSample = srcResv.pathSample;
    RcData dstRc;
    SurfaceInfo rcPrevSurf{}, rcSurf{};
    Material rcMat{}, rcPrevMat{};
    vec3 wi = V3s(0.0f);
    vec3 Li = V3s(0.0f);
    bool srcSampleValid = false;
    float dstJacobian = 0.0f, jacobian = 0.0f, dist = 0.0f, dstPHat = 0.0f, dstSamplePdf = 0.0f;

    if (pathSampleIsValid(srcSample)) {
        traceReplayPathForHybridShift(c, set, dstPrimaryIsec, dstPrimarySurf, primaryRay, srcSample.flags,
                                      srcSample.primaryRng, dstRc, gpx, unbiased);
        if (dstRc.rcPrevIsec.instanceIdx != InvalidHitIndex) {
            if (dstRc.rcPrevIsec.instanceIdx == SpecialHitIndex) rcPrevSurf = dstPrimarySurf;
            else loadSurfaceInfo(S, dstRc.rcPrevIsec, rcPrevSurf);
            loadSurfaceInfo(S, srcSample.rcIsec, rcSurf);
            rcMat = S.materials[rcSurf.matIndex];
            rcPrevMat = S.materials[rcPrevSurf.matIndex];
            dist = distance(rcPrevSurf.pos, rcSurf.pos);
            wi = normalize(rcSurf.pos - rcPrevSurf.pos);
            float cosTheta = -dot(rcSurf.norm, wi);
            const bool geoJ = unbiased && flagsRcVertexType(srcSample.flags) == RcVertexTypeSurface;
            dstJacobian = std::fabs(geoJ ? dot(geomNormal(S, srcSample.rcIsec), wi) : cosTheta) / square(dist);
            jacobian = dstJacobian / srcSample.rcJacobian;
            // unbiased reuse: a surface rc vertex may face away from its predecessor, as generated
            const bool facing = (unbiased && flagsRcVertexType(srcSample.flags) == RcVertexTypeSurface)
                                    ? cosTheta != 0.0f
                                    : cosTheta > 0.0f;
            if (dist > GRISDistanceThreshold && facing && !std::isnan(jacobian) && srcSample.rcJacobian > 0.0f &&
                isBSDFConnectible(rcPrevMat)) {
                if (traceVisibility(c, rcPrevSurf.pos, rcSurf.pos)) srcSampleValid = true;
            }
        }
    }
    if (srcSampleValid) {
        uint32_t rcType = flagsRcVertexType(srcSample.flags);
        vec3 LiL = V3s(0.0f);
        if (!std::isnan(srcSample.rcPrevSamplePdf) && srcSample.rcPrevSamplePdf > 1e-6f) {
            const float misW = unbiased ? rcLightMIS(S, rcType, rcPrevSurf, dstRc.rcPrevWo, rcSurf, wi, dstJacobian) : 1.0f;
            Li = reconnectLi(srcSample, unbiased ? L3(srcSample.rcLi) * misW : L3(srcSample.rcLi), rcMat, rcSurf, rcPrevMat,
                             rcPrevSurf, wi, dstRc, unbiased);
            if (learned)
                LiL = reconnectLi(srcSample, unbiased ? targetRcLi(srcSample, srcL) * misW : targetRcLi(srcSample, srcL),
                                  rcMat, rcSurf, rcPrevMat, rcPrevSurf, wi, dstRc, unbiased);
            if (!isBlack(Li) && !hasNan(Li)) dstPHat = luminance((learned ? LiL : Li) * jacobian);
            if (rcType == RcVertexTypeLightSampled) {
                float sumPower = S.lightTable[0].prob;
                dstSamplePdf = luminance(rcSurf.albedo) / sumPower / dstJacobian;
            } else {
                dstSamplePdf = guidedPdf(dstRc.rcPrevIsec.instanceIdx == SpecialHitIndex ? gpx : noGuide(), rcPrevMat,
                                         rcPrevSurf.norm, dstRc.rcPrevWo, wi);
            }
        }
        float srcPHat = learned ? srcL.pRef : luminance(L3(srcSample.F));
        if (unbiased) {
            // keep F consistent with the new reconnection pdf; reject a zero-pdf reconnection
            const float r = (dstSamplePdf > 1e-6f) ? srcSample.rcPrevSamplePdf / dstSamplePdf : 0.0f;
            if (r == 0.0f) dstPHat = 0.0f;
            Li = Li * r;
            LiL = LiL * r;
        }
        if (learned) srcL.pRef = luminance(LiL);
        srcSample.rcJacobian = dstJacobian;
        srcSample.rcPrevSamplePdf = dstSamplePdf;
        setV(srcSample.F, Li);
        if (srcSample.rcPrevSamplePdf < 1e-6f || std::isnan(srcSample.rcPrevSamplePdf)) srcSample.rcPrevSamplePdf = 0.0f;
        srcResv.pathSample = srcSample;
        srcResv.resampleWeight *= dstPHat / srcPHat;
    } else {
        srcResv.resampleWeight = 0.0f;
    }
    if (reservoirIsValid(srcResv) && reservoirMerge(dstResv, srcResv, sample1f(rng)) && dstL) *dstL = srcL;
    if (!unbiased) reservoirCap(dstResv, (float)set.cap);
}

// ---- unbiased reuse mode (no reference counterpart) ------------------------------
// The "1/Z" contribution weights of unbiased ReSTIR (Bitterli et al. 2020, Sec. 4) for
// the shift-mapped GRIS merges (Lin et al. 2022): after the merges the final weight
// divides by Z = sum of the M of the domains that could have produced the picked
// sample y instead of the total M. Domain j supports y when the hybrid shift of y
// onto j's primary vertex is valid under reuseAndMerge's conditions and the shifted
// sample's own contribution there (with its pdf at the reconnection predecessor)
// has a positive target value.
static bool shiftSupports(const Ctx& c, const GRISSettings& set, const GRISPathSample& y, const LearnedData& yL,
                          bool learned, const SurfaceInfo& primSurf, const Intersection& primIsec, const Ray& primRay,
                          const GuidePix& gpx) {
    const SceneArrays& S = *c.S;
    if (!pathSampleIsValid(y)) return false;
    RcData rc;
    traceReplayPathForHybridShift(c, set, primIsec, primSurf, primRay, y.flags, y.primaryRng, rc, gpx, true);
    if (rc.rcPrevIsec.instanceIdx == InvalidHitIndex) return false;
    SurfaceInfo rcPrevSurf{}, rcSurf{};
    if (rc.rcPrevIsec.instanceIdx == SpecialHitIndex) rcPrevSurf = primSurf;
    else loadSurfaceInfo(S, rc.rcPrevIsec, rcPrevSurf);
    loadSurfaceInfo(S, y.rcIsec, rcSurf);
    const Material& rcMat = S.materials[rcSurf.matIndex];
    const Material& rcPrevMat = S.materials[rcPrevSurf.matIndex];
    float dist = distance(rcPrevSurf.pos, rcSurf.pos);
    vec3 wi = normalize(rcSurf.pos - rcPrevSurf.pos);
    float cosTheta = -dot(rcSurf.norm, wi);
    float jac =
        std::fabs(flagsRcVertexType(y.flags) == RcVertexTypeSurface ? dot(geomNormal(S, y.rcIsec), wi) : cosTheta) /
        square(dist);
    const bool facing = flagsRcVertexType(y.flags) == RcVertexTypeSurface ? cosTheta != 0.0f : cosTheta > 0.0f;
    if (!(dist > GRISDistanceThreshold && facing && jac > 0.0f && !std::isnan(jac) && isBSDFConnectible(rcPrevMat)))
        return false;
    if (!traceVisibility(c, rcPrevSurf.pos, rcSurf.pos)) return false;
    GRISPathSample x = y;
    x.rcPrevSamplePdf = flagsRcVertexType(y.flags) == RcVertexTypeLightSampled
                            ? luminance(rcSurf.albedo) / S.lightTable[0].prob / jac
                            : guidedPdf(rc.rcPrevIsec.instanceIdx == SpecialHitIndex ? gpx : noGuide(), rcPrevMat,
                                        rcPrevSurf.norm, rc.rcPrevWo, wi);
    if (std::isnan(x.rcPrevSamplePdf) || !(x.rcPrevSamplePdf > 1e-6f)) return false;
    const float misW = rcLightMIS(S, flagsRcVertexType(y.flags), rcPrevSurf, rc.rcPrevWo, rcSurf, wi, jac);
    vec3 Li = reconnectLi(x, L3(x.rcLi) * misW, rcMat, rcSurf, rcPrevMat, rcPrevSurf, wi, rc, true);
    // x is a candidate of that domain iff its resampling weight luminance(F) is positive
    if (hasNan(Li) || !(luminance(Li) > 0.0f)) return false;
    if (learned)
        return luminance(reconnectLi(x, targetRcLi(x, yL) * misW, rcMat, rcSurf, rcPrevMat, rcPrevSurf, wi, rc, true)) >
               0.0f;
    return true;
}

// Primary vertex of pixel (x, y) as its own passes rebuilt it; false for background.
static bool primaryOf(const Camera& cam, const GBufView& g, int x, int y, SurfaceInfo& surf, Intersection& isec,
                      Ray& ray) {
    if (x < 0 || y < 0 || x >= (int)g.W || y >= (int)g.H) return false;
    vec2 uv = V2(((float)x + 0.5f) / (float)g.W, ((float)y + 0.5f) / (float)g.H);
    float depth; vec3 norm, albedo; int matMeshId;
    uint32_t am[2]; fetchAlbedoMatId(g, x, y, am);
    if (!unpackGBuffer(texDepthNormal(g, uv), am, depth, norm, albedo, matMeshId)) return false;
    ray = pinholeCameraSampleRay(cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    surf = SurfaceInfo{ray.ori + ray.dir * (depth - 1e-4f), norm, albedo, (uint32_t)(matMeshId >> 16), false};
    isec = Intersection{};
    isec.instanceIdx = SpecialHitIndex;
    isec.bary[0] = uv.x; isec.bary[1] = uv.y;
    return true;
}

static void normaliseZ(GRISReservoir& r, float z) {
    if (z > 0.0f) r.resampleWeight = r.resampleWeight * (r.sampleCount / z);
}

// Index of a reservoir read, or -1 where the read is out of range (reads zero).
static inline int64_t resvIndex(uint32_t W, uint32_t H, int x, int y) {
    uint64_t idx = (uint64_t)W * (uint32_t)y + (uint32_t)x;
    return (x < 0 || y < 0 || idx >= (uint64_t)W * H) ? -1 : (int64_t)idx;
}
static inline LearnedData learnedAt(const LearnedData* buf, int64_t idx) {
    LearnedData l{{0.0f, 0.0f, 0.0f}, 0.0f};
    if (buf && idx >= 0) l = buf[idx];
    return l;
}

static inline GRISReservoir resvAt(const GRISReservoir* buf, uint32_t W, uint32_t H, int x, int y) {
    // buffer index1D(pixel) as in the reference; out-of-range reads are zero
    uint64_t idx = (uint64_t)W * (uint32_t)y + (uint32_t)x;
    GRISReservoir r;
    if (x < 0 || y < 0 || idx >= (uint64_t)W * H) { std::memset(&r, 0, sizeof(r)); return r; }
    return buf[idx];
}

// gris_resample_temporal.glsl:11-34
static bool findPreviousReservoir(const Camera& cam, const Camera& prevCam, const GBufView& gPrev,
                                  const GRISReservoir* resvPrev, vec2 uv, vec3 pos, vec3 normal, int matMeshId,
                                  GRISReservoir& resv, int64_t& idx) {
    if (uv.x < 0.0f || uv.y < 0.0f || uv.x > 1.0f || uv.y > 1.0f) return false;
    // GLSL leaves ivec2(NaN) undefined: a non-finite reprojection is rejected
    if (std::isnan(uv.x) || std::isnan(uv.y)) return false;
    int px = (int)(uv.x * (float)cam.filmSize[0]);
    int py = (int)(uv.y * (float)cam.filmSize[1]);
    float depthPrev; vec3 normalPrev, albedoPrev; int matMeshIdPrev;
    uint32_t am[2]; fetchAlbedoMatId(gPrev, px, py, am);
    if (!unpackGBuffer(texDepthNormal(gPrev, uv), am, depthPrev, normalPrev, albedoPrev, matMeshIdPrev)) return false;
    Ray ray = pinholeCameraSampleRay(prevCam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    vec3 posPrev = ray.ori + ray.dir * (depthPrev - 1e-4f);
    if (matMeshIdPrev != matMeshId || dot(normalPrev, normal) < 0.95f || distance(pos, posPrev) > 0.5f) return false;
    resv = resvAt(resvPrev, cam.filmSize[0], cam.filmSize[1], px, py);
    idx = resvIndex(cam.filmSize[0], cam.filmSize[1], px, py);
    return true;
}

// gris_resample_temporal.glsl:36-83
bool temporalReuse(const Ctx& c, const Camera& cam, const Camera& prevCam, const GBufView& g, const GBufView& gPrev,
                   const GRISSettings& set, const GRISReservoir* resvThis, const GRISReservoir* resvPrev,
                   uint32_t x, uint32_t y, GRISReservoir& outTemp, const LearnedData* learnThis,
                   const LearnedData* learnPrev, LearnedData* outLearn) {
    const bool learned = set.targetFunction == 1u;
    const SceneArrays& S = *c.S;
    vec2 uv = V2(((float)x + 0.5f) / (float)g.W, ((float)y + 0.5f) / (float)g.H);
    float depth; vec3 norm, albedo; int matMeshId;
    uint32_t am[2]; fetchAlbedoMatId(g, (int)x, (int)y, am);
    if (!unpackGBuffer(texDepthNormal(g, uv), am, depth, norm, albedo, matMeshId)) return false;  // Temp not written
    vec2 motion = fetchMotion(g, (int)x, (int)y);
    int matId = matMeshId >> 16;
    Ray ray = pinholeCameraSampleRay(cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    uint32_t rng = makeSeedPix(cam.seed, x, y) ^ 1u;
    uint32_t resvRng = ~rng;
    vec3 pos = ray.ori + ray.dir * (depth - 1e-4f);
    (void)S;
    GRISReservoir resv = resvThis[index1D(g.W, x, y)];
    LearnedData L = learnedAt(learned ? learnThis : nullptr, (int64_t)index1D(g.W, x, y));
    Intersection dstPrimaryIsec{};
    dstPrimaryIsec.instanceIdx = SpecialHitIndex;
    dstPrimaryIsec.bary[0] = uv.x; dstPrimaryIsec.bary[1] = uv.y;
    const bool unbiased = set.unbiasedReuse != 0u;
    const float mCanon = resv.sampleCount;
    float mPrev = 0.0f;
    int prevX = -1, prevY = -1;
    if (set.temporalReuse) {
        GRISReservoir temporalResv; std::memset(&temporalResv, 0, sizeof(temporalResv));
        reservoirReset(temporalResv);
        SurfaceInfo dstPrimarySurf{pos, norm, albedo, (uint32_t)matId, false};
        int64_t pidx = -1;
        if (((cam.frameIndex & CameraClearFlag) == 0) &&
            findPreviousReservoir(cam, prevCam, gPrev, resvPrev, uv + motion, pos, norm, matMeshId, temporalResv, pidx)) {
            if (reservoirIsValid(temporalResv)) {
                reuseAndMerge(c, set, resv, learned ? &L : nullptr, dstPrimarySurf, dstPrimaryIsec, ray, temporalResv,
                              learnedAt(learned ? learnPrev : nullptr, pidx), learned, resvRng, unbiased,
                              set.guidedProposal ? guidePixel(c.guideLobes, g.W, g.H, x, y) : noGuide());
                mPrev = temporalResv.sampleCount;
                vec2 puv = uv + motion;
                prevX = (int)(puv.x * (float)cam.filmSize[0]);
                prevY = (int)(puv.y * (float)cam.filmSize[1]);
            }
        }
    }
    if (unbiased) {
        if (mPrev > 0.0f && reservoirIsValid(resv) && pathSampleIsValid(resv.pathSample)) {
            float z = mCanon;
            SurfaceInfo ps{}; Intersection pi{}; Ray pr{};
            const GuidePix prevGp = set.guidedProposal
                                        ? guidePixel(c.guideLobesPrev, g.W, g.H, (uint32_t)prevX, (uint32_t)prevY)
                                        : noGuide();
            if (primaryOf(prevCam, gPrev, prevX, prevY, ps, pi, pr) &&
                shiftSupports(c, set, resv.pathSample, L, learned, ps, pi, pr, prevGp))
                z += mPrev;
            normaliseZ(resv, z);
        }
        reservoirCap(resv, (float)set.cap);
    }
    if (!reservoirIsValid(resv)) reservoirReset(resv);
    outTemp = resv;
    if (outLearn) *outLearn = L;
    return true;
}

// gris_resample_spatial.glsl:11-34
static bool findNeighborReservoir(const Camera& cam, const GBufView& g, const GRISReservoir* resvTemp, vec2 uv, vec3 pos,
                                  vec3 normal, GRISReservoir& resv, int64_t& idx) {
    if (uv.x < 0.0f || uv.y < 0.0f || uv.x > 1.0f || uv.y > 1.0f) return false;
    int px = (int)(uv.x * (float)cam.filmSize[0]);
    int py = (int)(uv.y * (float)cam.filmSize[1]);
    float depthPrev; vec3 normalPrev, albedoPrev; int matMeshIdPrev;
    uint32_t am[2]; fetchAlbedoMatId(g, px, py, am);
    if (!unpackGBuffer(texDepthNormal(g, uv), am, depthPrev, normalPrev, albedoPrev, matMeshIdPrev)) return false;
    Ray ray = pinholeCameraSampleRay(cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    vec3 posPrev = ray.ori + ray.dir * (depthPrev - 1e-4f);
    if (dot(normalPrev, normal) < 0.9f || distance(pos, posPrev) > 0.4f) return false;
    resv = resvAt(resvTemp, cam.filmSize[0], cam.filmSize[1], px, py);
    idx = resvIndex(cam.filmSize[0], cam.filmSize[1], px, py);
    return true;
}

// gris_resample_spatial.glsl:36-135
vec3 spatialReuse(const Ctx& c, const Camera& cam, const GBufView& g, const GRISSettings& set,
                  const GRISReservoir* resvTemp, uint32_t x, uint32_t y, GRISReservoir& outThis, bool& written,
                  const LearnedData* learnTemp, LearnedData* outLearn) {
    written = false;
    const bool learned = set.targetFunction == 1u;
    const SceneArrays& S = *c.S;
    vec2 uv = V2(((float)x + 0.5f) / (float)g.W, ((float)y + 0.5f) / (float)g.H);
    float depth; vec3 norm, albedo; int matMeshId;
    uint32_t am[2]; fetchAlbedoMatId(g, (int)x, (int)y, am);
    if (!unpackGBuffer(texDepthNormal(g, uv), am, depth, norm, albedo, matMeshId)) return V3s(0.0f);  // not written
    uint32_t rng = makeSeedPix(cam.seed, x, y) ^ 2u;
    Ray ray = pinholeCameraSampleRay(cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
    vec3 pos = ray.ori + ray.dir * (depth - 1e-4f);
    int matId = matMeshId >> 16;
    vec3 radiance = V3s(0.0f);
    GRISReservoir resv = resvTemp[index1D(g.W, x, y)];
    LearnedData L = learnedAt(learned ? learnTemp : nullptr, (int64_t)index1D(g.W, x, y));
    Intersection dstPrimaryIsec{};
    dstPrimaryIsec.instanceIdx = SpecialHitIndex;
    dstPrimaryIsec.bary[0] = uv.x; dstPrimaryIsec.bary[1] = uv.y;
    SurfaceInfo dstPrimarySurf{pos, norm, albedo, (uint32_t)matId, false};
    const bool unbiased = set.unbiasedReuse != 0u;
    const float mCanon = resv.sampleCount;
    const bool guided = set.guidedProposal && c.guideLobes;
    const GuidePix gpx = guided ? guidePixel(c.guideLobes, g.W, g.H, x, y) : noGuide();
    std::vector<std::pair<int, int>> nbPix;
    std::vector<float> nbM;
    if (set.spatialReuse) {
        vec2 texelSize = V2(1.0f / (float)g.W, 1.0f / (float)g.H);
        for (uint32_t i = 0; i < set.spatialNeighbors; i++) {
            vec2 d = toConcentricDisk(sample2f(rng));
            vec2 neighbor = uv + (d * set.spatialRadius) * texelSize;
            GRISReservoir neighborResv; std::memset(&neighborResv, 0, sizeof(neighborResv));
            reservoirReset(neighborResv);
            int64_t nidx = -1;
            if (findNeighborReservoir(cam, g, resvTemp, neighbor, pos, norm, neighborResv, nidx)) {
                if (reservoirIsValid(neighborResv)) {
                    reuseAndMerge(c, set, resv, learned ? &L : nullptr, dstPrimarySurf, dstPrimaryIsec, ray, neighborResv,
                                  learnedAt(learned ? learnTemp : nullptr, nidx), learned, rng, unbiased, gpx);
                    if (unbiased && nbPix.size() < UnbiasedMaxNeighbors) {
                        nbPix.emplace_back((int)(neighbor.x * (float)cam.filmSize[0]),
                                           (int)(neighbor.y * (float)cam.filmSize[1]));
                        nbM.push_back(neighborResv.sampleCount);
                    }
                }
            }
        }
    }
    if (unbiased) {
        if (!nbPix.empty() && reservoirIsValid(resv) && pathSampleIsValid(resv.pathSample)) {
            float z = mCanon;
            for (size_t j = 0; j < nbPix.size(); j++) {
                if (!(nbM[j] > 0.0f)) continue;
                SurfaceInfo ns{}; Intersection ni{}; Ray nr{};
                const GuidePix ngp = guided ? guidePixel(c.guideLobes, g.W, g.H, (uint32_t)nbPix[j].first,
                                                         (uint32_t)nbPix[j].second)
                                            : noGuide();
                if (primaryOf(cam, g, nbPix[j].first, nbPix[j].second, ns, ni, nr) &&
                    shiftSupports(c, set, resv.pathSample, L, learned, ns, ni, nr, ngp))
                    z += nbM[j];
            }
            normaliseZ(resv, z);
        }
        reservoirCap(resv, (float)set.cap);
    }
    if (!reservoirIsValid(resv)) reservoirReset(resv);
    outThis = resv;
    if (outLearn) *outLearn = L;
    written = true;
    if (reservoirIsValid(resv) && resv.sampleCount > 0.0f) {
        GRISPathSample ps = resv.pathSample;
        if (pathSampleIsValid(ps)) {
            RcData rc;
            traceReplayPathForHybridShift(c, set, dstPrimaryIsec, dstPrimarySurf, ray, ps.flags, ps.primaryRng, rc, gpx);
            if (rc.rcPrevIsec.instanceIdx != InvalidHitIndex) {
                SurfaceInfo rcPrevSurf{}, rcSurf{};
                if (rc.rcPrevIsec.instanceIdx == SpecialHitIndex) rcPrevSurf = dstPrimarySurf;
                else loadSurfaceInfo(S, rc.rcPrevIsec, rcPrevSurf);
                loadSurfaceInfo(S, ps.rcIsec, rcSurf);
                const Material& rcMat = S.materials[rcSurf.matIndex];
                const Material& rcPrevMat = S.materials[rcPrevSurf.matIndex];
                vec3 wi = normalize(rcSurf.pos - rcPrevSurf.pos);
                float misW = 1.0f;
                if (unbiased) {
                    float gj = std::fabs(dot(rcSurf.norm, wi)) / square(distance(rcPrevSurf.pos, rcSurf.pos));
                    misW = rcLightMIS(S, flagsRcVertexType(ps.flags), rcPrevSurf, rc.rcPrevWo, rcSurf, wi, gj);
                }
                vec3 Li = reconnectLi(ps, unbiased ? L3(ps.rcLi) * misW : L3(ps.rcLi), rcMat, rcSurf, rcPrevMat, rcPrevSurf,
                                      wi, rc, unbiased);
                float pHat = learned ? luminance(reconnectLi(ps, unbiased ? targetRcLi(ps, L) * misW : targetRcLi(ps, L),
                                                             rcMat, rcSurf, rcPrevMat, rcPrevSurf, wi, rc, unbiased))
                                     : luminance(Li);
                if (!isBlack(Li) && !hasNan(Li)) {
                    radiance = Li / pHat * resv.resampleWeight / resv.sampleCount;
                }
            }
        }
    }
    return clampColor(radiance);
}

// Learned-target refresh (no reference counterpart): for a surface rc vertex
// the contribution with the predicted radiance is F_L = F * Lhat / rcLi per
// channel (F is linear in rcLi); resampleWeight, which after stream RIS refers
// to luminance(F), is rescaled to refer to luminance(F_L).
bool learnedRefresh(const Ctx& c, const GBufView& g, uint32_t x, uint32_t y, GRISReservoir& resv, const float Lhat[3],
                    LearnedData& out, bool& need, vec3& pos) {
    const size_t p = index1D(g.W, x, y);
    need = false;
    pos = V3s(0.0f);
    if (g.depthNormal[4 * p] == 0.0f) return false;
    const GRISPathSample& ps = resv.pathSample;
    need = reservoirIsValid(resv) && ps.rcIsec.instanceIdx != InvalidHitIndex &&
           flagsRcVertexType(ps.flags) == RcVertexTypeSurface;
    if (need) {
        SurfaceInfo sf{};
        loadSurfaceInfo(*c.S, ps.rcIsec, sf);
        pos = sf.pos;
    }
    vec3 L = V3s(0.0f), F = L3(ps.F), FL = F;
    const float pOld = luminance(F);
    if (need) {
        L = V3(Lhat[0], Lhat[1], Lhat[2]);
        FL.x = ps.rcLi[0] > 0.0f ? F.x * (L.x / ps.rcLi[0]) : F.x;
        FL.y = ps.rcLi[1] > 0.0f ? F.y * (L.y / ps.rcLi[1]) : F.y;
        FL.z = ps.rcLi[2] > 0.0f ? F.z * (L.z / ps.rcLi[2]) : F.z;
    }
    const float pNew = luminance(FL);
    if (need && pOld > 0.0f) resv.resampleWeight = resv.resampleWeight * (pNew / pOld);
    out.L[0] = L.x; out.L[1] = L.y; out.L[2] = L.z;
    out.pRef = need ? pNew : pOld;
    return true;
}

}  // namespace orc
