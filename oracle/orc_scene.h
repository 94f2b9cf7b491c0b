// ORACLE — test infrastructure only. Nothing in the product links or calls this.
//
// Scene data model (std430 layouts of src/shader/layouts.glsl:6-94), surface
// loading (src/shader/ray_layouts.glsl:50-104), BSDFs
// (src/shader/material.glsl:31-364) and light sampling
// (src/shader/light_sampling.glsl:6-53), restated in scalar FP32.
#pragma once
#include "orc_math.h"
#include <vector>

namespace orc {

// ---- layouts.glsl ------------------------------------------------------------
struct Material {           // layouts.glsl:6-25 (RESTIR_PT_MATERIAL = 1), 32 B
    float baseColor[3];
    uint32_t type;
    uint32_t textureIdx;
    float metallic, roughness, ior;
};
struct Texture {            // an R8G8B8A8_SRGB albedo image (uTextures, layouts.glsl:178)
    uint32_t width, height, filter, pad;   // filter 0 nearest, 1 linear (src/Scene.cpp:442)
    const uint8_t* rgba8;
};
struct MeshVertex {         // layouts.glsl:53-58, 32 B
    float pos[3]; float uvx; float norm[3]; float uvy;
};
struct ObjectInstance {     // layouts.glsl:60-72, 224 B (matrices column-major)
    float transform[16], transformInv[16], transformInvT[16];
    float radiance[3]; float pad0;
    uint32_t indexOffset, indexCount, matIndex; float pad2;
};
struct TriangleLight {      // layouts.glsl:74-83, 64 B
    float v0[3]; float nx; float v1[3]; float ny; float v2[3]; float nz; float radiance[3]; float area;
};
struct LightSampleTableElement { float prob; uint32_t failId; };  // layouts.glsl:85-88
struct Intersection { float bary[2]; uint32_t instanceIdx; uint32_t triangleIdx; };  // layouts.glsl:90-94
struct Camera {             // layouts.glsl:27-51, 352 B
    float view[16], proj[16], projView[16], lastProjView[16];
    float pos[3]; float FOV;
    float angle[3]; float nearZ;
    float front[3]; float farZ;
    float right[3]; float lensRadius;
    float up[3]; float focalDist;
    uint32_t filmSize[2]; uint32_t frameIndex; uint32_t seed;
};
struct GRISPathSample {     // layouts.glsl:132-147, 80 B
    Intersection rcIsec;
    float rcLi[3]; uint32_t rcRng;
    float rcWi[3]; uint32_t flags;
    float pad[2]; float rcPrevSamplePdf; float rcJacobian;
    float F[3]; uint32_t primaryRng;
};
struct GRISReservoir {      // layouts.glsl:149-156, 96 B
    GRISPathSample pathSample;
    float sampleCount, resampleWeight, contribWeight, pad0;
};
static_assert(sizeof(Camera) == 352, "camera layout");
static_assert(sizeof(GRISReservoir) == 96, "reservoir layout");
static_assert(sizeof(ObjectInstance) == 224, "instance layout");

// Scene arrays as the reference binds them in descriptor set 1
// (layouts.glsl:171-178).
struct SceneArrays {
    const Material* materials; uint32_t numMaterials;
    const int32_t* materialIndices; uint32_t numMaterialIndices;
    const MeshVertex* vertices; uint32_t numVertices;
    const uint32_t* indices; uint32_t numIndices;
    const ObjectInstance* instances; uint32_t numInstances;
    const TriangleLight* lights; uint32_t numLights;
    const LightSampleTableElement* lightTable; uint32_t numLightTable;
    const Texture* textures; uint32_t numTextures;
};

static inline vec3 L3(const float* p) { return V3(p[0], p[1], p[2]); }
static inline vec3 toV(const float* p) { return V3(p[0], p[1], p[2]); }

// GLSL mat4 * vec4 (column-major): c0*v.x + c1*v.y + c2*v.z + c3*v.w, xyz part.
static inline vec3 mat4MulPoint(const float* m, vec3 p, float w) {
    vec3 c0 = V3(m[0], m[1], m[2]), c1 = V3(m[4], m[5], m[6]), c2 = V3(m[8], m[9], m[10]), c3 = V3(m[12], m[13], m[14]);
    return c0 * p.x + c1 * p.y + c2 * p.z + c3 * w;
}
static inline vec4 mat4MulVec4(const float* m, vec4 v) {
    vec4 r;
    r.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
    r.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
    r.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
    r.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
    return r;
}

// ---- material.glsl -------------------------------------------------------------
enum : uint32_t { Diffuse = 1u << 0, Glossy = 1u << 1, Specular = 1u << 2, Reflection = 1u << 4,
                  Transmission = 1u << 5, InvalidBSDFSample = 0x80000000u };
enum : uint32_t { MatPrincipled = 0, MatLambert = 1, MatMetallicWorkflow = 2, MatMetal = 3,
                  MatDielectric = 4, MatThinDielectric = 5, MatFake = 6, MatLight = 7 };

struct BSDFSample { vec3 wi; float pdf; vec3 bsdf; uint32_t type; };

static inline float fresnelSchlickF(float cosTheta, float ior) {
    float f0 = std::fabs(1.0f - ior) / (1.0f + ior);
    return mixf(f0, 1.0f, pow5(1.0f - cosTheta));
}
static inline vec3 fresnelSchlick3(float cosTheta, vec3 f0) {
    return mix3s(f0, V3s(1.0f), pow5(1.0f - cosTheta));
}
static inline float schlickG(float cosTheta, float alpha) {
    float a = alpha * 0.5f;
    return cosTheta / (cosTheta * (1.0f - a) + a);
}
static inline float smithG(float cosWo, float cosWi, float alpha) {
    return schlickG(std::fabs(cosWo), alpha) * schlickG(std::fabs(cosWi), alpha);
}
static inline float GTR2Distrib(float cosTheta, float alpha) {
    if (cosTheta < 1e-6f) return 0.0f;
    float aa = alpha * alpha;
    float nom = aa;
    float denom = cosTheta * cosTheta * (aa - 1.0f) + 1.0f;
    denom = denom * denom * Pi;
    return nom / denom;
}
static inline float GTR2Pdf(vec3 n, vec3 m, vec3 wo, float alpha) {
    return GTR2Distrib(dot(n, m), alpha) * schlickG(dot(n, wo), alpha) * absDot(m, wo) / absDot(n, wo);
}
// material.glsl:84-101
static inline vec3 GTR2Sample(vec3 n, vec3 wo, float alpha, vec2 r) {
    mat3 transMat = matLocalToWorld(n);
    mat3 transInv = inverse3(transMat);
    vec3 vh = normalize(mul(transInv, wo) * V3(alpha, alpha, 1.0f));
    float lenSq = vh.x * vh.x + vh.y * vh.y;
    vec3 t = lenSq > 0.0f ? V3(-vh.y, vh.x, 0.0f) / std::sqrt(lenSq) : V3(1.0f, 0.0f, 0.0f);
    vec3 b = cross(vh, t);
    vec2 p = toConcentricDisk(r);
    float s = 0.5f * (vh.z + 1.0f);
    p.y = (1.0f - s) * std::sqrt(1.0f - p.x * p.x) + s * p.y;
    vec3 wh = t * p.x + b * p.y + vh * std::sqrt(gmax(0.0f, 1.0f - dot2(p, p)));
    wh = V3(wh.x * alpha, wh.y * alpha, gmax(0.0f, wh.z));
    return normalize(mul(transMat, wh));
}
static inline bool isGTR2Connectible(float roughness) { return roughness > 0.05f; }
static inline bool isGTR2Delta(float roughness) { return roughness < 0.01f; }

static inline bool refractDir(vec3 n, vec3 wi, float ior, vec3& wt) {
    float cosIn = dot(n, wi);
    if (cosIn < 0.0f) ior = 1.0f / ior;
    float sin2In = gmax(0.0f, 1.0f - cosIn * cosIn);
    float sin2Tr = sin2In / (ior * ior);
    if (sin2Tr >= 1.0f) return false;
    float cosTr = std::sqrt(1.0f - sin2Tr);
    if (cosIn < 0.0f) cosTr = -cosTr;
    wt = normalize(-wi / ior + n * (cosIn / ior - cosTr));
    return true;
}

static inline vec3 metallicWorkflowBSDF(const Material& mat, vec3 albedo, vec3 n, vec3 wo, vec3 wi) {
    float alpha = square(mat.roughness);
    vec3 wh = normalize(wo + wi);
    float cosO = dot(n, wo);
    float cosI = dot(n, wi);
    if (cosI * cosO < 1e-7f) return V3s(0.0f);
    vec3 f = fresnelSchlick3(dot(wh, wo), mix3s(V3s(0.08f), albedo, mat.metallic));
    float g = smithG(cosO, cosI, alpha);
    float d = GTR2Distrib(dot(n, wh), alpha);
    return mix3(albedo * PiInv * (1.0f - mat.metallic), V3s(g * d / (4.0f * cosI * cosO)), f);
}
static inline float metallicWorkflowPdf(const Material& mat, vec3 n, vec3 wo, vec3 wi) {
    vec3 wh = normalize(wo + wi);
    return mixf(satDot(n, wi) * PiInv, GTR2Pdf(n, wh, wo, square(mat.roughness)) / (4.0f * absDot(wh, wo)),
                1.0f / (2.0f - mat.metallic));
}
static inline bool metallicWorkflowSampleBSDF(const Material& mat, vec3 albedo, vec3 n, vec3 wo, vec3 r, BSDFSample& s) {
    float alpha = square(mat.roughness);
    s.type = Reflection;
    if (r.z > (1.0f / (2.0f - mat.metallic))) {
        s.wi = sampleCosineWeightedHemisphere(n, V2(r.x, r.y));
        s.type |= Diffuse;
    } else {
        vec3 wh = GTR2Sample(n, wo, alpha, V2(r.x, r.y));
        s.wi = -reflect(wo, wh);
        s.type |= isGTR2Delta(mat.roughness) ? Specular : Glossy;
    }
    if (dot(n, s.wi) < 0.0f) { s.type = InvalidBSDFSample; return false; }
    s.bsdf = metallicWorkflowBSDF(mat, albedo, n, wo, s.wi);
    s.pdf = metallicWorkflowPdf(mat, n, wo, s.wi);
    return true;
}
static inline vec3 metalBSDF(const Material& mat, vec3 albedo, vec3 n, vec3 wo, vec3 wi) {
    if (isGTR2Delta(mat.roughness)) return V3s(0.0f);
    float alpha = square(mat.roughness);
    vec3 wh = normalize(wo + wi);
    float cosO = dot(n, wo);
    float cosI = dot(n, wi);
    if (cosI * cosO < 1e-7f) return V3s(0.0f);
    float f = fresnelSchlickF(absDot(wh, wo), mat.ior);
    float g = smithG(cosO, cosI, alpha);
    float d = GTR2Distrib(dot(n, wh), alpha);
    return albedo * f * g * d / (4.0f * cosI * cosO);
}
static inline float metalPdf(const Material& mat, vec3 n, vec3 wo, vec3 wi) {
    if (isGTR2Delta(mat.roughness)) return 0.0f;
    vec3 wh = normalize(wo + wi);
    return GTR2Pdf(n, wh, wo, square(mat.roughness)) / (4.0f * absDot(wh, wo));
}
static inline bool metalSampleBSDF(const Material& mat, vec3 albedo, vec3 n, vec3 wo, vec3 r, BSDFSample& s) {
    float alpha = square(mat.roughness);
    bool isDelta = isGTR2Delta(mat.roughness);
    if (isDelta) s.wi = -reflect(wo, n);
    else { vec3 wh = GTR2Sample(n, wo, alpha, V2(r.x, r.y)); s.wi = -reflect(wo, wh); }
    if (dot(n, s.wi) < 0.0f) { s.type = InvalidBSDFSample; return false; }
    s.bsdf = isDelta ? albedo * fresnelSchlickF(absDot(n, wo), mat.ior) : metalBSDF(mat, albedo, n, wo, s.wi);
    s.pdf = isDelta ? 1.0f : metalPdf(mat, n, wo, s.wi);
    s.type = Reflection | (isDelta ? Specular : Glossy);
    return true;
}
static inline bool dielectricSampleBSDF(const Material& matIn, vec3 albedo, vec3 n, vec3 wo, vec3 r, BSDFSample& s) {
    Material mat = matIn;
    float pdfReflect = fresnelSchlickF(dot(n, wo), mat.ior);
    s.bsdf = albedo;
    if (r.z < pdfReflect) {
        s.wi = reflect(-wo, n);
        s.type = Specular | Reflection;
        s.pdf = 1.0f;
    } else {
        if (!refractDir(n, wo, mat.ior, s.wi)) { s.type = InvalidBSDFSample; return false; }
        if (dot(n, wo) < 0.0f) mat.ior = 1.0f / mat.ior;
        s.bsdf = s.bsdf / (mat.ior * mat.ior);
        s.type = Specular | Transmission;
        s.pdf = 1.0f;
    }
    return true;
}

// material.glsl:286-360
static inline vec3 evalBSDF(const Material& mat, vec3 albedo, vec3 n, vec3 wo, vec3 wi) {
    switch (mat.type) {
    case MatLambert: return albedo * PiInv;
    case MatMetallicWorkflow: return metallicWorkflowBSDF(mat, albedo, n, wo, wi);
    case MatMetal: return metalBSDF(mat, albedo, n, wo, wi);
    default: return V3s(0.0f);
    }
}
static inline float evalPdf(const Material& mat, vec3 n, vec3 wo, vec3 wi) {
    switch (mat.type) {
    case MatLambert: return absDot(n, wi) * PiInv;
    case MatMetallicWorkflow: return metallicWorkflowPdf(mat, n, wo, wi);
    case MatMetal: return metalPdf(mat, n, wo, wi);
    default: return 0.0f;
    }
}
static inline bool sampleBSDF(const Material& mat, vec3 albedo, vec3 n, vec3 wo, vec3 r, BSDFSample& s) {
    switch (mat.type) {
    case MatLambert:
        s.wi = sampleCosineWeightedHemisphere(n, V2(r.x, r.y));
        s.pdf = absDot(n, s.wi) * PiInv;
        s.bsdf = albedo * PiInv;
        s.type = Diffuse | Reflection;
        return true;
    case MatMetallicWorkflow: return metallicWorkflowSampleBSDF(mat, albedo, n, wo, r, s);
    case MatMetal: return metalSampleBSDF(mat, albedo, n, wo, r, s);
    case MatDielectric: return dielectricSampleBSDF(mat, albedo, n, wo, r, s);
    case MatFake: s.wi = -wo; s.bsdf = albedo; s.pdf = 1.0f; s.type = Specular | Transmission; return true;
    default: return false;
    }
}
static inline bool isBSDFDelta(const Material& mat) {
    switch (mat.type) {
    case MatLambert: return false;
    case MatMetallicWorkflow: return isGTR2Delta(mat.roughness) && mat.metallic > 0.9f;
    case MatMetal: return isGTR2Delta(mat.roughness);
    default: return true;
    }
}
static inline bool isBSDFConnectible(const Material& mat) {
    switch (mat.type) {
    case MatLambert: return true;
    case MatMetallicWorkflow: return isGTR2Connectible(mat.roughness) || mat.metallic < 0.9f;
    case MatMetal: return isGTR2Connectible(mat.roughness);
    default: return false;
    }
}
static inline bool isSampleTypeDelta(uint32_t type) { return (type & Specular) == Specular; }

// ---- ray_layouts.glsl ----------------------------------------------------------
static const float MinRayDistance = 1e-4f;
static const float MaxRayDistance = 1e7f;
static const uint32_t InvalidHitIndex = 0xffffffffu;
static const uint32_t SpecialHitIndex = 0xfffffffeu;
static const uint32_t InvalidResourceIdx = 0xffffffffu;

struct SurfaceInfo { vec3 pos; vec3 norm; vec3 albedo; uint32_t matIndex; bool isLight; };

// ray_layouts.glsl:50-57
static inline void loadLightSurfaceInfo(const SceneArrays& S, uint32_t tri, vec3 bary, SurfaceInfo& info) {
    const TriangleLight& L = S.lights[tri];
    info.pos = L3(L.v0) * bary.x + L3(L.v1) * bary.y + L3(L.v2) * bary.z;
    info.norm = V3(L.nx, L.ny, L.nz);
    info.albedo = L3(L.radiance);
    info.matIndex = 0;  // left unset by the reference; never read for lights
}
// ray_layouts.glsl:59-88
// ---- texture(uTextures[i], uv) (ray_layouts.glsl:80-86, GBuffer.frag:33-37) ------
// The sampler of zvk::Image::samplerCreateInfo (zvk/core/Memory.cpp:75-92): repeat wrap,
// nearest or linear filter, level 0 (no mip data is generated). The image format is
// R8G8B8A8_SRGB, so each texel is converted to linear before filtering (Vulkan spec,
// "Texel Filtering": sRGB conversion precedes filtering). Definitions fixed here where the
// API leaves the implementation free: the sRGB transfer function evaluated in double and
// rounded to float (a 256-entry table); texel indices from floor in float32, clamped to
// +-2^30 before the integer conversion; bilinear weights with 8 fractional bits
// (subTexelPrecisionBits = 8), rounded to nearest even; the blend
// (t00 (1-a) + t10 a) (1-b) + (t01 (1-a) + t11 a) b.
static inline const float* srgbTable() {
    static float t[256];
    static bool init = false;
    if (!init) {
        for (int c = 0; c < 256; c++) {
            double v = c / 255.0;
            t[c] = (float)(v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4));
        }
        init = true;
    }
    return t;
}
static inline int texIndex(float x) { return (int)std::fmax(std::fmin(x, 1073741824.0f), -1073741824.0f); }
static inline int wrapRepeat(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }
static inline vec3 texelLinear(const Texture& t, int x, int y) {
    const uint8_t* p = t.rgba8 + 4 * ((size_t)y * t.width + (size_t)x);
    const float* lut = srgbTable();
    return V3(lut[p[0]], lut[p[1]], lut[p[2]]);
}
static inline vec3 sampleTexture(const Texture& t, float u, float v) {
    const int W = (int)t.width, H = (int)t.height;
    if (t.filter == 0) {
        int ix = wrapRepeat(texIndex(std::floor(u * (float)W)), W), iy = wrapRepeat(texIndex(std::floor(v * (float)H)), H);
        return texelLinear(t, ix, iy);
    }
    float fu = u * (float)W - 0.5f, fv = v * (float)H - 0.5f;
    float i0f = std::floor(fu), j0f = std::floor(fv);
    float a = std::nearbyint((fu - i0f) * 256.0f) * (1.0f / 256.0f);
    float b = std::nearbyint((fv - j0f) * 256.0f) * (1.0f / 256.0f);
    int i0 = texIndex(i0f), j0 = texIndex(j0f);
    int x0 = wrapRepeat(i0, W), x1 = wrapRepeat(i0 + 1, W), y0 = wrapRepeat(j0, H), y1 = wrapRepeat(j0 + 1, H);
    vec3 t00 = texelLinear(t, x0, y0), t10 = texelLinear(t, x1, y0), t01 = texelLinear(t, x0, y1),
         t11 = texelLinear(t, x1, y1);
    return (t00 * (1.0f - a) + t10 * a) * (1.0f - b) + (t01 * (1.0f - a) + t11 * a) * b;
}
// A material's albedo at texture coordinate (u, v): baseColor without a texture.
static inline vec3 materialAlbedo(const SceneArrays& S, const Material& m, float u, float v) {
    if (m.textureIdx == InvalidResourceIdx) return L3(m.baseColor);
    return sampleTexture(S.textures[m.textureIdx], u, v);
}

static inline void loadObjectSurfaceInfo(const SceneArrays& S, uint32_t inst, uint32_t tri, vec3 bary, SurfaceInfo& info) {
    const ObjectInstance& I = S.instances[inst];
    info.matIndex = (uint32_t)S.materialIndices[I.indexOffset / 3 + tri];
    uint32_t i0 = S.indices[I.indexOffset + tri * 3 + 0];
    uint32_t i1 = S.indices[I.indexOffset + tri * 3 + 1];
    uint32_t i2 = S.indices[I.indexOffset + tri * 3 + 2];
    const MeshVertex &v0 = S.vertices[i0], &v1 = S.vertices[i1], &v2 = S.vertices[i2];
    vec3 pos = L3(v0.pos) * bary.x + L3(v1.pos) * bary.y + L3(v2.pos) * bary.z;
    vec3 norm = L3(v0.norm) * bary.x + L3(v1.norm) * bary.y + L3(v2.norm) * bary.z;
    info.pos = mat4MulPoint(I.transform, pos, 1.0f);
    info.norm = normalize(mat4MulPoint(I.transformInvT, norm, 1.0f));
    float uvx = v0.uvx * bary.x + v1.uvx * bary.y + v2.uvx * bary.z;
    float uvy = v0.uvy * bary.x + v1.uvy * bary.y + v2.uvy * bary.z;
    info.albedo = materialAlbedo(S, S.materials[info.matIndex], uvx, uvy);
}
// Geometric normal of a hit triangle in world space (sign arbitrary; callers
// take |cos|): the area measure of a reconnection vertex's Jacobian.
static inline vec3 geomNormal(const SceneArrays& S, const Intersection& isec) {
    if (isec.instanceIdx == 0) {
        const TriangleLight& L = S.lights[isec.triangleIdx];
        return V3(L.nx, L.ny, L.nz);
    }
    const ObjectInstance& I = S.instances[isec.instanceIdx - 1];
    const uint32_t io = I.indexOffset, tri = isec.triangleIdx;
    const vec3 p0 = mat4MulPoint(I.transform, L3(S.vertices[S.indices[io + tri * 3 + 0]].pos), 1.0f);
    const vec3 p1 = mat4MulPoint(I.transform, L3(S.vertices[S.indices[io + tri * 3 + 1]].pos), 1.0f);
    const vec3 p2 = mat4MulPoint(I.transform, L3(S.vertices[S.indices[io + tri * 3 + 2]].pos), 1.0f);
    return normalize(cross(p1 - p0, p2 - p0));
}
// ray_layouts.glsl:90-104
static inline void loadSurfaceInfo(const SceneArrays& S, const Intersection& isec, SurfaceInfo& info) {
    vec3 b = V3(1.0f - isec.bary[0] - isec.bary[1], isec.bary[0], isec.bary[1]);
    if (isec.instanceIdx == 0) { loadLightSurfaceInfo(S, isec.triangleIdx, b, info); info.isLight = true; }
    else { loadObjectSurfaceInfo(S, isec.instanceIdx - 1, isec.triangleIdx, b, info); info.isLight = false; }
}

// ---- light_sampling.glsl:6-53 --------------------------------------------------
static inline vec3 sampleTriangleLight(const TriangleLight& light, vec3 ref, vec3& wi, float& dist, float& pdf,
                                       float& jacobian, vec2& bary, vec2 r) {
    bary = uvToBary(r);
    vec3 pos = L3(light.v0) * (1.0f - bary.x - bary.y) + L3(light.v1) * bary.x + L3(light.v2) * bary.y;
    dist = distance(ref, pos);
    vec3 n = V3(light.nx, light.ny, light.nz);
    wi = (pos - ref) / dist;
    jacobian = absDot(n, wi) / square(dist);
    pdf = 1.0f / jacobian / light.area;
    vec3 radiance = L3(light.radiance);
    if (dot(n, wi) > 0.0f) radiance = V3s(0.0f);
    return radiance;
}
static inline vec3 sampleLightByPower(const SceneArrays& S, vec3 ref, vec3& wi, float& dist, float& pdf, float& jacobian,
                                      vec2& bary, uint32_t& id, vec4 r) {
    float sumPower = S.lightTable[0].prob;
    uint32_t numLights = S.lightTable[0].failId;
    id = (uint32_t)((float)numLights * r.x);
    // r.x can be exactly 1.0 (sample1f), which indexes one past the table.
    // Out-of-range buffer reads return zero (Vulkan robustBufferAccess).
    LightSampleTableElement e = (id + 1 < S.numLightTable) ? S.lightTable[id + 1] : LightSampleTableElement{0.0f, 0u};
    id = (r.y < e.prob) ? id : e.failId - 1;
    static const TriangleLight zeroLight = {};
    const TriangleLight& light = (id < S.numLights) ? S.lights[id] : zeroLight;
    vec3 radiance = sampleTriangleLight(light, ref, wi, dist, pdf, jacobian, bary, V2(r.z, r.w));
    pdf *= luminance(L3(light.radiance)) * light.area / sumPower;
    return radiance;
}

}  // namespace orc
