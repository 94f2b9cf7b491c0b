// Device scene view, BSDFs, light sampling and surface loading for the gfx950
// kernels. Restates src/shader/material.glsl:31-364,
// src/shader/light_sampling.glsl:6-53 and src/shader/ray_layouts.glsl:50-104
// over this build's HBM layout:
//   nodes     4 x float4 per internal LBVH node (64 B): both child boxes and
//             both child references (bit 31 = leaf)
//   leafTris  3 x float4 per leaf in Morton order (48 B): world-space
//             vertices; .w of the first two carry the instance custom index
//             and primitive index
//   instances 7 x float4 (112 B): transform columns 0-3 (xyz), transformInvT
//             columns 0-3 (xyz), indexOffset
#pragma once
#include "rt_dmath.h"
#include "../../include/restirpt.h"

namespace rtd {

struct DevInstance {
    float m[12];     // transform columns c0..c3 (xyz each)
    float n[12];     // transformInvT columns c0..c3 (xyz each)
    uint32_t indexOffset;
    uint32_t triCount;   // triangles of the instance (bounds of sample references)
    uint32_t pad[2];
};

struct DevScene {
    const float4* nodes;
    const float4* leafTris;
    const rt_material* materials;
    const int32_t* materialIndices;
    const rt_mesh_vertex* vertices;
    const uint32_t* indices;
    const DevInstance* instances;
    const rt_triangle_light* lights;
    const rt_light_table_entry* lightTable;
    const uint4* textures;        // per texture {first texel, width, height, filter}
    const uint32_t* texels;       // RGBA8
    const float* srgb;            // sRGB byte -> linear
    uint32_t numTextures;
    uint32_t numPrims;
    uint32_t numLights;
    uint32_t numLightTable;
    uint32_t numMaterials;
    uint32_t numInstances;
};

// ---- material.glsl --------------------------------------------------------------------
enum : uint32_t { kDiffuse = 1u << 0, kGlossy = 1u << 1, kSpecular = 1u << 2, kReflection = 1u << 4,
                  kTransmission = 1u << 5, kInvalidBSDFSample = 0x80000000u };
enum : uint32_t { kLambert = 1, kMetallicWorkflow = 2, kMetal = 3, kDielectric = 4, kFake = 6 };

struct Mat { f3 baseColor; uint32_t type; float metallic, roughness, ior; };

RT_DEV Mat loadMat(const DevScene& S, uint32_t i) {
    const rt_material* p = S.materials + i;
    Mat m;
    m.baseColor = mk3(p->baseColor[0], p->baseColor[1], p->baseColor[2]);
    m.type = p->type; m.metallic = p->metallic; m.roughness = p->roughness; m.ior = p->ior;
    return m;
}

struct BSDFSample { f3 wi; float pdf; f3 bsdf; uint32_t type; };

RT_DEV float fresnelSchlickF(float cosTheta, float ior) {
    float f0 = fabsf(1.0f - ior) / (1.0f + ior);
    return mixf(f0, 1.0f, pow5(1.0f - cosTheta));
}
RT_DEV f3 fresnelSchlick3(float cosTheta, f3 f0) { return mix3s(f0, s3(1.0f), pow5(1.0f - cosTheta)); }
RT_DEV float schlickG(float cosTheta, float alpha) {
    float a = alpha * 0.5f;
    return cosTheta / (cosTheta * (1.0f - a) + a);
}
RT_DEV float smithG(float cosWo, float cosWi, float alpha) {
    return schlickG(fabsf(cosWo), alpha) * schlickG(fabsf(cosWi), alpha);
}
RT_DEV float GTR2Distrib(float cosTheta, float alpha) {
    if (cosTheta < 1e-6f) return 0.0f;
    float aa = alpha * alpha;
    float denom = cosTheta * cosTheta * (aa - 1.0f) + 1.0f;
    denom = denom * denom * RT_PI;
    return aa / denom;
}
RT_DEV float GTR2Pdf(f3 n, f3 m, f3 wo, float alpha) {
    return GTR2Distrib(dot(n, m), alpha) * schlickG(dot(n, wo), alpha) * absDot(m, wo) / absDot(n, wo);
}
RT_DEV f3 GTR2Sample(f3 n, f3 wo, float alpha, f2 r) {
    m3 T = matLocalToWorld(n);
    m3 Ti = inverse3(T);
    f3 vh = normalize(mul(Ti, wo) * mk3(alpha, alpha, 1.0f));
    float lenSq = vh.x * vh.x + vh.y * vh.y;
    f3 t = lenSq > 0.0f ? mk3(-vh.y, vh.x, 0.0f) / sqrtf(lenSq) : mk3(1.0f, 0.0f, 0.0f);
    f3 b = cross(vh, t);
    f2 p = toConcentricDisk(r);
    float s = 0.5f * (vh.z + 1.0f);
    p.y = (1.0f - s) * sqrtf(1.0f - p.x * p.x) + s * p.y;
    f3 wh = t * p.x + b * p.y + vh * sqrtf(gmax(0.0f, 1.0f - dot2(p, p)));
    wh = mk3(wh.x * alpha, wh.y * alpha, gmax(0.0f, wh.z));
    return normalize(mul(T, wh));
}
RT_DEV bool isGTR2Connectible(float r) { return r > 0.05f; }
RT_DEV bool isGTR2Delta(float r) { return r < 0.01f; }

RT_DEV f3 metallicWorkflowBSDF(const Mat& mat, f3 albedo, f3 n, f3 wo, f3 wi) {
    float alpha = square(mat.roughness);
    f3 wh = normalize(wo + wi);
    float cosO = dot(n, wo);
    float cosI = dot(n, wi);
    if (cosI * cosO < 1e-7f) return s3(0.0f);
    f3 f = fresnelSchlick3(dot(wh, wo), mix3s(s3(0.08f), albedo, mat.metallic));
    float g = smithG(cosO, cosI, alpha);
    float d = GTR2Distrib(dot(n, wh), alpha);
    return mix3(albedo * RT_PI_INV * (1.0f - mat.metallic), s3(g * d / (4.0f * cosI * cosO)), f);
}
RT_DEV float metallicWorkflowPdf(const Mat& mat, f3 n, f3 wo, f3 wi) {
    f3 wh = normalize(wo + wi);
    return mixf(satDot(n, wi) * RT_PI_INV, GTR2Pdf(n, wh, wo, square(mat.roughness)) / (4.0f * absDot(wh, wo)),
                1.0f / (2.0f - mat.metallic));
}
RT_DEV f3 metalBSDF(const Mat& mat, f3 albedo, f3 n, f3 wo, f3 wi) {
    if (isGTR2Delta(mat.roughness)) return s3(0.0f);
    float alpha = square(mat.roughness);
    f3 wh = normalize(wo + wi);
    float cosO = dot(n, wo);
    float cosI = dot(n, wi);
    if (cosI * cosO < 1e-7f) return s3(0.0f);
    float f = fresnelSchlickF(absDot(wh, wo), mat.ior);
    float g = smithG(cosO, cosI, alpha);
    float d = GTR2Distrib(dot(n, wh), alpha);
    return albedo * f * g * d / (4.0f * cosI * cosO);
}
RT_DEV float metalPdf(const Mat& mat, f3 n, f3 wo, f3 wi) {
    if (isGTR2Delta(mat.roughness)) return 0.0f;
    f3 wh = normalize(wo + wi);
    return GTR2Pdf(n, wh, wo, square(mat.roughness)) / (4.0f * absDot(wh, wo));
}
RT_DEV bool refractDir(f3 n, f3 wi, float ior, f3& wt) {
    float cosIn = dot(n, wi);
    if (cosIn < 0.0f) ior = 1.0f / ior;
    float sin2In = gmax(0.0f, 1.0f - cosIn * cosIn);
    float sin2Tr = sin2In / (ior * ior);
    if (sin2Tr >= 1.0f) return false;
    float cosTr = sqrtf(1.0f - sin2Tr);
    if (cosIn < 0.0f) cosTr = -cosTr;
    wt = normalize(-wi / ior + n * (cosIn / ior - cosTr));
    return true;
}

RT_DEV f3 evalBSDF(const Mat& mat, f3 albedo, f3 n, f3 wo, f3 wi) {
    switch (mat.type) {
    case kLambert: return albedo * RT_PI_INV;
    case kMetallicWorkflow: return metallicWorkflowBSDF(mat, albedo, n, wo, wi);
    case kMetal: return metalBSDF(mat, albedo, n, wo, wi);
    default: return s3(0.0f);
    }
}
RT_DEV float evalPdf(const Mat& mat, f3 n, f3 wo, f3 wi) {
    switch (mat.type) {
    case kLambert: return absDot(n, wi) * RT_PI_INV;
    case kMetallicWorkflow: return metallicWorkflowPdf(mat, n, wo, wi);
    case kMetal: return metalPdf(mat, n, wo, wi);
    default: return 0.0f;
    }
}
RT_DEV bool sampleBSDF(const Mat& mat, f3 albedo, f3 n, f3 wo, f3 r, BSDFSample& s) {
    switch (mat.type) {
    case kLambert:
        s.wi = sampleCosineWeightedHemisphere(n, mk2(r.x, r.y));
        s.pdf = absDot(n, s.wi) * RT_PI_INV;
        s.bsdf = albedo * RT_PI_INV;
        s.type = kDiffuse | kReflection;
        return true;
    case kMetallicWorkflow: {
        float alpha = square(mat.roughness);
        s.type = kReflection;
        if (r.z > (1.0f / (2.0f - mat.metallic))) {
            s.wi = sampleCosineWeightedHemisphere(n, mk2(r.x, r.y));
            s.type |= kDiffuse;
        } else {
            f3 wh = GTR2Sample(n, wo, alpha, mk2(r.x, r.y));
            s.wi = -reflect(wo, wh);
            s.type |= isGTR2Delta(mat.roughness) ? kSpecular : kGlossy;
        }
        if (dot(n, s.wi) < 0.0f) { s.type = kInvalidBSDFSample; return false; }
        s.bsdf = metallicWorkflowBSDF(mat, albedo, n, wo, s.wi);
        s.pdf = metallicWorkflowPdf(mat, n, wo, s.wi);
        return true;
    }
    case kMetal: {
        float alpha = square(mat.roughness);
        bool isDelta = isGTR2Delta(mat.roughness);
        if (isDelta) s.wi = -reflect(wo, n);
        else { f3 wh = GTR2Sample(n, wo, alpha, mk2(r.x, r.y)); s.wi = -reflect(wo, wh); }
        if (dot(n, s.wi) < 0.0f) { s.type = kInvalidBSDFSample; return false; }
        s.bsdf = isDelta ? albedo * fresnelSchlickF(absDot(n, wo), mat.ior) : metalBSDF(mat, albedo, n, wo, s.wi);
        s.pdf = isDelta ? 1.0f : metalPdf(mat, n, wo, s.wi);
        s.type = kReflection | (isDelta ? kSpecular : kGlossy);
        return true;
    }
    case kDielectric: {
        float ior = mat.ior;
        float pdfReflect = fresnelSchlickF(dot(n, wo), ior);
        s.bsdf = albedo;
        if (r.z < pdfReflect) {
            s.wi = reflect(-wo, n);
            s.type = kSpecular | kReflection;
            s.pdf = 1.0f;
        } else {
            if (!refractDir(n, wo, ior, s.wi)) { s.type = kInvalidBSDFSample; return false; }
            if (dot(n, wo) < 0.0f) ior = 1.0f / ior;
            s.bsdf = s.bsdf / (ior * ior);
            s.type = kSpecular | kTransmission;
            s.pdf = 1.0f;
        }
        return true;
    }
    case kFake:
        s.wi = -wo; s.bsdf = albedo; s.pdf = 1.0f; s.type = kSpecular | kTransmission;
        return true;
    default:
        return false;
    }
}
RT_DEV bool isBSDFDelta(const Mat& mat) {
    switch (mat.type) {
    case kLambert: return false;
    case kMetallicWorkflow: return isGTR2Delta(mat.roughness) && mat.metallic > 0.9f;
    case kMetal: return isGTR2Delta(mat.roughness);
    default: return true;
    }
}
RT_DEV bool isBSDFConnectible(const Mat& mat) {
    switch (mat.type) {
    case kLambert: return true;
    case kMetallicWorkflow: return isGTR2Connectible(mat.roughness) || mat.metallic < 0.9f;
    case kMetal: return isGTR2Connectible(mat.roughness);
    default: return false;
    }
}
RT_DEV bool isSampleTypeDelta(uint32_t type) { return (type & kSpecular) == kSpecular; }

// ---- ray_layouts.glsl ---------------------------------------------------------------------
#define RT_MIN_RAY 1e-4f
#define RT_MAX_RAY 1e7f
#define RT_INVALID_HIT 0xffffffffu
#define RT_SPECIAL_HIT 0xfffffffeu

struct Isec { float b1, b2; uint32_t inst, prim; };
struct Surf { f3 pos; f3 norm; f3 albedo; uint32_t matIndex; bool isLight; };

RT_DEV f3 colMul4(const float* m, f3 p, float w) {  // GLSL mat4 * vec4(p, w), xyz
    return mk3(m[0], m[1], m[2]) * p.x + mk3(m[3], m[4], m[5]) * p.y + mk3(m[6], m[7], m[8]) * p.z +
           mk3(m[9], m[10], m[11]) * w;
}

// ---- texture(uTextures[i], uv) (ray_layouts.glsl:80-86, GBuffer.frag:33-37) ------------
// R8G8B8A8_SRGB image, repeat sampler, level 0 (zvk/core/Memory.cpp:75-92,
// src/Scene.cpp:428-445): texels converted to linear, then nearest or bilinear
// filtering with 8-bit weights rounded to nearest even; texel indices from
// floor in float32, clamped to +-2^30 before the integer conversion.
RT_DEV int texIndex(float x) { return (int)fmaxf(fminf(x, 1073741824.0f), -1073741824.0f); }
RT_DEV int wrapRepeat(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }
RT_DEV f3 texelLinear(const DevScene& S, const uint4& t, int x, int y) {
    const uint32_t c = S.texels[t.x + (uint32_t)y * t.y + (uint32_t)x];
    return mk3(S.srgb[c & 255u], S.srgb[(c >> 8) & 255u], S.srgb[(c >> 16) & 255u]);
}
RT_DEV f3 sampleTexture(const DevScene& S, uint32_t idx, float u, float v) {
    const uint4 t = S.textures[idx];
    const int W = (int)t.y, H = (int)t.z;
    if (t.w == 0u) {
        const int ix = wrapRepeat(texIndex(floorf(u * (float)W)), W), iy = wrapRepeat(texIndex(floorf(v * (float)H)), H);
        return texelLinear(S, t, ix, iy);
    }
    const float fu = u * (float)W - 0.5f, fv = v * (float)H - 0.5f;
    const float i0f = floorf(fu), j0f = floorf(fv);
    const float a = rintf((fu - i0f) * 256.0f) * (1.0f / 256.0f);
    const float b = rintf((fv - j0f) * 256.0f) * (1.0f / 256.0f);
    const int i0 = texIndex(i0f), j0 = texIndex(j0f);
    const int x0 = wrapRepeat(i0, W), x1 = wrapRepeat(i0 + 1, W), y0 = wrapRepeat(j0, H), y1 = wrapRepeat(j0 + 1, H);
    const f3 t00 = texelLinear(S, t, x0, y0), t10 = texelLinear(S, t, x1, y0), t01 = texelLinear(S, t, x0, y1),
             t11 = texelLinear(S, t, x1, y1);
    return (t00 * (1.0f - a) + t10 * a) * (1.0f - b) + (t01 * (1.0f - a) + t11 * a) * b;
}
// A material's albedo at texture coordinate (u, v): baseColor without a texture.
RT_DEV f3 materialAlbedo(const DevScene& S, const rt_material& m, float u, float v) {
    if (m.textureIdx == 0xffffffffu) return ld3(m.baseColor);
    return sampleTexture(S, m.textureIdx, u, v);
}

RT_DEV void loadSurfaceInfo(const DevScene& S, const Isec& is, Surf& info) {
    float b0 = 1.0f - is.b1 - is.b2;
    if (is.inst == 0) {
        const rt_triangle_light& L = S.lights[is.prim];
        info.pos = ld3(L.v0) * b0 + ld3(L.v1) * is.b1 + ld3(L.v2) * is.b2;
        info.norm = mk3(L.nx, L.ny, L.nz);
        info.albedo = ld3(L.radiance);
        info.matIndex = 0;
        info.isLight = true;
        return;
    }
    const DevInstance& I = S.instances[is.inst - 1];
    uint32_t io = I.indexOffset;
    info.matIndex = (uint32_t)S.materialIndices[io / 3 + is.prim];
    const rt_mesh_vertex& v0 = S.vertices[S.indices[io + is.prim * 3 + 0]];
    const rt_mesh_vertex& v1 = S.vertices[S.indices[io + is.prim * 3 + 1]];
    const rt_mesh_vertex& v2 = S.vertices[S.indices[io + is.prim * 3 + 2]];
    f3 pos = ld3(v0.pos) * b0 + ld3(v1.pos) * is.b1 + ld3(v2.pos) * is.b2;
    f3 nrm = ld3(v0.norm) * b0 + ld3(v1.norm) * is.b1 + ld3(v2.norm) * is.b2;
    info.pos = colMul4(I.m, pos, 1.0f);
    info.norm = normalize(colMul4(I.n, nrm, 1.0f));
    const rt_material& m = S.materials[info.matIndex];
    if (m.textureIdx == 0xffffffffu) {
        info.albedo = ld3(m.baseColor);
    } else {
        const float uvx = v0.uvx * b0 + v1.uvx * is.b1 + v2.uvx * is.b2;
        const float uvy = v0.uvy * b0 + v1.uvy * is.b1 + v2.uvy * is.b2;
        info.albedo = sampleTexture(S, m.textureIdx, uvx, uvy);
    }
    info.isLight = false;
}

// Geometric normal of a hit triangle in world space (its sign is arbitrary:
// callers take |cos|). The area-to-solid-angle factor of a reconnection
// vertex is |n_g . w| / d^2; the interpolated shading normal is not the
// surface's measure.
RT_DEV f3 geomNormal(const DevScene& S, const Isec& is) {
    if (is.inst == 0) { const rt_triangle_light& L = S.lights[is.prim]; return mk3(L.nx, L.ny, L.nz); }
    const DevInstance& I = S.instances[is.inst - 1];
    const uint32_t io = I.indexOffset;
    const f3 p0 = colMul4(I.m, ld3(S.vertices[S.indices[io + is.prim * 3 + 0]].pos), 1.0f);
    const f3 p1 = colMul4(I.m, ld3(S.vertices[S.indices[io + is.prim * 3 + 1]].pos), 1.0f);
    const f3 p2 = colMul4(I.m, ld3(S.vertices[S.indices[io + is.prim * 3 + 2]].pos), 1.0f);
    return normalize(cross(p1 - p0, p2 - p0));
}

// ---- light_sampling.glsl:6-53 -----------------------------------------------------------------
RT_DEV f3 sampleLightByPower(const DevScene& S, f3 ref, f3& wi, float& dist, float& pdf, float& jacobian, f2& bary,
                             uint32_t& id, f4 r) {
    float sumPower = S.lightTable[0].prob;
    uint32_t numLights = S.lightTable[0].failId;
    id = (uint32_t)((float)numLights * r.x);
    // r.x == 1.0 indexes one past the table: such reads return zero (robust access)
    float eProb = 0.0f; uint32_t eFail = 0u;
    if (id + 1 < S.numLightTable) { eProb = S.lightTable[id + 1].prob; eFail = S.lightTable[id + 1].failId; }
    id = (r.y < eProb) ? id : eFail - 1u;
    f3 v0 = s3(0.0f), v1 = s3(0.0f), v2 = s3(0.0f), n = s3(0.0f), rad = s3(0.0f);
    float area = 0.0f;
    if (id < S.numLights) {
        const rt_triangle_light& L = S.lights[id];
        v0 = ld3(L.v0); v1 = ld3(L.v1); v2 = ld3(L.v2); n = mk3(L.nx, L.ny, L.nz); rad = ld3(L.radiance); area = L.area;
    }
    bary = uvToBary(mk2(r.z, r.w));
    f3 pos = v0 * (1.0f - bary.x - bary.y) + v1 * bary.x + v2 * bary.y;
    dist = distance(ref, pos);
    wi = (pos - ref) / dist;
    jacobian = absDot(n, wi) / square(dist);
    pdf = 1.0f / jacobian / area;
    f3 radiance = (dot(n, wi) > 0.0f) ? s3(0.0f) : rad;
    pdf *= luminance(rad) * area / sumPower;
    return radiance;
}

}  // namespace rtd
