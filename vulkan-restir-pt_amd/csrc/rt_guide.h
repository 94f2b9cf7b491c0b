// Learned directional proposal at the primary vertex (no counterpart in the
// reference): an 8-lobe von Mises-Fisher mixture per pixel, predicted by the
// proposal network from the primary vertex's position and stored in lobe
// planes, sampled with probability kGuideFraction and otherwise the BSDF,
// with the one-sample MIS density
//   p(w) = kGuideFraction q(w) + (1 - kGuideFraction) p_bsdf(w),
//   q(w) = sum_k pi_k kappa_k / (2 pi (1 - e^{-2 kappa_k})) e^{kappa_k (mu_k . w - 1)}.
// Every function here is deterministic FP32 (own exp / log / sin / cos), so a
// path sampled on the GPU is reproduced bit for bit wherever it is replayed.
//
// The mixture is evaluated once per 2x2 pixel block (at the block's first
// non-background pixel in raster order) and shared by the block's pixels.
// Lobe planes (float4, plane stride = blocks): plane 2k = (mu_k, kappa_k),
// plane 2k + 1 = (pi_k kappa_k / (2 pi (1 - e^{-2 kappa_k})), e^{-2 kappa_k},
// cdf_k = pi_0 + ... + pi_k (cdf_7 = 1), pi_k).
#pragma once
#include "rt_dscene.h"

namespace rtg {

using namespace rtd;

constexpr int kLobes = 8;
constexpr int kPlanes = 2 * kLobes;
constexpr float kGuideFraction = 0.5f;

// One pixel's mixture: L = lobe planes (null: no proposal), n = plane stride.
struct GuidePix {
    const float4* L;
    uint32_t n, pix;
};
RT_DEV GuidePix noGuide() { GuidePix g; g.L = nullptr; g.n = 0; g.pix = 0; return g; }
RT_DEV GuidePix guideAt(const float4* L, uint32_t n, uint32_t pix) { GuidePix g; g.L = L; g.n = n; g.pix = pix; return g; }
RT_DEV uint32_t guideBlocksX(uint32_t W) { return (W + 1u) >> 1; }
RT_DEV uint32_t guideBlocks(uint32_t W, uint32_t H) { return guideBlocksX(W) * ((H + 1u) >> 1); }
// The mixture of pixel (x, y) of a W x H frame (null planes: no proposal).
RT_DEV GuidePix guidePixel(const float4* L, uint32_t W, uint32_t H, uint32_t x, uint32_t y) {
    if (!L) return noGuide();
    return guideAt(L, guideBlocks(W, H), (y >> 1) * guideBlocksX(W) + (x >> 1));
}
// The frame's proposal bindings: lobe planes of this and the previous frame, and the
// per-pixel training record (primary direction, path luminance) the path tracer writes.
struct GuideFrame {
    const float4* lobes;
    const float4* lobesPrev;
    float4* samples;
};

// Materials whose BSDF sampler has no delta component: the mixture is combined with those only.
RT_DEV bool guidable(const Mat& mat) {
    switch (mat.type) {
    case kLambert: return true;
    case kMetallicWorkflow:
    case kMetal: return !isGTR2Delta(mat.roughness);
    default: return false;
    }
}

RT_DEV float mixturePdf(const GuidePix& g, f3 w) {
    float p = 0.0f;
#pragma unroll
    for (int k = 0; k < kLobes; k++) {
        const float4 a = g.L[(size_t)(2 * k) * g.n + g.pix];
        const float4 b = g.L[(size_t)(2 * k + 1) * g.n + g.pix];
        p = p + b.x * rt_exp(a.w * ((a.x * w.x + a.y * w.y + a.z * w.z) - 1.0f));
    }
    return p;
}

// Lobe by inversion of the cdf; direction by the inversion of Jakob ("Numerically stable sampling
// of the von Mises-Fisher distribution on S^2", 2012): cos = 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa.
RT_DEV f3 mixtureSample(const GuidePix& g, float uLobe, float u1, float u2) {
    int k = kLobes - 1;
    for (int i = 0; i < kLobes - 1; i++) {
        if (uLobe < g.L[(size_t)(2 * i + 1) * g.n + g.pix].z) { k = i; break; }
    }
    const float4 a = g.L[(size_t)(2 * k) * g.n + g.pix];
    const float4 b = g.L[(size_t)(2 * k + 1) * g.n + g.pix];
    float wz = 1.0f + rt_log(u1 + (1.0f - u1) * b.y) / a.w;
    wz = gclamp(wz, -1.0f, 1.0f);
    const float r = sqrtf(gmax(1.0f - wz * wz, 0.0f));
    const float phi = 2.0f * RT_PI * u2;
    return normalize(mul(matLocalToWorld(mk3(a.x, a.y, a.z)), mk3(r * rt_cos(phi), r * rt_sin(phi), wz)));
}

// Density of sampling wi at a vertex with this material (the BSDF's own where g has no mixture).
RT_DEV float guidedPdf(const GuidePix& g, const Mat& mat, f3 n, f3 wo, f3 wi) {
    const float pb = evalPdf(mat, n, wo, wi);
    if (!g.L || !guidable(mat)) return pb;
    return kGuideFraction * mixturePdf(g, wi) + (1.0f - kGuideFraction) * pb;
}

// The primary vertex's direction: u = the proposal's four random numbers (drawn whenever the
// proposal is on, so that every domain consumes the same count), rb = the BSDF sampler's three.
RT_DEV bool guidedSampleBSDF(const GuidePix& g, const Mat& mat, f3 albedo, f3 n, f3 wo, f4 u, f3 rb, BSDFSample& s) {
    if (!g.L || !guidable(mat)) return sampleBSDF(mat, albedo, n, wo, rb, s);
    if (u.x < kGuideFraction) {
        s.wi = mixtureSample(g, u.y, u.z, u.w);
        if (!(dot(n, s.wi) > 0.0f)) { s.type = kInvalidBSDFSample; return false; }
        s.bsdf = evalBSDF(mat, albedo, n, wo, s.wi);
        s.type = kReflection | kGlossy;
    } else if (!sampleBSDF(mat, albedo, n, wo, rb, s)) {
        return false;
    }
    s.pdf = kGuideFraction * mixturePdf(g, s.wi) + (1.0f - kGuideFraction) * evalPdf(mat, n, wo, s.wi);
    return true;
}

}  // namespace rtg
