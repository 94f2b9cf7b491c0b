// ORACLE — test infrastructure only. Nothing in the product links or calls this.
//
// The learned directional proposal at the primary vertex (no counterpart in the
// reference), restated for the oracle's path tracer: an 8-lobe von Mises-Fisher
// mixture per 2x2 pixel block read from lobe planes (the kernels' layout: float4 plane
// 2k = (mu_k, kappa_k), plane 2k+1 = (pi_k kappa_k / (2 pi (1 - e^{-2 kappa_k})),
// e^{-2 kappa_k}, cdf_k, pi_k), plane stride = blocks), sampled with probability
// 1/2 and otherwise the BSDF, under the one-sample MIS density
//   p(w) = q(w) / 2 + p_bsdf(w) / 2.
// Lobe directions follow Jakob's inversion for the vMF distribution on S^2
// ("Numerically stable sampling of the von Mises-Fisher distribution on S^2",
// 2012): cos = 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa.
#pragma once
#include "orc_scene.h"

namespace orc {

static const int GuideLobes = 8;
static const float GuideFraction = 0.5f;

struct GuidePix {
    const float* L;  // float4 planes (null: no proposal)
    uint32_t n, pix;
};
static inline GuidePix noGuide() { return GuidePix{nullptr, 0, 0}; }
static inline GuidePix guideAt(const float* L, uint32_t n, uint32_t pix) { return GuidePix{L, n, pix}; }
// One mixture per 2x2 pixel block: plane stride = blocks, pixel (x, y) reads its block's.
static inline uint32_t guideBlocksX(uint32_t W) { return (W + 1u) >> 1; }
static inline uint32_t guideBlocks(uint32_t W, uint32_t H) { return guideBlocksX(W) * ((H + 1u) >> 1); }
static inline GuidePix guidePixel(const float* L, uint32_t W, uint32_t H, uint32_t x, uint32_t y) {
    if (!L) return noGuide();
    return guideAt(L, guideBlocks(W, H), (y >> 1) * guideBlocksX(W) + (x >> 1));
}
static inline const float* guidePlane(const GuidePix& g, int plane) {
    return g.L + 4 * ((size_t)plane * g.n + g.pix);
}

static inline bool guidable(const Material& mat) {
    switch (mat.type) {
    case MatLambert: return true;
    case MatMetallicWorkflow:
    case MatMetal: return !isGTR2Delta(mat.roughness);
    default: return false;
    }
}

static inline float mixturePdf(const GuidePix& g, vec3 w) {
    float p = 0.0f;
    for (int k = 0; k < GuideLobes; k++) {
        const float* a = guidePlane(g, 2 * k);
        const float* b = guidePlane(g, 2 * k + 1);
        p = p + b[0] * rt_exp(a[3] * ((a[0] * w.x + a[1] * w.y + a[2] * w.z) - 1.0f));
    }
    return p;
}

static inline vec3 mixtureSample(const GuidePix& g, float uLobe, float u1, float u2) {
    int k = GuideLobes - 1;
    for (int i = 0; i < GuideLobes - 1; i++) {
        if (uLobe < guidePlane(g, 2 * i + 1)[2]) { k = i; break; }
    }
    const float* a = guidePlane(g, 2 * k);
    const float* b = guidePlane(g, 2 * k + 1);
    float wz = 1.0f + rt_log(u1 + (1.0f - u1) * b[1]) / a[3];
    wz = gclamp(wz, -1.0f, 1.0f);
    const float r = std::sqrt(gmax(1.0f - wz * wz, 0.0f));
    const float phi = 2.0f * Pi * u2;
    return normalize(mul(matLocalToWorld(V3(a[0], a[1], a[2])), V3(r * rt_cos(phi), r * rt_sin(phi), wz)));
}

static inline float guidedPdf(const GuidePix& g, const Material& mat, vec3 n, vec3 wo, vec3 wi) {
    const float pb = evalPdf(mat, n, wo, wi);
    if (!g.L || !guidable(mat)) return pb;
    return GuideFraction * mixturePdf(g, wi) + (1.0f - GuideFraction) * pb;
}

static inline bool guidedSampleBSDF(const GuidePix& g, const Material& mat, vec3 albedo, vec3 n, vec3 wo, vec4 u, vec3 rb,
                                    BSDFSample& s) {
    if (!g.L || !guidable(mat)) return sampleBSDF(mat, albedo, n, wo, rb, s);
    if (u.x < GuideFraction) {
        s.wi = mixtureSample(g, u.y, u.z, u.w);
        if (!(dot(n, s.wi) > 0.0f)) { s.type = InvalidBSDFSample; return false; }
        s.bsdf = evalBSDF(mat, albedo, n, wo, s.wi);
        s.type = Reflection | Glossy;
    } else if (!sampleBSDF(mat, albedo, n, wo, rb, s)) {
        return false;
    }
    s.pdf = GuideFraction * mixturePdf(g, s.wi) + (1.0f - GuideFraction) * evalPdf(mat, n, wo, s.wi);
    return true;
}

}  // namespace orc
