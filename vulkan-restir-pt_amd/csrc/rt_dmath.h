// Device math for the gfx950 ReSTIR PT kernels.
//
// FP contract: every kernel is compiled with -ffp-contract=off and correctly
// rounded FP32 division and square root, so each expression below rounds
// exactly as written. Operation order follows the reference GLSL
// (src/shader/math.glsl); where GLSL leaves the implementation free, this
// build fixes one definition:
//   sin/cos   Cody-Waite reduction by pi/2 and the Cephes single-precision
//             polynomials (Moshier, Cephes sinf.c); tan = sin / cos
//   normalize v * (1 / sqrt(dot(v, v)))
//   inverse   3x3 adjugate / determinant
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RT_DEV __device__ __forceinline__

namespace rtd {

struct f2 { float x, y; };
struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };
struct m3 { f3 c0, c1, c2; };

RT_DEV f2 mk2(float a, float b) { f2 r; r.x = a; r.y = b; return r; }
RT_DEV f3 mk3(float a, float b, float c) { f3 r; r.x = a; r.y = b; r.z = c; return r; }
RT_DEV f3 s3(float a) { return mk3(a, a, a); }
RT_DEV f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

RT_DEV f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
RT_DEV f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
RT_DEV f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
RT_DEV f3 operator/(f3 a, f3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
RT_DEV f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
RT_DEV f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
RT_DEV f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
RT_DEV f2 operator+(f2 a, f2 b) { return mk2(a.x + b.x, a.y + b.y); }
RT_DEV f2 operator*(f2 a, f2 b) { return mk2(a.x * b.x, a.y * b.y); }
RT_DEV f2 operator*(f2 a, float s) { return mk2(a.x * s, a.y * s); }

RT_DEV float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RT_DEV float dot2(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
RT_DEV f3 cross(f3 a, f3 b) { return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
RT_DEV float length(f3 v) { return sqrtf(dot(v, v)); }
RT_DEV float distance(f3 a, f3 b) { return length(a - b); }
RT_DEV f3 normalize(f3 v) { float inv = 1.0f / sqrtf(dot(v, v)); return v * inv; }
RT_DEV float gmin(float a, float b) { return b < a ? b : a; }
RT_DEV float gmax(float a, float b) { return a < b ? b : a; }
RT_DEV float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
RT_DEV f3 reflect(f3 I, f3 N) { return I - N * (2.0f * dot(N, I)); }
RT_DEV float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
RT_DEV f3 mix3(f3 x, f3 y, f3 a) {
    return mk3(x.x * (1.0f - a.x) + y.x * a.x, x.y * (1.0f - a.y) + y.y * a.y, x.z * (1.0f - a.z) + y.z * a.z);
}
RT_DEV f3 mix3s(f3 x, f3 y, float a) { return x * (1.0f - a) + y * a; }
RT_DEV f3 mul(const m3& m, f3 v) { return m.c0 * v.x + m.c1 * v.y + m.c2 * v.z; }

// ---- trig ------------------------------------------------------------------------
RT_DEV float poly_sin(float r) {
    float z = r * r;
    float p = -1.9515295891e-4f;
    p = p * z + 8.3321608736e-3f;
    p = p * z + -1.6666654611e-1f;
    return r + r * (z * p);
}
RT_DEV float poly_cos(float r) {
    float z = r * r;
    float p = 2.443315711809948e-5f;
    p = p * z + -1.388731625493765e-3f;
    p = p * z + 4.166664568298827e-2f;
    float y = p * z * z;
    y = y - 0.5f * z;
    return y + 1.0f;
}
RT_DEV float reduce_pio2(float x, int& q) {
    float k = rintf(x * 0.636619772367581343f);
    float r = x - k * 1.5703125f;
    r = r - k * 4.837512969970703125e-4f;
    r = r - k * 7.54978995489188216e-8f;
    q = (int)k;
    return r;
}
RT_DEV float rt_sin(float x) {
    int q; float r = reduce_pio2(x, q);
    int m = q & 3;
    float v = (m & 1) ? poly_cos(r) : poly_sin(r);
    return (m & 2) ? -v : v;
}
RT_DEV float rt_cos(float x) {
    int q; float r = reduce_pio2(x, q);
    int m = (q + 1) & 3;
    float v = (m & 1) ? poly_cos(r) : poly_sin(r);
    return (m & 2) ? -v : v;
}
RT_DEV float rt_tan(float x) { return rt_sin(x) / rt_cos(x); }

// ---- exp / log -----------------------------------------------------------------------
// exp(x) = 2^k e^r with k = rint(x / ln 2) and r = x - k ln 2 taken in two parts (Cody and
// Waite's additive range reduction); e^r by its Taylor series through r^7 (|r| <= ln(2)/2:
// truncation below 2^-27 relative). Results below FLT_MIN are returned as 0.
RT_DEV float rt_exp(float x) {
    if (x != x) return x;
    if (x > 88.0f) return INFINITY;
    if (!(x > -87.0f)) return 0.0f;
    const float k = rintf(x * 1.44269504088896341f);
    float r = x - k * 0.693145751953125f;
    r = r - k * 1.42860682030941723212e-6f;
    float p = 1.984126984126984e-4f;
    p = p * r + 1.388888888888889e-3f;
    p = p * r + 8.333333333333333e-3f;
    p = p * r + 4.166666666666667e-2f;
    p = p * r + 1.666666666666667e-1f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return ldexpf(p, (int)k);
}
// log(x) = e ln 2 + log(m) with x = m 2^e, m in [sqrt(1/2), sqrt(2)); log(m) = 2 atanh(s),
// s = (m - 1) / (m + 1), |s| <= 0.1716, by the odd series 2 (s + s^3/3 + ... + s^11/11).
RT_DEV float rt_log(float x) {
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e;
    float m = frexpf(x, &e);
    if (m < 0.70710678118654752f) { m = m * 2.0f; e = e - 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float z = s * s;
    float p = 0.18181818181818182f;
    p = p * z + 0.22222222222222222f;
    p = p * z + 0.28571428571428571f;
    p = p * z + 0.4f;
    p = p * z + 0.66666666666666667f;
    p = p * z + 2.0f;
    const float fe = (float)e;
    return fe * 0.693145751953125f + (fe * 1.42860682030941723212e-6f + s * p);
}

// ---- math.glsl -----------------------------------------------------------------------
#define RT_PI 3.14159265358979323846f
#define RT_PI_INV (1.0f / 3.14159265358979323846f)

RT_DEV float square(float x) { return x * x; }
RT_DEV float satDot(f3 a, f3 b) { return gmax(dot(a, b), 0.0f); }
RT_DEV float absDot(f3 a, f3 b) { return fabsf(dot(a, b)); }
RT_DEV float pow5(float x) { float x2 = x * x; return x2 * x2 * x; }
RT_DEV float luminance(f3 c) { return c.x * 0.299f + c.y * 0.587f + c.z * 0.114f; }
RT_DEV bool isBlack(f3 c) { return luminance(c) < 1e-5f; }
RT_DEV bool hasNan(f3 c) { return isnan(c.x) || isnan(c.y) || isnan(c.z); }
RT_DEV f3 clampColor(f3 c) {
    if (hasNan(c)) return s3(0.0f);
    return mk3(gclamp(c.x, 0.0f, 1e4f), gclamp(c.y, 0.0f, 1e4f), gclamp(c.z, 0.0f, 1e4f));
}
RT_DEV float MISWeight(float f, float g) { return (f * f) / (f * f + g * g); }

RT_DEV f2 toConcentricDisk(f2 v) {
    if (v.x == 0.0f && v.y == 0.0f) return mk2(0.0f, 0.0f);
    v = mk2(v.x * 2.0f - 1.0f, v.y * 2.0f - 1.0f);
    float phi, r;
    if (v.x * v.x > v.y * v.y) {
        r = v.x;
        phi = RT_PI * v.y / v.x * 0.25f;
    } else {
        r = v.y;
        phi = RT_PI * 0.5f - RT_PI * v.x / v.y * 0.25f;
    }
    return mk2(r * rt_cos(phi), r * rt_sin(phi));
}
RT_DEV f3 getTangent(f3 n) { return (fabsf(n.z) > 0.999f) ? mk3(0, 1, 0) : mk3(0, 0, 1); }
RT_DEV m3 matLocalToWorld(f3 n) {
    f3 t = getTangent(n);
    f3 b = normalize(cross(n, t));
    t = cross(b, n);
    m3 m; m.c0 = t; m.c1 = b; m.c2 = n;
    return m;
}
RT_DEV f3 sampleCosineWeightedHemisphere(f3 n, f2 u) {
    f2 uv = toConcentricDisk(u);
    float z = sqrtf(1.0f - dot2(uv, uv));
    return normalize(mul(matLocalToWorld(n), mk3(uv.x, uv.y, z)));
}
RT_DEV f2 uvToBary(f2 uv) {
    float r = sqrtf(uv.y);
    return mk2(1.0f - r, uv.x * r);
}
RT_DEV m3 inverse3(const m3& m) {
    float a00 = m.c0.x, a01 = m.c1.x, a02 = m.c2.x;
    float a10 = m.c0.y, a11 = m.c1.y, a12 = m.c2.y;
    float a20 = m.c0.z, a21 = m.c1.z, a22 = m.c2.z;
    float c00 = a11 * a22 - a12 * a21;
    float c01 = a12 * a20 - a10 * a22;
    float c02 = a10 * a21 - a11 * a20;
    float det = a00 * c00 + a01 * c01 + a02 * c02;
    float inv = 1.0f / det;
    m3 r;
    r.c0 = mk3(c00 * inv, c01 * inv, c02 * inv);
    r.c1 = mk3((a02 * a21 - a01 * a22) * inv, (a00 * a22 - a02 * a20) * inv, (a01 * a20 - a00 * a21) * inv);
    r.c2 = mk3((a01 * a12 - a02 * a11) * inv, (a02 * a10 - a00 * a12) * inv, (a00 * a11 - a01 * a10) * inv);
    return r;
}

// ---- RNG (math.glsl:227-266) ------------------------------------------------------------
RT_DEV uint32_t hash2(uint32_t seed) {
    seed = (seed ^ 61u) ^ (seed >> 16u);
    seed *= 9u;
    seed = seed ^ (seed >> 4u);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15u);
    return seed;
}
RT_DEV uint32_t makeSeedPix(uint32_t seed, uint32_t ix, uint32_t iy) {
    return hash2((seed + ix) ^ (iy - 1u)) + hash2(iy * (ix - 2u));
}
RT_DEV float sample1f(uint32_t& rng) { rng = hash2(rng); return (float)rng / 4294967296.0f; }
RT_DEV f2 sample2f(uint32_t& rng) { float a = sample1f(rng); float b = sample1f(rng); return mk2(a, b); }
RT_DEV f3 sample3f(uint32_t& rng) { float a = sample1f(rng); float b = sample1f(rng); float c = sample1f(rng); return mk3(a, b, c); }
RT_DEV f4 sample4f(uint32_t& rng) {
    f4 r; r.x = sample1f(rng); r.y = sample1f(rng); r.z = sample1f(rng); r.w = sample1f(rng); return r;
}

// ---- bit casts ---------------------------------------------------------------------------
RT_DEV uint32_t fbits(float f) { return __float_as_uint(f); }
RT_DEV float bitsf(uint32_t u) { return __uint_as_float(u); }

}  // namespace rtd
