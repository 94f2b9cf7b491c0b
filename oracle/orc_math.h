// ORACLE — test infrastructure only. Nothing in the product links or calls this.
//
// Scalar FP32 restatement of the reference's GLSL math library
// (src/shader/math.glsl). Every operation is written out in the order the GLSL
// expression evaluates it, with no FMA contraction (built with
// -ffp-contract=off). Where GLSL leaves precision to the implementation
// (sin, cos, tan, normalize, inverse), the build fixes one definition and the
// oracle restates it:
//   * sin/cos: Cody-Waite reduction by pi/2 (3-part constant) and the Cephes
//     single-precision minimax polynomials (S. L. Moshier, Cephes Math Library,
//     sinf.c), evaluated by Horner's rule.  tan(x) := sin(x)/cos(x).
//   * normalize(v) := v * (1 / sqrt(dot(v, v)))
//   * inverse(mat3) := adjugate / determinant (cofactor expansion, row 0).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

struct vec2 { float x, y; };
struct vec3 { float x, y, z; };
struct vec4 { float x, y, z, w; };
struct mat3 { vec3 c[3]; };  // column-major, GLSL mat3(c0, c1, c2)

static inline vec2 V2(float a, float b) { return {a, b}; }
static inline vec3 V3(float a, float b, float c) { return {a, b, c}; }
static inline vec3 V3s(float a) { return {a, a, a}; }

static inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline vec3 operator/(vec3 a, vec3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
static inline vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline vec3 operator*(float s, vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline vec3 operator/(vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
static inline vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
static inline vec2 operator+(vec2 a, vec2 b) { return {a.x + b.x, a.y + b.y}; }
static inline vec2 operator-(vec2 a, vec2 b) { return {a.x - b.x, a.y - b.y}; }
static inline vec2 operator*(vec2 a, vec2 b) { return {a.x * b.x, a.y * b.y}; }
static inline vec2 operator*(vec2 a, float s) { return {a.x * s, a.y * s}; }
static inline vec2 operator/(vec2 a, vec2 b) { return {a.x / b.x, a.y / b.y}; }

static inline float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float dot2(vec2 a, vec2 b) { return a.x * b.x + a.y * b.y; }
static inline vec3 cross(vec3 a, vec3 b) {
    // GLSL spec order: (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
static inline float length(vec3 v) { return std::sqrt(dot(v, v)); }
static inline float distance(vec3 a, vec3 b) { return length(a - b); }
static inline vec3 normalize(vec3 v) { float inv = 1.0f / std::sqrt(dot(v, v)); return v * inv; }
static inline float fminN(float a, float b) { return std::fmin(a, b); }
static inline float fmaxN(float a, float b) { return std::fmax(a, b); }
// GLSL min/max/clamp on well-defined inputs.
static inline float gmin(float a, float b) { return b < a ? b : a; }
static inline float gmax(float a, float b) { return a < b ? b : a; }
static inline float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
static inline vec3 reflect(vec3 I, vec3 N) { return I - N * (2.0f * dot(N, I)); }
static inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
static inline vec3 mix3(vec3 x, vec3 y, vec3 a) {
    return {x.x * (1.0f - a.x) + y.x * a.x, x.y * (1.0f - a.y) + y.y * a.y, x.z * (1.0f - a.z) + y.z * a.z};
}
static inline vec3 mix3s(vec3 x, vec3 y, float a) { return x * (1.0f - a) + y * a; }
static inline vec3 mul(const mat3& m, vec3 v) {  // m * v = c0*v.x + c1*v.y + c2*v.z
    return m.c[0] * v.x + m.c[1] * v.y + m.c[2] * v.z;
}

// ---- trigonometry (fixed definition shared by the build) -------------------
static inline float poly_sin(float r) {  // r in [-pi/4, pi/4]
    float z = r * r;
    float p = -1.9515295891e-4f;
    p = p * z + 8.3321608736e-3f;
    p = p * z + -1.6666654611e-1f;
    return r + r * (z * p);
}
static inline float poly_cos(float r) {
    float z = r * r;
    float p = 2.443315711809948e-5f;
    p = p * z + -1.388731625493765e-3f;
    p = p * z + 4.166664568298827e-2f;
    float y = p * z * z;
    y = y - 0.5f * z;
    return y + 1.0f;
}
static inline void sincos_red(float x, float* r, int* q) {
    float k = std::nearbyint(x * 0.636619772367581343f);  // round-half-even
    float rr = x - k * 1.5703125f;
    rr = rr - k * 4.837512969970703125e-4f;
    rr = rr - k * 7.54978995489188216e-8f;
    *r = rr;
    *q = (int)k;
}
static inline float rt_sin(float x) {
    float r; int q; sincos_red(x, &r, &q);
    int m = q & 3;
    float v = (m & 1) ? poly_cos(r) : poly_sin(r);
    return (m & 2) ? -v : v;
}
static inline float rt_cos(float x) {
    float r; int q; sincos_red(x, &r, &q);
    int m = (q + 1) & 3;
    float v = (m & 1) ? poly_cos(r) : poly_sin(r);
    return (m & 2) ? -v : v;
}
static inline float rt_tan(float x) { return rt_sin(x) / rt_cos(x); }

// exp and log of the learned proposal (no counterpart in the reference), the definitions the
// kernels use: exp by Cody-Waite reduction to |r| <= ln(2)/2 and the Taylor series through r^7;
// log by x = m 2^e, m in [sqrt(1/2), sqrt(2)), and log(m) = 2 atanh((m-1)/(m+1)) through s^11.
static inline float rt_exp(float x) {
    if (x != x) return x;
    if (x > 88.0f) return INFINITY;
    if (!(x > -87.0f)) return 0.0f;
    const float k = std::nearbyint(x * 1.44269504088896341f);
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
    return std::ldexp(p, (int)k);
}
static inline float rt_log(float x) {
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e;
    float m = std::frexp(x, &e);
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

// ---- math.glsl ---------------------------------------------------------------
static const float Pi = 3.14159265358979323846f;
static const float PiInv = 1.0f / 3.14159265358979323846f;

static inline float sqr(float x) { return x * x; }
static inline float square(float x) { return x * x; }
static inline float satDot(vec3 a, vec3 b) { return gmax(dot(a, b), 0.0f); }
static inline float absDot(vec3 a, vec3 b) { return std::fabs(dot(a, b)); }
static inline float pow5(float x) { float x2 = x * x; return x2 * x2 * x; }
// math.glsl:171
static inline float luminance(vec3 c) { return c.x * 0.299f + c.y * 0.587f + c.z * 0.114f; }
static inline bool isBlack(vec3 c) { return luminance(c) < 1e-5f; }
static inline bool hasNan(vec3 c) { return std::isnan(c.x) || std::isnan(c.y) || std::isnan(c.z); }
// math.glsl:187-192
static inline vec3 clampColor(vec3 c) {
    if (hasNan(c)) return V3s(0.0f);
    return {gclamp(c.x, 0.0f, 1e4f), gclamp(c.y, 0.0f, 1e4f), gclamp(c.z, 0.0f, 1e4f)};
}
// math.glsl:268-270
static inline float MISWeight(float f, float g) { return (f * f) / (f * f + g * g); }

// math.glsl:23-39
static inline vec2 toConcentricDisk(vec2 v) {
    if (v.x == 0.0f && v.y == 0.0f) return V2(0.0f, 0.0f);
    v = V2(v.x * 2.0f - 1.0f, v.y * 2.0f - 1.0f);
    float phi, r;
    if (v.x * v.x > v.y * v.y) {
        r = v.x;
        phi = Pi * v.y / v.x * 0.25f;
    } else {
        r = v.y;
        phi = Pi * 0.5f - Pi * v.x / v.y * 0.25f;
    }
    return V2(r * rt_cos(phi), r * rt_sin(phi));
}

// math.glsl:69-83
static inline vec3 getTangent(vec3 n) { return (std::fabs(n.z) > 0.999f) ? V3(0, 1, 0) : V3(0, 0, 1); }
static inline mat3 matLocalToWorld(vec3 n) {
    vec3 t = getTangent(n);
    vec3 b = normalize(cross(n, t));
    t = cross(b, n);
    return mat3{{t, b, n}};
}
static inline vec3 localToWorld(vec3 n, vec3 v) { return normalize(mul(matLocalToWorld(n), v)); }
// math.glsl:85-89
static inline vec3 sampleCosineWeightedHemisphere(vec3 n, vec2 u) {
    vec2 uv = toConcentricDisk(u);
    float z = std::sqrt(1.0f - dot2(uv, uv));
    return localToWorld(n, V3(uv.x, uv.y, z));
}
// math.glsl:117-120
static inline vec2 uvToBary(vec2 uv) {
    float r = std::sqrt(uv.y);
    return V2(1.0f - r, uv.x * r);
}

// GLSL inverse(mat3): adjugate / determinant.
static inline mat3 inverse3(const mat3& m) {
    // m[c][r] = column c, row r
    float a00 = m.c[0].x, a01 = m.c[1].x, a02 = m.c[2].x;  // row 0
    float a10 = m.c[0].y, a11 = m.c[1].y, a12 = m.c[2].y;  // row 1
    float a20 = m.c[0].z, a21 = m.c[1].z, a22 = m.c[2].z;  // row 2
    float c00 = a11 * a22 - a12 * a21;
    float c01 = a12 * a20 - a10 * a22;
    float c02 = a10 * a21 - a11 * a20;
    float det = a00 * c00 + a01 * c01 + a02 * c02;
    float inv = 1.0f / det;
    // inverse = adj / det, adj[r][c] = cofactor[c][r]
    float i00 = c00 * inv;
    float i01 = (a02 * a21 - a01 * a22) * inv;
    float i02 = (a01 * a12 - a02 * a11) * inv;
    float i10 = c01 * inv;
    float i11 = (a00 * a22 - a02 * a20) * inv;
    float i12 = (a02 * a10 - a00 * a12) * inv;
    float i20 = c02 * inv;
    float i21 = (a01 * a20 - a00 * a21) * inv;
    float i22 = (a00 * a11 - a01 * a10) * inv;
    mat3 r;
    r.c[0] = V3(i00, i10, i20);
    r.c[1] = V3(i01, i11, i21);
    r.c[2] = V3(i02, i12, i22);
    return r;
}

// ---- RNG: math.glsl:227-266 ---------------------------------------------------
static inline uint32_t hash2(uint32_t seed) {
    seed = (seed ^ 61u) ^ (seed >> 16u);
    seed *= 9u;
    seed = seed ^ (seed >> 4u);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15u);
    return seed;
}
static inline uint32_t makeSeed2(uint32_t rand, uint32_t index) { return hash2(rand) + hash2(index); }
static inline uint32_t makeSeedPix(uint32_t seed, uint32_t ix, uint32_t iy) {
    return makeSeed2((seed + ix) ^ (iy - 1u), iy * (ix - 2u));
}
static inline uint32_t urand(uint32_t& rng) { return rng = hash2(rng); }
// float(u) / 4294967295.0: the FP32 literal rounds to 2^32.
static inline float sample1f(uint32_t& rng) { return (float)urand(rng) / 4294967296.0f; }
static inline vec2 sample2f(uint32_t& rng) { float a = sample1f(rng); float b = sample1f(rng); return V2(a, b); }
static inline vec3 sample3f(uint32_t& rng) { vec2 a = sample2f(rng); float b = sample1f(rng); return V3(a.x, a.y, b); }
static inline vec4 sample4f(uint32_t& rng) { vec2 a = sample2f(rng); vec2 b = sample2f(rng); return {a.x, a.y, b.x, b.y}; }

// ---- half conversion (round to nearest even), for the RG16F motion vectors ---
static inline uint16_t f32_to_f16(float f) {
    uint32_t x; std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t absx = x & 0x7fffffffu;
    if (absx >= 0x7f800000u) {  // inf or nan
        return (uint16_t)(sign | 0x7c00u | (absx > 0x7f800000u ? 0x200u : 0u));
    }
    if (absx >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // overflow -> inf
    if (absx < 0x38800000u) {  // subnormal half or zero
        // value = absx as float; half subnormal unit = 2^-24
        float a; std::memcpy(&a, &absx, 4);
        float scaled = a * 16777216.0f;  // exact (power of two)
        float r = std::nearbyint(scaled);  // round half even
        return (uint16_t)(sign | (uint32_t)r);
    }
    uint32_t mant = absx & 0x7fffffu;
    int32_t exp = (int32_t)(absx >> 23) - 127 + 15;
    uint32_t h = ((uint32_t)exp << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1;
    return (uint16_t)(sign | h);
}
static inline float f16_to_f32(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1f, m = h & 0x3ff;
    float v;
    if (e == 0) v = (float)m * (1.0f / 16777216.0f);
    else if (e == 31) { uint32_t bits = 0x7f800000u | (m << 13); std::memcpy(&v, &bits, 4); }
    else { uint32_t bits = ((e - 15 + 127) << 23) | (m << 13); std::memcpy(&v, &bits, 4); }
    uint32_t b; std::memcpy(&b, &v, 4); b |= sign; std::memcpy(&v, &b, 4);
    return v;
}

}  // namespace orc
