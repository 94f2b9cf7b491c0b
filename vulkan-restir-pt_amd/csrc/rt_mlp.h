// Device side of the learned target function: a multiresolution hash-grid
// encoding (16 levels x 2 features, FP16 table) feeding a 32-64-64-3 MLP with
// FP16 weights evaluated on the matrix cores (v_mfma_f32_32x32x8_f16, FP32
// accumulation). The network maps a world position to RGB radiance.
//
// Every layer is evaluated in transposed form, H'^T = W^T H^T, with 32
// samples per MFMA column block. In the 32x32x8 operand layout a lane holds
// column j = lane % 32 and, for row block g of its 16 accumulators, rows
// 8g + 4(lane / 32) + (0..3); the B operand of a k-step s holds rows
// 8s + 4(lane / 32) + (0..3) of column j. The accumulators of one layer are
// therefore, four at a time, exactly the B operand of the next layer: no data
// moves between layers. Weight gradients contract over samples and use an LDS
// transpose of the activations.
#pragma once
#include "rt_dmath.h"

namespace rtm {

constexpr int kLevels = 16;
constexpr int kFeatures = 2;
constexpr int kEnc = kLevels * kFeatures;  // 32
constexpr int kHidden = 64;
constexpr int kOut = 3;
// FP16 weight block, input-major: W1 [32][64], W2 [64][64], W3 [64][outputs]
constexpr int kW1 = 0, kW2 = kEnc * kHidden, kW3 = kW2 + kHidden * kHidden;
constexpr int weightsFor(int outputs) { return kW3 + kHidden * outputs; }
constexpr int kWeights = weightsFor(kOut);  // 6336
// The learned proposal's network: the same encoding and hidden layers with 32
// outputs, four per lobe of an 8-lobe von Mises-Fisher mixture (weight logit,
// log concentration, polar and azimuth parameters), rows 4k..4k+3 for lobe k.
constexpr int kGuideOut = 32;
constexpr int kGuideLobes = 8;
constexpr int kGuideWeights = weightsFor(kGuideOut);  // 8192
// Raw output -> lobe parameter, per lobe k: (a + 0, b + ln 4, c + cBias_k, d + dBias_k) with the
// offsets placing the untrained network's lobes (outputs near 0: the network has no biases) at the
// eight octant directions (+-1, +-1, +-1) / sqrt 3 with concentration ~4 (bit 0/1/2 of k: negative
// x/y/z). pi = softmax(a'), log kappa = 0.5 + 4.5 tanh((b' - 0.5) / 4.5) in (-4, 5),
// theta = pi sigmoid(c'), phi = 2 pi sigmoid(d'),  mu = (sin theta cos phi, sin theta sin phi, cos theta).
constexpr float kGuideKappaBias = 1.3862943611198906f;          // ln 4
constexpr float kGuideThetaBias = 0.8278732f;                   // logit(acos(1/sqrt 3) / pi), negated for +z
constexpr float kGuidePhiBias[4] = {-1.9459101f, -0.5108256f,   // logit(phi / 2 pi), phi = pi/4, 3pi/4,
                                    1.9459101f, 0.5108256f};     // 7pi/4, 5pi/4 (index: x sign + 2 y sign)
RT_DEV float guideBias(int k, int e) {
    if (e == 1) return kGuideKappaBias;
    if (e == 2) return (k & 4) ? kGuideThetaBias : -kGuideThetaBias;
    if (e == 3) return kGuidePhiBias[k & 3];
    return 0.0f;
}
// log kappa of the biased raw output b' and its derivative d log kappa / d b'
constexpr float kGuideLogKappaMid = 0.5f, kGuideLogKappaHalf = 4.5f;  // log kappa in (-4, 5)
RT_DEV float guideLogKappa(float b, float& dlk) {
    const float t = tanhf((b - kGuideLogKappaMid) * (1.0f / kGuideLogKappaHalf));
    dlk = 1.0f - t * t;
    return kGuideLogKappaMid + kGuideLogKappaHalf * t;
}
constexpr int kTileSamples = 32;
constexpr uint32_t kBatchCap = 1u << 16;  // training samples gathered per frame
constexpr float kLossEps = 0.01f;    // relative L2: (y^ - y)^2 / (sg(y^)^2 + eps)
constexpr float kGradClip = 64.0f;   // per-sample output gradient bound (FP16 range)
constexpr float kLogClamp = 20.0f;   // exp output activation input bound

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

struct MlpView {
    const uint32_t* grid;    // [levels][T] packed half2 features
    const _Float16* w;       // kWeights FP16 weights
    uint32_t tableMask;      // T - 1
    uint32_t denseMask;      // bit l: level l indexes densely
    uint32_t res[kLevels];   // grid resolution N_l (cells per unit cube edge)
    float lo[3], invExt[3];  // world -> unit cube
};

// C += A B on the matrix cores, 32 x 32 outputs. The operands of the kernels
// are built as k-steps of 8 (a lane holds 4 halves of its row / column: lanes
// 0-31 the step's k 0-3, lanes 32-63 k 4-7). mfma2 takes two such steps at
// once on gfx950's K = 16 instruction: each lane's two 4-half pieces are
// concatenated for A and for B alike, so every product of the two steps is
// summed exactly once whatever k the hardware assigns to a slot. mfma is the
// single-step form for an odd step count.
RT_DEV f16v mfma(h4 a, h4 b, f16v c) { return __builtin_amdgcn_mfma_f32_32x32x8f16(a, b, c, 0, 0, 0); }
RT_DEV h8 cat8(h4 a, h4 b) { return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7); }
RT_DEV f16v mfma2(h4 a0, h4 b0, h4 a1, h4 b1, f16v c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(cat8(a0, a1), cat8(b0, b1), c, 0, 0, 0);
}
RT_DEV f16v zero16() {
    f16v z;
#pragma unroll
    for (int i = 0; i < 16; i++) z[i] = 0.0f;
    return z;
}

// Unit-cube coordinate of a world position, clamped to [0, 1].
RT_DEV void normalizePos(const MlpView& M, float px, float py, float pz, float& u, float& v, float& w) {
    u = fminf(fmaxf((px - M.lo[0]) * M.invExt[0], 0.0f), 1.0f);
    v = fminf(fmaxf((py - M.lo[1]) * M.invExt[1], 0.0f), 1.0f);
    w = fminf(fmaxf((pz - M.lo[2]) * M.invExt[2], 0.0f), 1.0f);
}

// Table slot of grid corner (x, y, z) at level l: dense where (N+1)^3 fits
// the table, else the spatial hash of Teschner et al. (primes 1, 2654435761,
// 805459861) masked to the table size.
RT_DEV uint32_t cornerIndex(const MlpView& M, int l, uint32_t x, uint32_t y, uint32_t z) {
    if ((M.denseMask >> l) & 1u) {
        uint32_t r = M.res[l] + 1u;
        return x + r * (y + r * z);
    }
    return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & M.tableMask;
}

// Trilinear corner c (bit 0 -> x, 1 -> y, 2 -> z) of level l at unit-cube
// point (u, v, w): slot index and weight.
struct Corner { uint32_t slot; float wt; };
// The cell is clamped to [0, N-1] so that u = 1 interpolates to the last
// corner instead of addressing corner N+1.
RT_DEV void levelCell(const MlpView& M, int l, float u, float v, float w, uint32_t ci[3], float fr[3]) {
    float N = (float)M.res[l];
    float x = u * N, y = v * N, z = w * N;
    float fx = fminf(floorf(x), N - 1.0f), fy = fminf(floorf(y), N - 1.0f), fz = fminf(floorf(z), N - 1.0f);
    ci[0] = (uint32_t)fx; ci[1] = (uint32_t)fy; ci[2] = (uint32_t)fz;
    fr[0] = x - fx; fr[1] = y - fy; fr[2] = z - fz;
}
RT_DEV Corner corner(const MlpView& M, int l, const uint32_t ci[3], const float fr[3], int c) {
    Corner k;
    k.slot = (uint32_t)l * (M.tableMask + 1u) +
             cornerIndex(M, l, ci[0] + (c & 1), ci[1] + ((c >> 1) & 1), ci[2] + ((c >> 2) & 1));
    float wx = (c & 1) ? fr[0] : 1.0f - fr[0];
    float wy = (c & 2) ? fr[1] : 1.0f - fr[1];
    float wz = (c & 4) ? fr[2] : 1.0f - fr[2];
    k.wt = (wx * wy) * wz;
    return k;
}

RT_DEV void encodeLevel(const MlpView& M, int l, float u, float v, float w, float& f0, float& f1) {
    uint32_t ci[3]; float fr[3];
    levelCell(M, l, u, v, w, ci, fr);
    f0 = 0.0f; f1 = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        Corner k = corner(M, l, ci, fr, c);
        uint32_t p = M.grid[k.slot];
        _Float16 a = __builtin_bit_cast(_Float16, (unsigned short)(p & 0xffffu));
        _Float16 b = __builtin_bit_cast(_Float16, (unsigned short)(p >> 16));
        f0 = f0 + k.wt * (float)a;
        f1 = f1 + k.wt * (float)b;
    }
}

// B operands of layer 1 for this lane: k-step s holds features 8s + 4h + (0..3)
// = levels 4s + 2h and 4s + 2h + 1 of sample (u, v, w), h = lane / 32.
RT_DEV void encodeLane(const MlpView& M, uint32_t h, float u, float v, float w, h4 enc[4]) {
#pragma unroll
    for (int s = 0; s < 4; s++) {
        float a0, a1, b0, b1;
        encodeLevel(M, 4 * s + 2 * (int)h, u, v, w, a0, a1);
        encodeLevel(M, 4 * s + 2 * (int)h + 1, u, v, w, b0, b1);
        enc[s][0] = (_Float16)a0; enc[s][1] = (_Float16)a1;
        enc[s][2] = (_Float16)b0; enc[s][3] = (_Float16)b1;
    }
}

// Weight fragments, from an LDS (or global) copy of the FP16 weight block.
// Forward A operand of W^T (rows = outputs): A[o][i] = W[i][o].
RT_DEV h4 fwdA(const _Float16* W, int ld, int o, int iBase, int nOut) {
    h4 a;
#pragma unroll
    for (int e = 0; e < 4; e++) a[e] = (o < nOut) ? W[(iBase + e) * ld + o] : (_Float16)0.0f;
    return a;
}
// Backward A operand of W (rows = inputs): A[i][o] = W[i][o].
RT_DEV h4 bwdA(const _Float16* W, int ld, int i, int oBase, int nOut) {
    h4 a;
#pragma unroll
    for (int e = 0; e < 4; e++) a[e] = (oBase + e < nOut) ? W[i * ld + oBase + e] : (_Float16)0.0f;
    return a;
}

RT_DEV h4 reluHalf(const f16v& acc, int g) {
    h4 b;
#pragma unroll
    for (int e = 0; e < 4; e++) b[e] = (_Float16)fmaxf(acc[4 * g + e], 0.0f);
    return b;
}

// Forward pass of one 32-sample tile. W points at the FP16 weight block.
// Outputs: layer accumulators (pre-activation z1, z2 per 32-row block) and
// z3, whose output row 8g + 4h + e sits in register 4g + e of the lanes with
// lane / 32 = h (rows 0-2 of the radiance network: registers 0-2 of lanes 0-31).
struct Fwd { f16v z1[2], z2[2], z3; };
template <int kO>
RT_DEV void forwardTileT(const _Float16* W, uint32_t lane, const h4 enc[4], Fwd& F) {
    const int j = (int)(lane & 31u), h = (int)(lane >> 5);
#pragma unroll
    for (int t = 0; t < 2; t++) {
        f16v acc = zero16();
#pragma unroll
        for (int s = 0; s < 4; s += 2)
            acc = mfma2(fwdA(W + kW1, kHidden, 32 * t + j, 8 * s + 4 * h, kHidden), enc[s],
                        fwdA(W + kW1, kHidden, 32 * t + j, 8 * (s + 1) + 4 * h, kHidden), enc[s + 1], acc);
        F.z1[t] = acc;
    }
#pragma unroll
    for (int t = 0; t < 2; t++) {
        f16v acc = zero16();
#pragma unroll
        for (int s = 0; s < 8; s += 2)
            acc = mfma2(fwdA(W + kW2, kHidden, 32 * t + j, 8 * s + 4 * h, kHidden), reluHalf(F.z1[s >> 2], s & 3),
                        fwdA(W + kW2, kHidden, 32 * t + j, 8 * (s + 1) + 4 * h, kHidden),
                        reluHalf(F.z1[(s + 1) >> 2], (s + 1) & 3), acc);
        F.z2[t] = acc;
    }
    f16v acc = zero16();
#pragma unroll
    for (int s = 0; s < 8; s += 2)
        acc = mfma2(fwdA(W + kW3, kO, j, 8 * s + 4 * h, kO), reluHalf(F.z2[s >> 2], s & 3),
                    fwdA(W + kW3, kO, j, 8 * (s + 1) + 4 * h, kO), reluHalf(F.z2[(s + 1) >> 2], (s + 1) & 3), acc);
    F.z3 = acc;
}
RT_DEV void forwardTile(const _Float16* W, uint32_t lane, const h4 enc[4], Fwd& F) { forwardTileT<kOut>(W, lane, enc, F); }

RT_DEV float outAct(float z) { return expf(fminf(fmaxf(z, -kLogClamp), kLogClamp)); }

}  // namespace rtm
