// Learned target function: parameter set-up, inference, fused forward +
// backward over a training batch, deterministic weight-gradient reduction and
// Adam, behind the rt_mlp_* entry points of include/restirpt.h.
//
// Training data flow per step: one wave per 32-sample tile evaluates the
// network (rt_mlp.h), forms the relative-L2 output gradient, back-propagates
// through the three layers on the matrix cores, scatters the encoding
// gradient into the grid-gradient table with 64-bit integer atomics on 32.32
// fixed-point values (integer sums do not depend on the order of the atomics),
// and accumulates the weight gradients of all its tiles in registers. Per-wave
// partial sums are reduced in a fixed order. Training is therefore
// deterministic: the same batch gives bit-identical parameters.
#include "rt_internal.h"
#include <cmath>
#include <cstring>
#include <vector>

using namespace rtm;
using rti::launch;

namespace {

constexpr int kStats = 2;                 // loss sum, valid sample count (after the weight gradients)
constexpr int partialStride(int outputs) { return weightsFor(outputs) + kStats; }
constexpr int kPartialStride = partialStride(kOut);
constexpr int kMaxPartialStride = partialStride(kGuideOut);
constexpr uint32_t kMaxTrainWaves = 512;
constexpr int kInferWaves = 4;
constexpr float kFix = 4294967296.0f;        // 2^32: grid-gradient fixed-point scale
constexpr double kInvFix = 1.0 / 4294967296.0;

// ---- initialisation -------------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ float uniformPM(uint64_t seed, uint64_t i) {  // [-1, 1)
    return (float)(splitmix(seed * 0x100000001B3ull + i) >> 40) * (2.0f / 16777216.0f) - 1.0f;
}

__global__ void k_mlp_init_grid(float* g32, _Float16* g16, uint64_t n, uint32_t seed) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = uniformPM(seed, i) * 1e-4f;
    g32[i] = v;
    g16[i] = (_Float16)v;
}
template <int kO>
__global__ void k_mlp_init_weights(float* w32, _Float16* w16, uint32_t seed) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= weightsFor(kO)) return;
    float lim;  // Glorot uniform per layer
    if (i < kW2) lim = sqrtf(6.0f / (float)(kEnc + kHidden));
    else if (i < kW3) lim = sqrtf(6.0f / (float)(kHidden + kHidden));
    else lim = sqrtf(6.0f / (float)(kHidden + kO));
    float v = uniformPM((uint64_t)seed ^ 0xA5A5A5A5ull, (uint64_t)i) * lim;
    w32[i] = v;
    w16[i] = (_Float16)v;
}
__global__ void k_fill_f32(float* p, uint64_t n, float v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_i64(long long* p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
__global__ void k_to_f16(const float* src, _Float16* dst, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (_Float16)src[i];
}

template <int kO>
__device__ __forceinline__ void loadWeightsLDS(_Float16* sW, const _Float16* w) {
    for (int i = threadIdx.x; i < weightsFor(kO); i += blockDim.x) sW[i] = w[i];
    __syncthreads();
}

// ---- inference ------------------------------------------------------------------------------
// Radiance network: exp-activated RGB per sample. Proposal network: its 32 raw outputs per sample.
template <int kO>
__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(MlpView M, const float* __restrict__ pos, uint32_t n,
                                                                float* __restrict__ out) {
    __shared__ _Float16 sW[weightsFor(kO)];
    loadWeightsLDS<kO>(sW, M.w);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t s = (blockIdx.x * kInferWaves + wave) * kTileSamples + (lane & 31u);
    float u = 0.0f, v = 0.0f, w = 0.0f;
    if (s < n) normalizePos(M, pos[3 * (size_t)s], pos[3 * (size_t)s + 1], pos[3 * (size_t)s + 2], u, v, w);
    h4 enc[4];
    encodeLane(M, lane >> 5, u, v, w, enc);
    Fwd F;
    forwardTileT<kO>(sW, lane, enc, F);
    if (s >= n) return;
    if constexpr (kO == kOut) {
        if (lane < 32u) {
            out[3 * (size_t)s] = outAct(F.z3[0]);
            out[3 * (size_t)s + 1] = outAct(F.z3[1]);
            out[3 * (size_t)s + 2] = outAct(F.z3[2]);
        }
    } else {
        const int h = (int)(lane >> 5);
        for (int g = 0; g < 4; g++)
            for (int e = 0; e < 4; e++) out[(size_t)kO * s + 8 * g + 4 * h + e] = F.z3[4 * g + e];
    }
}

// ---- fused forward + backward -----------------------------------------------------------------
template <int kO>
struct TrainLDS {
    _Float16 w[weightsFor(kO)];
    _Float16 enc[kTileSamples][kEnc];     // encoding, [sample][feature]
    _Float16 h1[kTileSamples][kHidden];   // relu(z1)
    _Float16 h2[kTileSamples][kHidden];   // relu(z2)
    _Float16 d1[kTileSamples][kHidden];   // dL/dz1
    _Float16 d2[kTileSamples][kHidden];   // dL/dz2
    _Float16 d3[kTileSamples][kO < 4 ? 4 : kO];  // dL/dz3
};

// Row block g of an accumulator tile t of this lane: neuron 32t + 8g + 4h + e.
__device__ __forceinline__ int neuron(int t, int g, int h, int e) { return 32 * t + 8 * g + 4 * h + e; }

__device__ __forceinline__ float clipGrad(float g) { return fminf(fmaxf(g, -kGradClip), kGradClip); }

// Relative L2 output gradient of the radiance network: rows 0-2 of lanes 0-31 (k-step 0).
__device__ __forceinline__ void radianceLoss(const Fwd& F, int h, const float y[3], bool valid, float& lossSum,
                                             float& count, h4 d3b[1]) {
    for (int e = 0; e < 4; e++) d3b[0][e] = (_Float16)0.0f;
    if (h != 0) return;
    for (int c = 0; c < 3; c++) {
        float z = F.z3[c];
        float yh = outAct(z);
        float den = yh * yh + kLossEps;
        float diff = yh - y[c];
        float g = 0.0f;
        if (valid) {
            lossSum += diff * diff / den;
            // d/dz of (yh - y)^2 / sg(yh^2 + eps) with yh = exp(z)
            g = (fabsf(z) < kLogClamp) ? 2.0f * diff / den * yh : 0.0f;
            g = clipGrad(g);
        }
        d3b[0][c] = (_Float16)g;
    }
    if (valid) count += 1.0f;
}

// Weighted negative log-likelihood of the 8-lobe vMF mixture at direction wd, weight wv
// (the sample's weight over the batch mean). This lane holds lobes k = 2g + h in registers
// 4g..4g+3, mapped to (pi, kappa, mu) as guideBias / guideLogKappa state;
// v(w) = kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (mu.w - 1)); loss = -wv log sum pi v.
// Gradients with the responsibilities r = pi v / q: da = wv (pi - r);
// db = -wv r (1 - 2 kappa e^{-2 kappa} / (1 - e^{-2 kappa}) + kappa (mu.w - 1)) dlogkappa/db;
// dmu = -wv r kappa w, through dmu/dtheta and dmu/dphi to dc and dd.
__device__ __forceinline__ void vmfLoss(const Fwd& F, int h, const float wd[3], float wv, bool valid, float& lossSum,
                                        float& count, h4 d3b[4]) {
    float amax = fmaxf(fmaxf(F.z3[0], F.z3[4]), fmaxf(F.z3[8], F.z3[12]));
    amax = fmaxf(amax, __shfl_xor(amax, 32));
    float ea[4], esum = 0.0f;
    for (int g = 0; g < 4; g++) { ea[g] = expf(F.z3[4 * g] - amax); esum += ea[g]; }
    esum += __shfl_xor(esum, 32);
    float pi[4], kap[4], dlk[4], e2[4], dt[4], pv[4], sc[4], sd[4], st[4], ct[4], sp[4], cp[4];
    float q = 0.0f;
    for (int g = 0; g < 4; g++) {
        const int k = 2 * g + h;
        pi[g] = ea[g] / esum;
        kap[g] = expf(guideLogKappa(F.z3[4 * g + 1] + guideBias(k, 1), dlk[g]));
        sc[g] = 1.0f / (1.0f + expf(-(F.z3[4 * g + 2] + guideBias(k, 2))));
        sd[g] = 1.0f / (1.0f + expf(-(F.z3[4 * g + 3] + guideBias(k, 3))));
        const float th = RT_PI * sc[g], ph = 2.0f * RT_PI * sd[g];
        st[g] = sinf(th); ct[g] = cosf(th); sp[g] = sinf(ph); cp[g] = cosf(ph);
        dt[g] = st[g] * cp[g] * wd[0] + st[g] * sp[g] * wd[1] + ct[g] * wd[2];
        e2[g] = expf(-2.0f * kap[g]);
        const float v = kap[g] / (2.0f * RT_PI * (1.0f - e2[g])) * expf(kap[g] * (dt[g] - 1.0f));
        pv[g] = pi[g] * v;
        q += pv[g];
    }
    q += __shfl_xor(q, 32);
    q = fmaxf(q, 1e-20f);
    if (valid && h == 0) { lossSum += -wv * logf(q); count += 1.0f; }
    for (int g = 0; g < 4; g++) {
        float da = 0.0f, db = 0.0f, dc = 0.0f, dd = 0.0f;
        if (valid) {
            const float r = pv[g] / q;
            da = wv * (pi[g] - r);
            db = -wv * r * (1.0f - 2.0f * kap[g] * e2[g] / (1.0f - e2[g]) + kap[g] * (dt[g] - 1.0f)) * dlk[g];
            const float m = -wv * r * kap[g];  // dL/dmu = m * wd
            const float dth = m * (ct[g] * cp[g] * wd[0] + ct[g] * sp[g] * wd[1] - st[g] * wd[2]);
            const float dph = m * (-st[g] * sp[g] * wd[0] + st[g] * cp[g] * wd[1]);
            dc = dth * RT_PI * sc[g] * (1.0f - sc[g]);
            dd = dph * 2.0f * RT_PI * sd[g] * (1.0f - sd[g]);
        }
        d3b[g][0] = (_Float16)clipGrad(da);
        d3b[g][1] = (_Float16)clipGrad(db);
        d3b[g][2] = (_Float16)clipGrad(dc);
        d3b[g][3] = (_Float16)clipGrad(dd);
    }
}

// Targets: 3 floats per sample (radiance network: RGB, negative = masked) or 4 (proposal
// network: unit direction and weight; weight <= 0 or non-finite = masked). wmean: the mean
// weight of the valid samples (proposal network only).
template <int kO>
__global__ __launch_bounds__(64) void k_mlp_train(MlpView M, const float* __restrict__ pos,
                                                  const float* __restrict__ target, uint32_t n,
                                                  const float* __restrict__ wmean,
                                                  long long* __restrict__ gridGrad, float* __restrict__ partial) {
    constexpr int kSteps = (kO + 7) / 8;  // k-steps over the outputs
    constexpr int kTS = kO == kOut ? 3 : 4;  // target stride
    __shared__ TrainLDS<kO> L;
    loadWeightsLDS<kO>(L.w, M.w);
    const uint32_t lane = threadIdx.x;
    const int j = (int)(lane & 31u), h = (int)(lane >> 5);
    const uint32_t tiles = (n + kTileSamples - 1) / kTileSamples;
    f16v gW1[2], gW2[2][2], gW3[2];
    for (int t = 0; t < 2; t++) { gW1[t] = zero16(); gW3[t] = zero16(); gW2[t][0] = zero16(); gW2[t][1] = zero16(); }
    float lossSum = 0.0f, count = 0.0f;
    const float invMean = (kO == kOut || !wmean || !(*wmean > 0.0f)) ? 0.0f : 1.0f / *wmean;

    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t s = tile * kTileSamples + (uint32_t)j;
        float u = 0.0f, v = 0.0f, w = 0.0f;
        float y[4] = {-1.0f, 0.0f, 0.0f, -1.0f};
        if (s < n) {
            normalizePos(M, pos[3 * (size_t)s], pos[3 * (size_t)s + 1], pos[3 * (size_t)s + 2], u, v, w);
            for (int c = 0; c < kTS; c++) y[c] = target[kTS * (size_t)s + c];
        }
        bool valid;
        if constexpr (kO == kOut)
            valid = s < n && y[0] >= 0.0f && y[1] >= 0.0f && y[2] >= 0.0f && isfinite(y[0]) && isfinite(y[1]) &&
                    isfinite(y[2]);
        else
            valid = s < n && isfinite(y[3]) && y[3] > 0.0f && isfinite(y[0]) && isfinite(y[1]) && isfinite(y[2]) &&
                    invMean > 0.0f;
        h4 enc[4];
        encodeLane(M, (uint32_t)h, u, v, w, enc);
        Fwd F;
        forwardTileT<kO>(L.w, lane, enc, F);

        // output gradient: k-step g holds output rows 8g + 4h + (0..3) of this lane's sample
        h4 d3b[kSteps];
        if constexpr (kO == kOut) radianceLoss(F, h, y, valid, lossSum, count, d3b);
        else vmfLoss(F, h, y, y[3] * invMean, valid, lossSum, count, d3b);
        // dh2 = W3 dz3
        f16v dz2[2];
        for (int t = 0; t < 2; t++) {
            f16v acc = zero16();
            int g = 0;
            for (; g + 1 < kSteps; g += 2)
                acc = mfma2(bwdA(L.w + kW3, kO, 32 * t + j, 8 * g + 4 * h, kO), d3b[g],
                            bwdA(L.w + kW3, kO, 32 * t + j, 8 * (g + 1) + 4 * h, kO), d3b[g + 1], acc);
            if (g < kSteps) acc = mfma(bwdA(L.w + kW3, kO, 32 * t + j, 8 * g + 4 * h, kO), d3b[g], acc);
            for (int r = 0; r < 16; r++) acc[r] = F.z2[t][r] > 0.0f ? acc[r] : 0.0f;
            dz2[t] = acc;
        }
        // dh1 = W2 dz2
        f16v dz1[2];
        for (int t = 0; t < 2; t++) {
            f16v acc = zero16();
            for (int s8 = 0; s8 < 8; s8 += 2) {
                h4 b0, b1;
                for (int e = 0; e < 4; e++) {
                    b0[e] = (_Float16)dz2[s8 >> 2][4 * (s8 & 3) + e];
                    b1[e] = (_Float16)dz2[(s8 + 1) >> 2][4 * ((s8 + 1) & 3) + e];
                }
                acc = mfma2(bwdA(L.w + kW2, kHidden, 32 * t + j, 8 * s8 + 4 * h, kHidden), b0,
                            bwdA(L.w + kW2, kHidden, 32 * t + j, 8 * (s8 + 1) + 4 * h, kHidden), b1, acc);
            }
            for (int r = 0; r < 16; r++) acc[r] = F.z1[t][r] > 0.0f ? acc[r] : 0.0f;
            dz1[t] = acc;
        }
        // denc = W1 dz1: rows = the 32 encoding features, laid out like enc[]
        f16v denc = zero16();
        for (int s8 = 0; s8 < 8; s8 += 2) {
            h4 b0, b1;
            for (int e = 0; e < 4; e++) {
                b0[e] = (_Float16)dz1[s8 >> 2][4 * (s8 & 3) + e];
                b1[e] = (_Float16)dz1[(s8 + 1) >> 2][4 * ((s8 + 1) & 3) + e];
            }
            denc = mfma2(bwdA(L.w + kW1, kHidden, j, 8 * s8 + 4 * h, kHidden), b0,
                         bwdA(L.w + kW1, kHidden, j, 8 * (s8 + 1) + 4 * h, kHidden), b1, denc);
        }
        // scatter the encoding gradient: register group g holds levels 4g + 2h (+1)
        if (s < n && valid) {
            for (int g = 0; g < 4; g++) {
                for (int q = 0; q < 2; q++) {
                    const int lev = 4 * g + 2 * h + q;
                    const float d0 = denc[4 * g + 2 * q], d1 = denc[4 * g + 2 * q + 1];
                    uint32_t ci[3]; float fr[3];
                    levelCell(M, lev, u, v, w, ci, fr);
                    for (int c = 0; c < 8; c++) {
                        Corner k = corner(M, lev, ci, fr, c);
                        unsigned long long* gp = reinterpret_cast<unsigned long long*>(gridGrad + 2 * (size_t)k.slot);
                        atomicAdd(gp, (unsigned long long)__float2ll_rn(k.wt * d0 * kFix));
                        atomicAdd(gp + 1, (unsigned long long)__float2ll_rn(k.wt * d1 * kFix));
                    }
                }
            }
        }
        // stage activations and gradients as [sample][neuron] for the weight-gradient products
        for (int s4 = 0; s4 < 4; s4++)
            for (int e = 0; e < 4; e++) L.enc[j][8 * s4 + 4 * h + e] = enc[s4][e];
        for (int t = 0; t < 2; t++)
            for (int g = 0; g < 4; g++)
                for (int e = 0; e < 4; e++) {
                    const int nn = neuron(t, g, h, e);
                    L.h1[j][nn] = (_Float16)fmaxf(F.z1[t][4 * g + e], 0.0f);
                    L.h2[j][nn] = (_Float16)fmaxf(F.z2[t][4 * g + e], 0.0f);
                    L.d1[j][nn] = (_Float16)dz1[t][4 * g + e];
                    L.d2[j][nn] = (_Float16)dz2[t][4 * g + e];
                }
        if constexpr (kO == kOut) {
            if (h == 0)
                for (int e = 0; e < 4; e++) L.d3[j][e] = d3b[0][e];
        } else {
            for (int g = 0; g < kSteps; g++)
                for (int e = 0; e < 4; e++) L.d3[j][8 * g + 4 * h + e] = d3b[g][e];
        }
        __syncthreads();
        // dW^T[o][i] += sum over samples dz[o][j] * act[i][j]; k-step s4 covers samples 8 s4 + 4h + e,
        // two k-steps per instruction
        for (int s4 = 0; s4 < 4; s4 += 2) {
            h4 aD1[2][2], aD2[2][2], aD3[2], bEnc[2], bH1[2][2], bH2[2][2];
            for (int q = 0; q < 2; q++) {
                const int jb = 8 * (s4 + q) + 4 * h;
                for (int e = 0; e < 4; e++) {
                    for (int t = 0; t < 2; t++) {
                        aD1[q][t][e] = L.d1[jb + e][32 * t + j];
                        aD2[q][t][e] = L.d2[jb + e][32 * t + j];
                        bH1[q][t][e] = L.h1[jb + e][32 * t + j];
                        bH2[q][t][e] = L.h2[jb + e][32 * t + j];
                    }
                    aD3[q][e] = (j < kO) ? L.d3[jb + e][j] : (_Float16)0.0f;
                    bEnc[q][e] = L.enc[jb + e][j];
                }
            }
            for (int t = 0; t < 2; t++) {
                gW1[t] = mfma2(aD1[0][t], bEnc[0], aD1[1][t], bEnc[1], gW1[t]);
                for (int u2 = 0; u2 < 2; u2++)
                    gW2[t][u2] = mfma2(aD2[0][t], bH1[0][u2], aD2[1][t], bH1[1][u2], gW2[t][u2]);
                gW3[t] = mfma2(aD3[0], bH2[0][t], aD3[1], bH2[1][t], gW3[t]);
            }
        }
        __syncthreads();
    }

    // per-wave partials in the weight layout (input-major): C row = output neuron, column = input
    float* P = partial + (size_t)blockIdx.x * partialStride(kO);
    for (int t = 0; t < 2; t++)
        for (int g = 0; g < 4; g++)
            for (int e = 0; e < 4; e++) {
                const int o = neuron(t, g, h, e), r = 4 * g + e;
                P[kW1 + j * kHidden + o] = gW1[t][r];
                for (int u2 = 0; u2 < 2; u2++) P[kW2 + (32 * u2 + j) * kHidden + o] = gW2[t][u2][r];
            }
    for (int t = 0; t < 2; t++)
        for (int g = 0; g < 4; g++)
            for (int e = 0; e < 4; e++) {
                const int o = 8 * g + 4 * h + e;
                if (o < kO) P[kW3 + (32 * t + j) * kO + o] = gW3[t][4 * g + e];
            }
    // loss and count: sum over lanes 0-31
    for (int off = 16; off > 0; off >>= 1) {
        lossSum += __shfl_xor(lossSum, off);
        count += __shfl_xor(count, off);
    }
    if (lane == 0) { P[weightsFor(kO)] = lossSum; P[weightsFor(kO) + 1] = count; }
}

// Mean weight of the valid samples of a proposal batch (one workgroup, fixed-order sums).
constexpr uint32_t kMeanThreads = 1024;
__global__ __launch_bounds__(kMeanThreads) void k_guide_wmean(const float* __restrict__ target, uint32_t n,
                                                              float* __restrict__ out) {
    __shared__ float sw[kMeanThreads], sc[kMeanThreads];
    float w = 0.0f, c = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += kMeanThreads) {
        const float* y = target + 4 * (size_t)i;
        if (isfinite(y[3]) && y[3] > 0.0f && isfinite(y[0]) && isfinite(y[1]) && isfinite(y[2])) { w += y[3]; c += 1.0f; }
    }
    sw[threadIdx.x] = w; sc[threadIdx.x] = c;
    __syncthreads();
    for (int off = (int)kMeanThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) { sw[threadIdx.x] += sw[threadIdx.x + off]; sc[threadIdx.x] += sc[threadIdx.x + off]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sc[0] > 0.0f ? sw[0] / sc[0] : 0.0f;
}

// Weight gradients and stats: a fixed-order sum over the waves' partials, in
// two launches so that the reduction spreads over the whole GPU: each group of
// kReduceGroup consecutive waves is summed into the first row of the group,
// then the groups' rows are summed.
constexpr uint32_t kReduceGroup = 16;

// Sum of column k over `rows` rows rowStride apart, in a fixed order (eight
// interleaved accumulators for memory-level parallelism, combined pairwise).
__device__ __forceinline__ float columnSum(const float* __restrict__ partial, uint32_t rows, size_t rowStride, int k) {
    float a[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t w = 0;
    for (; w + 8u <= rows; w += 8u)
#pragma unroll
        for (int i = 0; i < 8; i++) a[i] += partial[(size_t)(w + i) * rowStride + k];
    for (; w < rows; w++) a[0] += partial[(size_t)w * rowStride + k];
    return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

// Group g (blockIdx.y) of the waves: rows [16 g, 16 g + 16) summed into row 16 g.
__global__ void k_mlp_reduce_groups(float* __restrict__ partial, uint32_t waves, int stride) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= stride) return;
    const uint32_t w0 = blockIdx.y * kReduceGroup;
    const uint32_t rows = min(kReduceGroup, waves - w0);
    float* p = partial + (size_t)w0 * stride;
    p[k] = columnSum(p, rows, (size_t)stride, k);
}

// The groups' sums (rows 16 g). stride: partial stride of the network; norm:
// outputs per sample the mean loss averages over (3 for the radiance network,
// 1 for the proposal network).
__global__ void k_mlp_reduce(const float* __restrict__ partial, uint32_t groups, int stride, float norm,
                             float* __restrict__ out, float* __restrict__ lossOut) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= stride) return;
    const size_t rowStride = (size_t)kReduceGroup * stride;
    const float acc = columnSum(partial, groups, rowStride, k);
    out[k] = acc;
    if (k == stride - kStats && lossOut) {
        const float cnt = columnSum(partial, groups, rowStride, stride - kStats + 1);
        *lossOut = cnt > 0.0f ? acc / (norm * cnt) : 0.0f;
    }
}

__device__ __forceinline__ float gradValue(float g) { return g; }
__device__ __forceinline__ float gradValue(long long g) { return (float)((double)g * kInvFix); }

// Adam on one parameter array; the gradient of the mean loss is grad / (norm count).
// Grid entries no sample touched (gradient exactly zero) are left unchanged.
template <class G>
__global__ void k_mlp_adam(float* __restrict__ p32, _Float16* __restrict__ p16, const G* __restrict__ grad,
                           float* __restrict__ m, float* __restrict__ v, uint64_t n, const float* __restrict__ stats,
                           float norm, float lr, float b1, float b2, float eps, float c1, float c2, int skipZero) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (skipZero && grad[i] == G(0)) return;
    float gr = gradValue(grad[i]);
    float cnt = stats[1];
    float g = cnt > 0.0f ? gr / (norm * cnt) : 0.0f;
    float mi = b1 * m[i] + (1.0f - b1) * g;
    float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi; v[i] = vi;
    float p = p32[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p32[i] = p;
    p16[i] = (_Float16)p;
}

inline unsigned blocks(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

rt_status invalid(const char* m) { rti::setError(m); return RT_ERR_INVALID_ARGUMENT; }

void freeMlp(rt_mlp* m) {
    (void)hipFree(m->gridF32); (void)hipFree(m->gridF16); (void)hipFree(m->gridGrad); (void)hipFree(m->gridM);
    (void)hipFree(m->gridV); (void)hipFree(m->wF32); (void)hipFree(m->wF16); (void)hipFree(m->wGrad);
    (void)hipFree(m->wM); (void)hipFree(m->wV); (void)hipFree(m->partial); (void)hipFree(m->batchPos);
    (void)hipFree(m->batchTarget); (void)hipFree(m->wmean);
    delete m;
}

hipError_t zeroAdam(rt_mlp* m, hipStream_t st) {
    const uint64_t nw = (uint64_t)m->weights;
    k_fill_f32<<<blocks(m->gridParams, 256), 256, 0, st>>>(m->gridM, m->gridParams, 0.0f);
    k_fill_f32<<<blocks(m->gridParams, 256), 256, 0, st>>>(m->gridV, m->gridParams, 0.0f);
    k_fill_f32<<<blocks(nw, 256), 256, 0, st>>>(m->wM, nw, 0.0f);
    k_fill_f32<<<blocks(nw, 256), 256, 0, st>>>(m->wV, nw, 0.0f);
    m->step = 0;
    return hipGetLastError();
}

}  // namespace

namespace rti {
// Gradients of a batch into m->gridGrad / m->wGrad (used by the ABI and the frame trainer).
hipError_t mlpGradients(rt_mlp* m, const float* pos, const float* target, uint32_t n, float* loss, hipStream_t st) {
    const bool guide = m->outputs == (uint32_t)kGuideOut;
    const int stride = guide ? partialStride(kGuideOut) : kPartialStride;
    (void)launch(k_fill_i64, dim3(blocks(m->gridParams, 256)), dim3(256), st, m->gridGrad, m->gridParams);
    const uint32_t tiles = (n + kTileSamples - 1) / kTileSamples;
    const uint32_t waves = std::max(1u, std::min(tiles, m->partialWaves));
    if (n && guide) {
        (void)launch(k_guide_wmean, dim3(1), dim3(kMeanThreads), st, target, n, m->wmean);
        (void)launch(k_mlp_train<kGuideOut>, dim3(waves), dim3(64), st, m->view, pos, target, n, (const float*)m->wmean,
                     m->gridGrad, m->partial);
    } else if (n) {
        (void)launch(k_mlp_train<kOut>, dim3(waves), dim3(64), st, m->view, pos, target, n, (const float*)nullptr,
                     m->gridGrad, m->partial);
    } else {
        (void)launch(k_fill_f32, dim3(blocks(stride, 256)), dim3(256), st, m->partial, (uint64_t)stride, 0.0f);
    }
    const uint32_t rows = n ? waves : 1u;
    const uint32_t groups = (rows + kReduceGroup - 1u) / kReduceGroup;
    (void)launch(k_mlp_reduce_groups, dim3(blocks(stride, 256), groups), dim3(256), st, m->partial, rows, stride);
    (void)launch(k_mlp_reduce, dim3(blocks(stride, 256)), dim3(256), st, (const float*)m->partial, groups, stride,
                 guide ? 1.0f : 3.0f, m->wGrad, loss);
    return hipGetLastError();
}
hipError_t mlpAdam(rt_mlp* m, hipStream_t st) {
    m->step++;
    const rt_mlp_config& c = m->cfg;
    const float c1 = 1.0f - (float)std::pow((double)c.beta1, (double)m->step);
    const float c2 = 1.0f - (float)std::pow((double)c.beta2, (double)m->step);
    const float* stats = m->wGrad + m->weights;
    const float norm = m->outputs == (uint32_t)kGuideOut ? 1.0f : 3.0f;
    (void)launch(k_mlp_adam<long long>, dim3(blocks(m->gridParams, 256)), dim3(256), st, m->gridF32, m->gridF16,
                 (const long long*)m->gridGrad, m->gridM, m->gridV, m->gridParams, stats, norm, c.learningRate, c.beta1,
                 c.beta2, c.epsilon, c1, c2, 1);
    (void)launch(k_mlp_adam<float>, dim3(blocks(m->weights, 256)), dim3(256), st, m->wF32, m->wF16,
                 (const float*)m->wGrad, m->wM, m->wV, (uint64_t)m->weights, stats, norm, c.learningRate, c.beta1,
                 c.beta2, c.epsilon, c1, c2, 0);
    return hipGetLastError();
}
}  // namespace rti

namespace {
rt_status createNet(const rt_mlp_config* c, uint32_t outputs, hipStream_t stream, rt_mlp** out) {
    if (!out) return invalid("out is null");
    *out = nullptr;
    if (!c) return invalid("config is null");
    if (c->levels != (uint32_t)kLevels || c->featuresPerLevel != (uint32_t)kFeatures || c->hiddenWidth != (uint32_t)kHidden ||
        c->hiddenLayers != 2u) {
        rti::setError("only 16 levels x 2 features and 2 hidden layers of 64 are supported");
        return RT_ERR_UNSUPPORTED;
    }
    if (c->log2TableSize < 10 || c->log2TableSize > 24) return invalid("log2TableSize must be in [10, 24]");
    if (c->baseResolution < 1 || c->maxResolution < c->baseResolution || c->maxResolution > 65535)
        return invalid("resolutions must satisfy 1 <= base <= max <= 65535");
    for (int a = 0; a < 3; a++)
        if (!(c->boundsHi[a] > c->boundsLo[a])) return invalid("bounds must have positive extent");
    if (!(c->learningRate > 0.0f) || !(c->beta1 >= 0.0f && c->beta1 < 1.0f) || !(c->beta2 >= 0.0f && c->beta2 < 1.0f) ||
        !(c->epsilon >= 0.0f))
        return invalid("invalid optimiser settings");
    rt_mlp* m = new rt_mlp();
    m->cfg = *c;
    m->outputs = outputs;
    const bool guide = outputs == (uint32_t)kGuideOut;
    m->weights = (uint32_t)weightsFor((int)outputs);
    m->targetStride = guide ? 4u : 3u;
    const int stride = partialStride((int)outputs);
    const uint64_t T = 1ull << c->log2TableSize;
    m->gridParams = (uint64_t)kLevels * T * kFeatures;
    MlpView& V = m->view;
    V.tableMask = (uint32_t)(T - 1);
    V.denseMask = 0;
    const double b = std::exp((std::log((double)c->maxResolution) - std::log((double)c->baseResolution)) / (kLevels - 1));
    for (int l = 0; l < kLevels; l++) {
        uint32_t N = (uint32_t)std::floor((double)c->baseResolution * std::pow(b, (double)l) + 1e-9);
        V.res[l] = N;
        if ((uint64_t)(N + 1) * (N + 1) * (N + 1) <= T) V.denseMask |= 1u << l;
    }
    for (int a = 0; a < 3; a++) { V.lo[a] = c->boundsLo[a]; V.invExt[a] = 1.0f / (c->boundsHi[a] - c->boundsLo[a]); }
    m->partialWaves = kMaxTrainWaves;
    const size_t G = m->gridParams, NW = m->weights;
    bool ok = rti::hipCheck(hipMalloc(&m->gridF32, 4 * G), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->gridF16, 2 * G), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->gridGrad, 8 * G), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->gridM, 4 * G), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->gridV, 4 * G), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->wF32, 4 * NW), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->wF16, 2 * NW), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->wGrad, 4 * (size_t)stride), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->wM, 4 * NW), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->wV, 4 * NW), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->partial, 4 * (size_t)stride * kMaxTrainWaves), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->batchPos, sizeof(float) * 3 * kBatchCap), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->batchTarget, sizeof(float) * m->targetStride * kBatchCap), "hipMalloc") &&
              rti::hipCheck(hipMalloc(&m->wmean, sizeof(float)), "hipMalloc");
    if (ok) {
        V.grid = reinterpret_cast<const uint32_t*>(m->gridF16);
        V.w = m->wF16;
        k_mlp_init_grid<<<blocks(G, 256), 256, 0, stream>>>(m->gridF32, m->gridF16, G, c->seed);
        if (guide) k_mlp_init_weights<kGuideOut><<<blocks(NW, 256), 256, 0, stream>>>(m->wF32, m->wF16, c->seed);
        else k_mlp_init_weights<kOut><<<blocks(NW, 256), 256, 0, stream>>>(m->wF32, m->wF16, c->seed);
        k_fill_i64<<<blocks(G, 256), 256, 0, stream>>>(m->gridGrad, G);
        k_fill_f32<<<blocks(stride, 256), 256, 0, stream>>>(m->wGrad, (uint64_t)stride, 0.0f);
        k_fill_f32<<<1, 1, 0, stream>>>(m->wmean, 1, 0.0f);
        ok = rti::hipCheck(zeroAdam(m, stream), "init kernels") &&
             rti::hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    if (!ok) { freeMlp(m); return RT_ERR_HIP; }
    *out = m;
    return RT_OK;
}
}  // namespace

extern "C" {

rt_status rt_mlp_create(const rt_mlp_config* c, hipStream_t stream, rt_mlp** out) {
    return createNet(c, (uint32_t)kOut, stream, out);
}

rt_status rt_guide_create(const rt_mlp_config* c, hipStream_t stream, rt_mlp** out) {
    return createNet(c, (uint32_t)kGuideOut, stream, out);
}

rt_status rt_mlp_destroy(rt_mlp* m) {
    if (!m) return invalid("mlp is null");
    freeMlp(m);
    return RT_OK;
}

rt_status rt_mlp_param_counts(const rt_mlp* m, uint64_t* gridParams, uint64_t* weightParams) {
    if (!m) return invalid("mlp is null");
    if (gridParams) *gridParams = m->gridParams;
    if (weightParams) *weightParams = m->weights;
    return RT_OK;
}

rt_status rt_mlp_get_params(const rt_mlp* m, float* grid, float* weights) {
    if (!m) return invalid("mlp is null");
    if (grid) RT_HIP(hipMemcpy(grid, m->gridF32, 4 * m->gridParams, hipMemcpyDeviceToHost));
    if (weights) RT_HIP(hipMemcpy(weights, m->wF32, 4 * (size_t)m->weights, hipMemcpyDeviceToHost));
    return RT_OK;
}

rt_status rt_mlp_set_params(rt_mlp* m, const float* grid, const float* weights) {
    if (!m) return invalid("mlp is null");
    if (grid) {
        RT_HIP(hipMemcpy(m->gridF32, grid, 4 * m->gridParams, hipMemcpyHostToDevice));
        k_to_f16<<<blocks(m->gridParams, 256), 256>>>(m->gridF32, m->gridF16, m->gridParams);
    }
    if (weights) {
        RT_HIP(hipMemcpy(m->wF32, weights, 4 * (size_t)m->weights, hipMemcpyHostToDevice));
        k_to_f16<<<blocks(m->weights, 256), 256>>>(m->wF32, m->wF16, (uint64_t)m->weights);
    }
    RT_HIP(zeroAdam(m, nullptr));
    RT_HIP(hipDeviceSynchronize());
    return RT_OK;
}

rt_status rt_mlp_infer(const rt_mlp* m, const float* pos, uint32_t n, float* rgb, hipStream_t stream) {
    if (!m) return invalid("mlp is null");
    if (n == 0) return RT_OK;
    if (!pos || !rgb) return invalid("null array");
    const uint32_t tiles = (n + kTileSamples - 1) / kTileSamples;
    const unsigned grid = (tiles + kInferWaves - 1) / kInferWaves;
    if (m->outputs == (uint32_t)kGuideOut) k_mlp_infer<kGuideOut><<<grid, 64 * kInferWaves, 0, stream>>>(m->view, pos, n, rgb);
    else k_mlp_infer<kOut><<<grid, 64 * kInferWaves, 0, stream>>>(m->view, pos, n, rgb);
    RT_HIP(hipGetLastError());
    return RT_OK;
}

rt_status rt_mlp_gradients(rt_mlp* m, const float* pos, const float* target, uint32_t n, float* loss,
                           hipStream_t stream) {
    if (!m) return invalid("mlp is null");
    if (n && (!pos || !target)) return invalid("null array");
    RT_HIP(rti::mlpGradients(m, pos, target, n, loss, stream));
    return RT_OK;
}

rt_status rt_mlp_read_gradients(const rt_mlp* m, float* gridGrad, float* weightGrad) {
    if (!m) return invalid("mlp is null");
    RT_HIP(hipDeviceSynchronize());
    const int NW = (int)m->weights;
    std::vector<float> w(NW + kStats);
    RT_HIP(hipMemcpy(w.data(), m->wGrad, 4 * (size_t)(NW + kStats), hipMemcpyDeviceToHost));
    const float cnt = w[NW + 1];
    const float norm = m->outputs == (uint32_t)kGuideOut ? 1.0f : 3.0f;
    const float scale = cnt > 0.0f ? 1.0f / (norm * cnt) : 0.0f;
    if (weightGrad)
        for (int i = 0; i < NW; i++) weightGrad[i] = w[i] * scale;
    if (gridGrad) {
        std::vector<long long> g(m->gridParams);
        RT_HIP(hipMemcpy(g.data(), m->gridGrad, 8 * m->gridParams, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < m->gridParams; i++) gridGrad[i] = (float)((double)g[i] * kInvFix) * scale;
    }
    return RT_OK;
}

rt_status rt_mlp_adam_step(rt_mlp* m, hipStream_t stream) {
    if (!m) return invalid("mlp is null");
    RT_HIP(rti::mlpAdam(m, stream));
    return RT_OK;
}

rt_status rt_mlp_train_step(rt_mlp* m, const float* pos, const float* target, uint32_t n, float* loss,
                            hipStream_t stream) {
    rt_status s = rt_mlp_gradients(m, pos, target, n, loss, stream);
    if (s != RT_OK) return s;
    return rt_mlp_adam_step(m, stream);
}

rt_status rt_mlp_batch_download(const rt_mlp* m, float* pos, float* target, uint32_t* count) {
    if (!m) return invalid("mlp is null");
    const uint32_t n = m->batchSize;
    if (count) *count = n;
    RT_HIP(hipDeviceSynchronize());
    if (n && pos) RT_HIP(hipMemcpy(pos, m->batchPos, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
    if (n && target) RT_HIP(hipMemcpy(target, m->batchTarget, sizeof(float) * m->targetStride * n, hipMemcpyDeviceToHost));
    return RT_OK;
}

rt_status rt_mlp_train_frame(rt_mlp* m, float* loss, hipStream_t stream) {
    if (!m) return invalid("mlp is null");
    if (m->batchSize == 0) return invalid("no training batch: call rt_gris_learned_refresh (rt_gris_guide_train for a proposal network) first");
    return rt_mlp_train_step(m, m->batchPos, m->batchTarget, m->batchSize, loss, stream);
}

}  // extern "C"
