// ORACLE — test infrastructure only. Only tests/, the smoke check and the CPU
// baseline leg of the benchmark load this library; the product never does.
//
// Scalar restatement of the learned target function (hash-grid encoding + MLP)
// and of its loss gradient, for checking the GPU kernels. The reference has no
// learned component; the method follows the multiresolution hash encoding of
// Mueller et al. 2022 (instant-ngp: trilinear interpolation of per-level
// feature tables, dense indexing where the level grid fits the table, else the
// hash x*1 ^ y*2654435761 ^ z*805459861 mod T) and the relative L2 loss of
// Mueller et al. 2021 (neural radiance caching). Sums are taken in double.
//
// mode 1 mirrors where the GPU rounds to FP16 (weights, encoding, hidden
// activations and back-propagated gradients at each matrix product); mode 0 is
// the same function in exact double arithmetic (for finite-difference checks).
#include "orc_math.h"
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kLevels = 16, kF = 2, kEnc = 32, kH = 64, kOut = 3;
constexpr int kW1 = 0, kW2 = kEnc * kH, kW3 = kW2 + kH * kH, kWeights = kW3 + kH * kOut;
constexpr double kEps = 0.01, kClip = 64.0, kLogClamp = 20.0;

struct Net {
    uint32_t T, mask, dense;
    uint32_t res[kLevels];
    float lo[3], invExt[3];
    const float* grid;  // [level][T][2]
    const float* w;     // kWeights
    int mode;
};

double rnd(double x, int mode) { return mode ? (double)orc::f16_to_f32(orc::f32_to_f16((float)x)) : x; }

void setup(Net& N, uint32_t log2T, uint32_t baseRes, uint32_t maxRes, const float lo[3], const float hi[3]) {
    N.T = 1u << log2T;
    N.mask = N.T - 1;
    N.dense = 0;
    const double b = std::exp((std::log((double)maxRes) - std::log((double)baseRes)) / (kLevels - 1));
    for (int l = 0; l < kLevels; l++) {
        uint32_t r = (uint32_t)std::floor((double)baseRes * std::pow(b, (double)l) + 1e-9);
        N.res[l] = r;
        if ((uint64_t)(r + 1) * (r + 1) * (r + 1) <= N.T) N.dense |= 1u << l;
    }
    for (int a = 0; a < 3; a++) { N.lo[a] = lo[a]; N.invExt[a] = 1.0f / (hi[a] - lo[a]); }
}

struct Corners { uint64_t slot[8]; float wt[8]; };

// Cell, corner slots and trilinear weights in FP32, operation for operation as the kernels.
void corners(const Net& N, int l, const float uvw[3], Corners& C) {
    float Nl = (float)N.res[l];
    float x = uvw[0] * Nl, y = uvw[1] * Nl, z = uvw[2] * Nl;
    // cell clamped to [0, N-1]: u = 1 lands on the last corner with weight 1
    float fx = std::fmin(std::floor(x), Nl - 1.0f), fy = std::fmin(std::floor(y), Nl - 1.0f),
          fz = std::fmin(std::floor(z), Nl - 1.0f);
    uint32_t ci[3] = {(uint32_t)fx, (uint32_t)fy, (uint32_t)fz};
    float fr[3] = {x - fx, y - fy, z - fz};
    for (int c = 0; c < 8; c++) {
        uint32_t cx = ci[0] + (c & 1), cy = ci[1] + ((c >> 1) & 1), cz = ci[2] + ((c >> 2) & 1);
        uint32_t idx;
        if ((N.dense >> l) & 1u) {
            uint32_t r = N.res[l] + 1u;
            idx = cx + r * (cy + r * cz);
        } else {
            idx = (cx ^ (cy * 2654435761u) ^ (cz * 805459861u)) & N.mask;
        }
        C.slot[c] = (uint64_t)l * N.T + idx;
        float wx = (c & 1) ? fr[0] : 1.0f - fr[0];
        float wy = (c & 2) ? fr[1] : 1.0f - fr[1];
        float wz = (c & 4) ? fr[2] : 1.0f - fr[2];
        C.wt[c] = (wx * wy) * wz;
    }
}

void unitPos(const Net& N, const float* p, float uvw[3]) {
    for (int a = 0; a < 3; a++) uvw[a] = std::fmin(std::fmax((p[a] - N.lo[a]) * N.invExt[a], 0.0f), 1.0f);
}

struct Act { double enc[kEnc], z1[kH], z2[kH], z3[kOut], h1[kH], h2[kH]; Corners C[kLevels]; };

double weight(const Net& N, int i) { return N.mode ? rnd(N.w[i], 1) : (double)N.w[i]; }

void forward(const Net& N, const float* p, Act& A) {
    float uvw[3];
    unitPos(N, p, uvw);
    for (int l = 0; l < kLevels; l++) {
        corners(N, l, uvw, A.C[l]);
        if (N.mode) {  // FP32 accumulation over FP16 table entries, as the kernels
            float f0 = 0.0f, f1 = 0.0f;
            for (int c = 0; c < 8; c++) {
                f0 = f0 + A.C[l].wt[c] * (float)rnd(N.grid[2 * A.C[l].slot[c]], 1);
                f1 = f1 + A.C[l].wt[c] * (float)rnd(N.grid[2 * A.C[l].slot[c] + 1], 1);
            }
            A.enc[2 * l] = rnd(f0, 1);
            A.enc[2 * l + 1] = rnd(f1, 1);
        } else {
            double f0 = 0.0, f1 = 0.0;
            for (int c = 0; c < 8; c++) {
                f0 += (double)A.C[l].wt[c] * N.grid[2 * A.C[l].slot[c]];
                f1 += (double)A.C[l].wt[c] * N.grid[2 * A.C[l].slot[c] + 1];
            }
            A.enc[2 * l] = f0;
            A.enc[2 * l + 1] = f1;
        }
    }
    for (int o = 0; o < kH; o++) {
        double s = 0.0;
        for (int i = 0; i < kEnc; i++) s += weight(N, kW1 + i * kH + o) * A.enc[i];
        A.z1[o] = N.mode ? (double)(float)s : s;
        A.h1[o] = rnd(std::fmax(A.z1[o], 0.0), N.mode);
    }
    for (int o = 0; o < kH; o++) {
        double s = 0.0;
        for (int i = 0; i < kH; i++) s += weight(N, kW2 + i * kH + o) * A.h1[i];
        A.z2[o] = N.mode ? (double)(float)s : s;
        A.h2[o] = rnd(std::fmax(A.z2[o], 0.0), N.mode);
    }
    for (int o = 0; o < kOut; o++) {
        double s = 0.0;
        for (int i = 0; i < kH; i++) s += weight(N, kW3 + i * kOut + o) * A.h2[i];
        A.z3[o] = N.mode ? (double)(float)s : s;
    }
}

double outAct(double z) { return std::exp(std::fmin(std::fmax(z, -kLogClamp), kLogClamp)); }

bool validTarget(const float* y) {
    return y[0] >= 0.0f && y[1] >= 0.0f && y[2] >= 0.0f && std::isfinite(y[0]) && std::isfinite(y[1]) &&
           std::isfinite(y[2]);
}

}  // namespace

extern "C" {

// rgb[n][3] = network(pos[n][3]).
void orc_mlp_forward(uint32_t log2T, uint32_t baseRes, uint32_t maxRes, const float* lo, const float* hi,
                     const float* grid, const float* weights, const float* pos, uint32_t n, float* rgb, int mode) {
    Net N;
    setup(N, log2T, baseRes, maxRes, lo, hi);
    N.grid = grid; N.w = weights; N.mode = mode;
    Act A;
    for (uint32_t s = 0; s < n; s++) {
        forward(N, pos + 3 * (size_t)s, A);
        for (int c = 0; c < kOut; c++) rgb[3 * (size_t)s + c] = (float)outAct(A.z3[c]);
    }
}

// Gradient of the mean relative L2 loss over the valid samples; gridGrad has
// levels * 2^log2T * 2 entries, weightGrad kWeights. Returns the mean loss.
double orc_mlp_gradients(uint32_t log2T, uint32_t baseRes, uint32_t maxRes, const float* lo, const float* hi,
                         const float* grid, const float* weights, const float* pos, const float* target, uint32_t n,
                         double* gridGrad, double* weightGrad, int mode) {
    Net N;
    setup(N, log2T, baseRes, maxRes, lo, hi);
    N.grid = grid; N.w = weights; N.mode = mode;
    const uint64_t G = (uint64_t)kLevels * N.T * kF;
    std::memset(gridGrad, 0, sizeof(double) * G);
    std::memset(weightGrad, 0, sizeof(double) * kWeights);
    uint32_t count = 0;
    for (uint32_t s = 0; s < n; s++) count += validTarget(target + 3 * (size_t)s);
    if (count == 0) return 0.0;
    const double scale = 1.0 / (3.0 * count);
    double loss = 0.0;
    Act A;
    for (uint32_t s = 0; s < n; s++) {
        const float* y = target + 3 * (size_t)s;
        if (!validTarget(y)) continue;
        forward(N, pos + 3 * (size_t)s, A);
        double d3[kOut], d2[kH], d1[kH], de[kEnc];
        for (int c = 0; c < kOut; c++) {
            double yh = outAct(A.z3[c]);
            double den = yh * yh + kEps;
            double diff = yh - y[c];
            loss += diff * diff / den;
            double g = (std::fabs(A.z3[c]) < kLogClamp) ? 2.0 * diff / den * yh : 0.0;
            if (mode) g = std::fmin(std::fmax(g, -kClip), kClip);
            d3[c] = rnd(g, mode);
        }
        for (int i = 0; i < kH; i++) {
            double t = 0.0;
            for (int o = 0; o < kOut; o++) t += weight(N, kW3 + i * kOut + o) * d3[o];
            d2[i] = A.z2[i] > 0.0 ? rnd(N.mode ? (double)(float)t : t, mode) : 0.0;
        }
        for (int i = 0; i < kH; i++) {
            double t = 0.0;
            for (int o = 0; o < kH; o++) t += weight(N, kW2 + i * kH + o) * d2[o];
            d1[i] = A.z1[i] > 0.0 ? rnd(N.mode ? (double)(float)t : t, mode) : 0.0;
        }
        for (int i = 0; i < kEnc; i++) {
            double t = 0.0;
            for (int o = 0; o < kH; o++) t += weight(N, kW1 + i * kH + o) * d1[o];
            de[i] = N.mode ? (double)(float)t : t;
        }
        for (int i = 0; i < kEnc; i++)
            for (int o = 0; o < kH; o++) weightGrad[kW1 + i * kH + o] += d1[o] * A.enc[i];
        for (int i = 0; i < kH; i++)
            for (int o = 0; o < kH; o++) weightGrad[kW2 + i * kH + o] += d2[o] * A.h1[i];
        for (int i = 0; i < kH; i++)
            for (int o = 0; o < kOut; o++) weightGrad[kW3 + i * kOut + o] += d3[o] * A.h2[i];
        for (int l = 0; l < kLevels; l++)
            for (int c = 0; c < 8; c++) {
                gridGrad[2 * A.C[l].slot[c]] += (double)A.C[l].wt[c] * de[2 * l];
                gridGrad[2 * A.C[l].slot[c] + 1] += (double)A.C[l].wt[c] * de[2 * l + 1];
            }
    }
    for (uint64_t i = 0; i < G; i++) gridGrad[i] *= scale;
    for (int i = 0; i < kWeights; i++) weightGrad[i] *= scale;
    return loss * scale;
}

// Mean loss in exact double arithmetic (mode 0 forward), for finite
// differences. The loss treats its denominators sg(y^)^2 + eps as constants:
// denOut (n x 3, may be null) receives them, and denIn (may be null) supplies
// frozen ones, so that differences taken with denIn from the base point match
// the analytic gradient.
double orc_mlp_loss(uint32_t log2T, uint32_t baseRes, uint32_t maxRes, const float* lo, const float* hi,
                    const float* grid, const float* weights, const float* pos, const float* target, uint32_t n,
                    const double* denIn, double* denOut) {
    Net N;
    setup(N, log2T, baseRes, maxRes, lo, hi);
    N.grid = grid; N.w = weights; N.mode = 0;
    double loss = 0.0;
    uint32_t count = 0;
    Act A;
    for (uint32_t s = 0; s < n; s++) {
        const float* y = target + 3 * (size_t)s;
        if (!validTarget(y)) continue;
        count++;
        forward(N, pos + 3 * (size_t)s, A);
        for (int c = 0; c < kOut; c++) {
            double yh = outAct(A.z3[c]);
            double diff = yh - y[c];
            double den = denIn ? denIn[3 * (size_t)s + c] : yh * yh + kEps;
            if (denOut) denOut[3 * (size_t)s + c] = yh * yh + kEps;
            loss += diff * diff / den;
        }
    }
    return count ? loss / (3.0 * count) : 0.0;
}

// Level resolutions and the dense-level mask, as the network set-up derives them.
uint32_t orc_mlp_levels(uint32_t log2T, uint32_t baseRes, uint32_t maxRes, uint32_t* res16) {
    Net N;
    float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    setup(N, log2T, baseRes, maxRes, lo, hi);
    for (int l = 0; l < kLevels; l++) res16[l] = N.res[l];
    return N.dense;
}

}  // extern "C"
