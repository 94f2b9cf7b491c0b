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

constexpr int kLevels = 16, kF = 2, kEnc = 32, kH = 64, kOut = 3, kGuideOut = 32, kLobes = 8;
constexpr int kW1 = 0, kW2 = kEnc * kH, kW3 = kW2 + kH * kH, kWeights = kW3 + kH * kOut;
constexpr int kGuideWeights = kW3 + kH * kGuideOut;
// Raw output -> lobe parameter offsets (the untrained network's lobes sit at the octant
// directions (+-1, +-1, +-1)/sqrt 3, bit 0/1/2 of k = negative x/y/z, with concentration ~4):
// b + ln 4, c + logit(theta_k / pi), d + logit(phi_k / 2 pi); log kappa = 0.5 + 4.5 tanh((b' - 0.5) /
// 4.5), in (-4, 5).
constexpr double kLogKappaMid = 0.5, kLogKappaHalf = 4.5;
double guideBias(int k, int e) {
    if (e == 1) return std::log(4.0);
    const double sx = (k & 1) ? -1.0 : 1.0, sy = (k & 2) ? -1.0 : 1.0, sz = (k & 4) ? -1.0 : 1.0;
    auto logit = [](double p) { return std::log(p / (1.0 - p)); };
    if (e == 2) return logit(std::acos(sz / std::sqrt(3.0)) / M_PI);
    if (e == 3) {
        double phi = std::atan2(sy, sx);
        if (phi < 0) phi += 2.0 * M_PI;
        return logit(phi / (2.0 * M_PI));
    }
    return 0.0;
}
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

struct Act { double enc[kEnc], z1[kH], z2[kH], z3[kGuideOut], h1[kH], h2[kH]; Corners C[kLevels]; };

double weight(const Net& N, int i) { return N.mode ? rnd(N.w[i], 1) : (double)N.w[i]; }

// nOut: 3 (radiance network) or 32 (proposal network); W3 is [64][nOut].
void forward(const Net& N, const float* p, Act& A, int nOut = kOut) {
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
    for (int o = 0; o < nOut; o++) {
        double s = 0.0;
        for (int i = 0; i < kH; i++) s += weight(N, kW3 + i * nOut + o) * A.h2[i];
        A.z3[o] = N.mode ? (double)(float)s : s;
    }
}

// ---- proposal network: 8-lobe vMF mixture ------------------------------------------------
// Lobe k of the raw outputs z[4k..4k+3] = (a, b, c, d), offset by guideBias: pi = softmax(a),
// log kappa = 0.5 + 4.5 tanh((b' - 0.5) / 4.5), theta = pi sigmoid(c'), phi = 2 pi sigmoid(d'),
// mu = (sin theta cos phi, sin theta sin phi, cos theta); v(w) = kappa / (2 pi (1 - e^{-2 kappa}))
// e^{kappa (mu.w - 1)}.
struct Mixture {
    double pi[kLobes], kap[kLobes], dlk[kLobes], th[kLobes], ph[kLobes], mu[kLobes][3], sc[kLobes], sd[kLobes];
};
void mixtureOf(const double* z, Mixture& M) {
    double amax = -1e300;
    for (int k = 0; k < kLobes; k++) amax = std::fmax(amax, z[4 * k]);
    double es = 0.0;
    for (int k = 0; k < kLobes; k++) { M.pi[k] = std::exp(z[4 * k] - amax); es += M.pi[k]; }
    for (int k = 0; k < kLobes; k++) {
        M.pi[k] /= es;
        const double t = std::tanh((z[4 * k + 1] + guideBias(k, 1) - kLogKappaMid) / kLogKappaHalf);
        M.kap[k] = std::exp(kLogKappaMid + kLogKappaHalf * t);
        M.dlk[k] = 1.0 - t * t;
        M.sc[k] = 1.0 / (1.0 + std::exp(-(z[4 * k + 2] + guideBias(k, 2))));
        M.sd[k] = 1.0 / (1.0 + std::exp(-(z[4 * k + 3] + guideBias(k, 3))));
        M.th[k] = M_PI * M.sc[k];
        M.ph[k] = 2.0 * M_PI * M.sd[k];
        M.mu[k][0] = std::sin(M.th[k]) * std::cos(M.ph[k]);
        M.mu[k][1] = std::sin(M.th[k]) * std::sin(M.ph[k]);
        M.mu[k][2] = std::cos(M.th[k]);
    }
}
double lobeValue(const Mixture& M, int k, const double w[3]) {
    const double kap = M.kap[k];
    const double dt = M.mu[k][0] * w[0] + M.mu[k][1] * w[1] + M.mu[k][2] * w[2];
    return kap / (2.0 * M_PI * -std::expm1(-2.0 * kap)) * std::exp(kap * (dt - 1.0));
}
bool validGuideTarget(const float* y) {
    return std::isfinite(y[3]) && y[3] > 0.0f && std::isfinite(y[0]) && std::isfinite(y[1]) && std::isfinite(y[2]);
}
double guideMeanWeight(const float* target, uint32_t n, uint32_t& count) {
    double w = 0.0;
    count = 0;
    for (uint32_t s = 0; s < n; s++)
        if (validGuideTarget(target + 4 * (size_t)s)) { w += target[4 * (size_t)s + 3]; count++; }
    return count ? w / count : 0.0;
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

// Proposal network: out[n][32] = raw outputs at pos[n][3].
void orc_guide_forward(uint32_t log2T, uint32_t baseRes, uint32_t maxRes, const float* lo, const float* hi,
                       const float* grid, const float* weights, const float* pos, uint32_t n, float* out, int mode) {
    Net N;
    setup(N, log2T, baseRes, maxRes, lo, hi);
    N.grid = grid; N.w = weights; N.mode = mode;
    Act A;
    for (uint32_t s = 0; s < n; s++) {
        forward(N, pos + 3 * (size_t)s, A, kGuideOut);
        for (int o = 0; o < kGuideOut; o++) out[(size_t)kGuideOut * s + o] = (float)A.z3[o];
    }
}

// Lobe planes of raw outputs: lobes[n][16][4] (per sample the 16 float4 of the kernels'
// planes: (mu, kappa), (pi kappa / (2 pi (1 - e^{-2 kappa})), e^{-2 kappa}, cdf, pi)).
void orc_guide_lobes(const float* raw, uint32_t n, float* lobes) {
    for (uint32_t s = 0; s < n; s++) {
        double z[kGuideOut];
        for (int o = 0; o < kGuideOut; o++) z[o] = raw[(size_t)kGuideOut * s + o];
        Mixture M;
        mixtureOf(z, M);
        double cdf = 0.0;
        float* L = lobes + (size_t)64 * s;
        for (int k = 0; k < kLobes; k++) {
            cdf += M.pi[k];
            const double kap = M.kap[k];
            L[8 * k + 0] = (float)M.mu[k][0]; L[8 * k + 1] = (float)M.mu[k][1]; L[8 * k + 2] = (float)M.mu[k][2];
            L[8 * k + 3] = (float)kap;
            L[8 * k + 4] = (float)(M.pi[k] * kap / (2.0 * M_PI * -std::expm1(-2.0 * kap)));
            L[8 * k + 5] = (float)std::exp(-2.0 * kap);
            L[8 * k + 6] = k == kLobes - 1 ? 1.0f : (float)cdf;
            L[8 * k + 7] = (float)M.pi[k];
        }
    }
}

// Gradient of the proposal loss: mean over valid samples of -(w / mean w) log q(dir), target
// [n][4] = (unit direction, weight); gridGrad as orc_mlp_gradients, weightGrad 8192 entries.
// mode 1 rounds the output gradients to FP16 (clipped to +-64) as the kernels do.
double orc_guide_gradients(uint32_t log2T, uint32_t baseRes, uint32_t maxRes, const float* lo, const float* hi,
                           const float* grid, const float* weights, const float* pos, const float* target, uint32_t n,
                           double* gridGrad, double* weightGrad, int mode) {
    Net N;
    setup(N, log2T, baseRes, maxRes, lo, hi);
    N.grid = grid; N.w = weights; N.mode = mode;
    const uint64_t G = (uint64_t)kLevels * N.T * kF;
    std::memset(gridGrad, 0, sizeof(double) * G);
    std::memset(weightGrad, 0, sizeof(double) * kGuideWeights);
    uint32_t count = 0;
    const double wmean = guideMeanWeight(target, n, count);
    if (count == 0 || !(wmean > 0.0)) return 0.0;
    const double scale = 1.0 / count;
    double loss = 0.0;
    Act A;
    for (uint32_t s = 0; s < n; s++) {
        const float* y = target + 4 * (size_t)s;
        if (!validGuideTarget(y)) continue;
        forward(N, pos + 3 * (size_t)s, A, kGuideOut);
        const double wv = mode ? (double)(float)((float)y[3] * (1.0f / (float)wmean)) : y[3] / wmean;
        const double wd[3] = {y[0], y[1], y[2]};
        Mixture M;
        mixtureOf(A.z3, M);
        double v[kLobes], q = 0.0;
        for (int k = 0; k < kLobes; k++) { v[k] = lobeValue(M, k, wd); q += M.pi[k] * v[k]; }
        q = std::fmax(q, 1e-20);
        loss += -wv * std::log(q);
        double d3[kGuideOut], d2[kH], d1[kH], de[kEnc];
        for (int k = 0; k < kLobes; k++) {
            const double r = M.pi[k] * v[k] / q, kap = M.kap[k];
            const double dt = M.mu[k][0] * wd[0] + M.mu[k][1] * wd[1] + M.mu[k][2] * wd[2];
            const double e2 = std::exp(-2.0 * kap);
            double g[4];
            g[0] = wv * (M.pi[k] - r);
            g[1] = -wv * r * (1.0 - 2.0 * kap * e2 / -std::expm1(-2.0 * kap) + kap * (dt - 1.0)) * M.dlk[k];
            const double m = -wv * r * kap;
            const double st = std::sin(M.th[k]), ct = std::cos(M.th[k]), sp = std::sin(M.ph[k]), cp = std::cos(M.ph[k]);
            const double dth = m * (ct * cp * wd[0] + ct * sp * wd[1] - st * wd[2]);
            const double dph = m * (-st * sp * wd[0] + st * cp * wd[1]);
            g[2] = dth * M_PI * M.sc[k] * (1.0 - M.sc[k]);
            g[3] = dph * 2.0 * M_PI * M.sd[k] * (1.0 - M.sd[k]);
            for (int e = 0; e < 4; e++) {
                double gg = g[e];
                if (mode) gg = std::fmin(std::fmax(gg, -kClip), kClip);
                d3[4 * k + e] = rnd(gg, mode);
            }
        }
        for (int i = 0; i < kH; i++) {
            double t = 0.0;
            for (int o = 0; o < kGuideOut; o++) t += weight(N, kW3 + i * kGuideOut + o) * d3[o];
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
            for (int o = 0; o < kGuideOut; o++) weightGrad[kW3 + i * kGuideOut + o] += d3[o] * A.h2[i];
        for (int l = 0; l < kLevels; l++)
            for (int c = 0; c < 8; c++) {
                gridGrad[2 * A.C[l].slot[c]] += (double)A.C[l].wt[c] * de[2 * l];
                gridGrad[2 * A.C[l].slot[c] + 1] += (double)A.C[l].wt[c] * de[2 * l + 1];
            }
    }
    for (uint64_t i = 0; i < G; i++) gridGrad[i] *= scale;
    for (int i = 0; i < kGuideWeights; i++) weightGrad[i] *= scale;
    return loss * scale;
}

// The proposal loss in exact double arithmetic (for finite differences); wmean (> 0) fixes the
// weight normalisation so that differences see the same constant as the analytic gradient.
double orc_guide_loss(uint32_t log2T, uint32_t baseRes, uint32_t maxRes, const float* lo, const float* hi,
                      const float* grid, const float* weights, const float* pos, const float* target, uint32_t n,
                      double wmean) {
    Net N;
    setup(N, log2T, baseRes, maxRes, lo, hi);
    N.grid = grid; N.w = weights; N.mode = 0;
    uint32_t count = 0;
    const double wm = guideMeanWeight(target, n, count);
    if (count == 0) return 0.0;
    if (!(wmean > 0.0)) wmean = wm;
    double loss = 0.0;
    Act A;
    for (uint32_t s = 0; s < n; s++) {
        const float* y = target + 4 * (size_t)s;
        if (!validGuideTarget(y)) continue;
        forward(N, pos + 3 * (size_t)s, A, kGuideOut);
        const double wd[3] = {y[0], y[1], y[2]};
        Mixture M;
        mixtureOf(A.z3, M);
        double q = 0.0;
        for (int k = 0; k < kLobes; k++) q += M.pi[k] * lobeValue(M, k, wd);
        loss += -(y[3] / wmean) * std::log(std::fmax(q, 1e-20));
    }
    return loss / count;
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
