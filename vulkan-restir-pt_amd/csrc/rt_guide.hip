// The learned directional proposal's per-frame kernels: the proposal network
// evaluated once per 2x2 pixel block into the frame's lobe planes
// (matrix-core forward of rt_mlp.h, 32 blocks per wave), and the gather of the
// frame's training batch from the path tracer's per-pixel records.
#include "rt_dpixel.h"
#include "rt_guide.h"

using namespace rtd;
using namespace rtp;
using namespace rtm;
using rti::launch;

namespace {

constexpr int kRefreshWaves = 4;
constexpr uint32_t kGuideBatch = 1u << 14;  // training samples gathered per frame

// Primary vertex of pixel p as the pass kernels rebuild it; false for background.
__device__ __forceinline__ bool primaryPos(const rt_camera& cam, const GB& g, uint32_t p, f3& pos) {
    const uint32_t x = p % g.W, y = p / g.W;
    GPix gp;
    if (!unpackGB(texDNCentre(g, x, y), fetchAM(g, (int)x, (int)y), gp)) return false;
    const f2 uv = pixelUV(x, y, g.W, g.H);
    const CamRay ray = pinholeRay(cam, mk2(uv.x, 1.0f - uv.y));
    pos = ray.ori + ray.dir * (gp.depth - 1e-4f);
    return true;
}

// One sample per 2x2 pixel block b: the block's first non-background pixel in raster order.
__global__ __launch_bounds__(64 * kRefreshWaves) void k_gris_guide_refresh(MlpView M, rt_camera cam, GB g,
                                                                          float4* __restrict__ lobes) {
    __shared__ _Float16 sW[kGuideWeights];
    for (int i = threadIdx.x; i < kGuideWeights; i += blockDim.x) sW[i] = M.w[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t bw = rtg::guideBlocksX(g.W), n = rtg::guideBlocks(g.W, g.H);
    const uint32_t p = (blockIdx.x * kRefreshWaves + wave) * kTileSamples + (lane & 31u);
    f3 pos = s3(0.0f);
    bool valid = false;
    if (p < n) {
        const uint32_t bx = 2u * (p % bw), by = 2u * (p / bw);
        for (uint32_t c = 0; c < 4u && !valid; c++) {
            const uint32_t x = bx + (c & 1u), y = by + (c >> 1);
            if (x < g.W && y < g.H) valid = primaryPos(cam, g, y * g.W + x, pos);
        }
    }
    float u = 0.0f, v = 0.0f, w = 0.0f;
    normalizePos(M, pos.x, pos.y, pos.z, u, v, w);
    h4 enc[4];
    encodeLane(M, lane >> 5, u, v, w, enc);
    Fwd F;
    forwardTileT<kGuideOut>(sW, lane, enc, F);
    // this lane holds lobes k = 2g + h (registers 4g .. 4g + 3)
    const int h = (int)(lane >> 5);
    float amax = fmaxf(fmaxf(F.z3[0], F.z3[4]), fmaxf(F.z3[8], F.z3[12]));
    amax = fmaxf(amax, __shfl_xor(amax, 32));
    float ea[4], esum = 0.0f;
    for (int q = 0; q < 4; q++) { ea[q] = expf(F.z3[4 * q] - amax); esum += ea[q]; }
    esum += __shfl_xor(esum, 32);
    float pi[4], piAll[kGuideLobes];
    for (int q = 0; q < 4; q++) {
        pi[q] = ea[q] / esum;
        const float other = __shfl_xor(pi[q], 32);
        piAll[2 * q + h] = pi[q];
        piAll[2 * q + 1 - h] = other;
    }
    float cdf[kGuideLobes];
    float c = 0.0f;
    for (int k = 0; k < kGuideLobes; k++) { c += piAll[k]; cdf[k] = c; }
    cdf[kGuideLobes - 1] = 1.0f;
    if (!valid) return;
    for (int q = 0; q < 4; q++) {
        const int k = 2 * q + h;
        float dlk;
        const float kap = expf(guideLogKappa(F.z3[4 * q + 1] + guideBias(k, 1), dlk));
        const float e2 = expf(-2.0f * kap);
        const float th = RT_PI / (1.0f + expf(-(F.z3[4 * q + 2] + guideBias(k, 2))));
        const float ph = 2.0f * RT_PI / (1.0f + expf(-(F.z3[4 * q + 3] + guideBias(k, 3))));
        const float st = sinf(th), ct = cosf(th);
        lobes[(size_t)(2 * k) * n + p] = make_float4(st * cosf(ph), st * sinf(ph), ct, kap);
        lobes[(size_t)(2 * k + 1) * n + p] =
            make_float4(pi[q] * (kap / (2.0f * RT_PI * (1.0f - e2))), e2, cdf[k], pi[q]);
    }
}

// Training sample i: pixel offset + i * stride; position of its primary vertex, target =
// (direction sampled there, luminance of the path's contributions) or weight -1 (masked)
// for background pixels and paths that ended at the primary vertex.
__global__ void k_gris_guide_gather(rt_camera cam, GB g, const float4* __restrict__ samples, uint32_t stride,
                                    uint32_t offset, uint32_t count, float* __restrict__ batchPos,
                                    float* __restrict__ batchTarget) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t p = offset + i * stride;
    f3 pos = s3(0.0f);
    float4 t = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    if (primaryPos(cam, g, p, pos)) {
        const float4 r = samples[p];
        if (r.x * r.x + r.y * r.y + r.z * r.z > 0.5f && r.w > 0.0f) t = r;
    }
    batchPos[3 * (size_t)i] = pos.x;
    batchPos[3 * (size_t)i + 1] = pos.y;
    batchPos[3 * (size_t)i + 2] = pos.z;
    batchTarget[4 * (size_t)i] = t.x;
    batchTarget[4 * (size_t)i + 1] = t.y;
    batchTarget[4 * (size_t)i + 2] = t.z;
    batchTarget[4 * (size_t)i + 3] = t.w;
}

}  // namespace

namespace rti {

hipError_t launchGuideRefresh(const rt_mlp* m, const rt_camera* cams, const rt_frame& f, hipStream_t stream) {
    const GB g = makeGB(f.depthNormal, f.albedoMatId, f.motionVector, f.width, f.height);
    const uint32_t n = ((f.width + 1u) >> 1) * ((f.height + 1u) >> 1);
    const uint32_t tiles = (n + kTileSamples - 1) / kTileSamples;
    (void)launch(k_gris_guide_refresh, dim3((tiles + kRefreshWaves - 1) / kRefreshWaves), dim3(64 * kRefreshWaves),
                 stream, m->view, cams[0], g, reinterpret_cast<float4*>(f.guideLobes));
    return hipGetLastError();
}

hipError_t launchGuideGather(rt_mlp* m, const rt_camera* cams, const rt_frame& f, hipStream_t stream) {
    const GB g = makeGB(f.depthNormal, f.albedoMatId, f.motionVector, f.width, f.height);
    const uint32_t n = f.width * f.height;
    const uint32_t stride = std::max(1u, (n + kGuideBatch - 1) / kGuideBatch);
    const uint32_t offset = cams[0].seed % stride;
    m->batchSize = offset < n ? (n - offset + stride - 1) / stride : 0u;
    if (m->batchSize)
        (void)launch(k_gris_guide_gather, dim3((m->batchSize + 255u) / 256u), dim3(256), stream, cams[0], g,
                     reinterpret_cast<const float4*>(f.guideSamples), stride, offset, m->batchSize, m->batchPos,
                     m->batchTarget);
    return hipGetLastError();
}

}  // namespace rti
