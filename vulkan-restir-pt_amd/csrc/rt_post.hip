// Post-processing of the two radiance outputs into the 8-bit image the
// reference presents and captures: post_proc.frag:16-43 (tone mapping: none,
// Filmic or ACES; gamma 1/2.2; the direct and indirect outputs switched off by
// the push constant), written as the screenshot bytes of
// Renderer::writeScreenshot (src/Renderer.cpp:758-793: the B8G8R8A8 UNORM
// swapchain image swizzled to R, G, B, A), row 0 at the top.
//
// The fragment at pixel (x, y) reads texel ivec2((size - 1) * uv) with
// uv = ((x + 0.5) / W, (y + 0.5) / H) (post_proc.frag:29; the full-screen quad
// of PostProcessFrag.cpp:20-27 maps uv (0, 0) to the top-left corner), so the
// last column and row are not shown and some texels are shown twice: the
// mapping is kept. pow is evaluated in double and rounded to float; the UNORM
// store rounds clamp(c, 0, 1) * 255 to nearest.
#include "rt_dpixel.h"

using namespace rtd;
using namespace rtp;

namespace {

// post_proc.frag:16-26
RT_DEV float calcFilmic(float c) {
    return (c * (c * 0.22f + 0.03f) + 0.002f) / (c * (c * 0.22f + 0.3f) + 0.06f) - 1.0f / 30.0f;
}
RT_DEV float filmic(float c) { return calcFilmic(c * 1.6f) / calcFilmic(11.2f); }
RT_DEV float aces(float c) { return (c * (c * 2.51f + 0.03f)) / (c * (c * 2.43f + 0.59f) + 0.14f); }

RT_DEV uint32_t unorm8(float c) { return (uint32_t)(gclamp(c, 0.0f, 1.0f) * 255.0f + 0.5f); }

RT_DEV float postChannel(float c, const rt_post_settings& s) {
    c = s.toneMapping == 0u ? c : (s.toneMapping == 1u ? filmic(c) : aces(c));
    if (s.correctGamma) c = (float)pow((double)c, 1.0 / 2.2);
    return c;
}

__global__ void k_post_process(const float4* __restrict__ direct, const float4* __restrict__ indirect, uint32_t W,
                               uint32_t H, rt_post_settings s, uint32_t* __restrict__ rgba8) {
    const uint32_t x = blockIdx.x * kTileX + (threadIdx.x & (kTileX - 1));
    const uint32_t y = blockIdx.y * kTileY + threadIdx.x / kTileX;
    if (x >= W || y >= H) return;
    const float u = ((float)x + 0.5f) / (float)W, v = ((float)y + 0.5f) / (float)H;
    const uint32_t cx = (uint32_t)((float)(W - 1) * u), cy = (uint32_t)((float)(H - 1) * v);
    const size_t t = (size_t)cy * W + cx;
    f3 c = s3(0.0f);
    if (s.noDirect == 0u) { const float4 d = direct[t]; c = c + mk3(d.x, d.y, d.z); }
    if (s.noIndirect == 0u) { const float4 d = indirect[t]; c = c + mk3(d.x, d.y, d.z); }
    const uint32_t r = unorm8(postChannel(c.x, s)), g = unorm8(postChannel(c.y, s)), b = unorm8(postChannel(c.z, s));
    rgba8[(size_t)y * W + x] = r | (g << 8) | (b << 16) | (255u << 24);
}

rt_status invalid(const char* m) { rti::setError(m); return RT_ERR_INVALID_ARGUMENT; }

}  // namespace

extern "C" rt_status rt_post_process(const rt_frame* f, const rt_post_settings* s, uint8_t* rgba8, hipStream_t stream) {
    if (!f || !s || !rgba8) return invalid("null argument");
    if (f->width == 0 || f->height == 0 || f->width >= 32768 || f->height >= 32768)
        return invalid("frame size must be in [1, 32767]");
    if (s->toneMapping > 2u) return invalid("toneMapping must be 0 (none), 1 (Filmic) or 2 (ACES)");
    if ((s->noDirect == 0u && !f->directOutput) || (s->noIndirect == 0u && !f->indirectOutput))
        return invalid("an enabled output is null");
    (void)rti::launch(k_post_process, tileGrid(f->width, f->height), dim3(kTileThreads), stream,
                      reinterpret_cast<const float4*>(f->directOutput), reinterpret_cast<const float4*>(f->indirectOutput),
                      f->width, f->height, *s, reinterpret_cast<uint32_t*>(rgba8));
    RT_HIP(hipGetLastError());
    return RT_OK;
}
