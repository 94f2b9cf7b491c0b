// ORACLE — test infrastructure only. Nothing in the product links or calls this.
//
// Post-processing restated from post_proc.frag:16-43 (tone mapping, gamma,
// direct/indirect switches, the (size - 1) * uv texel mapping) with the UNORM8
// store of the B8G8R8A8 swapchain and the byte swizzle of
// Renderer::writeScreenshot (src/Renderer.cpp:758-793): bytes R, G, B, A.
// pow is evaluated in double and rounded to float.
#include "orc_passes.h"
#include <cmath>

namespace {

float calcFilmic(float c) {
    return (c * (c * 0.22f + 0.03f) + 0.002f) / (c * (c * 0.22f + 0.3f) + 0.06f) - 1.0f / 30.0f;
}
float filmic(float c) { return calcFilmic(c * 1.6f) / calcFilmic(11.2f); }
float aces(float c) { return (c * (c * 2.51f + 0.03f)) / (c * (c * 2.43f + 0.59f) + 0.14f); }

float channel(float c, uint32_t tone, uint32_t gamma) {
    c = tone == 0 ? c : (tone == 1 ? filmic(c) : aces(c));
    if (gamma) c = (float)std::pow((double)c, 1.0 / 2.2);
    return c;
}
uint32_t unorm8(float c) {
    float v = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
    return (uint32_t)(v * 255.0f + 0.5f);
}

}  // namespace

extern "C" {

// settings4 = {toneMapping, correctGamma, noDirect, noIndirect}; direct/indirect RGBA32F (may be null when off)
void orc_post_process(const float* direct, const float* indirect, uint32_t W, uint32_t H, const uint32_t* settings4,
                      uint8_t* rgba8) {
    for (uint32_t y = 0; y < H; y++) {
        for (uint32_t x = 0; x < W; x++) {
            float u = ((float)x + 0.5f) / (float)W, v = ((float)y + 0.5f) / (float)H;
            uint32_t cx = (uint32_t)((float)(W - 1) * u), cy = (uint32_t)((float)(H - 1) * v);
            size_t t = 4 * ((size_t)cy * W + cx);
            float c[3] = {0.0f, 0.0f, 0.0f};
            for (int k = 0; k < 3; k++) {
                if (settings4[2] == 0) c[k] = c[k] + direct[t + k];
                if (settings4[3] == 0) c[k] = c[k] + indirect[t + k];
            }
            uint8_t* o = rgba8 + 4 * ((size_t)y * W + x);
            for (int k = 0; k < 3; k++) o[k] = (uint8_t)unorm8(channel(c[k], settings4[0], settings4[1]));
            o[3] = 255;
        }
    }
}

float orc_tonemap(float c, uint32_t tone) { return channel(c, tone, 0); }

}  // extern "C"
