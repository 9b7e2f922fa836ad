/*
 * or_post.c — ORACLE (test infrastructure only): CPU restatement of Reina's
 * post-processing chain in its dispatch order (src/Reina.cpp:472-577):
 *   blurX(rtImage -> ping, threshold)   shaders/postprocessing/bloom/blurX.comp.glsl:8
 *   blurY(ping -> pong, no threshold)   shaders/postprocessing/bloom/blurY.comp.glsl:8
 *   combine(rtImage, pong -> ping)      shaders/postprocessing/bloom/combine.comp.glsl:15-27
 *   tonemap(ping -> RGBA8)              shaders/postprocessing/tonemap/tonemapping.comp.glsl:41-84
 * Quirks kept: the Gaussian's sigma is the config's percent value while the
 * half-width is 3 * W * radius / 100 pixels, and skipped taps still add to the
 * weight sum (blurCommon.h.glsl:23-47); no sRGB encode of the UNORM output
 * (tonemapping.comp.glsl:79-83). UNORM8 conversion: floor(c*255+0.5).
 * Pinned by closed forms in tests/test_oracle_post.py: a constant image blurs
 * to c * (in-bounds weight / all weights) per axis, a single pixel spreads as
 * the product of the two axes' Gaussians of sigma = radius, an image below the
 * threshold has no bloom, ACES as row-vector products against float64.
 *
 * OR_PLANT (compile-time, 0 by default; 1-4 are in or_shade.c) builds a copy
 * with one deliberate mistake for the tests to catch: 5 ACES matrices applied
 * as column vectors, 6 weight sum over in-bounds taps only, 7 threshold in
 * both passes, 8 UNORM8 by truncation, 9 sigma = the radius in pixels.
 */
#include "oracle.h"

#ifndef OR_PLANT
#define OR_PLANT 0
#endif

/* blurCommon.h.glsl:10-13 */
static float gauss(float x, float sigma) { return or_expf(((-x) * x) / ((2.0f * sigma) * sigma)); }

/* blurCommon.h.glsl:15-56 */
static void blur_axis(const float* in, float* out, uint32_t W, uint32_t H, int ax, int thresh,
                      const BloomPushConsts* b) {
    for (uint32_t y = 0; y < H; ++y) {
        for (uint32_t x = 0; x < W; ++x) {
            float radiusPx = ((float)(int32_t)(ax ? W : H) * b->radius) / 100.0f;
            int32_t k = or_f2i(radiusPx * 3.0f + 0.5f);
            v3 c = V3(0.0f, 0.0f, 0.0f);
            float ws = 0.0f;
            float sigma = OR_PLANT == 9 ? radiusPx : b->radius;
            for (int32_t i = -k; i <= k; ++i) {
                float w = gauss((float)i, sigma);
                int32_t cx = (int32_t)x + (ax ? i : 0), cy = (int32_t)y + (ax ? 0 : i);
                int out_of_bounds = cx < 0 || cx >= (int32_t)W || cy < 0 || cy >= (int32_t)H;
                if (!(OR_PLANT == 6 && out_of_bounds)) ws = ws + w;
                if (out_of_bounds) continue;
                const float* p = &in[4ull * ((uint64_t)cy * W + (uint64_t)cx)];
                float lum = vdot(V3(p[0], p[1], p[2]), V3(0.299f, 0.587f, 0.114f));
                if (thresh && lum < b->threshold) continue;
                c = vadd(c, vscale(V3(p[0], p[1], p[2]), w));
            }
            if (ws < 0.0001f) c = V3(0.0f, 0.0f, 0.0f);
            else c = V3(c.x / ws, c.y / ws, c.z / ws);
            float* o = &out[4ull * ((uint64_t)y * W + x)];
            o[0] = c.x; o[1] = c.y; o[2] = c.z; o[3] = 1.0f;
        }
    }
}

/* tonemapping.comp.glsl:17-29: color * M is the row-vector product dot(color, column j) */
static v3 rowmul(v3 c, const float m[9]) {
#if OR_PLANT == 5
    return V3(vdot(c, V3(m[0], m[3], m[6])), vdot(c, V3(m[1], m[4], m[7])), vdot(c, V3(m[2], m[5], m[8])));
#endif
    return V3(vdot(c, V3(m[0], m[1], m[2])), vdot(c, V3(m[3], m[4], m[5])), vdot(c, V3(m[6], m[7], m[8])));
}
static const float ACES_IN[9] = {0.59719f, 0.35458f, 0.04823f, 0.07600f, 0.90834f, 0.01566f,
                                 0.02840f, 0.13383f, 0.83777f};
static const float ACES_OUT[9] = {1.60475f, -0.53108f, -0.07367f, -0.10208f, 1.10813f, -0.00605f,
                                  -0.00327f, -0.07276f, 1.07602f};

static float rrt1(float v) {                      /* tonemapping.comp.glsl:35-40 */
    float a = v * (v + 0.0245786f) - 0.000090537f;
    float b = v * (0.983729f * v + 0.4329510f) + 0.238081f;
    return a / b;
}

static uint8_t unorm8(float c) {
#if OR_PLANT == 8
    return (uint8_t)or_f2i(floorf(c * 255.0f));
#endif
    return (uint8_t)or_f2i(floorf(c * 255.0f + 0.5f));
}

void or_postprocess(const float* in, uint32_t W, uint32_t H, const BloomPushConsts* bloom,
                    const TonemappingPushConsts* tm, float* ping, float* pong, uint8_t* out) {
    blur_axis(in, ping, W, H, 1, 1, bloom);
    blur_axis(ping, pong, W, H, 0, OR_PLANT == 7, bloom);
    for (uint64_t i = 0; i < (uint64_t)W * H; ++i) {           /* combine.comp.glsl:15-27 */
        for (int k = 0; k < 3; ++k) ping[4 * i + k] = in[4 * i + k] + pong[4 * i + k] * bloom->intensity;
        ping[4 * i + 3] = 1.0f;
    }
    float e = or_exp2f(tm->exposure);
    for (uint64_t i = 0; i < (uint64_t)W * H; ++i) {           /* tonemapping.comp.glsl:64-84 */
        v3 c = V3(ping[4 * i] * e, ping[4 * i + 1] * e, ping[4 * i + 2] * e);
        c = rowmul(c, ACES_IN);
        c = V3(rrt1(c.x), rrt1(c.y), rrt1(c.z));
        c = rowmul(c, ACES_OUT);
        c = V3(or_clampf(c.x, 0.0f, 1.0f), or_clampf(c.y, 0.0f, 1.0f), or_clampf(c.z, 0.0f, 1.0f));
        out[4 * i + 0] = unorm8(c.x);
        out[4 * i + 1] = unorm8(c.y);
        out[4 * i + 2] = unorm8(c.z);
        out[4 * i + 3] = 255;
    }
}
