// post.hip — bloom and ACES-fitted tone mapping to RGBA8, in Reina's dispatch
// order (src/Reina.cpp:472-577):
//   blurX(rtImage -> ping, threshold)   shaders/postprocessing/bloom/blurX.comp.glsl:8
//   blurY(ping -> pong)                 shaders/postprocessing/bloom/blurY.comp.glsl:8
//   combine(rtImage, pong -> ping)      shaders/postprocessing/bloom/combine.comp.glsl:15-27
//   tonemap(ping -> out RGBA8)          shaders/postprocessing/tonemap/tonemapping.comp.glsl:41-84
// Kept from the reference: the Gaussian's sigma is the radius percentage
// itself while the half-width is 3 * size * radius / 100 pixels, and taps that
// are out of bounds or under the threshold still add to the weight sum; no
// sRGB encode. UNORM8 conversion: floor(c * 255 + 0.5).
//
// The blur is a 1-D stencil: each block stages the row (or column) segment it
// needs, plus the per-tap weights, in LDS, so HBM sees each input texel once
// per pass. The weights depend only on the tap offset and are evaluated once
// per block.
#include <hip/hip_runtime.h>

#include "rt_math.h"
#include "wavefront.h"

namespace rt {

constexpr int kTile = 256;        // output pixels per block along the blur axis
constexpr int kMaxHalf = 1024;    // LDS-staged half-width; larger kernels read HBM directly

__device__ __forceinline__ float gauss(float x, float sigma) { return expf_(((-x) * x) / ((2.0f * sigma) * sigma)); }

// axis 1 = X (along rows), 0 = Y (along columns)
template <int AXIS, bool THRESH>
__global__ __launch_bounds__(256) void k_blur(const float4* __restrict__ in, float4* __restrict__ out, uint32_t W,
                                              uint32_t H, BloomPushConsts b) {
    __shared__ float s_w[2 * kMaxHalf + 1];
    __shared__ float4 s_px[kTile + 2 * kMaxHalf];
    const uint32_t len = AXIS ? W : H;        // extent along the blur axis
    const uint32_t lines = AXIS ? H : W;
    const uint32_t line = blockIdx.y;
    const int32_t start = (int32_t)(blockIdx.x * kTile);
    if (line >= lines) return;
    float radiusPx = ((float)(int32_t)len * b.radius) / 100.0f;
    int32_t k = f2i_(radiusPx * 3.0f + 0.5f);
    bool staged = k <= kMaxHalf;
    if (staged) {
        for (int32_t t = threadIdx.x; t <= 2 * k; t += blockDim.x) s_w[t] = gauss((float)(t - k), b.radius);
        for (int32_t t = threadIdx.x; t < kTile + 2 * k; t += blockDim.x) {
            int32_t c = start - k + t;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (c >= 0 && c < (int32_t)len) {
                uint64_t idx = AXIS ? ((uint64_t)line * W + (uint32_t)c) : ((uint64_t)(uint32_t)c * W + line);
                v = in[idx];
            }
            s_px[t] = v;
        }
        __syncthreads();
    }
    int32_t pos = start + (int32_t)threadIdx.x;
    if (pos >= (int32_t)len) return;
    V3 c = v3(0.0f, 0.0f, 0.0f);
    float ws = 0.0f;
    for (int32_t i = -k; i <= k; ++i) {
        float w = staged ? s_w[i + k] : gauss((float)i, b.radius);
        ws = ws + w;
        int32_t cc = pos + i;
        if (cc < 0 || cc >= (int32_t)len) continue;
        float4 p;
        if (staged) {
            p = s_px[(int32_t)threadIdx.x + i + k];
        } else {
            uint64_t idx = AXIS ? ((uint64_t)line * W + (uint32_t)cc) : ((uint64_t)(uint32_t)cc * W + line);
            p = in[idx];
        }
        V3 rgb = v3(p.x, p.y, p.z);
        if (THRESH) {
            float lum = dot(rgb, v3(0.299f, 0.587f, 0.114f));
            if (lum < b.threshold) continue;
        }
        c = add(c, scale(rgb, w));
    }
    if (ws < 0.0001f) c = v3(0.0f, 0.0f, 0.0f);
    else c = divs(c, ws);
    uint64_t o = AXIS ? ((uint64_t)line * W + (uint32_t)pos) : ((uint64_t)(uint32_t)pos * W + line);
    out[o] = make_float4(c.x, c.y, c.z, 1.0f);
}

__device__ __forceinline__ float rrt1(float v) {
    float a = v * (v + 0.0245786f) - 0.000090537f;
    float bb = v * (0.983729f * v + 0.4329510f) + 0.238081f;
    return a / bb;
}

__device__ __forceinline__ uint32_t unorm8(float c) { return (uint32_t)f2i_(floorf(c * 255.0f + 0.5f)) & 0xFFu; }

// combine + tonemap fused: the combined image is also written (it is the
// reference's ping image at that point).
__global__ __launch_bounds__(256) void k_combine_tonemap(const float4* __restrict__ orig, const float4* __restrict__ bloom,
                                                         float4* __restrict__ combined, uchar4* __restrict__ out,
                                                         uint32_t n, float intensity, float exposure) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 o = orig[i], bl = bloom[i];
    V3 c = v3(o.x + bl.x * intensity, o.y + bl.y * intensity, o.z + bl.z * intensity);
    combined[i] = make_float4(c.x, c.y, c.z, 1.0f);
    float e = exp2f_(exposure);
    c = v3(c.x * e, c.y * e, c.z * e);
    c = v3(dot(c, v3(0.59719f, 0.35458f, 0.04823f)), dot(c, v3(0.07600f, 0.90834f, 0.01566f)),
           dot(c, v3(0.02840f, 0.13383f, 0.83777f)));
    c = v3(rrt1(c.x), rrt1(c.y), rrt1(c.z));
    c = v3(dot(c, v3(1.60475f, -0.53108f, -0.07367f)), dot(c, v3(-0.10208f, 1.10813f, -0.00605f)),
           dot(c, v3(-0.00327f, -0.07276f, 1.07602f)));
    c = v3(clampf_(c.x, 0.0f, 1.0f), clampf_(c.y, 0.0f, 1.0f), clampf_(c.z, 0.0f, 1.0f));
    out[i] = make_uchar4((unsigned char)unorm8(c.x), (unsigned char)unorm8(c.y), (unsigned char)unorm8(c.z), 255);
}

hipError_t launch_postprocess(const float4* in, float4* ping, float4* pong, uchar4* out, uint32_t W, uint32_t H,
                              const BloomPushConsts& b, const TonemappingPushConsts& t, hipStream_t st) {
    dim3 gx((W + kTile - 1) / kTile, H);
    hipLaunchKernelGGL((k_blur<1, true>), gx, dim3(kTile), 0, st, in, ping, W, H, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 gy((H + kTile - 1) / kTile, W);
    hipLaunchKernelGGL((k_blur<0, false>), gy, dim3(kTile), 0, st, ping, pong, W, H, b);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint32_t n = W * H;
    hipLaunchKernelGGL(k_combine_tonemap, dim3((n + 255) / 256), dim3(256), 0, st, in, pong, ping, out, n,
                       b.intensity, t.exposure);
    return hipGetLastError();
}

}  // namespace rt
