// integrator.hip — wavefront path tracer kernels.
//
// One vkCmdTraceRaysKHR of the reference (src/Reina.cpp:460-469) runs one
// raygen invocation per pixel that loops over samplesPerPixel samples and
// up to maxBounces segments each (shaders/raytrace/raytrace.rgen.glsl:97-184,
// 250-285). Here each pixel owns a path slot with its state in HBM (SoA,
// indexed by pixel) and the work proceeds in bounce iterations:
//
//   k_init    raygen of sample 0 for every pixel -> ray queue
//   k_trace   closest hit for every queued ray (BVH8 traversal, trace_kernel.h)
//   k_shade   hit/miss shading + the raygen bounce logic; a path that ends
//             finishes its sample and regenerates the pixel's next camera ray
//             (the RNG state runs sequentially through the pixel's samples,
//             exactly as in the reference); surviving rays are appended to the
//             next queue (wave64 ballot + block scan, one atomic per 256 rays)
//   k_finish  mean over the non-NaN samples, running-mean blend, store
//
// Queues are split into one partition per XCD (see wavefront.h). k_trace is
// persistent: a wave claims chunks of its XCD's partition and steals from the
// others when it runs dry; idle lanes are refilled mid-flight. k_shade is
// static: block b works on partition b % 8 in tiles of 256. Per-iteration
// control words are triple-buffered so no host round trip or memset sits
// between iterations.
#include <hip/hip_runtime.h>

#include "shade_step.h"
#include "traverse.h"
#include "wavefront.h"

namespace rt {

__global__ __launch_bounds__(256) void k_init(TraceParams P, PathState S, Queues Q, Ctrl* ctrl,
                                              unsigned long long* stats, uint32_t base, uint32_t npx) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;   // index in the group
    const bool anyrays = P.spp > 0 && P.maxBounces > 0;
    if (blockIdx.x == 0 && threadIdx.x < 3 * kParts) {
        uint32_t slot = threadIdx.x / kParts, p = threadIdx.x % kParts;
        uint32_t c = 0;
        if (slot == 0 && anyrays) {
            // pixels of partition p: full tiles t = p, p + 8, ... plus a partial last tile
            uint32_t tiles = (npx + 1023u) >> 10;
            for (uint32_t t = p; t < tiles; t += kParts) c += min(1024u, npx - (t << 10));
        }
        ctrl[slot].count[p][0] = c;
        ctrl[slot].fetch[p][0] = 0;
        ctrl[slot].sfetch[p][0] = 0;
        if (p == 0) {
            ctrl[slot].done[0] = 0;
            ctrl[slot].sdone[0] = 0;
        }
        if (threadIdx.x == 0 && base == 0) { stats[0] = 0; stats[1] = 0; stats[2] = 0; stats[3] = 0; stats[4] = 0; }
    }
    if (l >= npx) return;
    const uint32_t i = base + l;   // slot: group g = i / (W * H) of pixel i % (W * H)
    const uint32_t npix = P.W * P.H, g = i / npix, pix = i - g * npix;
    uint32_t x = pix % P.W, y = pix / P.W;
    // raytrace.rgen.glsl:259; group g of G takes the seed of batch sampleBatch * G + g
    uint32_t rng = ((P.sampleBatch * P.groups + g) * P.H + y) * P.W + x;
    if (!anyrays) {
        // no traced segment: every sample contributes clamp(0) = 0
        S.st0[i] = make_uint4(rng, P.spp, 0u, 0u);
        S.st3[i] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(P.spp));
        return;
    }
    V3 o, d;
    starting_ray(P.cam, (float)x, (float)y, rng, o, d, (P.fixes & REINA_RT_TRACE_FIX_LENS_RNG) != 0u);
    S.st0[i] = make_uint4(rng, 0u, 0u, 0u);
    S.st1[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);   // accumulatedDistance starts at 0; not reset per path
    S.st2[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    S.st3[i] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
    if (P.nee) S.st4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t p = (l >> 10) & (kParts - 1u);
    uint32_t j = p * Q.cap + (((l >> 13) << 10) | (l & 1023u));
    Q.ray[0][2 * j] = make_float4(o.x, o.y, o.z, 0.0f);
    Q.ray[0][2 * j + 1] = make_float4(d.x, d.y, d.z, 10000.0f);
    Q.pix[0][j] = i;
}

// Shade every ray of the current queue; block b serves partition b % 8 in
// tiles of 256 rays, and the survivors of a tile are appended to the same
// partition of the next queue with one atomic (wave ballot + LDS scan).
template <bool NEE>
__global__ __launch_bounds__(256) void k_shade(SceneView sv, TraceParams P, PathState S, Queues Q, uint32_t in_q,
                                               Ctrl* ctrl, uint32_t slot) {
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_base;
    const uint32_t nslot = (slot + 1) % 3, zslot = (slot + 2) % 3;
    if (blockIdx.x == 0 && threadIdx.x < kParts) {
        ctrl[zslot].count[threadIdx.x][0] = 0;
        ctrl[zslot].fetch[threadIdx.x][0] = 0;
        ctrl[zslot].sfetch[threadIdx.x][0] = 0;
        if (threadIdx.x == 0) {
            ctrl[zslot].done[0] = 0;
            ctrl[zslot].sdone[0] = 0;
        }
    }
    const uint32_t p = blockIdx.x % kParts;
    const uint32_t nblk = gridDim.x / kParts;       // blocks serving this partition
    const uint32_t bi = blockIdx.x / kParts;
    const uint32_t n = ctrl[slot].count[p][0];
    const uint32_t base_in = p * Q.cap;
    const float4* rin = Q.ray[in_q];
    const uint32_t* pin = Q.pix[in_q];
    float4* rout = Q.ray[in_q ^ 1];
    uint32_t* pout = Q.pix[in_q ^ 1];
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    for (uint32_t t0 = bi * 256u; t0 < n; t0 += nblk * 256u) {
        const uint32_t k = t0 + threadIdx.x;
        bool push = false;
        uint32_t pxl = 0;
        V3 no = v3(0.0f, 0.0f, 0.0f), nd = no;
        if (k < n) {
            const uint32_t i = base_in + k;
            pxl = pin[i];
            const ShadeResult res = shade_ray<NEE>(sv, P, S, pxl, pxl, rin[2 * i], rin[2 * i + 1], Q.hit[i]);
            push = res.push;
            no = res.no;
            nd = res.nd;
            if (NEE) {
                if (res.shadow) {
                    Q.sray[2 * i] = make_float4(res.so.x, res.so.y, res.so.z, 0.0f);
                    Q.sray[2 * i + 1] = make_float4(res.sd.x, res.sd.y, res.sd.z, res.stmax);
                    Q.scon[i] = make_float4(res.scontrib.x, res.scontrib.y, res.scontrib.z, __uint_as_float(pxl));
                } else {
                    Q.sray[2 * i + 1] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);   // no shadow ray
                }
            }
        }
        // compaction: wave ballot, block scan over the 4 waves, one atomic per tile
        uint64_t mask = __ballot(push ? 1 : 0);
        uint32_t before = (uint32_t)__popcll(mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
        if (lane == 0) s_wave[wave] = (uint32_t)__popcll(mask);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
            s_base = tot ? atomicAdd(&ctrl[nslot].count[p][0], tot) : 0u;
        }
        __syncthreads();
        if (push) {
            uint32_t off = s_base + before;
            for (uint32_t w = 0; w < wave; ++w) off += s_wave[w];
            uint32_t j = base_in + off;
            rout[2 * j] = make_float4(no.x, no.y, no.z, 0.0f);
            rout[2 * j + 1] = make_float4(nd.x, nd.y, nd.z, 10000.0f);
            pout[j] = pxl;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_finish(TraceParams P, PathState S, float4* img) {
    uint32_t npx = P.W * P.H;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    float4 s3 = S.st3[i];
    uint32_t cnt = __float_as_uint(s3.w);
    for (uint32_t g = 1; g < P.groups; ++g) {   // sample groups: sums in group order
        const float4 t = S.st3[i + g * npx];
        s3 = make_float4(s3.x + t.x, s3.y + t.y, s3.z + t.z, 0.0f);
        cnt += __float_as_uint(t.w);
    }
    float fa = (float)cnt;
    V3 fin = v3(s3.x / fa, s3.y / fa, s3.z / fa);
    if (P.sampleBatch > 0u) {                         // raytrace.rgen.glsl:277-282
        float4 prev = img[i];
        float b = (float)P.sampleBatch;
        float b1 = (float)(P.sampleBatch + 1u);
        fin = v3((prev.x * b + fin.x) / b1, (prev.y * b + fin.y) / b1, (prev.z * b + fin.z) / b1);
    }
    img[i] = make_float4(fin.x, fin.y, fin.z, 1.0f);
}

__global__ void k_debug_math(int fn, const float* in, const float* in2, float* out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = in[i], r;
    float s, c;
    switch (fn) {
        case 0: sincos_(x, &s, &c); r = s; break;
        case 1: sincos_(x, &s, &c); r = c; break;
        case 2: r = logf_(x); break;
        case 3: r = expf_(x); break;
        case 4: r = sqrtf_(x); break;
        case 5: r = powf_(x, in2[i]); break;
        case 6: r = exp2f_(x); break;
        default: r = __builtin_nanf("");
    }
    out[i] = r;
}

}  // namespace rt

// ---- launch wrappers (host) ------------------------------------------------------
namespace rt {

__global__ __launch_bounds__(256) void k_shade_records(SceneView sv, float4* __restrict__ out, uint32_t n) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    float4 r[kShadeRecF4];
    shade_record(sv, g, r);
#pragma unroll
    for (uint32_t k = 0; k < kShadeRecF4; ++k) out[(size_t)kShadeRecF4 * g + k] = r[k];
}

hipError_t launch_shade_records(const SceneView& sv, float4* out, uint32_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_shade_records, dim3((n + 255) / 256), dim3(256), 0, st, sv, out, n);
    return hipGetLastError();
}

hipError_t launch_init(const TraceParams& P, const PathState& S, const Queues& Q, Ctrl* ctrl,
                       unsigned long long* stats, uint32_t base, uint32_t n, hipStream_t st) {
    uint32_t blocks = (n + 255) / 256;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_init, dim3(blocks), dim3(256), 0, st, P, S, Q, ctrl, stats, base, n);
    return hipGetLastError();
}

hipError_t launch_trace(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                        unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu, hipStream_t st) {
    return sv.qnodes ? launch_trace_q8(sv, rays, hits, in, stats, counters, count, tu, st)
                     : launch_trace_f32(sv, rays, hits, in, stats, counters, count, tu, st);
}

hipError_t launch_shade(const SceneView& sv, const TraceParams& P, const PathState& S, const Queues& Q,
                        uint32_t in_q, Ctrl* ctrl, uint32_t slot, uint32_t grid, bool nee, hipStream_t st) {
    // k_shade needs a whole number (>= 1) of blocks per partition
    grid = grid < kParts ? kParts : grid - grid % kParts;
    if (nee)
        hipLaunchKernelGGL(k_shade<true>, dim3(grid), dim3(256), 0, st, sv, P, S, Q, in_q, ctrl, slot);
    else
        hipLaunchKernelGGL(k_shade<false>, dim3(grid), dim3(256), 0, st, sv, P, S, Q, in_q, ctrl, slot);
    return hipGetLastError();
}

hipError_t launch_connect(const SceneView& sv, const Queues& Q, const PathState& S, Ctrl* ctrl, uint32_t slot,
                          unsigned long long* stats, const Tuning& tu, hipStream_t st) {
    return sv.qnodes ? launch_connect_q8(sv, Q, S, ctrl, slot, stats, tu, st)
                     : launch_connect_f32(sv, Q, S, ctrl, slot, stats, tu, st);
}

hipError_t launch_finish(const TraceParams& P, const PathState& S, float4* img, hipStream_t st) {
    uint32_t npx = P.W * P.H;
    uint32_t blocks = (npx + 255) / 256;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_finish, dim3(blocks), dim3(256), 0, st, P, S, img);
    return hipGetLastError();
}

hipError_t launch_debug_math(int fn, const float* in, const float* in2, float* out, uint32_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_debug_math, dim3((n + 255) / 256), dim3(256), 0, st, fn, in, in2, out, n);
    return hipGetLastError();
}

}  // namespace rt
