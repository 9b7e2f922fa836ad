// server.hip — kernels of the multi-request wavefront: several render
// requests (each one vkCmdTraceRaysKHR of the reference, src/Reina.cpp:460-469,
// on its own scene, camera, resolution and sample budget) share the path-state
// slots, the partitioned ray queues and the bounce iterations.
//
//   k_init_job    raygen of sample 0 for every pixel of one newly admitted
//                 request; appends its rays to the current queue
//   k_trace       (trace_kernel.h, MULTI) closest hit in each ray's own scene
//   k_shade_multi shading + bounce logic with each ray's own scene and
//                 parameters; counts the pixels that finished all samples
//   k_finish_job  mean over the samples and running-mean blend into the
//                 request's image
//
// A pixel's arithmetic is exactly the single-scene kernels' (shade_ray is
// shared and the RNG stream is the pixel's own), so a request renders the
// same image alone, in a mix, or through reina_rt_trace.
#include <hip/hip_runtime.h>

#include "shade_step.h"
#include "wavefront.h"

namespace rt {

// Queue position base of partition p: p * cap. Request j's slots
// [px_base, px_base + npx) are tile aligned and a tile holds at least 256
// slots, so the 256 pixels of a block share one tile and one partition
// ((slot >> tile_log2) & 7): one atomic per block.
__global__ __launch_bounds__(256) void k_init_job(const JobRec* jobs, uint32_t j, PathState S, Queues Q,
                                                  Ctrl* ctrl, uint32_t cslot, uint32_t in_q, uint32_t* done,
                                                  uint32_t tile_log2) {
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_base;
    const JobRec& J = jobs[j];
    const TraceParams& P = J.P;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool anyrays = P.spp > 0 && P.maxBounces > 0;
    const bool valid = k < J.npx;
    const uint32_t slot = J.px_base + k;
    V3 o = v3(0.0f, 0.0f, 0.0f), d = o;
    if (valid) {
        if (P.nee) S.st4[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        // slot k: sample group g = k / (W * H) of pixel k % (W * H)
        const uint32_t npix = P.W * P.H, g = k / npix, pix = k - g * npix;
        uint32_t x = pix % P.W, y = pix / P.W;
        // raytrace.rgen.glsl:259; group g of G takes the seed of batch sampleBatch * G + g
        uint32_t rng = ((P.sampleBatch * P.groups + g) * P.H + y) * P.W + x;
        if (!anyrays) {
            S.st0[slot] = make_uint4(rng, P.spp, 0u, 0u);
            S.st3[slot] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(P.spp));
        } else {
            starting_ray(P.cam, (float)x, (float)y, rng, o, d, (P.fixes & REINA_RT_TRACE_FIX_LENS_RNG) != 0u);
            S.st0[slot] = make_uint4(rng, 0u, 0u, 0u);
            S.st1[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            S.st2[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            S.st3[slot] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
        }
    }
    const bool push = valid && anyrays;
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    const uint64_t mask = __ballot(push ? 1 : 0);
    const uint32_t before = (uint32_t)__popcll(mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    const uint32_t p = ((J.px_base + blockIdx.x * blockDim.x) >> tile_log2) & (kParts - 1u);
    if (threadIdx.x == 0) {
        uint32_t tot = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        s_base = tot ? atomicAdd(&ctrl[cslot].count[p][0], tot) : 0u;
        if (!anyrays) {
            uint32_t nv = min(256u, J.npx - blockIdx.x * blockDim.x);
            atomicAdd(done + j, nv);
        }
    }
    __syncthreads();
    if (push) {
        uint32_t off = s_base + before;
        for (uint32_t w = 0; w < wave; ++w) off += s_wave[w];
        uint32_t q = p * Q.cap + off;
        Q.ray[in_q][2 * q] = make_float4(o.x, o.y, o.z, 0.0f);
        Q.ray[in_q][2 * q + 1] = make_float4(d.x, d.y, d.z, 10000.0f);
        Q.pix[in_q][q] = slot | (j << kSlotBits);
    }
}

// k_shade of integrator.hip with per-ray request lookup; the pixels that
// finish their last sample are counted per request (one atomic per request
// present in the wave). NEE: the server holds shadow queues; requests with
// TraceParams.nee write their shadow rays, every other entry an empty one.
template <bool NEE>
__global__ __launch_bounds__(256) void k_shade_multi(const JobRec* jobs, PathState S, Queues Q, uint32_t in_q,
                                                     Ctrl* ctrl, uint32_t slot, uint32_t* done) {
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
    const uint32_t nblk = gridDim.x / kParts;
    const uint32_t bi = blockIdx.x / kParts;
    const uint32_t n = ctrl[slot].count[p][0];
    const uint32_t base_in = p * Q.cap;
    const float4* rin = Q.ray[in_q];
    const uint32_t* pin = Q.pix[in_q];
    float4* rout = Q.ray[in_q ^ 1];
    uint32_t* pout = Q.pix[in_q ^ 1];
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    const uint64_t ltmask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (uint32_t t0 = bi * 256u; t0 < n; t0 += nblk * 256u) {
        const uint32_t k = t0 + threadIdx.x;
        bool push = false, fin = false;
        uint32_t pw = 0;
        V3 no = v3(0.0f, 0.0f, 0.0f), nd = no;
        if (k < n) {
            const uint32_t i = base_in + k;
            pw = pin[i];
            const JobRec& J = jobs[pw >> kSlotBits];
            const uint32_t sl = pw & ((1u << kSlotBits) - 1u);
            const ShadeResult res = shade_ray<NEE>(J.sv, J.P, S, sl, sl - J.px_base, rin[2 * i], rin[2 * i + 1], Q.hit[i]);
            push = res.push;
            fin = res.finished;
            no = res.no;
            nd = res.nd;
            if (NEE) {
                if (res.shadow) {
                    Q.sray[2 * i] = make_float4(res.so.x, res.so.y, res.so.z, 0.0f);
                    Q.sray[2 * i + 1] = make_float4(res.sd.x, res.sd.y, res.sd.z, res.stmax);
                    Q.scon[i] = make_float4(res.scontrib.x, res.scontrib.y, res.scontrib.z, __uint_as_float(sl));
                } else {
                    Q.sray[2 * i + 1] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);   // no shadow ray
                }
            }
        }
        // finished pixels: one atomic per request present among the wave's finishers
        uint64_t fm = __ballot(fin ? 1 : 0);
        while (fm) {
            const uint32_t lead = (uint32_t)__builtin_ctzll(fm);
            const uint32_t jl = __shfl(pw >> kSlotBits, (int)lead);
            const uint64_t same = __ballot((fin && (pw >> kSlotBits) == jl) ? 1 : 0);
            if (lane == lead) atomicAdd(done + jl, (uint32_t)__popcll(same));
            fm &= ~same;
        }
        uint64_t mask = __ballot(push ? 1 : 0);
        uint32_t before = (uint32_t)__popcll(mask & ltmask);
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
            uint32_t q = base_in + off;
            rout[2 * q] = make_float4(no.x, no.y, no.z, 0.0f);
            rout[2 * q + 1] = make_float4(nd.x, nd.y, nd.z, 10000.0f);
            pout[q] = pw;
        }
        __syncthreads();
    }
}

// k_finish of integrator.hip for request j.
__global__ __launch_bounds__(256) void k_finish_job(const JobRec* jobs, uint32_t j, PathState S) {
    const JobRec& J = jobs[j];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t npix = J.P.W * J.P.H;   // J.npx = npix * groups slots
    if (k >= npix) return;
    float4 s3 = S.st3[J.px_base + k];
    uint32_t cnt = __float_as_uint(s3.w);
    for (uint32_t g = 1; g < J.P.groups; ++g) {   // sample groups: sums in group order
        const float4 t = S.st3[J.px_base + k + g * npix];
        s3 = make_float4(s3.x + t.x, s3.y + t.y, s3.z + t.z, 0.0f);
        cnt += __float_as_uint(t.w);
    }
    const float fa = (float)cnt;
    V3 fin = v3(s3.x / fa, s3.y / fa, s3.z / fa);
    if (J.P.sampleBatch > 0u) {                        // raytrace.rgen.glsl:277-282
        const float4 prev = J.img[k];
        const float b = (float)J.P.sampleBatch;
        const float b1 = (float)(J.P.sampleBatch + 1u);
        fin = v3((prev.x * b + fin.x) / b1, (prev.y * b + fin.y) / b1, (prev.z * b + fin.z) / b1);
    }
    J.img[k] = make_float4(fin.x, fin.y, fin.z, 1.0f);
}

hipError_t launch_init_job(const JobRec* jobs, uint32_t j, uint32_t npx, const PathState& S, const Queues& Q,
                           Ctrl* ctrl, uint32_t cslot, uint32_t in_q, uint32_t* done, uint32_t tile_log2,
                           hipStream_t st) {
    if (npx == 0) return hipSuccess;
    if (tile_log2 < 8 || tile_log2 > 12) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_init_job, dim3((npx + 255) / 256), dim3(256), 0, st, jobs, j, S, Q, ctrl, cslot, in_q, done,
                       tile_log2);
    return hipGetLastError();
}

hipError_t launch_shade_multi(const JobRec* jobs, const PathState& S, const Queues& Q, uint32_t in_q, Ctrl* ctrl,
                              uint32_t slot, uint32_t* done, uint32_t grid, bool nee, hipStream_t st) {
    grid = grid < kParts ? kParts : grid - grid % kParts;
    if (nee) hipLaunchKernelGGL(k_shade_multi<true>, dim3(grid), dim3(256), 0, st, jobs, S, Q, in_q, ctrl, slot, done);
    else hipLaunchKernelGGL(k_shade_multi<false>, dim3(grid), dim3(256), 0, st, jobs, S, Q, in_q, ctrl, slot, done);
    return hipGetLastError();
}

hipError_t launch_connect_multi(bool quantized, const JobRec* jobs, const Queues& Q, const PathState& S, Ctrl* ctrl,
                                uint32_t slot, const uint32_t* pix, unsigned long long* stats, const Tuning& tu,
                                hipStream_t st) {
    return quantized ? launch_connect_multi_q8(jobs, Q, S, ctrl, slot, pix, stats, tu, st)
                     : launch_connect_multi_f32(jobs, Q, S, ctrl, slot, pix, stats, tu, st);
}

hipError_t launch_finish_job(const JobRec* jobs, uint32_t j, uint32_t npx, const PathState& S, hipStream_t st) {
    if (npx == 0) return hipSuccess;
    hipLaunchKernelGGL(k_finish_job, dim3((npx + 255) / 256), dim3(256), 0, st, jobs, j, S);
    return hipGetLastError();
}

// Moves the live rays of queue 0 (counts in ctrl[0]) into the tail queue TQ
// (same partition, same position), sets tctrl's counts and clears its claim
// counters. Partition p = blockIdx.y.
__global__ __launch_bounds__(256) void k_tail_gather(Queues Q, const Ctrl* ctrl, Queues TQ, Ctrl* tctrl) {
    const uint32_t p = blockIdx.y;
    const uint32_t n = ctrl[0].count[p][0];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t i = p * Q.cap + j, o = p * TQ.cap + j;
        TQ.ray[0][2 * o] = Q.ray[0][2 * i];
        TQ.ray[0][2 * o + 1] = Q.ray[0][2 * i + 1];
        TQ.pix[0][o] = Q.pix[0][i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        tctrl->count[p][0] = n;
        tctrl->fetch[p][0] = 0;
        tctrl->sfetch[p][0] = 0;
    }
}

hipError_t launch_tail_gather(const Queues& Q, const Ctrl* ctrl, const Queues& TQ, Ctrl* tctrl, uint32_t max_count,
                              hipStream_t st) {
    if (max_count > TQ.cap) return hipErrorInvalidValue;
    uint32_t gx = (max_count + 255u) / 256u;
    gx = gx < 1u ? 1u : (gx > 512u ? 512u : gx);
    hipLaunchKernelGGL(k_tail_gather, dim3(gx, kParts), dim3(256), 0, st, Q, ctrl, TQ, tctrl);
    return hipGetLastError();
}

hipError_t launch_trace_multi(bool quantized, const float4* rays, float4* hits, const TraceInput& in,
                              unsigned long long* stats, const Tuning& tu, const uint32_t* pix, const JobRec* jobs,
                              hipStream_t st) {
    return quantized ? launch_trace_multi_q8(rays, hits, in, stats, tu, pix, jobs, st)
                     : launch_trace_multi_f32(rays, hits, in, stats, tu, pix, jobs, st);
}

}  // namespace rt
