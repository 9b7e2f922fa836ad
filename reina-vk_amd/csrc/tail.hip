// tail.hip — the end of a frame in one launch.
//
// A frame's bounce iterations shrink as pixels finish their samples, but a
// pixel's samples are sequential (one RNG state per pixel in the
// reference: shaders/raytrace/raytrace.rgen.glsl:259-264), so the last few
// thousand pixels with long paths need hundreds more iterations, each paying
// a launch and the latency of its slowest traversal. Once few paths remain,
// k_tail takes them over: each wave claims queue entries (one pixel each) and
// runs extend → shade steps for its 64 lanes, a lane taking the next entry
// when its pixel has finished all samples. Per pixel the arithmetic is
// exactly the wavefront kernels' (trav_step, shade_ray), so the image is
// bit-identical whatever order the lanes' steps interleave in.
//
// Without next-event estimation the wave does not wait for its slowest
// traversal: it keeps visiting nodes until `shade_min` lanes (or half the
// lanes holding a pixel, if fewer) have finished their traversal, shades
// those lanes only, and resumes the others where they stopped. A suspended
// lane keeps its ray (ra, rb), closest hit and stack position; its slab
// constants are recomputed from the ray when traversal resumes, so they are
// not live across the shading code. With next-event estimation every step
// runs in lockstep: extend, shade, then the shadow ray.
//
// The server's version (MULTI) does the same over the shared queues of many
// requests: each entry carries its request index, the lane reads that
// request's scene and parameters, and pixels that finish their last sample
// are counted per request as in k_shade_multi.
#include <hip/hip_runtime.h>

#include "shade_step.h"
#include "traverse.h"
#include "wavefront.h"

namespace rt {

// The queue entries for the lanes of a wave that want one (`want`): one
// atomic add for all of them on the partition the wave claimed from last (the
// counter is read first only when stealing from another partition: a read of
// a line that other waves' atomics hit costs several times the add). Every
// lane of the wave must be active. `part` and `used` (a mask of partitions
// found used up) are wave-uniform; `drained` is set once every partition is
// used up. Returns this lane's entry index, or 0xFFFFFFFF for none.
__device__ __forceinline__ uint32_t tail_claim(Ctrl* ctrl, uint32_t slot, uint32_t cap, bool want, uint32_t& part,
                                               uint32_t& used, bool& drained) {
    const uint64_t wm = __ballot(want);
    if (wm == 0ull || drained) return 0xFFFFFFFFu;
    const uint32_t nw = (uint32_t)__popcll(wm);
    uint32_t gs = 0xFFFFFFFFu, gn = 0, gp = part, gu = used;
    if (__lane_id() == 0) {
        for (uint32_t k = 0; k < kParts; ++k) {
            const uint32_t p = (part + k) % kParts;
            if (gu & (1u << p)) continue;
            const uint32_t n = ctrl[slot].count[p][0];
            uint32_t* f = &ctrl[slot].fetch[p][0];
            if (n == 0u || (k > 0u && __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= n)) {
                gu |= 1u << p;
                continue;
            }
            const uint32_t j = atomicAdd(f, nw);
            if (j < n) {
                gs = p * cap + j;
                gn = min(nw, n - j);
                gp = p;
                break;
            }
            gu |= 1u << p;
        }
    }
    gs = __builtin_amdgcn_readfirstlane(gs);
    gn = __builtin_amdgcn_readfirstlane(gn);
    part = __builtin_amdgcn_readfirstlane(gp);
    used = __builtin_amdgcn_readfirstlane(gu);
    if (gs == 0xFFFFFFFFu) {
        drained = true;
        return 0xFFFFFFFFu;
    }
    const uint64_t lt = (__lane_id() == 0) ? 0ull : (~0ull >> (64 - __lane_id()));
    const uint32_t rank = (uint32_t)__popcll(wm & lt);
    return (want && rank < gn) ? gs + rank : 0xFFFFFFFFu;
}

// Register budget: 3 waves per SIMD without next-event estimation (168
// VGPRs, a few bytes more scratch than the compiler's own 188 VGPRs at 2
// waves; the 1080p frame measured 2.6% faster); the NEE build keeps the
// compiler's choice (at 168 VGPRs it would spill ~200 bytes per lane).
template <int SD, bool Q, bool NEE, bool MULTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NEE ? 0 : 3))) void k_tail(SceneView sv, TraceParams P, PathState S, Queues Q_, uint32_t in_q,
                                              Ctrl* ctrl, uint32_t slot, unsigned long long* stats,
                                              const JobRec* __restrict__ jobs, uint32_t* done, uint32_t shade_min) {
    __shared__ uint2 stack[SD * 256];
    uint2* stk = stack + threadIdx.x;
    const float4* rin = Q_.ray[in_q];
    const uint32_t* pin = Q_.pix[in_q];
    uint32_t part = blockIdx.x % kParts;   // wave-uniform
    uint32_t used = 0;                      // wave-uniform: partitions found used up
    bool have = false, drained = false;     // drained: wave-uniform
    uint32_t pw = 0, pp = 0, px = 0;
    const JobRec* J = nullptr;
    const void* nodes = Q ? static_cast<const void*>(sv.qnodes) : static_cast<const void*>(sv.nodes);
    const float4* tris = sv.tris;
    uint64_t num_nodes = sv.num_nodes;
    float4 ra = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rb = ra;
    uint32_t nrays = 0, nshadow = 0, dummy_v = 0, dummy_t = 0;
    if constexpr (!NEE) {
        // Per lane: `have` a pixel; `fresh`: its next ray (ra, rb) is not yet
        // started; `tracing`: traversal in progress (h, ts, LDS stack);
        // `ready`: traversal done, h holds the closest hit.
        bool fresh = false, tracing = false, ready = false;
        HitRec h;
        TravState ts;
        h.t = 0.0f; h.prim = 0u; h.u = 0.0f; h.v = 0.0f;
        trav_begin(ts);
        for (;;) {
            const uint32_t i = tail_claim(ctrl, slot, Q_.cap, !have, part, used, drained);
            if (i != 0xFFFFFFFFu) {
                pw = pin[i];
                ra = rin[2 * i];
                rb = rin[2 * i + 1];
                have = true;
                fresh = true;
                if (MULTI) {
                    J = jobs + (pw >> kSlotBits);
                    pp = pw & ((1u << kSlotBits) - 1u);
                    px = pp - J->px_base;
                    nodes = Q ? static_cast<const void*>(J->sv.qnodes) : static_cast<const void*>(J->sv.nodes);
                    tris = J->sv.tris;
                    num_nodes = J->sv.num_nodes;
                } else {
                    pp = px = pw;
                }
            }
            const uint64_t holding = __ballot(have);
            if (holding == 0ull) break;
            // extend: fresh rays start, suspended ones resume, until enough lanes are ready
            {
                TRay r;
                if (fresh || tracing) tray_init(r, ra, rb);
                if (fresh) {
                    h = miss_record(r);
                    trav_begin(ts);
                    tracing = r.valid && num_nodes != 0;
                    ready = !tracing;
                    fresh = false;
                    nrays++;
                }
                const uint32_t half = ((uint32_t)__popcll(holding) + 1u) >> 1;
                uint32_t need = shade_min < half ? shade_min : half;
                if (need == 0u) need = 1u;   // every round shades at least one lane
                for (;;) {
                    if (__ballot(tracing) == 0ull) break;
                    if ((uint32_t)__popcll(__ballot(ready)) >= need) break;
                    if (tracing) {
                        tracing = trav_step<false, Q>(nodes, tris, r, h, ts, stk, 256, dummy_v, dummy_t);
                        ready = !tracing;
                    }
                }
            }
            // shade the lanes whose traversal finished
            ShadeResult res;
            res.push = false;
            res.finished = false;
            const bool shading = ready;
            if (shading) {
                const float4 hh = make_float4(h.t, __uint_as_float(h.prim), h.u, h.v);
                res = MULTI ? shade_ray<false>(J->sv, J->P, S, pp, px, ra, rb, hh)
                            : shade_ray<false>(sv, P, S, pp, px, ra, rb, hh);
                ready = false;
            }
            if (MULTI) {
                const bool fin = shading && res.finished;
                uint64_t fm = __ballot(fin ? 1 : 0);
                while (fm) {
                    const uint32_t lead = (uint32_t)__builtin_ctzll(fm);
                    const uint32_t jl = __shfl(pw >> kSlotBits, (int)lead);
                    const uint64_t same = __ballot((fin && (pw >> kSlotBits) == jl) ? 1 : 0);
                    if (__lane_id() == lead) atomicAdd(done + jl, (uint32_t)__popcll(same));
                    fm &= ~same;
                }
            }
            if (shading) {
                if (res.push) {
                    ra = make_float4(res.no.x, res.no.y, res.no.z, 0.0f);
                    rb = make_float4(res.nd.x, res.nd.y, res.nd.z, 10000.0f);
                    fresh = true;
                } else {
                    have = false;
                }
            }
        }
    } else for (;;) {
        // lanes without a pixel take the next queue entries (own XCD's partition first)
        const uint32_t i = tail_claim(ctrl, slot, Q_.cap, !have, part, used, drained);
        if (i != 0xFFFFFFFFu) {
            pw = pin[i];
            ra = rin[2 * i];
            rb = rin[2 * i + 1];
            have = true;
            if (MULTI) {
                J = jobs + (pw >> kSlotBits);
                pp = pw & ((1u << kSlotBits) - 1u);
                px = pp - J->px_base;
                nodes = Q ? static_cast<const void*>(J->sv.qnodes) : static_cast<const void*>(J->sv.nodes);
                tris = J->sv.tris;
                num_nodes = J->sv.num_nodes;
            } else {
                pp = px = pw;
            }
        }
        if (__ballot(have) == 0ull) break;
        // extend
        TRay r;
        HitRec h;
        TravState ts;
        bool tracing = false;
        if (have) {
            tray_init(r, ra, rb);
            h = miss_record(r);
            trav_begin(ts);
            tracing = r.valid && num_nodes != 0;
            nrays++;
        }
        while (__ballot(tracing)) {
            if (tracing) tracing = trav_step<false, Q>(nodes, tris, r, h, ts, stk, 256, dummy_v, dummy_t);
        }
        // shade
        ShadeResult res;
        res.push = false;
        res.shadow = false;
        res.finished = false;
        if (have) {
            const float4 hh = make_float4(h.t, __uint_as_float(h.prim), h.u, h.v);
            res = MULTI ? shade_ray<NEE>(J->sv, J->P, S, pp, px, ra, rb, hh)
                        : shade_ray<NEE>(sv, P, S, pp, px, ra, rb, hh);
        }
        // shadow ray (next-event estimation)
        if (NEE) {
            bool sh = have && res.shadow;
            HitRec hs;
            if (sh) {
                tray_init(r, make_float4(res.so.x, res.so.y, res.so.z, 0.0f),
                          make_float4(res.sd.x, res.sd.y, res.sd.z, res.stmax));
                hs = miss_record(r);
                trav_begin(ts);
                nshadow++;
                if (!r.valid || num_nodes == 0) sh = false;
            }
            const uint8_t* tflags = MULTI ? (J ? J->sv.tri_flags : nullptr) : sv.tri_flags;
            while (__ballot(sh)) {
                if (sh) sh = trav_step<false, Q, true>(nodes, tris, r, hs, ts, stk, 256, dummy_v, dummy_t, tflags);
            }
            if (have && res.shadow && hs.prim == 0xFFFFFFFFu) {
                float4& L = S.st2[pp];
                L = make_float4(L.x + res.scontrib.x, L.y + res.scontrib.y, L.z + res.scontrib.z, L.w);
            }
        }
        if (MULTI) {
            // pixels that finished their last sample: one atomic per request among the wave's finishers
            const bool fin = have && res.finished;
            uint64_t fm = __ballot(fin ? 1 : 0);
            while (fm) {
                const uint32_t lead = (uint32_t)__builtin_ctzll(fm);
                const uint32_t jl = __shfl(pw >> kSlotBits, (int)lead);
                const uint64_t same = __ballot((fin && (pw >> kSlotBits) == jl) ? 1 : 0);
                if (__lane_id() == lead) atomicAdd(done + jl, (uint32_t)__popcll(same));
                fm &= ~same;
            }
        }
        if (have) {
            if (res.push) {
                ra = make_float4(res.no.x, res.no.y, res.no.z, 0.0f);
                rb = make_float4(res.nd.x, res.nd.y, res.nd.z, 10000.0f);
            } else {
                have = false;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        nrays += __shfl_xor(nrays, o);
        nshadow += __shfl_xor(nshadow, o);
    }
    if (__lane_id() == 0) {
        if (nrays) { atomicAdd(stats, (unsigned long long)nrays); atomicAdd(stats + 4, (unsigned long long)nrays); }
        if (nshadow) atomicAdd(stats + 3, (unsigned long long)nshadow);
    }
}

template <bool Q, bool NEE, bool MULTI>
static void launch_tail_qn(const SceneView& sv, const TraceParams& P, const PathState& S, const Queues& Q_,
                           uint32_t in_q, Ctrl* ctrl, uint32_t slot, unsigned long long* stats, const JobRec* jobs,
                           uint32_t* done, const Tuning& tu, uint32_t blocks, hipStream_t st) {
    switch (tu.stack_depth) {
        case 8: hipLaunchKernelGGL((k_tail<8, Q, NEE, MULTI>), dim3(blocks), dim3(256), 0, st, sv, P, S, Q_, in_q, ctrl, slot, stats, jobs, done, tu.refill_min); break;
        case 16: hipLaunchKernelGGL((k_tail<16, Q, NEE, MULTI>), dim3(blocks), dim3(256), 0, st, sv, P, S, Q_, in_q, ctrl, slot, stats, jobs, done, tu.refill_min); break;
        case 24: hipLaunchKernelGGL((k_tail<24, Q, NEE, MULTI>), dim3(blocks), dim3(256), 0, st, sv, P, S, Q_, in_q, ctrl, slot, stats, jobs, done, tu.refill_min); break;
        default: hipLaunchKernelGGL((k_tail<32, Q, NEE, MULTI>), dim3(blocks), dim3(256), 0, st, sv, P, S, Q_, in_q, ctrl, slot, stats, jobs, done, tu.refill_min); break;
    }
}

// one lane per live path, at least one wave per partition
static uint32_t tail_blocks(uint32_t live) {
    const uint32_t blocks = (live + 255u) / 256u;
    return blocks < kParts ? kParts : blocks;
}

hipError_t launch_tail(const SceneView& sv, const TraceParams& P, const PathState& S, const Queues& Q,
                       uint32_t in_q, Ctrl* ctrl, uint32_t slot, unsigned long long* stats, const Tuning& tu,
                       uint32_t live, hipStream_t st) {
    const uint32_t blocks = tail_blocks(live);
    if (sv.qnodes != nullptr) {
        if (P.nee) launch_tail_qn<true, true, false>(sv, P, S, Q, in_q, ctrl, slot, stats, nullptr, nullptr, tu, blocks, st);
        else launch_tail_qn<true, false, false>(sv, P, S, Q, in_q, ctrl, slot, stats, nullptr, nullptr, tu, blocks, st);
    } else {
        if (P.nee) launch_tail_qn<false, true, false>(sv, P, S, Q, in_q, ctrl, slot, stats, nullptr, nullptr, tu, blocks, st);
        else launch_tail_qn<false, false, false>(sv, P, S, Q, in_q, ctrl, slot, stats, nullptr, nullptr, tu, blocks, st);
    }
    return hipGetLastError();
}

hipError_t launch_tail_multi(bool quantized, bool nee, const JobRec* jobs, const PathState& S, const Queues& Q,
                             uint32_t in_q, Ctrl* ctrl, uint32_t slot, unsigned long long* stats, uint32_t* done,
                             const Tuning& tu, uint32_t live, hipStream_t st, uint32_t max_blocks) {
    uint32_t blocks = tail_blocks(live);
    if (max_blocks != 0 && blocks > max_blocks) blocks = max_blocks < kParts ? kParts : max_blocks;
    const SceneView sv{};
    const TraceParams P{};
    if (quantized) {
        if (nee) launch_tail_qn<true, true, true>(sv, P, S, Q, in_q, ctrl, slot, stats, jobs, done, tu, blocks, st);
        else launch_tail_qn<true, false, true>(sv, P, S, Q, in_q, ctrl, slot, stats, jobs, done, tu, blocks, st);
    } else {
        if (nee) launch_tail_qn<false, true, true>(sv, P, S, Q, in_q, ctrl, slot, stats, jobs, done, tu, blocks, st);
        else launch_tail_qn<false, false, true>(sv, P, S, Q, in_q, ctrl, slot, stats, jobs, done, tu, blocks, st);
    }
    return hipGetLastError();
}

}  // namespace rt
