// trace_kernel.h — the persistent closest-hit kernel k_trace and its launch
// helpers, templated on the node format; trace_f32.hip and trace_q8.hip
// instantiate it for FP32 and 8-bit quantized BVH8 nodes.
#pragma once

#include <hip/hip_runtime.h>

#include "traverse.h"
#include "wavefront.h"

namespace rt {

__device__ __forceinline__ void part_range(const TraceInput& in, uint32_t p, uint32_t& start, uint32_t& size) {
    if (in.counts) {
        start = p * in.cap;
        size = in.counts[p * kLine];
    } else {
        uint32_t a = (uint32_t)(((uint64_t)in.n * p) / kParts);
        uint32_t b = (uint32_t)(((uint64_t)in.n * (p + 1)) / kParts);
        start = a;
        size = b - a;
    }
}

// Closest-hit kernel, persistent with per-lane refill: a lane whose ray is
// finished goes idle; once at least `refill_min` lanes of the wave are idle
// they take the next rays from the wave's current chunk (`chunk` rays claimed
// with one atomic add on its partition's counter, own XCD's partition first).
// A claim reads in.done, a mask of the used-up partitions on a line of its
// own, and skips them. On the partition it claimed from last it adds without
// reading the counter first: a read of a line that other waves' atomics hit
// costs several times the add. On the others it reads the counter first, which
// keeps failed adds (as costly) rare when rays run out. The claim that takes a
// partition's last rays sets its bit, so waves that start after every
// partition is used up stop after one read.
// Each loop iteration advances every active lane by one node visit.
// MULTI: the rays belong to several requests (possibly different scenes):
// pix[idx] >> 24 names the request whose scene the ray is traced in; sv is
// unused. Otherwise every ray is traced in sv.
// ANY: shadow rays (next-event estimation): any hit ends a ray; a ray that
// reaches its tmax unoccluded adds ao.scon[idx].rgb to the radiance of pixel
// ao.scon[idx].w; rays with a zero direction are empty entries. No hit
// records are written; stats[3] += traced shadow rays.
struct AnyHitOut {
    float4* st2;
    const float4* scon;
};

template <int SD, int WAVES, bool COUNT, bool Q, bool MULTI, bool ANY = false, bool H16 = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_trace(
    SceneView sv, const float4* __restrict__ rays, float4* __restrict__ hits, TraceInput in,
    unsigned long long* stats, uint32_t* counters, uint32_t refill_min, uint32_t chunk,
    const uint32_t* __restrict__ pix, const JobRec* __restrict__ jobs, AnyHitOut ao) {
    __shared__ uint2 stack[SD * 256];
    uint32_t nshadow = 0;
    if (!ANY && stats && blockIdx.x == 0 && threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t p = 0; p < kParts; ++p) {
            uint32_t a, b;
            part_range(in, p, a, b);
            tot += b;
        }
        atomicAdd(stats, (unsigned long long)tot);
    }
    uint2* stk = stack + threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint64_t ltmask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t part = blockIdx.x % kParts;    // wave-uniform: partition to claim from next
    bool active = false;
    bool exhausted = false;                 // wave-uniform
    uint32_t used_up = 0;                   // wave-uniform: partitions found used up
    uint32_t pool_next = 0, pool_end = 0;   // wave-uniform
    uint32_t idx = 0;
    TRay r;
    HitRec h;
    TravState ts;
    uint32_t nv = 0, nt = 0, wvis = 0, wtst = 0;
    const void* nodes = Q ? static_cast<const void*>(sv.qnodes) : static_cast<const void*>(sv.nodes);
    const float4* tris = sv.tris;
    uint64_t num_nodes = sv.num_nodes;
    const uint8_t* tflags = sv.tri_flags;
    for (;;) {
        uint64_t idle = __ballot(!active);
        uint32_t nidle = (uint32_t)__popcll(idle);
        if (!exhausted && (nidle >= refill_min || nidle == 64u)) {
            if (pool_next >= pool_end) {
                // claim a chunk: own partition first, then the others; a plain
                // read skips partitions that are already used up
                uint32_t got_start = 0xFFFFFFFFu, got_end = 0, got_part = part, got_used = used_up;
                if (lane == 0) {
                    if (in.done) got_used |= __hip_atomic_load(in.done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    for (uint32_t k = 0; k < kParts; ++k) {
                        const uint32_t p = (part + k) % kParts;
                        if (got_used & (1u << p)) continue;
                        uint32_t a, sz;
                        part_range(in, p, a, sz);
                        uint32_t* f = in.fetch + p * kLine;
                        if (sz == 0u ||
                            (k > 0u && __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= sz)) {
                            got_used |= 1u << p;
                            continue;
                        }
                        const uint32_t b = atomicAdd(f, chunk);
                        if (b < sz) {
                            got_start = a + b;
                            got_end = a + min(b + chunk, sz);
                            got_part = p;
                            // the claim that takes the partition's last rays marks it used up
                            if (in.done && sz - b <= chunk) atomicOr(in.done, 1u << p);
                            break;
                        }
                        got_used |= 1u << p;
                    }
                }
                // every lane is active here, so the first active lane is lane 0
                // and its values come back in scalar registers
                got_start = __builtin_amdgcn_readfirstlane(got_start);
                got_end = __builtin_amdgcn_readfirstlane(got_end);
                part = __builtin_amdgcn_readfirstlane(got_part);
                used_up = __builtin_amdgcn_readfirstlane(got_used);
                if (got_start == 0xFFFFFFFFu) {
                    exhausted = true;
                } else {
                    pool_next = got_start;
                    pool_end = got_end;
                }
            }
            if (!exhausted) {
                uint32_t take = min(nidle, pool_end - pool_next);
                if (!active) {
                    uint32_t rank = (uint32_t)__popcll(idle & ltmask);
                    if (rank < take) {
                        idx = pool_next + rank;
                        if constexpr (MULTI) {
                            const JobRec& J = jobs[pix[idx] >> kSlotBits];
                            nodes = Q ? static_cast<const void*>(J.sv.qnodes) : static_cast<const void*>(J.sv.nodes);
                            tris = J.sv.tris;
                            num_nodes = J.sv.num_nodes;
                            tflags = J.sv.tri_flags;
                        }
                        tray_init<H16>(r, rays[2 * idx], rays[2 * idx + 1]);
                        h = miss_record(r);
                        trav_begin(ts);
                        nv = 0;
                        nt = 0;
                        active = true;
                        if constexpr (ANY) {
                            if (!r.valid) {
                                active = false;
                            } else {
                                nshadow++;
                                if (num_nodes == 0) {
                                    const float4 c = ao.scon[idx];
                                    float4& L = ao.st2[__float_as_uint(c.w)];
                                    L = make_float4(L.x + c.x, L.y + c.y, L.z + c.z, L.w);
                                    active = false;
                                }
                            }
                        } else if (!r.valid || num_nodes == 0) {
                            hits[idx] = make_float4(h.t, __uint_as_float(h.prim), h.u, h.v);
                            if (COUNT && counters) { counters[2 * idx] = 0; counters[2 * idx + 1] = 0; }
                            active = false;
                        }
                    }
                }
                pool_next += take;
            }
        }
        if (__ballot(active) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (active) {
            if (!trav_step<COUNT, Q, ANY, H16>(nodes, tris, r, h, ts, stk, 256, nv, nt, tflags)) {
                if constexpr (ANY) {
                    if (h.prim == 0xFFFFFFFFu) {   // unoccluded
                        const float4 c = ao.scon[idx];
                        float4& L = ao.st2[__float_as_uint(c.w)];
                        L = make_float4(L.x + c.x, L.y + c.y, L.z + c.z, L.w);
                    }
                    active = false;
                    continue;
                }
                hits[idx] = make_float4(h.t, __uint_as_float(h.prim), h.u, h.v);
                if (COUNT) {
                    if (counters) { counters[2 * idx] = nv; counters[2 * idx + 1] = nt; }
                    wvis += nv;
                    wtst += nt;
                }
                active = false;
            }
        }
    }
    if (ANY && stats) {
        for (int o = 32; o > 0; o >>= 1) nshadow += __shfl_xor(nshadow, o);
        if (lane == 0 && nshadow) atomicAdd(stats + 3, (unsigned long long)nshadow);
    }
    if (COUNT && stats) {
        for (int o = 32; o > 0; o >>= 1) {
            wvis += __shfl_xor(wvis, o);
            wtst += __shfl_xor(wtst, o);
        }
        if (lane == 0) {
            atomicAdd(stats + 1, (unsigned long long)wvis);
            atomicAdd(stats + 2, (unsigned long long)wtst);
        }
    }
}

template <bool Q, bool MULTI, int SD, int WAVES>
static void launch_trace_sdw(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                             unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu,
                             const uint32_t* pix, const JobRec* jobs, hipStream_t st) {
    dim3 g(tu.grid), b(256);
    if constexpr (MULTI) {
        hipLaunchKernelGGL((k_trace<SD, WAVES, false, Q, true>), g, b, 0, st, sv, rays, hits, in, stats, counters,
                           tu.refill_min, tu.claim, pix, jobs, AnyHitOut{});
    } else {
        if (count)
            hipLaunchKernelGGL((k_trace<SD, WAVES, true, Q, false>), g, b, 0, st, sv, rays, hits, in, stats,
                               counters, tu.refill_min, tu.claim, pix, jobs, AnyHitOut{});
        else
            hipLaunchKernelGGL((k_trace<SD, WAVES, false, Q, false>), g, b, 0, st, sv, rays, hits, in, stats,
                               counters, tu.refill_min, tu.claim, pix, jobs, AnyHitOut{});
    }
}

template <bool Q, bool MULTI, int SD>
static void launch_trace_sd(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                            unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu,
                            const uint32_t* pix, const JobRec* jobs, hipStream_t st) {
    switch (tu.waves_per_simd) {
        case 4: launch_trace_sdw<Q, MULTI, SD, 4>(sv, rays, hits, in, stats, counters, count, tu, pix, jobs, st); break;
        case 5: launch_trace_sdw<Q, MULTI, SD, 5>(sv, rays, hits, in, stats, counters, count, tu, pix, jobs, st); break;
        case 6: launch_trace_sdw<Q, MULTI, SD, 6>(sv, rays, hits, in, stats, counters, count, tu, pix, jobs, st); break;
        case 8: launch_trace_sdw<Q, MULTI, SD, 8>(sv, rays, hits, in, stats, counters, count, tu, pix, jobs, st); break;
        default: launch_trace_sdw<Q, MULTI, SD, 1>(sv, rays, hits, in, stats, counters, count, tu, pix, jobs, st); break;
    }
}

template <bool Q, bool MULTI>
static hipError_t launch_trace_fmt(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                                   unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu,
                                   const uint32_t* pix, const JobRec* jobs, hipStream_t st) {
    switch (tu.stack_depth) {
        case 8: launch_trace_sd<Q, MULTI, 8>(sv, rays, hits, in, stats, counters, count, tu, pix, jobs, st); break;
        case 16: launch_trace_sd<Q, MULTI, 16>(sv, rays, hits, in, stats, counters, count, tu, pix, jobs, st); break;
        case 24: launch_trace_sd<Q, MULTI, 24>(sv, rays, hits, in, stats, counters, count, tu, pix, jobs, st); break;
        default: launch_trace_sd<Q, MULTI, 32>(sv, rays, hits, in, stats, counters, count, tu, pix, jobs, st); break;
    }
    return hipGetLastError();
}

template <bool Q, int SD>
static void launch_connect_sd(const SceneView& sv, const Queues& Q_, const PathState& S, Ctrl* ctrl, uint32_t slot,
                              unsigned long long* stats, const Tuning& tu, hipStream_t st) {
    TraceInput tin{&ctrl[slot].count[0][0], Q_.cap, 0, &ctrl[slot].sfetch[0][0], &ctrl[slot].sdone[0]};
    hipLaunchKernelGGL((k_trace<SD, 1, false, Q, false, true>), dim3(tu.grid), dim3(256), 0, st, sv, Q_.sray,
                       nullptr, tin, stats, nullptr, tu.refill_min, tu.claim, nullptr, nullptr, AnyHitOut{S.st2, Q_.scon});
}

template <bool Q>
static hipError_t launch_connect_fmt(const SceneView& sv, const Queues& Q_, const PathState& S, Ctrl* ctrl,
                                     uint32_t slot, unsigned long long* stats, const Tuning& tu, hipStream_t st) {
    switch (tu.stack_depth) {
        case 8: launch_connect_sd<Q, 8>(sv, Q_, S, ctrl, slot, stats, tu, st); break;
        case 16: launch_connect_sd<Q, 16>(sv, Q_, S, ctrl, slot, stats, tu, st); break;
        case 24: launch_connect_sd<Q, 24>(sv, Q_, S, ctrl, slot, stats, tu, st); break;
        default: launch_connect_sd<Q, 32>(sv, Q_, S, ctrl, slot, stats, tu, st); break;
    }
    return hipGetLastError();
}

// Shadow rays of the shared queues (server): entry i's ray is sray[i], its
// request is pix[i] >> kSlotBits.
template <bool Q, int SD>
static void launch_connect_multi_sd(const JobRec* jobs, const Queues& Q_, const PathState& S, Ctrl* ctrl,
                                    uint32_t slot, const uint32_t* pix, unsigned long long* stats, const Tuning& tu,
                                    hipStream_t st) {
    TraceInput tin{&ctrl[slot].count[0][0], Q_.cap, 0, &ctrl[slot].sfetch[0][0], &ctrl[slot].sdone[0]};
    hipLaunchKernelGGL((k_trace<SD, 1, false, Q, true, true>), dim3(tu.grid), dim3(256), 0, st, SceneView{}, Q_.sray,
                       nullptr, tin, stats, nullptr, tu.refill_min, tu.claim, pix, jobs, AnyHitOut{S.st2, Q_.scon});
}

template <bool Q>
static hipError_t launch_connect_multi_fmt(const JobRec* jobs, const Queues& Q_, const PathState& S, Ctrl* ctrl,
                                           uint32_t slot, const uint32_t* pix, unsigned long long* stats,
                                           const Tuning& tu, hipStream_t st) {
    switch (tu.stack_depth) {
        case 8: launch_connect_multi_sd<Q, 8>(jobs, Q_, S, ctrl, slot, pix, stats, tu, st); break;
        case 16: launch_connect_multi_sd<Q, 16>(jobs, Q_, S, ctrl, slot, pix, stats, tu, st); break;
        case 24: launch_connect_multi_sd<Q, 24>(jobs, Q_, S, ctrl, slot, pix, stats, tu, st); break;
        default: launch_connect_multi_sd<Q, 32>(jobs, Q_, S, ctrl, slot, pix, stats, tu, st); break;
    }
    return hipGetLastError();
}

}  // namespace rt
