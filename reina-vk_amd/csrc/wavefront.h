// wavefront.h — per-trace buffers of the wavefront integrator and the launch
// wrappers shared by the host driver (api.cpp) and the kernels (integrator.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "rt_layout.h"
#include "shade.h"

namespace rt {

struct TraceParams {
    Camera cam;
    uint32_t W, H;
    uint32_t spp, maxBounces;
    uint32_t sampleBatch;
    float directClamp;
    uint32_t nee;   // next-event estimation on (not a reference push constant)
    uint32_t fixes; // REINA_RT_TRACE_FIX_* corrections of reference behaviours (0: as the reference)
    uint32_t groups;  // sample groups per pixel (1: as the reference); spp above is per group
};

// The ray queues are split into kParts partitions, one per XCD: workgroup b
// runs on XCD b % 8 (workgroups are dispatched to the 8 XCDs round-robin),
// and each XCD's blocks work on their own partition, so the queue counters
// that are hit by atomics are spread over 8 addresses (device-scope atomics
// are resolved beyond the per-XCD L2 and serialise per address) and a
// partition's rays stay close to the XCD that wrote them.
constexpr uint32_t kParts = 8;
constexpr uint32_t kLine = 32;   // uint32 words per 128-byte line

// Control words of one bounce iteration: per partition the queue length and
// the extend kernel's claim counter, each on its own 128-byte line, and the
// mask of the partitions whose rays are all claimed. Three slots rotate so
// that iteration i's shade zeroes the slot of iteration i+2.
struct Ctrl {
    uint32_t count[kParts][kLine];
    uint32_t fetch[kParts][kLine];
    uint32_t sfetch[kParts][kLine];   // claim counters of the shadow-ray kernel
    // word 0: bit p is set by the claim that takes the last rays of the
    // extend kernel's partition p (sdone: of the shadow-ray kernel's)
    uint32_t done[kLine];
    uint32_t sdone[kLine];
};

// Pixel i starts in partition ((i >> 10) & 7) at position
// ((i >> 13) << 10) | (i & 1023); `cap` entries per partition.
__host__ __device__ inline uint32_t part_capacity(uint32_t npx) {
    uint32_t tiles = (npx + 1023u) >> 10;
    return ((tiles + kParts - 1u) / kParts) << 10;
}

// Path state per pixel (SoA, 64 bytes per pixel, 80 with next-event estimation).
struct PathState {
    uint4* st0;     // rng, sample index, segment index, insideDielectric
    float4* st1;    // throughput.xyz, accumulatedDistance
    float4* st2;    // radiance L.xyz, -
    float4* st3;    // sum.xyz, actualSamples (uint bits)
    float4* st4;    // NEE: origin of the last Lambertian continuation, its cosine pdf (0: no MIS)
};

// Double-buffered ray queues (ReinaRay layout) with the owning pixel, and the
// hit records of the current queue; kParts partitions of `cap` entries each.
struct Queues {
    float4* ray[2];
    uint32_t* pix[2];
    float4* hit;
    float4* sray;   // NEE: shadow ray of queue entry i (ReinaRay layout; zero direction = none)
    float4* scon;   // NEE: its unoccluded contribution rgb, pixel (uint bits)
    uint32_t cap;
};

// ---- several render requests in one wavefront ----------------------------
// A request owns a contiguous run of path-state slots starting at px_base (a
// multiple of 1024, so each 1024-slot tile stays in one queue partition).
// Queue entries carry slot | request << kSlotBits.
constexpr uint32_t kSlotBits = 24;
constexpr uint32_t kMaxJobs = 256;

struct JobRec {
    SceneView sv;
    TraceParams P;
    float4 RT_GLOBAL* img;
    uint32_t px_base;   // first path-state slot
    uint32_t npx;       // W * H
};

// Fills the shading records (SceneView::srec layout) of flattened triangles 0..n-1.
hipError_t launch_shade_records(const SceneView& sv, float4* out, uint32_t n, hipStream_t st);
// Pixels [base, base + n) start in queue Q (partition (l >> 10) & 7 for
// local index l = pixel - base).
hipError_t launch_init(const TraceParams& P, const PathState& S, const Queues& Q, Ctrl* ctrl,
                       unsigned long long* stats, uint32_t base, uint32_t n, hipStream_t st);
// Launch parameters of the traversal kernel (chosen per scene; see the autotuner).
struct Tuning {
    uint32_t stack_depth;   // LDS group-stack entries per lane: 8, 16, 24 or 32 (>= BVH depth - 1)
    uint32_t refill_min;    // idle lanes of a wave that trigger a refill (1..64)
    uint32_t grid;          // persistent blocks of 256 threads
    uint32_t shade_grid;    // blocks of the shade kernel
    uint32_t waves_per_simd;  // register budget of the traversal kernel: 0 (compiler's choice), 4, 5, 6 or 8
    uint32_t tail_paths;      // live paths at which the frame ends in one tail launch (0: never)
    uint32_t claim = 128;     // rays a traversal wave claims with one atomic: 64, 128, 256 or 512
};

// Rays to trace: kParts partitions. Partitioned mode (counts != null):
// partition p holds counts[p * kLine] rays starting at p * cap. Flat mode:
// n rays split into kParts contiguous ranges. fetch: kParts claim counters
// (stride kLine), zero before the launch; done (optional): a word on a line of
// its own, zero before the launch, whose bit p the claim of partition p's last
// rays sets.
struct TraceInput {
    const uint32_t* counts;
    uint32_t cap;
    uint32_t n;
    uint32_t* fetch;
    uint32_t* done;
};

// Closest hit for every ray of `in`. stats (optional): stats[0] += rays;
// when counting also stats[1] += node visits, stats[2] += triangle tests.
// counters (optional, counting): per-ray {visits, tests}.
hipError_t launch_trace(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                        unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu, hipStream_t st);
// Closest hit of partitioned queue entries that belong to several requests
// (pix: the queue's slot | request words).
hipError_t launch_trace_multi(bool quantized, const float4* rays, float4* hits, const TraceInput& in,
                              unsigned long long* stats, const Tuning& tu, const uint32_t* pix, const JobRec* jobs,
                              hipStream_t st);
hipError_t launch_trace_multi_f32(const float4* rays, float4* hits, const TraceInput& in, unsigned long long* stats,
                                  const Tuning& tu, const uint32_t* pix, const JobRec* jobs, hipStream_t st);
hipError_t launch_trace_multi_q8(const float4* rays, float4* hits, const TraceInput& in, unsigned long long* stats,
                                 const Tuning& tu, const uint32_t* pix, const JobRec* jobs, hipStream_t st);
// tile_log2: log2 of the server's slot tile (8..12); slot l starts in
// partition (l >> tile_log2) & 7.
hipError_t launch_init_job(const JobRec* jobs, uint32_t j, uint32_t npx, const PathState& S, const Queues& Q,
                           Ctrl* ctrl, uint32_t cslot, uint32_t in_q, uint32_t* done, uint32_t tile_log2,
                           hipStream_t st);
// nee: the server holds next-event-estimation state (S.st4, Q.sray, Q.scon);
// each request's TraceParams.nee then selects NEE for its rays, and every
// queue entry gets a shadow ray or an empty one.
hipError_t launch_shade_multi(const JobRec* jobs, const PathState& S, const Queues& Q, uint32_t in_q, Ctrl* ctrl,
                              uint32_t slot, uint32_t* done, uint32_t grid, bool nee, hipStream_t st);
// Shadow rays of the shared queues (pix: the iteration's slot | request words).
hipError_t launch_connect_multi(bool quantized, const JobRec* jobs, const Queues& Q, const PathState& S, Ctrl* ctrl,
                                uint32_t slot, const uint32_t* pix, unsigned long long* stats, const Tuning& tu,
                                hipStream_t st);
hipError_t launch_connect_multi_f32(const JobRec* jobs, const Queues& Q, const PathState& S, Ctrl* ctrl, uint32_t slot,
                                    const uint32_t* pix, unsigned long long* stats, const Tuning& tu, hipStream_t st);
hipError_t launch_connect_multi_q8(const JobRec* jobs, const Queues& Q, const PathState& S, Ctrl* ctrl, uint32_t slot,
                                   const uint32_t* pix, unsigned long long* stats, const Tuning& tu, hipStream_t st);
hipError_t launch_finish_job(const JobRec* jobs, uint32_t j, uint32_t npx, const PathState& S, hipStream_t st);
// Server tail launch on its own stream: the live rays of queue 0 move to the
// tail queue TQ (cap entries per partition; max_count = the largest partition
// count, at most TQ.cap) with their counts in tctrl.
hipError_t launch_tail_gather(const Queues& Q, const Ctrl* ctrl, const Queues& TQ, Ctrl* tctrl, uint32_t max_count,
                              hipStream_t st);
hipError_t launch_trace_f32_h16(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                                unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu,
                                hipStream_t st);
hipError_t launch_trace_f32(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                            unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu,
                            hipStream_t st);
hipError_t launch_trace_q8(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                           unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu,
                           hipStream_t st);
hipError_t launch_shade(const SceneView& sv, const TraceParams& P, const PathState& S, const Queues& Q,
                        uint32_t in_q, Ctrl* ctrl, uint32_t slot, uint32_t grid, bool nee, hipStream_t st);
// Shadow rays of the iteration's queue (any hit; back faces of culling
// instances are passed through); an unoccluded ray adds its contribution to
// its pixel's radiance. stats[3] += shadow rays.
hipError_t launch_connect(const SceneView& sv, const Queues& Q, const PathState& S, Ctrl* ctrl, uint32_t slot,
                          unsigned long long* stats, const Tuning& tu, hipStream_t st);
hipError_t launch_connect_f32(const SceneView& sv, const Queues& Q, const PathState& S, Ctrl* ctrl, uint32_t slot,
                              unsigned long long* stats, const Tuning& tu, hipStream_t st);
hipError_t launch_connect_q8(const SceneView& sv, const Queues& Q, const PathState& S, Ctrl* ctrl, uint32_t slot,
                             unsigned long long* stats, const Tuning& tu, hipStream_t st);
hipError_t launch_finish(const TraceParams& P, const PathState& S, float4* img, hipStream_t st);
// Finishes every live path of the current queue in one launch (tail.hip);
// stats[0] and stats[4] += its extension rays, stats[3] += its shadow rays.
hipError_t launch_tail(const SceneView& sv, const TraceParams& P, const PathState& S, const Queues& Q,
                       uint32_t in_q, Ctrl* ctrl, uint32_t slot, unsigned long long* stats, const Tuning& tu,
                       uint32_t live, hipStream_t st);
// Server version: finishes every request in the shared queues (pixels that
// finish their last sample are counted in done[request]). max_blocks caps
// the grid (0: one block per 256 live paths), so a tail launch can leave
// room on the GPU for kernels of another stream.
hipError_t launch_tail_multi(bool quantized, bool nee, const JobRec* jobs, const PathState& S, const Queues& Q,
                             uint32_t in_q, Ctrl* ctrl, uint32_t slot, unsigned long long* stats, uint32_t* done,
                             const Tuning& tu, uint32_t live, hipStream_t st, uint32_t max_blocks = 0);
hipError_t launch_debug_math(int fn, const float* in, const float* in2, float* out, uint32_t n, hipStream_t st);
hipError_t launch_postprocess(const float4* in, float4* ping, float4* pong, uchar4* out, uint32_t W, uint32_t H,
                              const BloomPushConsts& b, const TonemappingPushConsts& t, hipStream_t st);

}  // namespace rt
