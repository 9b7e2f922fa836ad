// trace_multi_q8.hip — multi-request k_trace instantiations for 8-bit quantized BVH8 nodes (80 bytes).
#include "trace_kernel.h"

namespace rt {

hipError_t launch_trace_multi_q8(const float4* rays, float4* hits, const TraceInput& in, unsigned long long* stats,
                               const Tuning& tu, const uint32_t* pix, const JobRec* jobs, hipStream_t st) {
    SceneView none{};
    return launch_trace_fmt<true, true>(none, rays, hits, in, stats, nullptr, false, tu, pix, jobs, st);
}

hipError_t launch_connect_multi_q8(const JobRec* jobs, const Queues& Q, const PathState& S, Ctrl* ctrl, uint32_t slot,
                                    const uint32_t* pix, unsigned long long* stats, const Tuning& tu, hipStream_t st) {
    return launch_connect_multi_fmt<true>(jobs, Q, S, ctrl, slot, pix, stats, tu, st);
}

}  // namespace rt
