// trace_q8.hip — single-scene k_trace instantiations for 8-bit quantized BVH8 nodes (80 bytes).
#include "trace_kernel.h"

namespace rt {

hipError_t launch_trace_q8(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                         unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu,
                         hipStream_t st) {
    return launch_trace_fmt<true, false>(sv, rays, hits, in, stats, counters, count, tu, nullptr, nullptr, st);
}

hipError_t launch_connect_q8(const SceneView& sv, const Queues& Q, const PathState& S, Ctrl* ctrl, uint32_t slot,
                           unsigned long long* stats, const Tuning& tu, hipStream_t st) {
    return launch_connect_fmt<true>(sv, Q, S, ctrl, slot, stats, tu, st);
}

}  // namespace rt
