// trace_f32.hip — single-scene k_trace instantiations for FP32 BVH8 nodes (256 bytes).
#include "trace_kernel.h"

namespace rt {

hipError_t launch_trace_f32(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                         unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu,
                         hipStream_t st) {
    return launch_trace_fmt<false, false>(sv, rays, hits, in, stats, counters, count, tu, nullptr, nullptr, st);
}

hipError_t launch_connect_f32(const SceneView& sv, const Queues& Q, const PathState& S, Ctrl* ctrl, uint32_t slot,
                           unsigned long long* stats, const Tuning& tu, hipStream_t st) {
    return launch_connect_fmt<false>(sv, Q, S, ctrl, slot, stats, tu, st);
}

}  // namespace rt
