// trace_f32.hip — single-scene k_trace instantiations for FP32 BVH8 nodes (256 bytes).
#include "trace_kernel.h"

namespace rt {

hipError_t launch_trace_f32(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                         unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu,
                         hipStream_t st) {
    return launch_trace_fmt<false, false>(sv, rays, hits, in, stats, counters, count, tu, nullptr, nullptr, st);
}

// FP16 inverse directions (REINA_RT_TRACE_F16_INV, closest-hit queries only)
template <int SD>
static void launch_h16_sd(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                          unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu,
                          hipStream_t st) {
    if (count)
        hipLaunchKernelGGL((k_trace<SD, 1, true, false, false, false, true>), dim3(tu.grid), dim3(256), 0, st, sv, rays,
                           hits, in, stats, counters, tu.refill_min, tu.claim, nullptr, nullptr, AnyHitOut{});
    else
        hipLaunchKernelGGL((k_trace<SD, 1, false, false, false, false, true>), dim3(tu.grid), dim3(256), 0, st, sv,
                           rays, hits, in, stats, counters, tu.refill_min, tu.claim, nullptr, nullptr, AnyHitOut{});
}

hipError_t launch_trace_f32_h16(const SceneView& sv, const float4* rays, float4* hits, const TraceInput& in,
                                unsigned long long* stats, uint32_t* counters, bool count, const Tuning& tu,
                                hipStream_t st) {
    switch (tu.stack_depth) {
        case 8: launch_h16_sd<8>(sv, rays, hits, in, stats, counters, count, tu, st); break;
        case 16: launch_h16_sd<16>(sv, rays, hits, in, stats, counters, count, tu, st); break;
        case 24: launch_h16_sd<24>(sv, rays, hits, in, stats, counters, count, tu, st); break;
        default: launch_h16_sd<32>(sv, rays, hits, in, stats, counters, count, tu, st); break;
    }
    return hipGetLastError();
}

hipError_t launch_connect_f32(const SceneView& sv, const Queues& Q, const PathState& S, Ctrl* ctrl, uint32_t slot,
                           unsigned long long* stats, const Tuning& tu, hipStream_t st) {
    return launch_connect_fmt<false>(sv, Q, S, ctrl, slot, stats, tu, st);
}

}  // namespace rt
