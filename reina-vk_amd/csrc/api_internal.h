// api_internal.h — host-side state shared by the C ABI translation units
// (api.cpp: scenes and single-request tracing; server.cpp: the request loop).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "rt_layout.h"
#include "wavefront.h"

namespace rt {

// Records msg as the calling thread's last error (reina_rt_last_error) and
// returns code.
int fail(int code, const std::string& msg);

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(REINA_RT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Live paths at or below which reina_rt_trace ends the frame in one tail
// launch (ReinaTuning.tail_paths default).
constexpr uint32_t kDefaultTailPaths = 1u << 20;

// Kernel parameters of one vkCmdTraceRaysKHR-equivalent render.
inline TraceParams make_trace_params(const RtPushConsts* pc, uint32_t W, uint32_t H) {
    TraceParams P;
    std::memcpy(P.cam.invView, pc->invView, 64);
    std::memcpy(P.cam.invProj, pc->invProjection, 64);
    P.cam.focusDist = pc->focusDist;
    P.cam.defocus = pc->defocusMultiplier;
    P.cam.W = (float)(int32_t)W;
    P.cam.H = (float)(int32_t)H;
    P.W = W;
    P.H = H;
    P.spp = pc->samplesPerPixel;
    P.maxBounces = pc->maxBounces;
    P.sampleBatch = pc->sampleBatch;
    P.directClamp = pc->directClamp;
    P.nee = 0;
    P.fixes = 0;
    P.groups = 1;
    return P;
}

}  // namespace rt

// A scene resident in HBM (reina_rt_scene_create).
struct ReinaScene {
    int device = 0;
    rt::SceneView sv{};
    void* arena = nullptr;
    uint64_t arena_bytes = 0;
    ReinaSceneInfo info{};
    std::vector<rt::Bvh8Node> host_nodes;   // kept for export
    std::vector<rt::QBvh8Node> host_qnodes; // quantized copy (REINA_RT_NODES_QUANTIZED)
    bool quantized = false;
    float device_build_ms = 0.0f;       // GPU time of a REINA_RT_BUILD_DEVICE build
    std::vector<float> host_tris;
    // light-sampling tables (next-event estimation)
    std::vector<rt::EmissiveInst> emi;
    std::vector<float> tri_cdf;
    std::vector<int32_t> inst_emi;
    std::vector<uint8_t> tri_flags;
    // trace buffers (grown on demand)
    uint64_t cap_px = 0;
    void* tbuf = nullptr;
    rt::PathState S{};
    rt::Queues Q{};
    rt::Ctrl* ctrl = nullptr;
    unsigned long long* stats = nullptr;
    uint32_t* fetch = nullptr;
    uint32_t* h_count = nullptr;        // pinned: [2 readbacks][partition * kLine] queue lengths
    rt::Tuning tune{16, 16, 0, 0, 0, 0};
    hipEvent_t cev[2] = {};             // queue-length readbacks
    uint32_t tail_count = 0;            // 1 if the last trace ended with a tail launch
    hipEvent_t tev0 = nullptr, tev1 = nullptr;   // around the tail launch
    uint32_t min_stack = 8;             // smallest legal stack depth for this BVH
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool traced = false;
    uint32_t options = 0;
    std::vector<hipEvent_t> xev;        // extend event pairs (REINA_RT_TRACE_TIME_EXTEND)
    uint32_t xev_used = 0;
    uint32_t last_iters = 0;
    bool trace_ready = false;           // control words and events created
    hipStream_t last_stream = nullptr;
};

using rt::fail;
