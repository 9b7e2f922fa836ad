// bvh_host.h — host BVH8 builder interface.
#pragma once

#include <mutex>
#include <string>
#include <vector>

#include "rt_layout.h"

namespace rt {

struct Bvh8Result {
    std::vector<Bvh8Node> nodes;
    std::vector<float> tris;     // 12 floats per triangle (3 x float4), leaf order
    uint32_t max_depth = 0;
    float sah_cost = 0.0f;
};

// tri: 9 floats per triangle (world-space v0, v1, v2), flattened id = index.
bool build_bvh8_host(const float* tri, uint64_t ntri, Bvh8Result& out, std::string& err);

// The same build on the GPU of the current HIP device (tri: host array,
// 9 floats per triangle): binned SAH over the Morton-sorted order, collapsed
// to BVH8 with the host builder's node and triangle format. device_ms: GPU
// time of the build (optional).
bool build_bvh8_device(const float* tri, uint64_t ntri, Bvh8Result& out, std::string& err, float* device_ms);

// Surface-area-heuristic cost of a BVH8 (FP64 sum over the child boxes of
// every node, relative to the root's area; leaf boxes weighted by their
// triangle count).
float bvh8_sah_cost(const std::vector<Bvh8Node>& nodes);

// 8-bit quantized copy of a BVH8 (same topology, header and meta). Every
// decoded child box contains the FP32 child box. Fails only for non-finite
// or astronomically large boxes (grid step above 2^127).
bool quantize_bvh8(const std::vector<Bvh8Node>& nodes, std::vector<QBvh8Node>& out, std::string& err);

}  // namespace rt
