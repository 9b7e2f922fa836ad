// rt_layout.h — data layouts in HBM shared by the host code and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/reina_rt.h"

namespace rt {

// One BVH8 node, 256 bytes (two 128-byte cache lines). Child boxes are SoA
// so each coordinate of all 8 children is one 32-byte run:
//   lo_x[8] hi_x[8] lo_y[8] hi_y[8] lo_z[8] hi_z[8]  (192 B)
//   word 48: child_base (bits 0-23): index of the first internal child
//            (internal children of a node are contiguous, in slot order);
//            bits 24-31: imask, the slots holding internal children
//   word 49: tri_base (bits 0-23): first triangle of this node's leaves;
//            bits 24-31: leafmask, the slots holding leaves
//   meta[8]: 0 = empty slot; 0x80 | rank = internal child child_base+rank;
//            otherwise a leaf of (meta >> 5) & 3 (1..3) triangles starting
//            at tri_base + (meta & 31)
// At most 2^24 nodes and 2^24 triangles per scene.
// Useful bytes per visit: 192 + 8 + 8 = 208 (the remaining 48 are padding).
struct alignas(256) Bvh8Node {
    float lo_x[8], hi_x[8], lo_y[8], hi_y[8], lo_z[8], hi_z[8];
    uint32_t child_base;   // | imask << 24
    uint32_t tri_base;     // | leafmask << 24
    uint8_t meta[8];
    uint32_t pad[12];
};
static_assert(sizeof(Bvh8Node) == 256, "node size");

constexpr uint32_t kNodeUsefulBytes = 208;
constexpr uint32_t kTriUsefulBytes = 36;   // 3 x float3 per tested triangle
constexpr uint32_t kMaxLeafTris = 3;
constexpr uint32_t kMaxStackDepth = 32;    // largest LDS group stack per lane (entries)

// Triangles in leaf order: 3 x float4 = (v0.xyz, id bits), (v1.xyz, 0), (v2.xyz, 0).

// Per-instance record used by the shading kernels (64 bytes).
struct alignas(16) InstanceRec {
    float c0[3], c1[3], c2[3], c3[3];   // gl_ObjectToWorldEXT columns
    uint32_t props_id;
    uint32_t material;
    uint32_t model;
    uint32_t tri_base;                  // first flattened triangle id
};
static_assert(sizeof(InstanceRec) == 64, "instance record size");

// Device-resident scene (pointers into one HBM arena).
struct SceneView {
    const Bvh8Node* nodes;
    const float4* tris;          // 3 per triangle
    const uint32_t* tri_inst;    // flattened id -> instance
    const uint32_t* tri_prim;    // flattened id -> gl_PrimitiveID
    const InstanceRec* inst;
    const ReinaInstanceProperties* props;
    const float4* vertices;
    const uint32_t* indices;
    const float* tbns;
    const uint32_t* tbn_indices;
    const float* tex_coords;
    const uint32_t* tex_indices;
    uint64_t num_tbns;
    uint64_t num_tbn_indices;
    uint64_t num_tex_coords;
    uint64_t num_nodes;
    uint64_t num_tris;
};

}  // namespace rt
