// rt_layout.h — data layouts in HBM shared by the host code and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/reina_rt.h"

namespace rt {

// Pointers that kernels read from memory rather than from their arguments
// (the scene views of a server's job table) are typed as global-memory
// pointers in device code. Otherwise the compiler cannot tell global from LDS
// or private memory and emits flat loads, which also count against the LDS
// counter, so the traversal's LDS stack operations would wait for them. The
// layout is the same for the host compiler, which sees plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
#define RT_GLOBAL __attribute__((address_space(1)))
#else
#define RT_GLOBAL
#endif

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

// Quantized BVH8 node, 80 bytes (five 16-byte loads). Same topology, header
// and meta as Bvh8Node; each child box is stored as 8-bit offsets from the
// node origin p on a power-of-two grid per axis:
//   word 0-2: p.x, p.y, p.z (float)
//   word 3:   ex | ey << 8 | ez << 16: grid step s = 2^(e - 127) per axis
//             (the float with exponent field e and zero mantissa, 1 <= e <= 254)
//   word 4:   child_base | imask << 24      word 5: tri_base | leafmask << 24
//   word 6-7: meta[8]
//   bytes 32-79: qlo_x[8] qhi_x[8] qlo_y[8] qhi_y[8] qlo_z[8] qhi_z[8]
// Decoded plane = fmaf(float(q), s, p), one rounding (q * s is exact). The
// builder picks q_lo / q_hi so that the decoded box contains the child's
// FP32 box (checked with this exact decode), so traversal stays
// conservative and the closest hit still equals brute force.
struct alignas(16) QBvh8Node {
    float px, py, pz;
    uint32_t exps;
    uint32_t child_base;   // | imask << 24
    uint32_t tri_base;     // | leafmask << 24
    uint8_t meta[8];
    uint8_t qlo_x[8], qhi_x[8], qlo_y[8], qhi_y[8], qlo_z[8], qhi_z[8];
};
static_assert(sizeof(QBvh8Node) == 80, "quantized node size");
constexpr uint32_t kQNodeUsefulBytes = 80;
constexpr uint32_t kTriUsefulBytes = 36;   // 3 x float3 per tested triangle
constexpr uint32_t kMaxLeafTris = 3;
constexpr uint32_t kMaxStackDepth = 32;    // largest LDS group stack per lane (entries)
// Surface-area costs of the BVH8 collapse (both builders): a node visit, and a
// triangle test relative to it (the traversal tests a leaf's triangles while
// the next node is fetched, so a triangle costs about half a node).
constexpr float kCollapseNodeCost = 1.0f;
constexpr float kCollapseTriCost = 0.5f;

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

// Light sampling for next-event estimation: one record per emissive instance
// (emission with |e|^2 > 1e-10, src/scene/Instance.cpp:75-77). Instances are
// picked by the CDF of weight = sum of world-space triangle area x luminance
// (0.2126 r + 0.7152 g + 0.0722 b), triangles within an instance by area,
// points uniformly on the triangle, so the area density of a point of
// instance k is pdf_area = (weight_k / total) / area_k.
struct alignas(16) EmissiveInst {
    uint32_t inst;       // instance index
    uint32_t ntri;       // triangles of the instance
    uint32_t cdf_base;   // first entry of the instance's triangle CDF in tri_cdf
    float cdf;           // cumulative instance weight / total weight (last = 1)
    float pdf_area;
    float pad[3];
};
static_assert(sizeof(EmissiveInst) == 32, "emissive record size");

// Per flattened triangle flags for shadow rays: bit 0: its instance culls
// back faces (closest-hit shaders pass through them, shaders/raytrace/lambertian.rchit.glsl:15-19);
// bit 1: the instance transform mirrors (det < 0), so the front face is the
// clockwise one in world space.
constexpr uint8_t kTriCull = 1, kTriMirror = 2;

// Device-resident scene (pointers into one HBM arena).
// A texture in the scene arena: w x h RGBA8 texels (packed r | g << 8 |
// b << 16 | a << 24) starting at texel index `base`.
struct TexInfo {
    uint32_t w, h;
    uint64_t base;
};

struct SceneView {
    const Bvh8Node RT_GLOBAL* nodes;       // FP32 nodes (null when quantized)
    const QBvh8Node RT_GLOBAL* qnodes;     // quantized nodes (null when FP32)
    const float4 RT_GLOBAL* tris;          // 3 per triangle
    const uint32_t RT_GLOBAL* tri_inst;    // flattened id -> instance
    const uint32_t RT_GLOBAL* tri_prim;    // flattened id -> gl_PrimitiveID
    const InstanceRec RT_GLOBAL* inst;
    const ReinaInstanceProperties RT_GLOBAL* props;
    const float4 RT_GLOBAL* vertices;
    const uint32_t RT_GLOBAL* indices;
    const float RT_GLOBAL* tbns;
    const uint32_t RT_GLOBAL* tbn_indices;
    const float RT_GLOBAL* tex_coords;
    const uint32_t RT_GLOBAL* tex_indices;
    uint64_t num_tbns;
    uint64_t num_tbn_indices;
    uint64_t num_tex_coords;
    uint64_t num_nodes;
    uint64_t num_tris;
    // next-event estimation
    const EmissiveInst RT_GLOBAL* emi;
    const float RT_GLOBAL* tri_cdf;
    const int32_t RT_GLOBAL* inst_emi;     // instance -> emissive record, -1 if not emissive
    const uint8_t RT_GLOBAL* tri_flags;
    uint32_t num_emi;
    uint32_t num_inst;
    // shading record of flattened triangle g, kShadeRecF4 float4 at srec + kShadeRecF4 * g:
    // (v0.xyz, instance), (v1.xyz, primitive), (v2.xyz, properties index), then T, B, N of
    // vertex 0, 1, 2 (27 floats, 1 pad): the object-space values hit_info reads, gathered
    // once per scene
    const float4 RT_GLOBAL* srec;
    // textures of InstanceProperties.textureID / normalMapTexID / bumpMapTexID
    const TexInfo RT_GLOBAL* tex;
    const uint32_t RT_GLOBAL* texels;
    uint32_t num_tex;
};
constexpr uint32_t kShadeRecF4 = 10;   // 160 bytes

}  // namespace rt
