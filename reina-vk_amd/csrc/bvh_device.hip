// bvh_device.hip — BVH8 construction on the GPU (REINA_RT_BUILD_DEVICE), the
// replacement of the driver's BLAS/TLAS build (src/graphics/Blas.cpp:8-124,
// src/tools/vktools.cpp:460-596) for flattened scenes.
//
//   1. triangle boxes and centroids; centroid bounds
//   2. 63-bit Morton codes of the centroids, radix-sorted with the triangle
//      index as payload (ties keep index order)
//   3. a binary BVH, bottom-up, by parallel locally-ordered clustering
//      (PLOC: Meister and Bittner, "Parallel Locally-Ordered Clustering for
//      Bounding Volume Hierarchy Construction", IEEE TVCG 2018): the clusters
//      (at first one per triangle, in Morton order) each find the neighbour
//      within kRadius positions whose union box has the smallest surface
//      area; mutual nearest neighbours merge into a new node, which takes the
//      lower position; the array is compacted by a scan, and this repeats
//      until one cluster is left. Ties are broken by the distance in the
//      array, then by the parity of the lower position, then by the index,
//      so every round merges at least one pair and the tree is deterministic
//      (equal boxes pair up as (0,1), (2,3), ...).
//   4. collapse to BVH8 level by level: a subtree of <= 3 triangles is a
//      leaf unless splitting it is cheaper by the SAH; a node opens its largest-area internal child until it has 8
//      children, children take octant slots as in the host builder, and node
//      / triangle offsets come from exclusive scans, so the result is
//      deterministic.
// The output has the host builder's format (Bvh8Node, leaf-ordered triangles,
// boxes padded outward by one ulp), so the traversal, the quantizer and the
// oracle's restatement apply unchanged.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bvh_host.h"

namespace rt {
namespace {

constexpr uint32_t kLeafMax = 3;
constexpr int kRadius = 32;
constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t ord_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_ord(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

struct Box6 {   // ordered-integer keys: lo (min), hi (max)
    uint32_t lo[3], hi[3];
};

__device__ __forceinline__ void tri_box(const float* w, float* lo, float* hi) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = fminf(fminf(w[a], w[3 + a]), w[6 + a]);
        hi[a] = fmaxf(fmaxf(w[a], w[3 + a]), w[6 + a]);
    }
}

__global__ void k_prep(const float* __restrict__ wtri, uint32_t n, float4* __restrict__ cent, Box6* cb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float lo[3], hi[3], c[3] = {0.0f, 0.0f, 0.0f};
    bool ok = i < n;
    if (ok) {
        tri_box(wtri + 9ull * i, lo, hi);
#pragma unroll
        for (int a = 0; a < 3; ++a) c[a] = 0.5f * lo[a] + 0.5f * hi[a];
        cent[i] = make_float4(c[0], c[1], c[2], 0.0f);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        uint32_t kmin = ok ? ord_key(c[a]) : 0xFFFFFFFFu, kmax = ok ? ord_key(c[a]) : 0u;
        for (int o = 32; o > 0; o >>= 1) {
            kmin = min(kmin, (uint32_t)__shfl_xor(kmin, o));
            kmax = max(kmax, (uint32_t)__shfl_xor(kmax, o));
        }
        if (__lane_id() == 0) {
            atomicMin(&cb->lo[a], kmin);
            atomicMax(&cb->hi[a], kmax);
        }
    }
}

__device__ __forceinline__ uint64_t spread21(uint32_t x) {
    uint64_t v = x & 0x1FFFFFu;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ void k_morton(const float4* __restrict__ cent, uint32_t n, const Box6* cb, uint64_t* __restrict__ keys,
                         uint32_t* __restrict__ ids) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = cent[i];
    const float cc[3] = {c.x, c.y, c.z};
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float lo = key_ord(cb->lo[a]), hi = key_ord(cb->hi[a]);
        const float ext = hi - lo;
        float t = ext > 0.0f ? (cc[a] - lo) / ext : 0.0f;
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        q[a] = min((uint32_t)(t * 2097152.0f), 2097151u);
    }
    keys[i] = spread21(q[0]) | (spread21(q[1]) << 1) | (spread21(q[2]) << 2);
    ids[i] = i;
}

// ---- binary build (PLOC) -------------------------------------------------------
struct BNode {
    uint32_t left, right;   // children; left = kNone for a triangle
    uint32_t count;         // triangles below
    uint32_t tri;           // triangle index (leaves)
    float lo[3], hi[3];
};

__global__ void k_leaves(const float* __restrict__ wtri, const uint32_t* __restrict__ ids, uint32_t n,
                         BNode* __restrict__ nodes, uint32_t* __restrict__ cl) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    BNode b;
    b.left = b.right = kNone;
    b.count = 1;
    b.tri = ids[k];
    tri_box(wtri + 9ull * b.tri, b.lo, b.hi);
    nodes[k] = b;
    cl[k] = k;
}

__device__ __forceinline__ float half_area_f(float x0, float y0, float z0, float x1, float y1, float z1) {
    const float dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
    return dx * dy + dy * dz + dz * dx;
}

// Nearest neighbour of every cluster within kRadius array positions: the
// smallest key (union half-area bits, |i - j|, parity of min(i, j), j). The
// block stages the boxes of its 256 clusters and kRadius on each side in LDS.
__global__ __launch_bounds__(256) void k_nn(const BNode* __restrict__ nodes, const uint32_t* __restrict__ cl,
                                            uint32_t N, uint32_t* __restrict__ nn) {
    __shared__ float sb[6][256 + 2 * kRadius];
    const int b0 = (int)(blockIdx.x * 256u) - kRadius;
    for (int t = threadIdx.x; t < 256 + 2 * kRadius; t += 256) {
        const int j = b0 + t;
        if (j >= 0 && j < (int)N) {
            const BNode& x = nodes[cl[j]];
            sb[0][t] = x.lo[0]; sb[1][t] = x.lo[1]; sb[2][t] = x.lo[2];
            sb[3][t] = x.hi[0]; sb[4][t] = x.hi[1]; sb[5][t] = x.hi[2];
        }
    }
    __syncthreads();
    const int i = (int)(blockIdx.x * 256u + threadIdx.x);
    if (i >= (int)N) return;
    const int ti = i - b0;
    const float lx = sb[0][ti], ly = sb[1][ti], lz = sb[2][ti], hx = sb[3][ti], hy = sb[4][ti], hz = sb[5][ti];
    uint64_t best = ~0ull;
    const int j0 = max(0, i - kRadius), j1 = min((int)N - 1, i + kRadius);
    for (int j = j0; j <= j1; ++j) {
        if (j == i) continue;
        const int tj = j - b0;
        const float a = half_area_f(fminf(lx, sb[0][tj]), fminf(ly, sb[1][tj]), fminf(lz, sb[2][tj]),
                                    fmaxf(hx, sb[3][tj]), fmaxf(hy, sb[4][tj]), fmaxf(hz, sb[5][tj]));
        const uint32_t d = (uint32_t)abs(i - j);
        const uint32_t par = (uint32_t)min(i, j) & 1u;
        const uint64_t key = ((uint64_t)__float_as_uint(a) << 32) | ((uint64_t)d << 25) | ((uint64_t)par << 24) |
                             (uint64_t)j;
        best = key < best ? key : best;
    }
    nn[i] = (uint32_t)(best & 0xFFFFFFu);
}

// flags[i] = (leads a merge) << 32 | (keeps a position)
__global__ void k_flags(const uint32_t* __restrict__ nn, uint32_t N, uint64_t* __restrict__ flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > N) return;
    if (i == N) { flags[N] = 0; return; }
    const uint32_t j = nn[i];
    const bool mutual = nn[j] == i;
    const uint64_t lead = (mutual && i < j) ? 1ull : 0ull;
    const uint64_t keep = (mutual && i > j) ? 0ull : 1ull;
    flags[i] = (lead << 32) | keep;
}

__global__ void k_emit(BNode* __restrict__ nodes, const uint32_t* __restrict__ cl, const uint32_t* __restrict__ nn,
                       const uint64_t* __restrict__ flags, const uint64_t* __restrict__ scan, uint32_t N,
                       uint32_t node_base, uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint64_t f = flags[i];
    if (!(f & 1ull)) return;
    const uint32_t pos = (uint32_t)(scan[i] & 0xFFFFFFFFull);
    if (f >> 32) {
        const uint32_t id = node_base + (uint32_t)(scan[i] >> 32);
        const BNode& l = nodes[cl[i]];
        const BNode& r = nodes[cl[nn[i]]];
        BNode b;
        b.left = cl[i];
        b.right = cl[nn[i]];
        b.count = l.count + r.count;
        b.tri = kNone;
#pragma unroll
        for (int a = 0; a < 3; ++a) { b.lo[a] = fminf(l.lo[a], r.lo[a]); b.hi[a] = fmaxf(l.hi[a], r.hi[a]); }
        nodes[id] = b;
        out[pos] = id;
    } else {
        out[pos] = cl[i];
    }
}

// ---- collapse to BVH8 ---------------------------------------------------------
struct Item {
    uint32_t bnode, out;
};

__device__ __forceinline__ float area_f(const BNode& b) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// A subtree of at most kLeafMax triangles becomes one leaf slot when that is
// no dearer than splitting it by the surface area heuristic (triangle test
// and node visit cost 1 each), as in the host builder.
__device__ __forceinline__ bool inner(const BNode* nodes, const BNode& x) {
    if (x.left == kNone) return false;
    if (x.count > kLeafMax) return true;
    const float pa = area_f(x);
    if (!(pa > 0.0f)) return false;
    const BNode& l = nodes[x.left];
    const BNode& r = nodes[x.right];
    const float split = 1.0f + (area_f(l) * (float)l.count + area_f(r) * (float)r.count) / pa;
    return (float)x.count > split;
}

// Children of one BVH8 node in slot order (kNone = empty slot).
__device__ void collapse_children(const BNode* nodes, uint32_t bn, uint32_t* slot_child) {
    uint32_t ch[8];
    int nc = 0;
    const BNode& b = nodes[bn];
    if (!inner(nodes, b)) {
        ch[nc++] = bn;   // a leaf root
    } else {
        ch[nc++] = b.left;
        ch[nc++] = b.right;
        while (nc < 8) {
            int bi = -1;
            float ba = -1.0f;
            for (int c = 0; c < nc; ++c) {
                const BNode& x = nodes[ch[c]];
                if (inner(nodes, x) && area_f(x) > ba) { ba = area_f(x); bi = c; }
            }
            if (bi < 0) break;
            const BNode& o = nodes[ch[bi]];
            ch[bi] = o.left;
            ch[nc++] = o.right;
        }
    }
    // octant slots: greedy on the signed centroid offsets from the node centre
    float pc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) pc[a] = 0.5f * (b.lo[a] + b.hi[a]);
    float cost[8][8];
    for (int c = 0; c < nc; ++c) {
        const BNode& x = nodes[ch[c]];
        float d[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) d[a] = 0.5f * (x.lo[a] + x.hi[a]) - pc[a];
        for (int s = 0; s < 8; ++s)
            cost[c][s] = ((s & 1) ? d[0] : -d[0]) + ((s & 2) ? d[1] : -d[1]) + ((s & 4) ? d[2] : -d[2]);
    }
    for (int s = 0; s < 8; ++s) slot_child[s] = kNone;
    uint32_t cused = 0, sused = 0;
    for (int it = 0; it < nc; ++it) {
        int bc = -1, bs = -1;
        float bv = -INFINITY;
        for (int c = 0; c < nc; ++c) {
            if (cused >> c & 1u) continue;
            for (int s = 0; s < 8; ++s) {
                if (sused >> s & 1u) continue;
                if (bc < 0 || cost[c][s] > bv) { bv = cost[c][s]; bc = c; bs = s; }
            }
        }
        cused |= 1u << bc;
        sused |= 1u << bs;
        slot_child[bs] = ch[bc];
    }
}

__global__ void k_collapse_count(const BNode* __restrict__ nodes, const Item* __restrict__ items, uint32_t nitems,
                                 uint32_t* __restrict__ kids, uint32_t* __restrict__ n_int, uint32_t* __restrict__ n_tri) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitems) return;
    uint32_t sc[8];
    collapse_children(nodes, items[i].bnode, sc);
    uint32_t ni = 0, nt = 0;
    for (int s = 0; s < 8; ++s) {
        kids[8 * i + s] = sc[s];
        if (sc[s] == kNone) continue;
        const BNode& x = nodes[sc[s]];
        if (inner(nodes, x)) ni++;
        else nt += x.count;
    }
    n_int[i] = ni;
    n_tri[i] = nt;
}

__global__ void k_collapse_write(const BNode* __restrict__ nodes, const Item* __restrict__ items, uint32_t nitems,
                                 const uint32_t* __restrict__ kids, const uint32_t* __restrict__ scan_int,
                                 const uint32_t* __restrict__ scan_tri, uint32_t node_base, uint32_t tri_base0,
                                 const float* __restrict__ wtri,
                                 Bvh8Node* __restrict__ out, float4* __restrict__ out_tris, Item* __restrict__ next) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitems) return;
    Bvh8Node nd;
    const uint32_t child_base = node_base + scan_int[i];
    const uint32_t tri_base = tri_base0 + scan_tri[i];
    uint32_t rank = 0, toff = 0, imask = 0, lmask = 0;
    uint8_t meta[8];
    for (int s = 0; s < 8; ++s) {
        const uint32_t c = kids[8 * i + s];
        if (c == kNone) {
            nd.lo_x[s] = nd.lo_y[s] = nd.lo_z[s] = 0.0f;
            nd.hi_x[s] = nd.hi_y[s] = nd.hi_z[s] = 0.0f;
            meta[s] = 0;
            continue;
        }
        const BNode& x = nodes[c];
        nd.lo_x[s] = nextafterf(x.lo[0], -INFINITY);
        nd.lo_y[s] = nextafterf(x.lo[1], -INFINITY);
        nd.lo_z[s] = nextafterf(x.lo[2], -INFINITY);
        nd.hi_x[s] = nextafterf(x.hi[0], INFINITY);
        nd.hi_y[s] = nextafterf(x.hi[1], INFINITY);
        nd.hi_z[s] = nextafterf(x.hi[2], INFINITY);
        if (inner(nodes, x)) {
            meta[s] = (uint8_t)(0x80u | rank);
            next[scan_int[i] + rank] = Item{c, child_base + rank};
            imask |= 1u << s;
            rank++;
        } else {
            meta[s] = (uint8_t)((x.count << 5) | toff);
            // the subtree's triangles, left first (at most kLeafMax leaves, depth <= kLeafMax - 1)
            uint32_t stk[kLeafMax + 1];
            int sp = 0;
            stk[sp++] = c;
            uint32_t j = 0;
            while (sp > 0) {
                const BNode& y = nodes[stk[--sp]];
                if (y.left != kNone) {
                    stk[sp++] = y.right;
                    stk[sp++] = y.left;
                    continue;
                }
                const uint32_t t = y.tri;
                const float* w = wtri + 9ull * t;
                float4* o = out_tris + 3ull * (tri_base + toff + j);
                o[0] = make_float4(w[0], w[1], w[2], __uint_as_float(t));
                o[1] = make_float4(w[3], w[4], w[5], 0.0f);
                o[2] = make_float4(w[6], w[7], w[8], 0.0f);
                j++;
            }
            lmask |= 1u << s;
            toff += x.count;
        }
    }
    nd.child_base = (rank ? child_base : 0u) | (imask << 24);
    nd.tri_base = tri_base | (lmask << 24);
    for (int s = 0; s < 8; ++s) nd.meta[s] = meta[s];
    for (int k = 0; k < 12; ++k) nd.pad[k] = 0u;
    out[items[i].out] = nd;
}

template <typename T>
struct DBuf {
    T* p = nullptr;
    ~DBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t n) { return hipMalloc(&p, (n ? n : 1) * sizeof(T)); }
};

inline uint32_t blocks(uint64_t n, uint32_t b = 256) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

#define D_TRY(expr)                                                               \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);              \
            return false;                                                         \
        }                                                                         \
    } while (0)

bool build_bvh8_device(const float* tri, uint64_t ntri, Bvh8Result& out, std::string& err, float* device_ms) {
    out = Bvh8Result();
    if (ntri >= (1ull << 24)) {
        err = "more than 2^24 triangles";
        return false;
    }
    const uint32_t n = (uint32_t)ntri;
    if (n == 0) {   // one empty root node, as the host builder
        out.nodes.resize(1);
        std::memset(out.nodes.data(), 0, sizeof(Bvh8Node));
        out.max_depth = 1;
        return true;
    }
    hipStream_t st = nullptr;
    D_TRY(hipStreamCreate(&st));
    struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } sg{st};
    hipEvent_t e0, e1;
    D_TRY(hipEventCreate(&e0));
    D_TRY(hipEventCreate(&e1));
    struct EvGuard { hipEvent_t a, b; ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } eg{e0, e1};
    DBuf<float> d_w;
    D_TRY(d_w.alloc(9ull * n));
    D_TRY(hipMemcpyAsync(d_w.p, tri, 36ull * n, hipMemcpyHostToDevice, st));
    D_TRY(hipEventRecord(e0, st));
    DBuf<float4> d_cent;
    DBuf<Box6> d_cb;
    D_TRY(d_cent.alloc(n));
    D_TRY(d_cb.alloc(1));
    {
        Box6 init;
        for (int a = 0; a < 3; ++a) { init.lo[a] = 0xFFFFFFFFu; init.hi[a] = 0u; }
        D_TRY(hipMemcpyAsync(d_cb.p, &init, sizeof init, hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_prep, dim3(blocks(n)), dim3(256), 0, st, d_w.p, n, d_cent.p, d_cb.p);
    D_TRY(hipGetLastError());
    DBuf<uint64_t> d_k0, d_k1;
    DBuf<uint32_t> d_i0, d_i1;
    D_TRY(d_k0.alloc(n)); D_TRY(d_k1.alloc(n)); D_TRY(d_i0.alloc(n)); D_TRY(d_i1.alloc(n));
    hipLaunchKernelGGL(k_morton, dim3(blocks(n)), dim3(256), 0, st, d_cent.p, n, d_cb.p, d_k0.p, d_i0.p);
    D_TRY(hipGetLastError());
    size_t tmp_bytes = 0;
    D_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, d_k0.p, d_k1.p, d_i0.p, d_i1.p, (int)n, 0, 63, st));
    DBuf<uint8_t> d_tmp;
    size_t scan_bytes = 0;
    {
        uint64_t* a = nullptr;
        D_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, a, a, (int)n + 1, st));
    }
    D_TRY(d_tmp.alloc(std::max(tmp_bytes, scan_bytes)));
    D_TRY(hipcub::DeviceRadixSort::SortPairs(d_tmp.p, tmp_bytes, d_k0.p, d_k1.p, d_i0.p, d_i1.p, (int)n, 0, 63, st));
    const uint32_t* ids = d_i1.p;

    // binary build: clustering rounds until one cluster is left
    DBuf<BNode> d_nodes;
    DBuf<uint32_t> d_cl0, d_cl1, d_nn;
    DBuf<uint64_t> d_flags, d_fscan;
    D_TRY(d_nodes.alloc(2ull * n)); D_TRY(d_cl0.alloc(n)); D_TRY(d_cl1.alloc(n)); D_TRY(d_nn.alloc(n));
    D_TRY(d_flags.alloc(n + 1ull)); D_TRY(d_fscan.alloc(n + 1ull));
    hipLaunchKernelGGL(k_leaves, dim3(blocks(n)), dim3(256), 0, st, d_w.p, ids, n, d_nodes.p, d_cl0.p);
    D_TRY(hipGetLastError());
    uint32_t N = n, node_count = n, rounds = 0;
    uint32_t* cl = d_cl0.p;
    uint32_t* cl_out = d_cl1.p;
    while (N > 1) {
        hipLaunchKernelGGL(k_nn, dim3(blocks(N)), dim3(256), 0, st, d_nodes.p, cl, N, d_nn.p);
        hipLaunchKernelGGL(k_flags, dim3(blocks(N + 1ull)), dim3(256), 0, st, d_nn.p, N, d_flags.p);
        D_TRY(hipGetLastError());
        D_TRY(hipcub::DeviceScan::ExclusiveSum(d_tmp.p, scan_bytes, d_flags.p, d_fscan.p, (int)N + 1, st));
        uint64_t tot = 0;
        D_TRY(hipMemcpyAsync(&tot, d_fscan.p + N, 8, hipMemcpyDeviceToHost, st));
        D_TRY(hipStreamSynchronize(st));
        const uint32_t merges = (uint32_t)(tot >> 32), keep = (uint32_t)tot;
        if (merges == 0 || keep + merges != N || node_count + merges > 2 * n - 1 || ++rounds > 1000000) {
            err = "device BVH build did not converge";
            return false;
        }
        hipLaunchKernelGGL(k_emit, dim3(blocks(N)), dim3(256), 0, st, d_nodes.p, cl, d_nn.p, d_flags.p, d_fscan.p, N,
                           node_count, cl_out);
        D_TRY(hipGetLastError());
        node_count += merges;
        N = keep;
        std::swap(cl, cl_out);
    }
    uint32_t root = 0;
    D_TRY(hipMemcpyAsync(&root, cl, 4, hipMemcpyDeviceToHost, st));
    D_TRY(hipStreamSynchronize(st));
    // collapse
    const uint32_t n_internal = n - 1;
    std::vector<Bvh8Node> nodes_h;
    DBuf<Bvh8Node> d_out;
    DBuf<float4> d_otri;
    DBuf<Item> d_it0, d_it1;
    DBuf<uint32_t> d_ck, d_ni, d_nt, d_si, d_stri;
    const uint32_t out_cap = std::max(1u, n_internal) + 1;
    D_TRY(d_out.alloc(out_cap)); D_TRY(d_otri.alloc(3ull * n));
    D_TRY(d_it0.alloc(out_cap)); D_TRY(d_it1.alloc(out_cap));
    D_TRY(d_ck.alloc(8ull * out_cap)); D_TRY(d_ni.alloc(out_cap + 1)); D_TRY(d_nt.alloc(out_cap + 1));
    D_TRY(d_si.alloc(out_cap + 1)); D_TRY(d_stri.alloc(out_cap + 1));
    {
        Item r0{root, 0};
        D_TRY(hipMemcpyAsync(d_it0.p, &r0, sizeof r0, hipMemcpyHostToDevice, st));
    }
    uint32_t nitems = 1, out_nodes = 1, out_tris = 0, depth = 0;
    Item* items = d_it0.p;
    Item* nitems_buf = d_it1.p;
    size_t scan2 = 0;
    {
        uint32_t* a = nullptr;
        D_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan2, a, a, (int)out_cap + 1, st));
    }
    DBuf<uint8_t> d_tmp2;
    D_TRY(d_tmp2.alloc(scan2));
    while (nitems > 0) {
        depth++;
        hipLaunchKernelGGL(k_collapse_count, dim3(blocks(nitems, 64)), dim3(64), 0, st, d_nodes.p, items, nitems,
                           d_ck.p, d_ni.p, d_nt.p);
        D_TRY(hipGetLastError());
        D_TRY(hipMemsetAsync(d_ni.p + nitems, 0, 4, st));
        D_TRY(hipMemsetAsync(d_nt.p + nitems, 0, 4, st));
        D_TRY(hipcub::DeviceScan::ExclusiveSum(d_tmp2.p, scan2, d_ni.p, d_si.p, (int)nitems + 1, st));
        D_TRY(hipcub::DeviceScan::ExclusiveSum(d_tmp2.p, scan2, d_nt.p, d_stri.p, (int)nitems + 1, st));
        uint32_t tot[2] = {0, 0};
        D_TRY(hipMemcpyAsync(&tot[0], d_si.p + nitems, 4, hipMemcpyDeviceToHost, st));
        D_TRY(hipMemcpyAsync(&tot[1], d_stri.p + nitems, 4, hipMemcpyDeviceToHost, st));
        D_TRY(hipStreamSynchronize(st));
        if (out_nodes + tot[0] > out_cap || out_tris + tot[1] > n) {
            err = "device BVH collapse overflow";
            return false;
        }
        hipLaunchKernelGGL(k_collapse_write, dim3(blocks(nitems, 64)), dim3(64), 0, st, d_nodes.p, items, nitems,
                           d_ck.p, d_si.p, d_stri.p, out_nodes, out_tris, d_w.p, d_out.p, d_otri.p, nitems_buf);
        D_TRY(hipGetLastError());
        out_nodes += tot[0];
        out_tris += tot[1];
        nitems = tot[0];
        std::swap(items, nitems_buf);
    }
    D_TRY(hipEventRecord(e1, st));
    if (out_tris != n) {
        err = "device BVH lost triangles";
        return false;
    }
    if (out_nodes >= (1u << 24)) {
        err = "too many BVH8 nodes for the index encoding (max 2^24)";
        return false;
    }
    out.nodes.resize(out_nodes);
    out.tris.resize(12ull * n);
    D_TRY(hipMemcpyAsync(out.nodes.data(), d_out.p, sizeof(Bvh8Node) * out_nodes, hipMemcpyDeviceToHost, st));
    D_TRY(hipMemcpyAsync(out.tris.data(), d_otri.p, 48ull * n, hipMemcpyDeviceToHost, st));
    D_TRY(hipStreamSynchronize(st));
    float ms = 0.0f;
    D_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (device_ms) *device_ms = ms;
    out.max_depth = depth;
    out.sah_cost = bvh8_sah_cost(out.nodes);
    return true;
}

}  // namespace rt
