// bvh_device.hip — BVH8 construction on the GPU (REINA_RT_BUILD_DEVICE), the
// replacement of the driver's BLAS/TLAS build (src/graphics/Blas.cpp:8-124,
// src/tools/vktools.cpp:460-596) for flattened scenes.
//
//   1. triangle boxes and centroids; centroid bounds
//   2. 63-bit Morton codes of the centroids, radix-sorted with the triangle
//      index as payload (ties keep index order)
//   3. a binary BVH, top-down and level-synchronous: each node covers a
//      contiguous range of the sorted order and is split at the best of up to
//      32 equal-count bins along that order by the surface area heuristic
//      (bin boxes by per-chunk reduction + ordered-integer atomics; left and
//      right prefix areas and costs summed in FP64); ranges of <= 3
//      triangles are leaves; from level 40 on, splits are at the median
//   4. collapse to BVH8 level by level: a node opens its largest-area
//      internal child until it has 8 children, children take octant slots as
//      in the host builder, and node / triangle offsets come from exclusive
//      scans, so the result is deterministic.
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

constexpr uint32_t kBins = 32;
constexpr uint32_t kLeafMax = 3;
constexpr uint32_t kMedianLevel = 40;
constexpr uint32_t kChunk = 32;
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

// ---- binary build ----------------------------------------------------------
struct BNode {
    uint32_t first, count;
    uint32_t left;      // first of two consecutive children, kNone for a leaf
    uint32_t pad;
    float lo[3], hi[3];
};

__device__ __forceinline__ uint32_t nbins(uint32_t count) { return count < kBins ? count : kBins; }

__device__ __forceinline__ uint32_t bin_of(uint32_t off, uint32_t count) {
    return (uint32_t)(((uint64_t)off * nbins(count)) / count);
}

__global__ void k_bins_reset(Box6* bins, uint32_t nact) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nact * kBins) return;
#pragma unroll
    for (int a = 0; a < 3; ++a) { bins[i].lo[a] = 0xFFFFFFFFu; bins[i].hi[a] = 0u; }
}

// Each thread folds kChunk consecutive sorted positions into running boxes,
// flushing with ordered-integer atomics when the (node, bin) changes.
__global__ void k_bins_accum(const float* __restrict__ wtri, const uint32_t* __restrict__ ids,
                             const uint32_t* __restrict__ pos_node, const uint32_t* __restrict__ node_slot,
                             const BNode* __restrict__ nodes, uint32_t n, Box6* bins) {
    const uint32_t c0 = (blockIdx.x * blockDim.x + threadIdx.x) * kChunk;
    if (c0 >= n) return;
    const uint32_t c1 = min(c0 + kChunk, n);
    uint32_t cur = kNone;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    auto flush = [&]() {
        if (cur == kNone) return;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            atomicMin(&bins[cur].lo[a], ord_key(lo[a]));
            atomicMax(&bins[cur].hi[a], ord_key(hi[a]));
        }
    };
    for (uint32_t p = c0; p < c1; ++p) {
        const uint32_t nd = pos_node[p];
        if (nd == kNone) continue;
        const BNode& b = nodes[nd];
        const uint32_t key = node_slot[nd] * kBins + bin_of(p - b.first, b.count);
        float tl[3], th[3];
        tri_box(wtri + 9ull * ids[p], tl, th);
        if (key != cur) {
            flush();
            cur = key;
#pragma unroll
            for (int a = 0; a < 3; ++a) { lo[a] = tl[a]; hi[a] = th[a]; }
        } else {
#pragma unroll
            for (int a = 0; a < 3; ++a) { lo[a] = fminf(lo[a], tl[a]); hi[a] = fmaxf(hi[a], th[a]); }
        }
    }
    flush();
}

__device__ __forceinline__ double half_area(const float* lo, const float* hi) {
    const double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// One thread per active node: node box, leaf or split position (SAH over the
// bins, prefix / suffix areas in FP64), and whether it has children.
__global__ void k_split(BNode* nodes, const uint32_t* __restrict__ act, uint32_t nact, const Box6* __restrict__ bins,
                        uint32_t level, uint32_t* __restrict__ split_pos, uint32_t* __restrict__ has_kids) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nact) return;
    BNode& b = nodes[act[s]];
    const uint32_t K = nbins(b.count);
    float blo[kBins][3], bhi[kBins][3];
    float nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = 0; k < K; ++k) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            blo[k][a] = key_ord(bins[s * kBins + k].lo[a]);
            bhi[k][a] = key_ord(bins[s * kBins + k].hi[a]);
            nlo[a] = fminf(nlo[a], blo[k][a]);
            nhi[a] = fmaxf(nhi[a], bhi[k][a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) { b.lo[a] = nlo[a]; b.hi[a] = nhi[a]; }
    if (b.count <= kLeafMax) {
        b.left = kNone;
        has_kids[s] = 0;
        split_pos[s] = 0;
        return;
    }
    uint32_t nl_best = b.count / 2;
    if (level < kMedianLevel) {
        // suffix areas of bins k..K-1
        double rarea[kBins];
        float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = (int)K - 1; k >= 1; --k) {
#pragma unroll
            for (int a = 0; a < 3; ++a) { rlo[a] = fminf(rlo[a], blo[k][a]); rhi[a] = fmaxf(rhi[a], bhi[k][a]); }
            rarea[k] = half_area(rlo, rhi);
        }
        float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        double best = INFINITY;
        for (uint32_t k = 0; k + 1 < K; ++k) {
#pragma unroll
            for (int a = 0; a < 3; ++a) { llo[a] = fminf(llo[a], blo[k][a]); lhi[a] = fmaxf(lhi[a], bhi[k][a]); }
            // positions in bins 0..k (bin_of: floor(off * K / count) <= k)
            const uint32_t nl = (uint32_t)(((uint64_t)(k + 1) * b.count + K - 1) / K);
            const double cost = half_area(llo, lhi) * (double)nl + rarea[k + 1] * (double)(b.count - nl);
            if (cost < best) { best = cost; nl_best = nl; }
        }
    }
    split_pos[s] = b.first + nl_best;
    has_kids[s] = 1;
}

__global__ void k_make_kids(BNode* nodes, const uint32_t* __restrict__ act, uint32_t nact,
                            const uint32_t* __restrict__ split_pos, const uint32_t* __restrict__ has_kids,
                            const uint32_t* __restrict__ scan, uint32_t node_base, uint32_t* __restrict__ next_act,
                            uint32_t* __restrict__ node_slot) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nact || !has_kids[s]) return;
    BNode& b = nodes[act[s]];
    const uint32_t l = node_base + 2 * scan[s];
    b.left = l;
    BNode L{}, R{};
    L.first = b.first;
    L.count = split_pos[s] - b.first;
    R.first = split_pos[s];
    R.count = b.first + b.count - split_pos[s];
    L.left = R.left = kNone;
    nodes[l] = L;
    nodes[l + 1] = R;
    next_act[2 * scan[s]] = l;
    next_act[2 * scan[s] + 1] = l + 1;
    node_slot[l] = 2 * scan[s];
    node_slot[l + 1] = 2 * scan[s] + 1;
}

__global__ void k_assign(const BNode* __restrict__ nodes, uint32_t* __restrict__ pos_node, uint32_t n) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t nd = pos_node[p];
    if (nd == kNone) return;
    const BNode& b = nodes[nd];
    if (b.left == kNone) { pos_node[p] = kNone; return; }
    pos_node[p] = (p < nodes[b.left + 1].first) ? b.left : b.left + 1;
}

// ---- collapse to BVH8 ---------------------------------------------------------
struct Item {
    uint32_t bnode, out;
};

__device__ __forceinline__ float area_f(const BNode& b) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// Children of one BVH8 node in slot order (kNone = empty slot).
__device__ void collapse_children(const BNode* nodes, uint32_t bn, uint32_t* slot_child) {
    uint32_t ch[8];
    int nc = 0;
    const BNode& b = nodes[bn];
    if (b.left == kNone) {
        ch[nc++] = bn;   // a leaf root
    } else {
        ch[nc++] = b.left;
        ch[nc++] = b.left + 1;
        while (nc < 8) {
            int bi = -1;
            float ba = -1.0f;
            for (int c = 0; c < nc; ++c) {
                const BNode& x = nodes[ch[c]];
                if (x.left != kNone && area_f(x) > ba) { ba = area_f(x); bi = c; }
            }
            if (bi < 0) break;
            const uint32_t l = nodes[ch[bi]].left;
            ch[bi] = l;
            ch[nc++] = l + 1;
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
        if (x.left != kNone) ni++;
        else nt += x.count;
    }
    n_int[i] = ni;
    n_tri[i] = nt;
}

__global__ void k_collapse_write(const BNode* __restrict__ nodes, const Item* __restrict__ items, uint32_t nitems,
                                 const uint32_t* __restrict__ kids, const uint32_t* __restrict__ scan_int,
                                 const uint32_t* __restrict__ scan_tri, uint32_t node_base, uint32_t tri_base0,
                                 const float* __restrict__ wtri, const uint32_t* __restrict__ ids,
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
        if (x.left != kNone) {
            meta[s] = (uint8_t)(0x80u | rank);
            next[scan_int[i] + rank] = Item{c, child_base + rank};
            imask |= 1u << s;
            rank++;
        } else {
            meta[s] = (uint8_t)((x.count << 5) | toff);
            for (uint32_t j = 0; j < x.count; ++j) {
                const uint32_t t = ids[x.first + j];
                const float* w = wtri + 9ull * t;
                float4* o = out_tris + 3ull * (tri_base + toff + j);
                o[0] = make_float4(w[0], w[1], w[2], __uint_as_float(t));
                o[1] = make_float4(w[3], w[4], w[5], 0.0f);
                o[2] = make_float4(w[6], w[7], w[8], 0.0f);
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
        uint32_t* a = nullptr;
        D_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, a, a, (int)n + 2, st));
    }
    D_TRY(d_tmp.alloc(std::max(tmp_bytes, scan_bytes)));
    D_TRY(hipcub::DeviceRadixSort::SortPairs(d_tmp.p, tmp_bytes, d_k0.p, d_k1.p, d_i0.p, d_i1.p, (int)n, 0, 63, st));
    const uint32_t* ids = d_i1.p;

    // binary build
    const uint32_t cap = 2 * n + 1;
    DBuf<BNode> d_nodes;
    DBuf<uint32_t> d_pos, d_slot, d_act0, d_act1, d_split, d_kids, d_scan;
    DBuf<Box6> d_bins;
    D_TRY(d_nodes.alloc(cap)); D_TRY(d_pos.alloc(n)); D_TRY(d_slot.alloc(cap));
    D_TRY(d_act0.alloc(n + 2)); D_TRY(d_act1.alloc(n + 2)); D_TRY(d_split.alloc(n + 2)); D_TRY(d_kids.alloc(n + 2));
    D_TRY(d_scan.alloc(n + 2));
    {
        BNode root{};
        root.first = 0; root.count = n; root.left = kNone;
        D_TRY(hipMemcpyAsync(d_nodes.p, &root, sizeof root, hipMemcpyHostToDevice, st));
        D_TRY(hipMemsetAsync(d_pos.p, 0, 4ull * n, st));
        D_TRY(hipMemsetAsync(d_slot.p, 0, 4, st));
        D_TRY(hipMemsetAsync(d_act0.p, 0, 4, st));
    }
    uint32_t nact = 1, node_count = 1, level = 0;
    uint32_t* act = d_act0.p;
    uint32_t* nxt = d_act1.p;
    size_t bins_cap = 0;
    while (nact > 0) {
        if ((size_t)nact * kBins > bins_cap) {
            bins_cap = std::max((size_t)nact * kBins, (size_t)2 * bins_cap);
            if (d_bins.p) { (void)hipStreamSynchronize(st); (void)hipFree(d_bins.p); d_bins.p = nullptr; }
            D_TRY(d_bins.alloc(bins_cap));
        }
        hipLaunchKernelGGL(k_bins_reset, dim3(blocks((uint64_t)nact * kBins)), dim3(256), 0, st, d_bins.p, nact);
        hipLaunchKernelGGL(k_bins_accum, dim3(blocks((n + kChunk - 1) / kChunk)), dim3(256), 0, st, d_w.p, ids,
                           d_pos.p, d_slot.p, d_nodes.p, n, d_bins.p);
        hipLaunchKernelGGL(k_split, dim3(blocks(nact, 64)), dim3(64), 0, st, d_nodes.p, act, nact, d_bins.p, level,
                           d_split.p, d_kids.p);
        D_TRY(hipGetLastError());
        D_TRY(hipMemsetAsync(d_kids.p + nact, 0, 4, st));
        D_TRY(hipcub::DeviceScan::ExclusiveSum(d_tmp.p, scan_bytes, d_kids.p, d_scan.p, (int)nact + 1, st));
        uint32_t tot = 0;
        D_TRY(hipMemcpyAsync(&tot, d_scan.p + nact, 4, hipMemcpyDeviceToHost, st));
        hipLaunchKernelGGL(k_make_kids, dim3(blocks(nact)), dim3(256), 0, st, d_nodes.p, act, nact, d_split.p,
                           d_kids.p, d_scan.p, node_count, nxt, d_slot.p);
        hipLaunchKernelGGL(k_assign, dim3(blocks(n)), dim3(256), 0, st, d_nodes.p, d_pos.p, n);
        D_TRY(hipGetLastError());
        D_TRY(hipStreamSynchronize(st));
        node_count += 2 * tot;
        nact = 2 * tot;
        std::swap(act, nxt);
        level++;
        if (level > 4096 || node_count > cap) {
            err = "device BVH build did not converge";
            return false;
        }
    }
    // collapse
    const uint32_t n_internal = (node_count - 1) / 2;
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
        Item root{0, 0};
        D_TRY(hipMemcpyAsync(d_it0.p, &root, sizeof root, hipMemcpyHostToDevice, st));
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
                           d_ck.p, d_si.p, d_stri.p, out_nodes, out_tris, d_w.p, ids, d_out.p, d_otri.p, nitems_buf);
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
