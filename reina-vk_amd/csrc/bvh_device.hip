// bvh_device.hip — BVH8 construction on the GPU (REINA_RT_BUILD_DEVICE), the
// replacement of the driver's BLAS/TLAS build (src/graphics/Blas.cpp:8-124,
// src/tools/vktools.cpp:460-596) for flattened scenes.
//
//   1. triangle boxes and centroids; scene box and centroid bounds
//   2. a binary BVH, top-down and breadth-first by the surface area heuristic:
//      - a task (a contiguous range of the triangle reference array with its
//        box and centroid bounds) of more than kSmall triangles is one
//        workgroup: 32 bins per axis over the centroid bounds (counts and bin
//        boxes by LDS atomics), prefix and suffix sums over the bins with
//        the SAH cost area(L)·N(L) + area(R)·N(R) evaluated in FP64, the
//        cheapest split over the three axes, then a stable partition of the
//        range into the other of two reference buffers (block-wide scans of
//        the side flags) and the exact box and centroid bounds of both sides;
//        as the host builder, but every level in one launch;
//      - a task of at most kSmall triangles is one wave, which builds its
//        whole subtree in registers: per range and axis a bitonic sort of the
//        centroids across the 64 lanes, prefix and suffix boxes by lane
//        shuffles, and the exact (unbinned) SAH split, FP64 cost.
//      A binary tree over positions [0, n) cuts each position boundary m
//      exactly once, so the internal node that splits its range at m gets the
//      id n + m - 1 and the leaf at position p the id p: ids are
//      deterministic without counters, and a child writes its id into its
//      parent when it splits. Nodes that cannot be split (all centroids equal)
//      are halved by position.
//   3. collapse to BVH8 level by level: a subtree of <= 3 triangles is a
//      leaf unless splitting it is cheaper by the SAH; a node opens its
//      largest-area internal child until it has 8 children, children take
//      octant slots as in the host builder, and node / triangle offsets come
//      from exclusive scans (this file's own), so the result is deterministic.
// The output has the host builder's format (Bvh8Node, leaf-ordered triangles,
// boxes padded outward by one ulp), so the traversal, the quantizer and the
// oracle's restatement apply unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bvh_host.h"

namespace rt {
namespace {

constexpr uint32_t kLeafMax = 3;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kBins = 32;
constexpr uint32_t kSmall = 64;   // tasks of at most this many triangles: one wave each

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

// scene box (b) and centroid bounds (c)
struct SceneBounds {
    Box6 b, c;
};

__device__ __forceinline__ void tri_box(const float* w, float* lo, float* hi) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = fminf(fminf(w[a], w[3 + a]), w[6 + a]);
        hi[a] = fmaxf(fmaxf(w[a], w[3 + a]), w[6 + a]);
    }
}

__device__ __forceinline__ float half_area(float x0, float y0, float z0, float x1, float y1, float z1) {
    const float dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
    return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Triangle boxes (plo = lo.xyz, phi = hi.xyz), the identity reference order
// and the scene's box and centroid bounds.
__global__ void k_prep(const float* __restrict__ wtri, uint32_t n, float4* __restrict__ plo, float4* __restrict__ phi,
                       uint32_t* __restrict__ ref, SceneBounds* sb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = i < n;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (ok) {
        tri_box(wtri + 9ull * i, lo, hi);
        plo[i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        phi[i] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        ref[i] = i;
#pragma unroll
        for (int a = 0; a < 3; ++a) clo[a] = chi[a] = 0.5f * lo[a] + 0.5f * hi[a];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float bl = wave_min(lo[a]), bh = wave_max(hi[a]), cl = wave_min(clo[a]), ch = wave_max(chi[a]);
        if (__lane_id() == 0 && bl <= bh) {
            atomicMin(&sb->b.lo[a], ord_key(bl));
            atomicMax(&sb->b.hi[a], ord_key(bh));
            atomicMin(&sb->c.lo[a], ord_key(cl));
            atomicMax(&sb->c.hi[a], ord_key(ch));
        }
    }
}

// ---- binary build --------------------------------------------------------------
struct BNode {
    uint32_t left, right;   // children; left = kNone for a triangle
    uint32_t count;         // triangles below
    uint32_t tri;           // triangle index (leaves)
    float lo[3], hi[3];
};

// A range [begin, begin + count) of reference buffer `buf` to split; its node
// id is known only once it splits, so it tells its parent then.
struct Task {
    uint32_t begin, count, parent, side;   // side: 0 left child, 1 right child
    uint32_t buf, pad0, pad1, pad2;
    float lo[3], hi[3];     // box
    float clo[3], chi[3];   // centroid bounds
};

struct SmallTask {
    uint32_t begin, count, parent, meta;   // meta: side | buf << 1
};

struct BuildCtl {
    uint32_t n_next;    // tasks appended for the next level
    uint32_t n_small;   // small tasks appended
    uint32_t root;
    uint32_t pad;
};

__device__ __forceinline__ void set_child(BNode* nodes, uint32_t parent, uint32_t side, uint32_t id, BuildCtl* ctl) {
    if (parent == kNone) ctl->root = id;
    else if (side == 0u) nodes[parent].left = id;
    else nodes[parent].right = id;
}

__device__ __forceinline__ void write_leaf(BNode* nodes, uint32_t pos, uint32_t tri, float4 lo, float4 hi) {
    BNode b;
    b.left = b.right = kNone;
    b.count = 1;
    b.tri = tri;
    b.lo[0] = lo.x; b.lo[1] = lo.y; b.lo[2] = lo.z;
    b.hi[0] = hi.x; b.hi[1] = hi.y; b.hi[2] = hi.z;
    nodes[pos] = b;
}

// Appends a child range as a task of the next level (> kSmall triangles) or
// a small task; a single triangle was already written as a leaf.
__device__ void emit_child(uint32_t begin, uint32_t count, uint32_t parent, uint32_t side, uint32_t buf,
                           const float* lo, const float* hi, const float* clo, const float* chi, Task* next,
                           SmallTask* small, BuildCtl* ctl) {
    if (count < 2u) return;
    if (count <= kSmall) {
        const uint32_t k = atomicAdd(&ctl->n_small, 1u);
        small[k] = SmallTask{begin, count, parent, side | (buf << 1)};
        return;
    }
    Task t;
    t.begin = begin; t.count = count; t.parent = parent; t.side = side;
    t.buf = buf; t.pad0 = t.pad1 = t.pad2 = 0u;
#pragma unroll
    for (int a = 0; a < 3; ++a) { t.lo[a] = lo[a]; t.hi[a] = hi[a]; t.clo[a] = clo[a]; t.chi[a] = chi[a]; }
    next[atomicAdd(&ctl->n_next, 1u)] = t;
}

// One workgroup per task of more than kSmall triangles: binned SAH split,
// stable partition into the other buffer, the node and its children.
__global__ __launch_bounds__(256) void k_split_tasks(const Task* __restrict__ tasks, uint32_t n,
                                                     const float4* __restrict__ plo, const float4* __restrict__ phi,
                                                     uint32_t* __restrict__ ref0, uint32_t* __restrict__ ref1,
                                                     BNode* __restrict__ nodes, Task* __restrict__ next,
                                                     SmallTask* __restrict__ small, BuildCtl* ctl) {
    __shared__ uint32_t s_cnt[3][kBins];
    __shared__ uint32_t s_key[3][kBins][6];
    __shared__ double s_cost[3];
    __shared__ int s_split[3];
    __shared__ uint32_t s_wl[4];
    __shared__ uint32_t s_side[2][12];   // exact bounds of both sides (ordered keys)
    const Task T = tasks[blockIdx.x];
    const uint32_t tid = threadIdx.x, lane = __lane_id(), w = tid >> 6;
    const uint32_t* src = T.buf ? ref1 : ref0;
    uint32_t* dst = T.buf ? ref0 : ref1;
    float sc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float ext = T.chi[a] - T.clo[a];
        sc[a] = ext > 0.0f ? (float)kBins / ext : 0.0f;
    }
    for (uint32_t k = tid; k < 3u * kBins; k += 256u) {
        s_cnt[k / kBins][k % kBins] = 0u;
        for (int q = 0; q < 3; ++q) {
            s_key[k / kBins][k % kBins][q] = 0xFFFFFFFFu;
            s_key[k / kBins][k % kBins][3 + q] = 0u;
        }
    }
    if (tid < 24u) s_side[tid / 12u][tid % 12u] = (tid % 12u) % 6u < 3u ? 0xFFFFFFFFu : 0u;
    __syncthreads();
    // binning
    for (uint32_t i = tid; i < T.count; i += 256u) {
        const uint32_t t = src[T.begin + i];
        const float4 lo = plo[t], hi = phi[t];
        const float b[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (sc[a] == 0.0f) continue;
            const float c = 0.5f * b[a] + 0.5f * b[3 + a];
            int k = (int)((c - T.clo[a]) * sc[a]);
            k = min(max(k, 0), kBins - 1);
            atomicAdd(&s_cnt[a][k], 1u);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                atomicMin(&s_key[a][k][q], ord_key(b[q]));
                atomicMax(&s_key[a][k][3 + q], ord_key(b[3 + q]));
            }
        }
    }
    __syncthreads();
    // the cheapest split per axis: wave a, lane = bin
    if (w < 3u) {
        const int a = (int)w;
        const bool bin = lane < (uint32_t)kBins && sc[a] != 0.0f;
        uint32_t cnt = bin ? s_cnt[a][lane] : 0u;
        float lo[3], hi[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            lo[q] = (bin && cnt) ? key_ord(s_key[a][lane][q]) : INFINITY;
            hi[q] = (bin && cnt) ? key_ord(s_key[a][lane][3 + q]) : -INFINITY;
        }
        uint32_t pc = cnt, qc = cnt;
        float plo_[3], phi_[3], qlo[3], qhi[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) { plo_[q] = qlo[q] = lo[q]; phi_[q] = qhi[q] = hi[q]; }
        for (int d = 1; d < kBins; d <<= 1) {   // prefix over bins 0..lane, suffix over lane..31
            const uint32_t uc = __shfl_up(pc, d), dc = __shfl_down(qc, d);
            float ul[3], uh[3], dl[3], dh[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                ul[q] = __shfl_up(plo_[q], d); uh[q] = __shfl_up(phi_[q], d);
                dl[q] = __shfl_down(qlo[q], d); dh[q] = __shfl_down(qhi[q], d);
            }
            if (lane >= (uint32_t)d) {
                pc += uc;
#pragma unroll
                for (int q = 0; q < 3; ++q) { plo_[q] = fminf(plo_[q], ul[q]); phi_[q] = fmaxf(phi_[q], uh[q]); }
            }
            if (lane + (uint32_t)d < (uint32_t)kBins) {
                qc += dc;
#pragma unroll
                for (int q = 0; q < 3; ++q) { qlo[q] = fminf(qlo[q], dl[q]); qhi[q] = fmaxf(qhi[q], dh[q]); }
            }
        }
        // split after bin `lane`: left = bins 0..lane, right = bins lane+1..31
        const uint32_t rc = __shfl_down(qc, 1);
        float rl[3], rh[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) { rl[q] = __shfl_down(qlo[q], 1); rh[q] = __shfl_down(qhi[q], 1); }
        double cost = INFINITY;
        if (bin && lane + 1u < (uint32_t)kBins && pc > 0u && rc > 0u)
            cost = (double)half_area(plo_[0], plo_[1], plo_[2], phi_[0], phi_[1], phi_[2]) * (double)pc +
                   (double)half_area(rl[0], rl[1], rl[2], rh[0], rh[1], rh[2]) * (double)rc;
        int k = (int)lane;
        for (int o = 32; o > 0; o >>= 1) {   // argmin, ties to the lower bin
            const double oc = __shfl_xor(cost, o);
            const int ok = __shfl_xor(k, o);
            if (oc < cost || (oc == cost && ok < k)) { cost = oc; k = ok; }
        }
        if (lane == 0) { s_cost[a] = cost; s_split[a] = k; }
    }
    __syncthreads();
    int axis = -1, split = 0;
    double best = INFINITY;
    for (int a = 0; a < 3; ++a)
        if (s_cost[a] < best) { best = s_cost[a]; axis = a; split = s_split[a]; }
    // left count
    uint32_t nl = 0;
    if (axis >= 0) {
        for (int k = 0; k <= split; ++k) nl += s_cnt[axis][k];
    } else {
        nl = T.count / 2u;   // all centroids equal: halve by position
    }
    const uint32_t nr = T.count - nl;
    const uint32_t id = n + T.begin + nl - 1u;
    // stable partition into dst, exact bounds of both sides
    float sb[2][12];
#pragma unroll
    for (int q = 0; q < 12; ++q) { sb[0][q] = sb[1][q] = (q % 6) < 3 ? INFINITY : -INFINITY; }
    uint32_t offl = 0, offr = 0;
    for (uint32_t base = 0; base < T.count; base += 256u) {
        const uint32_t i = base + tid;
        const bool valid = i < T.count;
        uint32_t t = 0;
        float4 lo = make_float4(0.0f, 0.0f, 0.0f, 0.0f), hi = lo;
        bool left = false;
        if (valid) {
            t = src[T.begin + i];
            lo = plo[t];
            hi = phi[t];
            if (axis >= 0) {
                const float c = 0.5f * (axis == 0 ? lo.x : axis == 1 ? lo.y : lo.z) +
                                0.5f * (axis == 0 ? hi.x : axis == 1 ? hi.y : hi.z);
                int k = (int)((c - T.clo[axis]) * sc[axis]);
                k = min(max(k, 0), kBins - 1);
                left = k <= split;
            } else {
                left = i < nl;
            }
        }
        const uint64_t bal = __ballot(left);
        const uint32_t below = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) s_wl[w] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t wb = 0, tot = 0;
        for (uint32_t q = 0; q < 4; ++q) {
            if (q < w) wb += s_wl[q];
            tot += s_wl[q];
        }
        __syncthreads();
        if (valid) {
            const uint32_t rl = wb + below;
            const uint32_t pos = left ? T.begin + offl + rl : T.begin + nl + offr + (tid - rl);
            dst[pos] = t;
            if (left && nl == 1u) write_leaf(nodes, pos, t, lo, hi);
            if (!left && nr == 1u) write_leaf(nodes, pos, t, lo, hi);
            const int sd = left ? 0 : 1;
            const float b[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const float c = 0.5f * b[q] + 0.5f * b[3 + q];
                sb[sd][q] = fminf(sb[sd][q], b[q]);
                sb[sd][3 + q] = fmaxf(sb[sd][3 + q], b[3 + q]);
                sb[sd][6 + q] = fminf(sb[sd][6 + q], c);
                sb[sd][9 + q] = fmaxf(sb[sd][9 + q], c);
            }
        }
        const uint32_t valid_n = min(256u, T.count - base);
        offl += tot;
        offr += valid_n - tot;
    }
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            const bool is_min = (q % 6) < 3;
            const float v = is_min ? wave_min(sb[sd][q]) : wave_max(sb[sd][q]);
            if (lane == 0 && v == v && fabsf(v) != INFINITY) {
                if (is_min) atomicMin(&s_side[sd][q], ord_key(v));
                else atomicMax(&s_side[sd][q], ord_key(v));
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        BNode b;
        b.left = nl == 1u ? T.begin : kNone;
        b.right = nr == 1u ? T.begin + nl : kNone;
        b.count = T.count;
        b.tri = kNone;
#pragma unroll
        for (int a = 0; a < 3; ++a) { b.lo[a] = T.lo[a]; b.hi[a] = T.hi[a]; }
        nodes[id] = b;
        set_child(nodes, T.parent, T.side, id, ctl);
        float v[2][12];
        for (int sd = 0; sd < 2; ++sd)
            for (int q = 0; q < 12; ++q) v[sd][q] = key_ord(s_side[sd][q]);
        emit_child(T.begin, nl, id, 0u, T.buf ^ 1u, v[0], v[0] + 3, v[0] + 6, v[0] + 9, next, small, ctl);
        emit_child(T.begin + nl, nr, id, 1u, T.buf ^ 1u, v[1], v[1] + 3, v[1] + 6, v[1] + 9, next, small, ctl);
    }
}

// Lexicographic (key, payload) bitonic sort across the 64 lanes, ascending.
__device__ __forceinline__ void wave_sort(float& key, uint32_t& pay, uint32_t lane) {
    for (uint32_t k = 2; k <= 64u; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const float ok = __shfl_xor(key, (int)j);
            const uint32_t op = __shfl_xor(pay, (int)j);
            const bool other_less = ok < key || (ok == key && op < pay);
            const bool up = (lane & k) == 0u;
            const bool lower = (lane & j) == 0u;
            const bool take = (lower == up) ? other_less : !other_less;
            if (take) { key = ok; pay = op; }
        }
    }
}

// One wave per task of 2..kSmall triangles: the whole subtree by the exact
// SAH over centroid-sorted orders, lane i holding the triangle at position
// begin + i; pending ranges are a stack held one entry per lane.
__global__ __launch_bounds__(256) void k_small_tasks(const SmallTask* __restrict__ small, uint32_t nsmall, uint32_t n,
                                                     const float4* __restrict__ plo, const float4* __restrict__ phi,
                                                     const uint32_t* __restrict__ ref0,
                                                     const uint32_t* __restrict__ ref1, BNode* __restrict__ nodes,
                                                     BuildCtl* ctl) {
    const uint32_t ti = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (ti >= nsmall) return;   // whole waves only; no block barrier below
    const uint32_t lane = __lane_id();
    const SmallTask T = small[ti];
    const uint32_t* src = (T.meta >> 1) ? ref1 : ref0;
    uint32_t id = kNone;
    float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
    if (lane < T.count) {
        id = src[T.begin + lane];
        lo = plo[id];
        hi = phi[id];
    }
    // stack entry k lives in lane k: (r0, size, parent, side)
    uint32_t s_r0 = 0, s_sz = 0, s_par = 0, s_side = 0;
    if (lane == 0) { s_r0 = 0; s_sz = T.count; s_par = T.parent; s_side = T.meta & 1u; }
    int sp = 1;
    while (sp > 0) {
        --sp;
        const uint32_t r0 = __shfl(s_r0, sp), sz = __shfl(s_sz, sp), par = __shfl(s_par, sp), side = __shfl(s_side, sp);
        const bool in = lane >= r0 && lane < r0 + sz;
        // node box
        const float bx0 = wave_min(in ? lo.x : INFINITY), by0 = wave_min(in ? lo.y : INFINITY);
        const float bz0 = wave_min(in ? lo.z : INFINITY);
        const float bx1 = wave_max(in ? hi.x : -INFINITY), by1 = wave_max(in ? hi.y : -INFINITY);
        const float bz1 = wave_max(in ? hi.z : -INFINITY);
        double best = INFINITY;
        uint32_t best_k = 0, best_perm = lane;
        for (int a = 0; a < 3; ++a) {
            const float la = a == 0 ? lo.x : a == 1 ? lo.y : lo.z, ha = a == 0 ? hi.x : a == 1 ? hi.y : hi.z;
            float key = in ? 0.5f * la + 0.5f * ha : INFINITY;
            uint32_t pay = lane;
            wave_sort(key, pay, lane);
            // lane k (< sz) now holds the k-th range triangle in centroid order
            float l3[3] = {__shfl(lo.x, (int)pay), __shfl(lo.y, (int)pay), __shfl(lo.z, (int)pay)};
            float h3[3] = {__shfl(hi.x, (int)pay), __shfl(hi.y, (int)pay), __shfl(hi.z, (int)pay)};
            const bool mine = lane < sz;
            float pl[3], ph[3], ql[3], qh[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                pl[q] = ql[q] = mine ? l3[q] : INFINITY;
                ph[q] = qh[q] = mine ? h3[q] : -INFINITY;
            }
            for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const float ul = __shfl_up(pl[q], d), uh = __shfl_up(ph[q], d);
                    const float dl = __shfl_down(ql[q], d), dh = __shfl_down(qh[q], d);
                    if (lane >= (uint32_t)d) { pl[q] = fminf(pl[q], ul); ph[q] = fmaxf(ph[q], uh); }
                    if (lane + (uint32_t)d < 64u) { ql[q] = fminf(ql[q], dl); qh[q] = fmaxf(qh[q], dh); }
                }
            }
            float rl[3], rh[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) { rl[q] = __shfl_down(ql[q], 1); rh[q] = __shfl_down(qh[q], 1); }
            // split after sorted position `lane`: left = 0..lane, right = lane+1..sz-1
            double cost = INFINITY;
            if (lane + 1u < sz)
                cost = (double)half_area(pl[0], pl[1], pl[2], ph[0], ph[1], ph[2]) * (double)(lane + 1u) +
                       (double)half_area(rl[0], rl[1], rl[2], rh[0], rh[1], rh[2]) * (double)(sz - lane - 1u);
            uint32_t k = lane;
            for (int o = 32; o > 0; o >>= 1) {
                const double oc = __shfl_xor(cost, o);
                const uint32_t okk = __shfl_xor(k, o);
                if (oc < cost || (oc == cost && okk < k)) { cost = oc; k = okk; }
            }
            if (cost < best) {   // ties keep the lower axis
                best = cost;
                best_k = k;
                best_perm = pay;
            }
        }
        // reorder the range: position r0 + k takes the k-th triangle of the best
        // order (best_perm holds the lane it came from); every lane takes part
        // in each shuffle, so no source lane is inactive
        const uint32_t sel = __shfl(best_perm, (int)(in ? lane - r0 : lane));
        const uint32_t from = in ? sel : lane;
        const uint32_t nid = __shfl(id, (int)from);
        const float4 nlo = make_float4(__shfl(lo.x, (int)from), __shfl(lo.y, (int)from), __shfl(lo.z, (int)from), 0.0f);
        const float4 nhi = make_float4(__shfl(hi.x, (int)from), __shfl(hi.y, (int)from), __shfl(hi.z, (int)from), 0.0f);
        id = nid;
        lo = nlo;
        hi = nhi;
        const uint32_t nl = best_k + 1u, nr = sz - nl;
        const uint32_t node = n + T.begin + r0 + nl - 1u;
        if (nl == 1u && lane == r0) write_leaf(nodes, T.begin + r0, id, lo, hi);
        if (nr == 1u && lane == r0 + nl) write_leaf(nodes, T.begin + r0 + nl, id, lo, hi);
        if (lane == 0) {
            BNode b;
            b.left = nl == 1u ? T.begin + r0 : kNone;
            b.right = nr == 1u ? T.begin + r0 + nl : kNone;
            b.count = sz;
            b.tri = kNone;
            b.lo[0] = bx0; b.lo[1] = by0; b.lo[2] = bz0;
            b.hi[0] = bx1; b.hi[1] = by1; b.hi[2] = bz1;
            nodes[node] = b;
            set_child(nodes, par, side, node, ctl);
        }
        // push right, then left (the left range is split next)
        if (nr >= 2u) {
            if (lane == (uint32_t)sp) { s_r0 = r0 + nl; s_sz = nr; s_par = node; s_side = 1u; }
            ++sp;
        }
        if (nl >= 2u) {
            if (lane == (uint32_t)sp) { s_r0 = r0; s_sz = nl; s_par = node; s_side = 0u; }
            ++sp;
        }
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

// SAH-optimal collapse (Ylitie, Karras and Laine, "Efficient incoherent ray
// traversal on GPUs through compressed wide BVHs", HPG 2017): for every
// binary node, C[i-1] is the least surface-area cost of its subtree spread
// over i child slots of a BVH8 node (i = 1..8), computed bottom-up:
//   C(x, 1) = min(leaf: A x kCollapseTriCost x count (count <= kLeafMax),
//                 node: A x kCollapseNodeCost + D(x, 8))
//   C(x, i) = min(C(x, i - 1), D(x, i)),  D(x, i) = min_k C(l, k) + C(r, i - k).
// k[i-1] keeps the left child's slots of D(x, i) when that is the minimum
// (0: C(x, i) = C(x, i - 1)); split8 is the argmin of D(x, 8), the 8 slots of
// x's own BVH8 node. A triangle (id < n) is a leaf of cost A x kCollapseTriCost.
struct DpRec {
    float c[8];
    uint8_t k[8];
    uint8_t leaf, split8, pad0, pad1;
};

__device__ __forceinline__ float dp_cost(const BNode* nodes, const DpRec* dp, uint32_t n, uint32_t id, int slots) {
    return id < n ? area_f(nodes[id]) * kCollapseTriCost : dp[id - n].c[slots - 1];
}

// Parent of every binary node (the root's stays kNone) and cleared arrival counters.
__global__ void k_parents(const BNode* __restrict__ nodes, uint32_t n, uint32_t* __restrict__ parent,
                          uint32_t* __restrict__ arrive) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1u) return;
    const uint32_t x = n + i;
    parent[nodes[x].left] = x;
    parent[nodes[x].right] = x;
    arrive[i] = 0u;
}

// One thread per triangle walks up; the second thread to reach a node (both
// children done) computes its record and goes on, the first stops. The
// fences publish a record before its arrival is counted and order the
// sibling's record before it is read (device scope: the XCDs' L2s).
__global__ void k_collapse_dp(const BNode* __restrict__ nodes, uint32_t n, const uint32_t* __restrict__ parent,
                              uint32_t* __restrict__ arrive, DpRec* dp) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t x = parent[t];
    while (x != kNone) {
        __threadfence();
        if (atomicAdd(&arrive[x - n], 1u) == 0u) return;
        __threadfence();
        const BNode& b = nodes[x];
        float cl[8], cr[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            cl[j] = dp_cost(nodes, dp, n, b.left, j + 1);
            cr[j] = dp_cost(nodes, dp, n, b.right, j + 1);
        }
        DpRec r;
        float dist[8];
        uint8_t dk[8];
        dist[0] = INFINITY;
        dk[0] = 0;
#pragma unroll
        for (int j = 2; j <= 8; ++j) {
            float best = INFINITY;
            int bk = 1;
            for (int k = 1; k < j; ++k) {
                const float v = cl[k - 1] + cr[j - k - 1];
                if (v < best) { best = v; bk = k; }
            }
            dist[j - 1] = best;
            dk[j - 1] = (uint8_t)bk;
        }
        const float A = area_f(b);
        const float cint = A * kCollapseNodeCost + dist[7];
        const float cleaf = b.count <= kLeafMax ? A * kCollapseTriCost * (float)b.count : INFINITY;
        r.leaf = cleaf <= cint ? 1 : 0;
        r.split8 = dk[7];
        r.pad0 = r.pad1 = 0;
        r.c[0] = fminf(cleaf, cint);
        r.k[0] = 0;
#pragma unroll
        for (int j = 2; j <= 8; ++j) {
            if (dist[j - 1] < r.c[j - 2]) { r.c[j - 1] = dist[j - 1]; r.k[j - 1] = dk[j - 1]; }
            else { r.c[j - 1] = r.c[j - 2]; r.k[j - 1] = 0; }
        }
        dp[x - n] = r;
        x = parent[x];
    }
}

// Whether binary node x becomes a BVH8 node (else a leaf slot).
__device__ __forceinline__ bool inner(const DpRec* dp, uint32_t n, uint32_t x) {
    return x >= n && dp[x - n].leaf == 0;
}

// Children of one BVH8 node in slot order (kNone = empty slot): the 8-slot
// distribution of binary node bn, expanded.
__device__ void collapse_children(const BNode* nodes, const DpRec* dp, uint32_t n, uint32_t bn, uint32_t* slot_child) {
    uint32_t ch[8];
    int nc = 0;
    const BNode& b = nodes[bn];
    if (!inner(dp, n, bn)) {
        ch[nc++] = bn;   // a leaf root
    } else {
        uint32_t sx[8];
        int sj[8], sp = 0;
        const int k8 = dp[bn - n].split8;
        sx[sp] = b.right; sj[sp++] = 8 - k8;
        sx[sp] = b.left; sj[sp++] = k8;
        while (sp > 0) {
            const uint32_t m = sx[--sp];
            int j = sj[sp];
            if (m < n) j = 1;
            while (j > 1 && dp[m - n].k[j - 1] == 0) j--;
            if (j == 1) { ch[nc++] = m; continue; }
            const int k = dp[m - n].k[j - 1];
            const BNode& o = nodes[m];
            sx[sp] = o.right; sj[sp++] = j - k;
            sx[sp] = o.left; sj[sp++] = k;
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

__global__ void k_collapse_count(const BNode* __restrict__ nodes, const DpRec* __restrict__ dp, uint32_t n,
                                 const Item* __restrict__ items, uint32_t nitems,
                                 uint32_t* __restrict__ kids, uint32_t* __restrict__ n_int, uint32_t* __restrict__ n_tri) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitems) return;
    uint32_t sc[8];
    collapse_children(nodes, dp, n, items[i].bnode, sc);
    uint32_t ni = 0, nt = 0;
    for (int s = 0; s < 8; ++s) {
        kids[8 * i + s] = sc[s];
        if (sc[s] == kNone) continue;
        if (inner(dp, n, sc[s])) ni++;
        else nt += nodes[sc[s]].count;
    }
    n_int[i] = ni;
    n_tri[i] = nt;
}

__global__ void k_collapse_write(const BNode* __restrict__ nodes, const DpRec* __restrict__ dp, uint32_t n,
                                 const Item* __restrict__ items, uint32_t nitems,
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
        if (inner(dp, n, c)) {
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
struct DBuf {   // owns one device allocation; not copyable (a copy would free it twice)
    T* p = nullptr;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t n) { return hipMalloc(&p, (n ? n : 1) * sizeof(T)); }
};

inline uint32_t blocks(uint64_t n, uint32_t b = 256) { return (uint32_t)((n + b - 1) / b); }

// ---- exclusive scan of uint32 (node and triangle offsets of the collapse) ------
constexpr uint32_t kScanBlock = 1024;   // 256 threads x 4 elements

__global__ __launch_bounds__(256) void k_scan_blocks(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                     uint32_t n, uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_w[4];
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * kScanBlock + 4ull * threadIdx.x;
    uint32_t v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = base + q < n ? in[base + q] : 0u;
    const uint32_t t0 = v[0], t1 = t0 + v[1], t2 = t1 + v[2], t3 = t2 + v[3];
    uint32_t x = t3;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63u) s_w[w] = x;
    __syncthreads();
    uint32_t woff = 0, total = 0;
    for (uint32_t q = 0; q < 4u; ++q) {
        if (q < w) woff += s_w[q];
        total += s_w[q];
    }
    const uint32_t ex = woff + x - t3;
    const uint32_t r[4] = {ex, ex + t0, ex + t1, ex + t2};
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (base + q < n) out[base + q] = r[q];
    if (threadIdx.x == 0 && sums) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* __restrict__ out, uint32_t n,
                                                  const uint32_t* __restrict__ offs) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanBlock + 4ull * threadIdx.x;
    const uint32_t o = offs[blockIdx.x];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (base + q < n) out[base + q] += o;
}

// Scratch for the per-block sums of every recursion level (2^32 elements
// need at most 4 levels of 1024-element blocks).
struct ScanTmp {
    static constexpr int kLevels = 4;
    DBuf<uint32_t> lv[2 * kLevels];
    hipError_t ensure(uint32_t n) {
        size_t k = 0;
        for (uint32_t m = n; m > kScanBlock; m = (m + kScanBlock - 1) / kScanBlock, ++k) {
            const uint32_t nb = (m + kScanBlock - 1) / kScanBlock;
            if (k >= (size_t)kLevels) return hipErrorInvalidValue;
            if (!lv[2 * k].p) {
                hipError_t e = lv[2 * k].alloc(nb);
                if (e != hipSuccess) return e;
                e = lv[2 * k + 1].alloc(nb);
                if (e != hipSuccess) return e;
            }
        }
        return hipSuccess;
    }
};

// out[i] = in[0] + ... + in[i - 1]; tmp must have been sized for n by ensure().
hipError_t scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, ScanTmp& tmp, size_t level, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(256), 0, st, in, out, n, nullptr);
        return hipGetLastError();
    }
    uint32_t* sums = tmp.lv[2 * level].p;
    uint32_t* offs = tmp.lv[2 * level + 1].p;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, st, in, out, n, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = scan_u32(sums, offs, nb, tmp, level + 1, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(256), 0, st, out, n, offs);
    return hipGetLastError();
}

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
    DBuf<float4> d_plo, d_phi;
    DBuf<uint32_t> d_ref0, d_ref1;
    DBuf<SceneBounds> d_sb;
    DBuf<BuildCtl> d_ctl;
    DBuf<BNode> d_nodes;
    D_TRY(d_plo.alloc(n)); D_TRY(d_phi.alloc(n)); D_TRY(d_ref0.alloc(n)); D_TRY(d_ref1.alloc(n));
    D_TRY(d_sb.alloc(1)); D_TRY(d_ctl.alloc(1)); D_TRY(d_nodes.alloc(2ull * n));
    {
        SceneBounds init;
        for (int a = 0; a < 3; ++a) {
            init.b.lo[a] = init.c.lo[a] = 0xFFFFFFFFu;
            init.b.hi[a] = init.c.hi[a] = 0u;
        }
        D_TRY(hipMemcpyAsync(d_sb.p, &init, sizeof init, hipMemcpyHostToDevice, st));
        BuildCtl c{0u, 0u, kNone, 0u};
        D_TRY(hipMemcpyAsync(d_ctl.p, &c, sizeof c, hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_prep, dim3(blocks(n)), dim3(256), 0, st, d_w.p, n, d_plo.p, d_phi.p, d_ref0.p, d_sb.p);
    D_TRY(hipGetLastError());
    SceneBounds sb;
    D_TRY(hipMemcpyAsync(&sb, d_sb.p, sizeof sb, hipMemcpyDeviceToHost, st));
    D_TRY(hipStreamSynchronize(st));
    auto unkey = [](uint32_t k) {
        const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    };
    // binary build: levels of large tasks, then every small task
    const uint32_t cap_large = n / (kSmall + 1) + 2;
    const uint32_t cap_small = n / 2 + 2;
    DBuf<Task> d_t0, d_t1;
    DBuf<SmallTask> d_small;
    D_TRY(d_t0.alloc(cap_large)); D_TRY(d_t1.alloc(cap_large)); D_TRY(d_small.alloc(cap_small));
    uint32_t root = kNone;
    uint32_t n_small = 0;
    if (n == 1) {
        root = 0;   // a single triangle: the root is its leaf
        float4 lo = make_float4(unkey(sb.b.lo[0]), unkey(sb.b.lo[1]), unkey(sb.b.lo[2]), 0.0f);
        float4 hi = make_float4(unkey(sb.b.hi[0]), unkey(sb.b.hi[1]), unkey(sb.b.hi[2]), 0.0f);
        BNode b;
        b.left = b.right = kNone;
        b.count = 1;
        b.tri = 0;
        b.lo[0] = lo.x; b.lo[1] = lo.y; b.lo[2] = lo.z;
        b.hi[0] = hi.x; b.hi[1] = hi.y; b.hi[2] = hi.z;
        D_TRY(hipMemcpyAsync(d_nodes.p, &b, sizeof b, hipMemcpyHostToDevice, st));
    } else if (n <= kSmall) {
        SmallTask t{0u, n, kNone, 0u};
        D_TRY(hipMemcpyAsync(d_small.p, &t, sizeof t, hipMemcpyHostToDevice, st));
        n_small = 1;
    } else {
        Task t;
        t.begin = 0; t.count = n; t.parent = kNone; t.side = 0; t.buf = 0; t.pad0 = t.pad1 = t.pad2 = 0;
        for (int a = 0; a < 3; ++a) {
            t.lo[a] = unkey(sb.b.lo[a]); t.hi[a] = unkey(sb.b.hi[a]);
            t.clo[a] = unkey(sb.c.lo[a]); t.chi[a] = unkey(sb.c.hi[a]);
        }
        D_TRY(hipMemcpyAsync(d_t0.p, &t, sizeof t, hipMemcpyHostToDevice, st));
        uint32_t ntasks = 1, levels = 0;
        Task* cur = d_t0.p;
        Task* nxt = d_t1.p;
        while (ntasks > 0) {
            if (++levels > n) {   // every level shrinks the largest task
                err = "device BVH build did not converge";
                return false;
            }
            D_TRY(hipMemsetAsync(&d_ctl.p->n_next, 0, 4, st));
            hipLaunchKernelGGL(k_split_tasks, dim3(ntasks), dim3(256), 0, st, cur, n, d_plo.p, d_phi.p, d_ref0.p,
                               d_ref1.p, d_nodes.p, nxt, d_small.p, d_ctl.p);
            D_TRY(hipGetLastError());
            BuildCtl c;
            D_TRY(hipMemcpyAsync(&c, d_ctl.p, sizeof c, hipMemcpyDeviceToHost, st));
            D_TRY(hipStreamSynchronize(st));
            if (c.n_next > cap_large || c.n_small > cap_small) {
                err = "device BVH build task list overflow";
                return false;
            }
            ntasks = c.n_next;
            n_small = c.n_small;
            std::swap(cur, nxt);
        }
    }
    if (n_small) {
        hipLaunchKernelGGL(k_small_tasks, dim3((n_small + 3) / 4), dim3(256), 0, st, d_small.p, n_small, n, d_plo.p,
                           d_phi.p, d_ref0.p, d_ref1.p, d_nodes.p, d_ctl.p);
        D_TRY(hipGetLastError());
    }
    if (n > 1) {
        BuildCtl c;
        D_TRY(hipMemcpyAsync(&c, d_ctl.p, sizeof c, hipMemcpyDeviceToHost, st));
        D_TRY(hipStreamSynchronize(st));
        root = c.root;
        if (root < n || root >= 2 * n - 1) {
            err = "device BVH build produced no root";
            return false;
        }
    }
    // collapse: the SAH-optimal slot distribution of every binary node, bottom-up
    DBuf<uint32_t> d_parent, d_arrive;
    DBuf<DpRec> d_dp;
    D_TRY(d_parent.alloc(2ull * n)); D_TRY(d_arrive.alloc(n)); D_TRY(d_dp.alloc(n));
    D_TRY(hipMemsetAsync(d_parent.p, 0xFF, sizeof(uint32_t) * (2ull * n), st));
    if (n > 1) {
        hipLaunchKernelGGL(k_parents, dim3(blocks(n - 1)), dim3(256), 0, st, d_nodes.p, n, d_parent.p, d_arrive.p);
        D_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_collapse_dp, dim3(blocks(n)), dim3(256), 0, st, d_nodes.p, n, d_parent.p, d_arrive.p,
                           d_dp.p);
        D_TRY(hipGetLastError());
    }
    const uint32_t n_internal = n - 1;
    DBuf<Bvh8Node> d_out;
    DBuf<float4> d_otri;
    DBuf<Item> d_it0, d_it1;
    DBuf<uint32_t> d_ck, d_ni, d_nt, d_si, d_stri;
    const uint32_t out_cap = std::max(1u, n_internal) + 1;
    D_TRY(d_out.alloc(out_cap)); D_TRY(d_otri.alloc(3ull * n));
    D_TRY(d_it0.alloc(out_cap)); D_TRY(d_it1.alloc(out_cap));
    D_TRY(d_ck.alloc(8ull * out_cap)); D_TRY(d_ni.alloc(out_cap + 1)); D_TRY(d_nt.alloc(out_cap + 1));
    D_TRY(d_si.alloc(out_cap + 1)); D_TRY(d_stri.alloc(out_cap + 1));
    ScanTmp scan_tmp;
    D_TRY(scan_tmp.ensure(out_cap + 1));
    {
        Item r0{root, 0};
        D_TRY(hipMemcpyAsync(d_it0.p, &r0, sizeof r0, hipMemcpyHostToDevice, st));
    }
    uint32_t nitems = 1, out_nodes = 1, out_tris = 0, depth = 0;
    Item* items = d_it0.p;
    Item* nitems_buf = d_it1.p;
    while (nitems > 0) {
        depth++;
        hipLaunchKernelGGL(k_collapse_count, dim3(blocks(nitems, 64)), dim3(64), 0, st, d_nodes.p, d_dp.p, n, items,
                           nitems, d_ck.p, d_ni.p, d_nt.p);
        D_TRY(hipGetLastError());
        D_TRY(hipMemsetAsync(d_ni.p + nitems, 0, 4, st));
        D_TRY(hipMemsetAsync(d_nt.p + nitems, 0, 4, st));
        D_TRY(scan_u32(d_ni.p, d_si.p, nitems + 1, scan_tmp, 0, st));
        D_TRY(scan_u32(d_nt.p, d_stri.p, nitems + 1, scan_tmp, 0, st));
        uint32_t tot[2] = {0, 0};
        D_TRY(hipMemcpyAsync(&tot[0], d_si.p + nitems, 4, hipMemcpyDeviceToHost, st));
        D_TRY(hipMemcpyAsync(&tot[1], d_stri.p + nitems, 4, hipMemcpyDeviceToHost, st));
        D_TRY(hipStreamSynchronize(st));
        if (out_nodes + tot[0] > out_cap || out_tris + tot[1] > n) {
            err = "device BVH collapse overflow";
            return false;
        }
        hipLaunchKernelGGL(k_collapse_write, dim3(blocks(nitems, 64)), dim3(64), 0, st, d_nodes.p, d_dp.p, n, items,
                           nitems, d_ck.p, d_si.p, d_stri.p, out_nodes, out_tris, d_w.p, d_out.p, d_otri.p, nitems_buf);
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
