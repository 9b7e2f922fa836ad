// bvh_host.cpp — BVH8 construction on the host CPU (replaces the driver's
// BLAS/TLAS build, src/graphics/Blas.cpp:8-124 and src/tools/vktools.cpp:460-596,
// for the flattened world-space triangle soup).
//
//   1. Binned SAH BVH2 (32 bins per axis, all three axes), leaves of at most
//      3 triangles; nodes of one subtree are built in parallel on std::threads
//      below the top levels.
//   2. Collapse to BVH8 by the surface-area-optimal dynamic program of
//      Ylitie, Karras and Laine ("Efficient Incoherent Ray Traversal on GPUs
//      Through Compressed Wide BVHs", HPG 2017): bottom-up, the least cost of
//      every BVH2 subtree in 1..8 slots of a BVH8 node (a leaf of at most 3
//      triangles, a BVH8 node, or a split of the slots between its children;
//      costs kCollapseNodeCost per node and kCollapseTriCost per triangle,
//      times surface area), then each BVH8 node's children are its 8-slot
//      distribution, expanded.
//   3. Children are placed in slots by octant: slot s stands for the direction
//      (+-1,+-1,+-1) of its bits, and a greedy assignment gives each slot the
//      child whose centroid lies furthest in that direction from the parent's
//      centroid, so visiting slots i ^ oct approximates front-to-back order.
//   4. Boxes are padded outward by one ulp; internal children are contiguous
//      (breadth-first), and each node's leaf triangles are contiguous.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "bvh_host.h"

namespace rt {

namespace {

struct Aabb {
    float lo[3], hi[3];
    void reset() {
        lo[0] = lo[1] = lo[2] = INFINITY;
        hi[0] = hi[1] = hi[2] = -INFINITY;
    }
    void grow(const Aabb& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct B2Node {
    Aabb box;
    int32_t left = -1, right = -1;   // children (internal)
    uint32_t first = 0, count = 0;   // leaf range into ref[]
};

constexpr int kBins = 32;

struct Builder {
    const float* tri;          // 9 floats per triangle
    std::vector<Aabb> tbox;
    std::vector<float> cen;    // 3 per triangle
    std::vector<uint32_t> ref;
    std::vector<B2Node> nodes;
    std::mutex* mtx = nullptr;

    int32_t alloc_node() {
        std::lock_guard<std::mutex> g(*mtx);
        nodes.emplace_back();
        return (int32_t)nodes.size() - 1;
    }
};

// Build the subtree over ref[first, first+count); returns node index.
void build_rec(Builder& b, int32_t ni, uint32_t first, uint32_t count, int depth, std::atomic<int>* threads,
               std::vector<std::thread>* spawned, std::mutex* spawn_mtx);

void split_node(Builder& b, int32_t ni, uint32_t first, uint32_t count, int depth, std::atomic<int>* threads,
                std::vector<std::thread>* spawned, std::mutex* spawn_mtx) {
    Aabb box, cb;
    box.reset();
    cb.reset();
    for (uint32_t i = first; i < first + count; ++i) {
        uint32_t t = b.ref[i];
        box.grow(b.tbox[t]);
        cb.grow(&b.cen[3ull * t]);
    }
    {
        std::lock_guard<std::mutex> g(*b.mtx);
        b.nodes[ni].box = box;
    }
    if (count <= 1) {
        std::lock_guard<std::mutex> g(*b.mtx);
        b.nodes[ni].first = first;
        b.nodes[ni].count = count;
        return;
    }
    // binned SAH over the three axes
    float best_cost = INFINITY;
    int best_axis = -1, best_split = -1;
    for (int ax = 0; ax < 3; ++ax) {
        float ext = cb.hi[ax] - cb.lo[ax];
        if (!(ext > 0.0f)) continue;
        Aabb bins[kBins];
        uint32_t cnt[kBins] = {0};
        for (int k = 0; k < kBins; ++k) bins[k].reset();
        float sc = (float)kBins / ext;
        for (uint32_t i = first; i < first + count; ++i) {
            uint32_t t = b.ref[i];
            int k = (int)((b.cen[3ull * t + ax] - cb.lo[ax]) * sc);
            k = std::min(std::max(k, 0), kBins - 1);
            cnt[k]++;
            bins[k].grow(b.tbox[t]);
        }
        float rarea[kBins];
        uint32_t rcnt[kBins];
        Aabb acc;
        acc.reset();
        uint32_t c = 0;
        for (int k = kBins - 1; k > 0; --k) {
            acc.grow(bins[k]);
            c += cnt[k];
            rarea[k] = acc.area();
            rcnt[k] = c;
        }
        acc.reset();
        c = 0;
        for (int k = 0; k < kBins - 1; ++k) {
            acc.grow(bins[k]);
            c += cnt[k];
            if (c == 0 || rcnt[k + 1] == 0) continue;
            float cost = acc.area() * (float)c + rarea[k + 1] * (float)rcnt[k + 1];
            if (cost < best_cost) {
                best_cost = cost;
                best_axis = ax;
                best_split = k;
            }
        }
    }
    float parea = box.area();
    float leaf_cost = (float)count;                                   // Ct = 1
    float split_cost = 1.0f + (parea > 0.0f ? best_cost / parea : INFINITY);   // Ci = 1
    if (count <= kMaxLeafTris && (best_axis < 0 || leaf_cost <= split_cost)) {
        std::lock_guard<std::mutex> g(*b.mtx);
        b.nodes[ni].first = first;
        b.nodes[ni].count = count;
        return;
    }
    uint32_t mid;
    if (best_axis < 0) {
        mid = first + count / 2;   // all centroids coincide: split by index
    } else {
        float ext = cb.hi[best_axis] - cb.lo[best_axis];
        float sc = (float)kBins / ext;
        auto it = std::partition(b.ref.begin() + first, b.ref.begin() + first + count, [&](uint32_t t) {
            int k = (int)((b.cen[3ull * t + best_axis] - cb.lo[best_axis]) * sc);
            k = std::min(std::max(k, 0), kBins - 1);
            return k <= best_split;
        });
        mid = (uint32_t)(it - b.ref.begin());
        if (mid == first || mid == first + count) mid = first + count / 2;
    }
    int32_t l = b.alloc_node(), r = b.alloc_node();
    {
        std::lock_guard<std::mutex> g(*b.mtx);
        b.nodes[ni].left = l;
        b.nodes[ni].right = r;
    }
    uint32_t lc = mid - first, rc = first + count - mid;
    // hand large subtrees to another thread near the top
    if (depth < 6 && lc > 65536 && threads->fetch_add(1) < (int)std::thread::hardware_concurrency()) {
        std::lock_guard<std::mutex> g(*spawn_mtx);
        spawned->emplace_back([&b, l, first, lc, depth, threads, spawned, spawn_mtx]() {
            build_rec(b, l, first, lc, depth + 1, threads, spawned, spawn_mtx);
        });
    } else {
        build_rec(b, l, first, lc, depth + 1, threads, spawned, spawn_mtx);
    }
    build_rec(b, r, mid, rc, depth + 1, threads, spawned, spawn_mtx);
}

void build_rec(Builder& b, int32_t ni, uint32_t first, uint32_t count, int depth, std::atomic<int>* threads,
               std::vector<std::thread>* spawned, std::mutex* spawn_mtx) {
    split_node(b, ni, first, count, depth, threads, spawned, spawn_mtx);
}

struct Child {
    int32_t b2;       // BVH2 node
    bool leaf;
};

uint8_t leaf_meta(uint32_t count, uint32_t off) { return (uint8_t)(((count & 3u) << 5) | (off & 31u)); }

}  // namespace

bool build_bvh8_host(const float* tri, uint64_t ntri, Bvh8Result& out, std::string& err) {
    out.nodes.clear();
    out.tris.clear();
    out.max_depth = 0;
    if (ntri >= (1ull << 24)) {
        err = "too many triangles for the BVH8 index encoding (max 2^24 per scene)";
        return false;
    }
    if (ntri == 0) {
        Bvh8Node root;
        std::memset(&root, 0, sizeof root);
        out.nodes.push_back(root);
        out.max_depth = 1;
        return true;
    }
    std::mutex mtx;
    Builder b;
    b.tri = tri;
    b.mtx = &mtx;
    b.tbox.resize(ntri);
    b.cen.resize(3 * ntri);
    b.ref.resize(ntri);
    for (uint64_t t = 0; t < ntri; ++t) {
        Aabb bx;
        bx.reset();
        bx.grow(&tri[9 * t]);
        bx.grow(&tri[9 * t + 3]);
        bx.grow(&tri[9 * t + 6]);
        b.tbox[t] = bx;
        for (int k = 0; k < 3; ++k) b.cen[3 * t + k] = 0.5f * (bx.lo[k] + bx.hi[k]);
        b.ref[t] = (uint32_t)t;
    }
    b.nodes.reserve(2 * ntri + 1);   // never reallocates: at most 2n-1 nodes
    b.nodes.emplace_back();
    std::atomic<int> threads{0};
    std::vector<std::thread> spawned;
    std::mutex spawn_mtx;
    build_rec(b, 0, 0, (uint32_t)ntri, 0, &threads, &spawned, &spawn_mtx);
    for (;;) {
        std::thread th;
        {
            std::lock_guard<std::mutex> g(spawn_mtx);
            if (spawned.empty()) break;
            th = std::move(spawned.back());
            spawned.pop_back();
        }
        th.join();
    }

    // ---- SAH-optimal collapse (dynamic program over the binary tree) ----------
    // Ylitie, Karras and Laine, HPG 2017, as in bvh_device.hip: C(n, i) is the
    // least surface-area cost of subtree n spread over i slots of a BVH8 node;
    // dk keeps the left child's slots of the best split (0: as with i - 1 slots).
    const size_t nb = b.nodes.size();
    std::vector<uint32_t> sub_first(nb), sub_cnt(nb);
    std::vector<float> C(nb * 8);      // C[n * 8 + i - 1]: cost of subtree n in i slots
    std::vector<int8_t> dk(nb * 8);    // i >= 2: the best split's left slots, or 0 = as with i - 1
    std::vector<uint8_t> as_leaf(nb);  // the choice of C(n, 1): one leaf, else one BVH8 node
    const float cN = kCollapseNodeCost, cP = kCollapseTriCost;
    {
        // nodes in pre-order; visited in reverse, children come before parents
        std::vector<int32_t> order;
        order.reserve(nb);
        std::vector<int32_t> st{0};
        while (!st.empty()) {
            const int32_t n = st.back();
            st.pop_back();
            order.push_back(n);
            if (b.nodes[n].left >= 0) {
                st.push_back(b.nodes[n].left);
                st.push_back(b.nodes[n].right);
            }
        }
        for (size_t q = order.size(); q-- > 0;) {
            const int32_t n = order[q];
            const B2Node& x = b.nodes[n];
            const float A = x.box.area();
            float* c = &C[(size_t)n * 8];
            if (x.left < 0) {
                sub_first[n] = x.first;
                sub_cnt[n] = x.count;
                as_leaf[n] = 1;
                for (int i = 0; i < 8; ++i) {
                    c[i] = A * cP * (float)x.count;
                    dk[(size_t)n * 8 + i] = 0;
                }
                continue;
            }
            sub_first[n] = std::min(sub_first[x.left], sub_first[x.right]);
            sub_cnt[n] = sub_cnt[x.left] + sub_cnt[x.right];
            const float* cl = &C[(size_t)x.left * 8];
            const float* cr = &C[(size_t)x.right * 8];
            float dist[8];
            int8_t dsplit[8];
            for (int j = 2; j <= 8; ++j) {
                float best = INFINITY;
                int bk = 1;
                for (int k = 1; k < j; ++k) {
                    const float v = cl[k - 1] + cr[j - k - 1];
                    if (v < best) {
                        best = v;
                        bk = k;
                    }
                }
                dist[j - 1] = best;
                dsplit[j - 1] = (int8_t)bk;
            }
            const float cint = A * cN + dist[7];
            const float cleaf = sub_cnt[n] <= kMaxLeafTris ? A * cP * (float)sub_cnt[n] : INFINITY;
            as_leaf[n] = cleaf <= cint ? 1 : 0;
            c[0] = std::min(cleaf, cint);
            dk[(size_t)n * 8] = 0;
            for (int j = 2; j <= 8; ++j) {
                const bool split = dist[j - 1] < c[j - 2];
                c[j - 1] = split ? dist[j - 1] : c[j - 2];
                dk[(size_t)n * 8 + j - 1] = split ? dsplit[j - 1] : 0;
            }
        }
    }
    // children of a BVH8 node at binary node n: expand its 8-slot distribution
    auto expand = [&](int32_t n, std::vector<Child>& ch) {
        // the node's own 8 slots: k to the left child, 8 - k to the right
        const int32_t l = b.nodes[n].left, r = b.nodes[n].right;
        int k = 1;
        float best = INFINITY;
        for (int kk = 1; kk < 8; ++kk) {
            const float v = C[(size_t)l * 8 + kk - 1] + C[(size_t)r * 8 + 8 - kk - 1];
            if (v < best) {
                best = v;
                k = kk;
            }
        }
        std::vector<std::pair<int32_t, int>> st{{r, 8 - k}, {l, k}};
        while (!st.empty()) {
            auto [m, j] = st.back(); st.pop_back();
            while (j > 1 && dk[(size_t)m * 8 + j - 1] == 0) j--;
            if (j == 1) { ch.push_back({m, as_leaf[m] != 0}); continue; }
            const int kk = dk[(size_t)m * 8 + j - 1];
            st.push_back({b.nodes[m].right, j - kk});
            st.push_back({b.nodes[m].left, kk});
        }
    };

    // ---- collapse to BVH8, breadth-first ---------------------------------------
    struct Pending { int32_t b2; uint32_t node; uint32_t depth; };
    std::vector<Pending> queue;
    out.nodes.reserve(ntri / 2 + 8);
    out.nodes.emplace_back();
    queue.push_back({0, 0, 1});
    size_t qh = 0;
    double sah = 0.0;
    float root_area = b.nodes[0].box.area();
    while (qh < queue.size()) {
        Pending pd = queue[qh++];
        out.max_depth = std::max(out.max_depth, pd.depth);
        const B2Node& src = b.nodes[pd.b2];
        std::vector<Child> ch;
        if (src.left < 0 || (pd.b2 == 0 && as_leaf[0])) ch.push_back({pd.b2, true});
        else expand(pd.b2, ch);
        // octant slot assignment
        const Aabb& pb = src.box;
        float pc[3] = {0.5f * (pb.lo[0] + pb.hi[0]), 0.5f * (pb.lo[1] + pb.hi[1]), 0.5f * (pb.lo[2] + pb.hi[2])};
        int slot_of[8];
        for (int k = 0; k < 8; ++k) slot_of[k] = -1;
        {
            struct Cand { float cost; int c, s; };
            std::vector<Cand> cands;
            for (size_t c = 0; c < ch.size(); ++c) {
                const Aabb& cb2 = b.nodes[ch[c].b2].box;
                float d[3];
                for (int k = 0; k < 3; ++k) d[k] = 0.5f * (cb2.lo[k] + cb2.hi[k]) - pc[k];
                for (int s = 0; s < 8; ++s) {
                    float cost = ((s & 1) ? d[0] : -d[0]) + ((s & 2) ? d[1] : -d[1]) + ((s & 4) ? d[2] : -d[2]);
                    cands.push_back({cost, (int)c, s});
                }
            }
            std::stable_sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& c2) { return a.cost > c2.cost; });
            bool slot_used[8] = {false};
            for (const Cand& cd : cands) {
                if (slot_of[cd.c] >= 0 || slot_used[cd.s]) continue;
                slot_of[cd.c] = cd.s;
                slot_used[cd.s] = true;
            }
        }
        int child_at[8];
        for (int s = 0; s < 8; ++s) child_at[s] = -1;
        for (size_t c = 0; c < ch.size(); ++c) child_at[slot_of[c]] = (int)c;
        Bvh8Node nd;
        std::memset(&nd, 0, sizeof nd);
        nd.tri_base = (uint32_t)(out.tris.size() / 12);
        uint32_t toff = 0, rank = 0;
        uint32_t first_internal = (uint32_t)out.nodes.size();
        uint32_t n_internal = 0;
        for (int s = 0; s < 8; ++s)
            if (child_at[s] >= 0 && !ch[child_at[s]].leaf) n_internal++;
        nd.child_base = n_internal ? first_internal : 0;
        for (int s = 0; s < 8; ++s) {
            int c = child_at[s];
            if (c < 0) {
                nd.lo_x[s] = nd.lo_y[s] = nd.lo_z[s] = 0.0f;
                nd.hi_x[s] = nd.hi_y[s] = nd.hi_z[s] = 0.0f;
                nd.meta[s] = 0;
                continue;
            }
            const B2Node& cn = b.nodes[ch[c].b2];
            nd.lo_x[s] = std::nextafter(cn.box.lo[0], -INFINITY);
            nd.lo_y[s] = std::nextafter(cn.box.lo[1], -INFINITY);
            nd.lo_z[s] = std::nextafter(cn.box.lo[2], -INFINITY);
            nd.hi_x[s] = std::nextafter(cn.box.hi[0], INFINITY);
            nd.hi_y[s] = std::nextafter(cn.box.hi[1], INFINITY);
            nd.hi_z[s] = std::nextafter(cn.box.hi[2], INFINITY);
            float ca = cn.box.area();
            if (ch[c].leaf) {
                const uint32_t lcount = sub_cnt[ch[c].b2], lfirst = sub_first[ch[c].b2];
                nd.meta[s] = leaf_meta(lcount, toff);
                for (uint32_t j = 0; j < lcount; ++j) {
                    uint32_t t = b.ref[lfirst + j];
                    const float* v = &tri[9ull * t];
                    float id;
                    std::memcpy(&id, &t, 4);
                    float rec[12] = {v[0], v[1], v[2], id, v[3], v[4], v[5], 0.0f, v[6], v[7], v[8], 0.0f};
                    out.tris.insert(out.tris.end(), rec, rec + 12);
                }
                toff += lcount;
                if (root_area > 0.0f) sah += (double)ca / root_area * lcount;
            } else {
                nd.meta[s] = (uint8_t)(0x80u | rank);
                out.nodes.emplace_back();
                queue.push_back({ch[c].b2, first_internal + rank, pd.depth + 1});
                rank++;
            }
        }
        if (root_area > 0.0f) sah += (double)pb.area() / root_area;
        uint32_t imask = 0, lmask = 0;
        for (int k = 0; k < 8; ++k) {
            if (nd.meta[k] & 0x80) imask |= 1u << k;
            else if (nd.meta[k]) lmask |= 1u << k;
        }
        nd.child_base |= imask << 24;
        nd.tri_base |= lmask << 24;
        out.nodes[pd.node] = nd;
    }
    (void)sah;
    out.sah_cost = bvh8_sah_cost(out.nodes);
    if (out.nodes.size() >= (1ull << 24)) {
        err = "too many BVH8 nodes for the index encoding (max 2^24)";
        return false;
    }
    return true;
}

// Per axis: origin p = smallest child lo, grid step s = 2^(e-127) with the
// smallest e for which every child interval [lo, hi] can be covered by
// decoded planes fmaf(q, s, p), q in 0..255, rounded outward.
static bool quantize_axis(const float* lo, const float* hi, uint32_t valid, float& p, uint32_t& e,
                          uint8_t* qlo, uint8_t* qhi) {
    float pmin = INFINITY, pmax = -INFINITY;
    for (int k = 0; k < 8; ++k) {
        qlo[k] = 0;
        qhi[k] = 0;
        if (!(valid >> k & 1u)) continue;
        if (!std::isfinite(lo[k]) || !std::isfinite(hi[k])) return false;
        pmin = std::min(pmin, lo[k]);
        pmax = std::max(pmax, hi[k]);
    }
    if (!valid) {
        p = 0.0f;
        e = 127;
        return true;
    }
    p = pmin;
    double ext = (double)pmax - (double)pmin;
    int ex = -126;
    if (ext > 0.0) std::frexp(ext / 255.0, &ex);   // ext / 255 <= 2^ex
    int e0 = std::max(1, std::min(254, ex + 127));
    for (int ee = e0; ee <= 254; ++ee) {
        const float s = std::ldexp(1.0f, ee - 127);
        bool ok = true;
        for (int k = 0; k < 8 && ok; ++k) {
            if (!(valid >> k & 1u)) continue;
            double a = std::floor(((double)lo[k] - (double)p) / (double)s);
            int ql = (int)std::max(0.0, std::min(255.0, a));
            while (ql > 0 && std::fma((float)ql, s, p) > lo[k]) ql--;
            double b = std::ceil(((double)hi[k] - (double)p) / (double)s);
            int qh = (int)std::max(0.0, std::min(255.0, b));
            while (qh < 255 && std::fma((float)qh, s, p) < hi[k]) qh++;
            if (std::fma((float)ql, s, p) > lo[k] || std::fma((float)qh, s, p) < hi[k]) {
                ok = false;
                break;
            }
            qlo[k] = (uint8_t)ql;
            qhi[k] = (uint8_t)qh;
        }
        if (ok) {
            e = (uint32_t)ee;
            return true;
        }
    }
    return false;
}

// One node: per axis origin and grid step, then the child codes.
static bool quantize_node(const Bvh8Node& a, QBvh8Node& q) {
    std::memset(&q, 0, sizeof q);
    uint32_t valid = 0;
    for (int k = 0; k < 8; ++k)
        if (a.meta[k]) valid |= 1u << k;
    uint32_t ex = 0, ey = 0, ez = 0;
    if (!quantize_axis(a.lo_x, a.hi_x, valid, q.px, ex, q.qlo_x, q.qhi_x) ||
        !quantize_axis(a.lo_y, a.hi_y, valid, q.py, ey, q.qlo_y, q.qhi_y) ||
        !quantize_axis(a.lo_z, a.hi_z, valid, q.pz, ez, q.qlo_z, q.qhi_z))
        return false;
    q.exps = ex | ey << 8 | ez << 16;
    q.child_base = a.child_base;
    q.tri_base = a.tri_base;
    std::memcpy(q.meta, a.meta, 8);
    return true;
}

bool quantize_bvh8(const std::vector<Bvh8Node>& nodes, std::vector<QBvh8Node>& out, std::string& err) {
    out.resize(nodes.size());
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    std::atomic<int64_t> bad{-1};
    for (unsigned t = 0; t < nth; ++t)
        pool.emplace_back([&, t]() {
            const size_t a = nodes.size() * t / nth, b = nodes.size() * (t + 1) / nth;
            for (size_t n = a; n < b; ++n)
                if (!quantize_node(nodes[n], out[n])) {
                    int64_t expect = -1;
                    bad.compare_exchange_strong(expect, (int64_t)n);
                }
        });
    for (auto& th : pool) th.join();
    if (bad.load() >= 0) {
        err = "BVH8 node " + std::to_string(bad.load()) + " has a box that 8-bit quantization cannot cover";
        return false;
    }
    return true;
}

float bvh8_sah_cost(const std::vector<Bvh8Node>& nodes) {
    if (nodes.empty()) return 0.0f;
    auto area = [](double lx, double ly, double lz, double hx, double hy, double hz) {
        const double dx = hx - lx, dy = hy - ly, dz = hz - lz;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    };
    double rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 8; ++k) {
        if (!nodes[0].meta[k]) continue;
        rl[0] = std::min(rl[0], (double)nodes[0].lo_x[k]); rh[0] = std::max(rh[0], (double)nodes[0].hi_x[k]);
        rl[1] = std::min(rl[1], (double)nodes[0].lo_y[k]); rh[1] = std::max(rh[1], (double)nodes[0].hi_y[k]);
        rl[2] = std::min(rl[2], (double)nodes[0].lo_z[k]); rh[2] = std::max(rh[2], (double)nodes[0].hi_z[k]);
    }
    const double ra = area(rl[0], rl[1], rl[2], rh[0], rh[1], rh[2]);
    if (!(ra > 0.0)) return 0.0f;
    double sah = 1.0;   // the root node itself
    for (const Bvh8Node& n : nodes)
        for (int k = 0; k < 8; ++k) {
            const uint8_t m = n.meta[k];
            if (!m) continue;
            const double a = area(n.lo_x[k], n.lo_y[k], n.lo_z[k], n.hi_x[k], n.hi_y[k], n.hi_z[k]) / ra;
            sah += (m & 0x80) ? a : a * ((m >> 5) & 3);
        }
    return (float)sah;
}

}  // namespace rt
