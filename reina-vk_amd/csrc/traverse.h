// traverse.h — closest-hit query over the BVH8 (the software replacement of
// traceRayEXT with gl_RayFlagsOpaqueEXT, shaders/raytrace/raytrace.rgen.glsl:110-122).
//
// One ray per lane. Per visited node the lane tests the 8 child boxes (slabs,
// near/far plane chosen by the direction sign, interval widened by 2^-16 of
// |t| and of |o/d| so no box holding a hit triangle is rejected), tests the
// triangles of every hit leaf before the next node's boxes, one leaf per step
// with the next node's fetch overlapping the last leaf's (watertight test of
// Woop, Benthin and Wald, JCGT 2013), and turns the hit internal children into one "group"
// {child_base, internal-slot mask, ordered hit mask}. Children are visited in
// slot order i ^ oct (oct = ray direction sign bits; the builder places each
// child in the slot of its octant), so a group's lowest ordered bit is the
// approximately nearest child. At most one group per tree level is live, so
// the per-lane stack in LDS needs only max-depth entries.
//
// The closest hit is the argmin of (t, flattened triangle id) over triangles
// with t in [tmin, tmax]; the result therefore does not depend on traversal
// order and equals a brute-force scan of all triangles.
#pragma once

#include "rt_layout.h"
#include "rt_math.h"

namespace rt {

constexpr float kBoxEps = 1.52587890625e-05f;  // 2^-16
constexpr float kBoxEpsH16 = 9.765625e-04f;    // 2^-10: slab widening with FP16 inverse directions
constexpr float kFmaInvMax = 1099511627776.0f;  // 2^40
constexpr float kSlabBig = 3.0e38f;

struct TRay {
    float ox, oy, oz, dx, dy, dz, tmin, tmax;
    float ix, iy, iz;
    // slab constants: plane distance = fma(plane, 1/d, c), c = -o/d - E (near
    // planes) or -o/d + E (far planes), E = 2^-16 |o/d|
    float cnx, cny, cnz, cfx, cfy, cfz;
    bool fma_slabs;   // every |1/d| <= 2^40 and o/d finite; else the exact (plane - o) * (1/d) form
    float Sx, Sy, Sz;
    int kx, ky, kz;
    uint32_t oct;
    bool valid;
};

RT_D float comp3(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

RT_D bool finite_(float x) { return (x == x) && (x - x == 0.0f); }

// H16: the inverse direction is rounded to FP16 (relative error <= 2^-11,
// so every slab distance is off by at most 2^-11 |t| plus FP32 roundings,
// which the 2^-10 widening of slab_hit<true> covers); a ray whose inverse
// components are not normal FP16 numbers keeps FP32 inverses.
template <bool H16 = false>
RT_D void tray_init(TRay& r, float4 a, float4 b) {
    r.ox = a.x; r.oy = a.y; r.oz = a.z; r.tmin = a.w;
    r.dx = b.x; r.dy = b.y; r.dz = b.z; r.tmax = b.w;
    r.valid = finite_(r.ox) && finite_(r.oy) && finite_(r.oz) && finite_(r.dx) && finite_(r.dy) &&
              finite_(r.dz) && !(r.dx == 0.0f && r.dy == 0.0f && r.dz == 0.0f);
    float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = (kz == 2) ? 0 : kz + 1;
    int ky = (kx == 2) ? 0 : kx + 1;
    float dkz = comp3(r.dx, r.dy, r.dz, kz);
    if (dkz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = comp3(r.dx, r.dy, r.dz, kx) / dkz;
    r.Sy = comp3(r.dx, r.dy, r.dz, ky) / dkz;
    r.Sz = 1.0f / dkz;
    // A zero component gives an infinite inverse (see node_hits).
    r.ix = 1.0f / r.dx; r.iy = 1.0f / r.dy; r.iz = 1.0f / r.dz;
    if constexpr (H16) {
        auto normal16 = [](float v) { const float m = fabsf(v); return m >= 6.103515625e-05f && m <= 65504.0f; };
        if (normal16(r.ix) && normal16(r.iy) && normal16(r.iz)) {
            r.ix = (float)(_Float16)r.ix;
            r.iy = (float)(_Float16)r.iy;
            r.iz = (float)(_Float16)r.iz;
        }
    }
    const float oix = -(r.ox * r.ix), oiy = -(r.oy * r.iy), oiz = -(r.oz * r.iz);
    const float ex = fabsf(oix) * kBoxEps, ey = fabsf(oiy) * kBoxEps, ez = fabsf(oiz) * kBoxEps;
    r.cnx = oix - ex; r.cny = oiy - ey; r.cnz = oiz - ez;
    r.cfx = oix + ex; r.cfy = oiy + ey; r.cfz = oiz + ez;
    r.fma_slabs = finite_(oix) && finite_(oiy) && finite_(oiz) && fabsf(r.ix) <= kFmaInvMax &&
                  fabsf(r.iy) <= kFmaInvMax && fabsf(r.iz) <= kFmaInvMax;
    r.oct = (r.dx < 0.0f ? 1u : 0u) | (r.dy < 0.0f ? 2u : 0u) | (r.dz < 0.0f ? 4u : 0u);
}

// Watertight ray/triangle test; true if hit with t in [tmin, tbest].
RT_D bool wtest(const TRay& r, float4 p0, float4 p1, float4 p2, float tbest, float& to, float& uo,
                float& vo) {
    float Ax_ = p0.x - r.ox, Ay_ = p0.y - r.oy, Az_ = p0.z - r.oz;
    float Bx_ = p1.x - r.ox, By_ = p1.y - r.oy, Bz_ = p1.z - r.oz;
    float Cx_ = p2.x - r.ox, Cy_ = p2.y - r.oy, Cz_ = p2.z - r.oz;
    float Akz = comp3(Ax_, Ay_, Az_, r.kz), Bkz = comp3(Bx_, By_, Bz_, r.kz), Ckz = comp3(Cx_, Cy_, Cz_, r.kz);
    float Ax = comp3(Ax_, Ay_, Az_, r.kx) - r.Sx * Akz;
    float Ay = comp3(Ax_, Ay_, Az_, r.ky) - r.Sy * Akz;
    float Bx = comp3(Bx_, By_, Bz_, r.kx) - r.Sx * Bkz;
    float By = comp3(Bx_, By_, Bz_, r.ky) - r.Sy * Bkz;
    float Cx = comp3(Cx_, Cy_, Cz_, r.kx) - r.Sx * Ckz;
    float Cy = comp3(Cx_, Cy_, Cz_, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = (U + V) + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    float T = (U * Az + V * Bz) + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= r.tmin && t <= tbest)) return false;
    to = t;
    uo = V * rcp;
    vo = W * rcp;
    return true;
}

struct HitRec {
    float t;
    uint32_t prim;
    float u, v;
};

// Hit test of one slab interval [tn, tf]: widened by 2^-16 of |t| and
// clipped to [tmin, tb].
//
// Why the box test never rejects a box whose triangle passes the watertight
// test. FMA form (node_hits with fma_slabs): a plane distance is
// fma(plane, 1/d, -o/d -+ E), whose error is at most 2^-24 |t| plus the
// roundings of -o/d and of the E term, together under 2^-23 |o/d|; E =
// 2^-16 |o/d| covers the latter, the 2^-16 |t| widening the former and the
// relative error of the triangle test's t. Exact form (rays with a direction
// component below 2^-40, or o/d not finite): (plane - o) * (1/d) has error
// 2^-23 |t|; with d = 0 on an axis the distances are +-inf, or NaN (0 * inf)
// for a ray inside a plane, which fmaxf/fminf ignore (closed slab), and an
// interval end at +-inf is clamped to a finite value so the widening keeps it
// a miss.
template <bool H16 = false>
RT_D bool slab_hit(float tn, float tf, float tmin, float tb) {
    constexpr float eps = H16 ? kBoxEpsH16 : kBoxEps;
    tn = fmaxf(__builtin_fmaf(-fabsf(tn), eps, tn), tmin);
    tf = fminf(__builtin_fmaf(fabsf(tf), eps, tf), tb);
    return tn <= tf;
}

// The 8 children's intervals from their near / far planes, both forms.
template <bool H16 = false>
RT_D uint32_t slab8(const TRay& r, const float* NX, const float* FX, const float* NY, const float* FY,
                    const float* NZ, const float* FZ, float tb) {
    uint32_t hitm = 0;
    if (r.fma_slabs) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float tn = fmaxf(fmaxf(__builtin_fmaf(NX[k], r.ix, r.cnx), __builtin_fmaf(NY[k], r.iy, r.cny)),
                                   __builtin_fmaf(NZ[k], r.iz, r.cnz));
            const float tf = fminf(fminf(__builtin_fmaf(FX[k], r.ix, r.cfx), __builtin_fmaf(FY[k], r.iy, r.cfy)),
                                   __builtin_fmaf(FZ[k], r.iz, r.cfz));
            hitm |= (slab_hit<H16>(tn, tf, r.tmin, tb) ? 1u : 0u) << k;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float tn = fmaxf(fmaxf((NX[k] - r.ox) * r.ix, (NY[k] - r.oy) * r.iy), (NZ[k] - r.oz) * r.iz);
            float tf = fminf(fminf((FX[k] - r.ox) * r.ix, (FY[k] - r.oy) * r.iy), (FZ[k] - r.oz) * r.iz);
            tn = fminf(tn, kSlabBig);
            tf = fmaxf(tf, -kSlabBig);
            hitm |= (slab_hit<H16>(tn, tf, r.tmin, tb) ? 1u : 0u) << k;
        }
    }
    return hitm;
}

// Byte j of w as float times s plus p, one rounding (q * s is exact).
RT_D float dq8(uint32_t w, int j, float s, float p) {
    return __builtin_fmaf((float)((w >> (8 * j)) & 0xFFu), s, p);
}

// node_hits for a quantized node: q[0] = (p.xyz, exponents), q[2..4] =
// (qlo[0..3], qlo[4..7], qhi[0..3], qhi[4..7]) for x, y, z. The decoded
// planes go through the same slab test as the FP32 node.
RT_D uint32_t node_hits_q8(const TRay& r, uint4 h0, uint4 qx, uint4 qy, uint4 qz, float tb) {
    const float px = __uint_as_float(h0.x), py = __uint_as_float(h0.y), pz = __uint_as_float(h0.z);
    const float sx = __uint_as_float((h0.w & 0xFFu) << 23);
    const float sy = __uint_as_float(((h0.w >> 8) & 0xFFu) << 23);
    const float sz = __uint_as_float(((h0.w >> 16) & 0xFFu) << 23);
    const bool bx = r.ix >= 0.0f, by = r.iy >= 0.0f, bz = r.iz >= 0.0f;
    const uint32_t nx0 = bx ? qx.x : qx.z, nx1 = bx ? qx.y : qx.w, fx0 = bx ? qx.z : qx.x, fx1 = bx ? qx.w : qx.y;
    const uint32_t ny0 = by ? qy.x : qy.z, ny1 = by ? qy.y : qy.w, fy0 = by ? qy.z : qy.x, fy1 = by ? qy.w : qy.y;
    const uint32_t nz0 = bz ? qz.x : qz.z, nz1 = bz ? qz.y : qz.w, fz0 = bz ? qz.z : qz.x, fz1 = bz ? qz.w : qz.y;
    float NX[8], FX[8], NY[8], FY[8], NZ[8], FZ[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int j = k & 3;
        NX[k] = dq8(k < 4 ? nx0 : nx1, j, sx, px); FX[k] = dq8(k < 4 ? fx0 : fx1, j, sx, px);
        NY[k] = dq8(k < 4 ? ny0 : ny1, j, sy, py); FY[k] = dq8(k < 4 ? fy0 : fy1, j, sy, py);
        NZ[k] = dq8(k < 4 ? nz0 : nz1, j, sz, pz); FZ[k] = dq8(k < 4 ? fz0 : fz1, j, sz, pz);
    }
    return slab8(r, NX, FX, NY, FY, NZ, FZ, tb);
}

// Bit i of the result is bit (i ^ oct) of the 8-bit mask m.
RT_D uint32_t permute8(uint32_t m, uint32_t oct) {
    m = (oct & 1u) ? (((m & 0x55u) << 1) | ((m >> 1) & 0x55u)) : m;
    m = (oct & 2u) ? (((m & 0x33u) << 2) | ((m >> 2) & 0x33u)) : m;
    m = (oct & 4u) ? (((m & 0x0Fu) << 4) | ((m >> 4) & 0x0Fu)) : m;
    return m;
}

constexpr uint32_t kNoNode = 0xFFFFFFFFu;

// Traversal state of one lane between steps.
struct TravState {
    uint32_t cur;    // next node to visit (kNoNode: none left)
    int sp;          // LDS stack depth
    uint32_t pm;     // hit leaves of the last visited node not yet tested (ordered slot mask)
    uint32_t pb;     // that node's first triangle
    uint32_t pml, pmh;   // its leaf records, 8 bits per slot: count << 5 | offset
};

RT_D void trav_begin(TravState& ts) {
    ts.cur = 0;
    ts.sp = 0;
    ts.pm = 0;
    ts.pb = 0;
    ts.pml = 0;
    ts.pmh = 0;
}

// Triangle test of one leaf triangle against the closest hit so far;
// ANY: returns true when the hit ends a shadow ray (see trav_step).
template <bool ANY>
RT_D bool leaf_tri(const TRay& r, float4 p0, float4 p1, float4 p2, HitRec& h, const uint8_t* tflags) {
    const uint32_t id = __float_as_uint(p0.w);
    float t, u, v;
    if constexpr (ANY) {
        if (wtest(r, p0, p1, p2, h.t, t, u, v)) {
            const uint8_t fl = tflags[id];
            bool pass = false;
            if (fl & kTriCull) {
                const float ex = p1.x - p0.x, ey = p1.y - p0.y, ez = p1.z - p0.z;
                const float fx = p2.x - p0.x, fy = p2.y - p0.y, fz = p2.z - p0.z;
                float g = (r.dx * (ey * fz - ez * fy) + r.dy * (ez * fx - ex * fz)) + r.dz * (ex * fy - ey * fx);
                if (fl & kTriMirror) g = -g;
                pass = !(g < 0.0f);   // back face of a culling instance
            }
            if (!pass) {
                h.t = t; h.prim = id; h.u = u; h.v = v;
                return true;
            }
        }
        return false;
    } else {
        if (wtest(r, p0, p1, p2, h.t, t, u, v) && ((t < h.t) || (t == h.t && id < h.prim))) {
            h.t = t; h.prim = id; h.u = u; h.v = v;
        }
        return false;
    }
}

// One traversal step of a lane: the triangles of one hit leaf of the last
// visited node (at most 2, fetched together), and, once no other leaf is
// pending, the next node (box tests of its 8 children). Both are
// fetched before either is tested, so the triangle and node fetches overlap;
// the leaf is tested first, so its hit shortens the node's box tests exactly
// as when a node's leaves are tested before the next node is visited (the
// node visits are the same). The next node is the first child of the node's
// group, else the top group of the stack. Returns false when the traversal
// is finished (no node and no leaf left, or an ANY hit).
// `stk` points at this lane's first LDS stack entry; entries are `stride`
// uint2 apart.
// Q selects the node format: false = FP32 Bvh8Node, true = QBvh8Node.
// `nodes` points at the scene's Bvh8Node or QBvh8Node array, `tris` at its
// leaf-ordered triangles.
// ANY: shadow-ray mode: the first triangle hit with t in [tmin, tmax] ends
// the traversal, except back faces of triangles whose flags (kTriCull,
// kTriMirror) say their instance culls them.
template <bool COUNT, bool Q, bool ANY = false, bool H16 = false>
RT_D bool trav_step(const void* nodes, const float4* tris, const TRay& r, HitRec& h, TravState& ts, uint2* stk,
                    uint32_t stride, uint32_t& nvis, uint32_t& ntst, const uint8_t* tflags = nullptr) {
    // Every load of the previous step has been consumed by its test, so this
    // wait costs nothing; it tells the compiler so. Without it, the compiler
    // sees a path on which a node fetch is never tested (the test is under its
    // own do_node branch) and waits for all loads before this step's node
    // fetch, which then no longer overlaps the leaf fetch issued before it.
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0), expcnt and lgkmcnt unconstrained
    // fetch the first two triangles of the lowest pending leaf; a third one
    // stays pending as a one-triangle leaf (its record rewritten) for the next step
    const bool do_leaf = ts.pm != 0u;
    uint32_t cnt = 0;
    float4 a0, a1, a2, b0, b1, b2;
    if (do_leaf) {
        const uint32_t k = (uint32_t)__builtin_ctz(ts.pm) ^ r.oct;
        const uint32_t sh = 8u * (k & 3u);
        const uint32_t m = ((k < 4 ? ts.pml : ts.pmh) >> sh) & 0xFFu;
        const uint32_t off = m & 31u;
        cnt = (m >> 5) & 3u;
        if (cnt > 2u) {
            const uint32_t rest = (1u << 5) | (off + 2u);   // offsets stay below 24 (8 leaves x 3)
            if (k < 4) ts.pml = (ts.pml & ~(0xFFu << sh)) | (rest << sh);
            else ts.pmh = (ts.pmh & ~(0xFFu << sh)) | (rest << sh);
            cnt = 2u;
        } else {
            ts.pm &= ts.pm - 1u;
        }
        const float4* tv = tris + 3ull * (ts.pb + off);
        if (cnt > 0u) { a0 = tv[0]; a1 = tv[1]; a2 = tv[2]; }
        if (cnt > 1u) { b0 = tv[3]; b1 = tv[4]; b2 = tv[5]; }
    }
    // the next node, once no leaf is left pending
    const bool do_node = ts.cur != kNoNode && ts.pm == 0u;
    // fetch the node
    uint4 hdr = make_uint4(0u, 0u, 0u, 0u);
    uint4 q0 = hdr, qx = hdr, qy = hdr, qz = hdr;
    float4 pl[12];
    if (do_node) {
        if constexpr (Q) {
            const uint4* qn = reinterpret_cast<const uint4*>(static_cast<const QBvh8Node*>(nodes) + ts.cur);
            q0 = qn[0];
            hdr = qn[1];
            qx = qn[2];
            qy = qn[3];
            qz = qn[4];
        } else {
            const float* base = reinterpret_cast<const float*>(static_cast<const Bvh8Node*>(nodes) + ts.cur);
            hdr = *reinterpret_cast<const uint4*>(base + 48);
            // near/far planes per axis by direction sign (lo <= hi, so this equals min/max)
            const float4* nx = reinterpret_cast<const float4*>(base + (r.ix >= 0.0f ? 0 : 8));
            const float4* fx = reinterpret_cast<const float4*>(base + (r.ix >= 0.0f ? 8 : 0));
            const float4* ny = reinterpret_cast<const float4*>(base + (r.iy >= 0.0f ? 16 : 24));
            const float4* fy = reinterpret_cast<const float4*>(base + (r.iy >= 0.0f ? 24 : 16));
            const float4* nz = reinterpret_cast<const float4*>(base + (r.iz >= 0.0f ? 32 : 40));
            const float4* fz = reinterpret_cast<const float4*>(base + (r.iz >= 0.0f ? 40 : 32));
            pl[0] = nx[0]; pl[1] = nx[1]; pl[2] = fx[0]; pl[3] = fx[1];
            pl[4] = ny[0]; pl[5] = ny[1]; pl[6] = fy[0]; pl[7] = fy[1];
            pl[8] = nz[0]; pl[9] = nz[1]; pl[10] = fz[0]; pl[11] = fz[1];
        }
    }
    // test the leaf, in triangle order
    if (do_leaf) {
        if (COUNT) ntst += cnt;
        if (cnt > 0u && leaf_tri<ANY>(r, a0, a1, a2, h, tflags)) return false;
        if (cnt > 1u && leaf_tri<ANY>(r, b0, b1, b2, h, tflags)) return false;
    }
    // test the node; its hit leaves become pending
    if (do_node) {
        if (COUNT) nvis++;
        uint32_t hitm;
        if constexpr (Q) {
            hitm = node_hits_q8(r, q0, qx, qy, qz, h.t);
        } else {
            float NX[8], FX[8], NY[8], FY[8], NZ[8], FZ[8];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const float4 a = pl[j], b = pl[2 + j], c = pl[4 + j], d = pl[6 + j], e = pl[8 + j], f = pl[10 + j];
                NX[4 * j + 0] = a.x; NX[4 * j + 1] = a.y; NX[4 * j + 2] = a.z; NX[4 * j + 3] = a.w;
                FX[4 * j + 0] = b.x; FX[4 * j + 1] = b.y; FX[4 * j + 2] = b.z; FX[4 * j + 3] = b.w;
                NY[4 * j + 0] = c.x; NY[4 * j + 1] = c.y; NY[4 * j + 2] = c.z; NY[4 * j + 3] = c.w;
                FY[4 * j + 0] = d.x; FY[4 * j + 1] = d.y; FY[4 * j + 2] = d.z; FY[4 * j + 3] = d.w;
                NZ[4 * j + 0] = e.x; NZ[4 * j + 1] = e.y; NZ[4 * j + 2] = e.z; NZ[4 * j + 3] = e.w;
                FZ[4 * j + 0] = f.x; FZ[4 * j + 1] = f.y; FZ[4 * j + 2] = f.z; FZ[4 * j + 3] = f.w;
            }
            hitm = slab8<H16>(r, NX, FX, NY, FY, NZ, FZ, h.t);
        }
        const uint32_t child_base = hdr.x & 0xFFFFFFu;
        const uint32_t imask = hdr.x >> 24, leafmask = hdr.y >> 24;
        ts.pm = permute8(hitm & leafmask, r.oct);
        ts.pb = hdr.y & 0xFFFFFFu;
        ts.pml = hdr.z;
        ts.pmh = hdr.w;
        uint32_t om = permute8(hitm & imask, r.oct);
        if (om) {
            const uint32_t i = __builtin_ctz(om);
            om &= om - 1u;
            const uint32_t k = i ^ r.oct;
            if (om) { stk[ts.sp * stride] = make_uint2(child_base, imask | (om << 8)); ts.sp++; }
            ts.cur = child_base + __builtin_popcount(imask & ((1u << k) - 1u));
        } else if (ts.sp == 0) {
            ts.cur = kNoNode;
        } else {
            const uint2 g = stk[(ts.sp - 1) * stride];
            uint32_t gm = g.y >> 8;
            const uint32_t gi = g.y & 0xFFu;
            const uint32_t i = __builtin_ctz(gm);
            gm &= gm - 1u;
            const uint32_t k = i ^ r.oct;
            ts.cur = g.x + __builtin_popcount(gi & ((1u << k) - 1u));
            if (gm) stk[(ts.sp - 1) * stride] = make_uint2(g.x, gi | (gm << 8));
            else ts.sp--;
        }
    }
    return ts.cur != kNoNode || ts.pm != 0u;
}

RT_D HitRec miss_record(const TRay& r) {
    HitRec h;
    h.t = r.tmax;
    h.prim = 0xFFFFFFFFu;
    h.u = 0.0f;
    h.v = 0.0f;
    return h;
}

}  // namespace rt
