// shade.h — device restatement of Reina's closest-hit / miss shaders and the
// raygen camera model (shaders/raytrace/*.glsl), one function per shader
// stage, operating on a per-path payload held in registers.
#pragma once

#include "rt_layout.h"
#include "rt_math.h"

namespace rt {

constexpr float kPi = 3.14159265f;      // shaderCommon.h.glsl:47
constexpr float kInvPi = 0.31830989f;   // shaderCommon.h.glsl:48

// shaderCommon.h.glsl:39-45 (PCG-RXS-M-XS variant; 4294967295.0f rounds to 2^32)
RT_D float rnd(uint32_t& s) {
    s = s * 747796405u + 1u;
    uint32_t w = ((s >> ((s >> 28) + 4u)) ^ s) * 277803737u;
    w = (w >> 22) ^ w;
    return (float)w / 4294967295.0f;
}

// The part of HitPayload (shaderCommon.h.glsl:18-36) that the active path reads.
struct Payload {
    V3 color, rayOrigin, rayDirection, emission;
    V3 albedo, surfaceNormal;   // the BSDF's colour and shading normal (read by next-event estimation)
    uint32_t rng;
    float accumulatedDistance;
    bool rayHitSky, skip, insideDielectric;
};

struct HitInfo {
    V3 worldPosition, worldNormal, worldNormalGeometry;
    float uvx, uvy;
    bool frontFace;
    M3 tbn;
};

RT_D V3 ldv3(const float* p) { return v3(p[0], p[1], p[2]); }

RT_D V3 tbn_col(const SceneView& sv, uint32_t pos, int col) {
    if (pos >= sv.num_tbn_indices) return v3(0.0f, 0.0f, 0.0f);
    uint32_t ti = sv.tbn_indices[pos];
    if ((uint64_t)ti >= sv.num_tbns) return v3(0.0f, 0.0f, 0.0f);
    return ldv3(sv.tbns + 9ull * ti + 3 * col);
}

// The shading record of flattened triangle g (SceneView::srec): the same
// loads as closestHitCommon.h.glsl:52-148 makes through the instance's index
// and TBN offsets, done once per scene.
RT_D void shade_record(const SceneView& sv, uint32_t g, float4* r) {
    const uint32_t ii = sv.tri_inst[g], prim = sv.tri_prim[g];
    const InstanceRec in = sv.inst[ii];
    const ReinaInstanceProperties& p = sv.props[in.props_id];
    const uint32_t io = p.indicesOffset, to = p.tbnsIndicesOffset;
    float f[28];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float4 v = sv.vertices[sv.indices[3u * prim + io + k]];
        r[k] = make_float4(v.x, v.y, v.z, __uint_as_float(k == 0 ? ii : (k == 1 ? prim : in.props_id)));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const V3 col = tbn_col(sv, 3u * prim + to + k, c);
            f[9 * k + 3 * c + 0] = col.x;
            f[9 * k + 3 * c + 1] = col.y;
            f[9 * k + 3 * c + 2] = col.z;
        }
    }
    f[27] = 0.0f;
#pragma unroll
    for (int q = 0; q < 7; ++q) r[3 + q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
}

// closestHitCommon.h.glsl:52-148, from the triangle's shading record
// fix_normals: world normals by the inverse-transpose (REINA_RT_TRACE_FIX_NORMALS)
RT_D HitInfo hit_info(const SceneView& sv, const InstanceRec& in, const ReinaInstanceProperties& p,
                      const float4* rec, uint32_t prim, float bu, float bv, V3 D, bool fix_normals) {
    HitInfo r;
    float4 q[kShadeRecF4];
#pragma unroll
    for (int k = 0; k < (int)kShadeRecF4; ++k) q[k] = rec[k];
    float tb[28];
#pragma unroll
    for (int k = 0; k < 7; ++k) { tb[4 * k] = q[3 + k].x; tb[4 * k + 1] = q[3 + k].y; tb[4 * k + 2] = q[3 + k].z; tb[4 * k + 3] = q[3 + k].w; }
    // T, B, N of vertex k: tb[9k + 0..2], tb[9k + 3..5], tb[9k + 6..8]
    V3 v0 = v3(q[0].x, q[0].y, q[0].z), v1 = v3(q[1].x, q[1].y, q[1].z), v2 = v3(q[2].x, q[2].y, q[2].z);
    float bx = (1.0f - bu) - bv, by = bu, bz = bv;
    V3 op = add(add(scale(v0, bx), scale(v1, by)), scale(v2, bz));
    V3 c0 = ldv3(in.c0), c1 = ldv3(in.c1), c2 = ldv3(in.c2), c3 = ldv3(in.c3);
    r.worldPosition = add(add(add(scale(c0, op.x), scale(c1, op.y)), scale(c2, op.z)), c3);
    V3 ong = normalize(cross(sub(v1, v0), sub(v2, v0)));
    V3 on;
    if (!p.interpNormals) {
        on = ong;
    } else {
        V3 n0 = normalize(v3(tb[6], tb[7], tb[8]));
        V3 n1 = normalize(v3(tb[15], tb[16], tb[17]));
        V3 n2 = normalize(v3(tb[24], tb[25], tb[26]));
        on = normalize(add(add(scale(n0, bx), scale(n1, by)), scale(n2, bz)));
    }
    if (p.texIndicesOffset == 0xFFFFFFFFu) {
        r.uvx = 0.0f; r.uvy = 0.0f;
    } else {
        uint32_t tio = p.texIndicesOffset;
        float tu[3], tv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            uint32_t ti = sv.tex_indices[3u * prim + tio + k];
            bool ok = (uint64_t)ti < sv.num_tex_coords;   // 0xFFFFFFFF entries read as (0, 0)
            tu[k] = ok ? sv.tex_coords[2ull * ti] : 0.0f;
            tv[k] = ok ? sv.tex_coords[2ull * ti + 1] : 0.0f;
        }
        r.uvx = (tu[0] * bx + tu[1] * by) + tu[2] * bz;
        r.uvy = (tv[0] * bx + tv[1] * by) + tv[2] * bz;
    }
    M3 M;
    M.c0 = c0; M.c1 = c1; M.c2 = c2;
    float det = dot(M.c0, cross(M.c1, M.c2));
    float invdet = 1.0f / det;
    M3 Ni;
    Ni.c0 = scale(cross(M.c1, M.c2), invdet);
    Ni.c1 = scale(cross(M.c2, M.c0), invdet);
    Ni.c2 = scale(cross(M.c0, M.c1), invdet);
    const M3& NM = fix_normals ? Ni : M;
    r.worldNormal = normalize(mmul(NM, on));
    r.worldNormalGeometry = normalize(mmul(NM, ong));
    r.frontFace = dot(D, r.worldNormalGeometry) < 0.0f;
    r.worldNormal = faceforward(r.worldNormal, D, r.worldNormalGeometry);
    r.worldNormalGeometry = faceforward(r.worldNormalGeometry, D, r.worldNormalGeometry);
    // TBN, read even when interpNormals is false (closestHitCommon.h.glsl:120-145)
    V3 T = v3(0.0f, 0.0f, 0.0f), B = T, N = T;
    const float w3[3] = {bx, by, bz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        V3 tk = v3(tb[9 * k], tb[9 * k + 1], tb[9 * k + 2]);
        V3 bk = v3(tb[9 * k + 3], tb[9 * k + 4], tb[9 * k + 5]);
        V3 nk = v3(tb[9 * k + 6], tb[9 * k + 7], tb[9 * k + 8]);
        if (k == 0) { T = scale(tk, w3[0]); B = scale(bk, w3[0]); N = scale(nk, w3[0]); }
        else { T = add(T, scale(tk, w3[k])); B = add(B, scale(bk, w3[k])); N = add(N, scale(nk, w3[k])); }
    }
    V3 tangent = normalize(T), bitangent = normalize(B), normal = normalize(N);
    V3 wT = normalize(mmul(M, tangent));
    V3 wB = normalize(mmul(M, bitangent));
    V3 wN = normalize(mmul(Ni, normal));
    wT = normalize(sub(wT, scale(wN, dot(wN, wT))));
    wB = normalize(sub(wB, scale(wN, dot(wN, wB))));
    wB = scale(wB, -1.0f);
    wN = scale(wN, r.frontFace ? 1.0f : -1.0f);
    r.tbn.c0 = wT; r.tbn.c1 = wB; r.tbn.c2 = wN;
    return r;
}

// closestHitCommon.h.glsl:156-177: integer ULP offset, float fallback near 0.
RT_D float offc(float p, float n) {
    int32_t of = f2i_(256.0f * n);
    uint32_t nb = __float_as_uint(p) + (uint32_t)((p < 0.0f) ? -of : of);
    float pi = __uint_as_float(nb);
    return (fabsf(p) < (1.0f / 32.0f)) ? p + (1.0f / 65536.0f) * n : pi;
}
RT_D V3 offset_along(V3 p, V3 n) { return v3(offc(p.x, n.x), offc(p.y, n.y), offc(p.z, n.z)); }
// dielectric.rchit.glsl:14-38 / disney.rchit.glsl:10-34
RT_D V3 offset_dielectric(V3 p, V3 n, V3 dir) { return offset_along(p, (dot(n, dir) < 0.0f) ? neg(n) : n); }

// closestHitCommon.h.glsl:195-205
RT_D V3 random_unit_vec(uint32_t& rng) {
    for (;;) {
        float a = rnd(rng);
        float b = rnd(rng);
        float c = rnd(rng);
        V3 v = sub(scale(v3(a, b, c), 2.0f), v3(1.0f, 1.0f, 1.0f));
        float l2 = dot(v, v);
        if (0.0001f < l2 && l2 < 1.0f) return normalize(v);
    }
}
// closestHitCommon.h.glsl:207-222
RT_D V3 fuzzy_reflection(V3 rayIn, V3 n, float fuzz, uint32_t& rng) {
    return add(reflect(normalize(rayIn), normalize(n)), scale(random_unit_vec(rng), fuzz));
}
RT_D V3 diffuse_reflection(V3 n, uint32_t& rng) {
    float theta = (2.0f * kPi) * rnd(rng);
    float u = 2.0f * rnd(rng) - 1.0f;
    float r = sqrtf_(1.0f - u * u);
    float s, c;
    sincos_(theta, &s, &c);
    return normalize(add(n, v3(r * c, r * s, u)));
}

// raytrace.rmiss.glsl:10-19
RT_D V3 sky_color(V3 dir) {
    float t = 0.5f * (normalize(dir).y + 1.0f);
    return scale(mix(v3(0.1f, 0.1f, 0.1f), v3(0.4f, 1.7f, 2.0f), t), 0.07f);
}

// ---- Disney BSDF (brdfDisney.h.glsl) ----------------------------------------
RT_D V3 sample_ggx_vndf(V3 V, float ax, float ay, uint32_t& rng) {         // :9-40
    bool flip = V.z < 0.0f;
    if (flip) V.z = V.z * -1.0f;
    float r1 = rnd(rng);
    float r2 = rnd(rng);
    V3 Vh = normalize(v3(ax * V.x, ay * V.y, V.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    V3 T1 = lensq > 0.0f ? scale(v3(-Vh.y, Vh.x, 0.0f), 1.0f / sqrtf_(lensq)) : v3(1.0f, 0.0f, 0.0f);
    V3 T2 = cross(Vh, T1);
    float r = sqrtf_(r1);
    float phi = (2.0f * kPi) * r2;
    float sp, cp;
    sincos_(phi, &sp, &cp);
    float t1 = r * cp;
    float t2 = r * sp;
    float s = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - s) * sqrtf_(1.0f - t1 * t1) + s * t2;
    V3 Nh = add(add(scale(T1, t1), scale(T2, t2)), scale(Vh, sqrtf_(maxf_(0.0f, (1.0f - t1 * t1) - t2 * t2))));
    if (flip) Nh.z = Nh.z * -1.0f;
    return normalize(v3(ax * Nh.x, ay * Nh.y, maxf_(0.0f, Nh.z)));
}

RT_D float d_ggx_aniso(V3 m, float ax, float ay) {                          // :42-48
    float NoM = maxf_(m.z, 0.0f);
    float inv = 1.0f / (((m.x / ax) * (m.x / ax) + (m.y / ay) * (m.y / ay)) + NoM * NoM);
    return (inv * inv) / ((kPi * ax) * ay);
}

// bounded: the bounded-VNDF pdf as the reference; else k = 1, the density of
// sample_ggx_vndf (REINA_RT_TRACE_FIX_METAL_PDF)
RT_D float pdf_ggx_reflection(V3 i, V3 o, float ax, float ay, bool bounded) {   // :55-73
    V3 m = normalize(add(i, o));
    float ndf = d_ggx_aniso(m, ax, ay);
    float aix = ax * i.x, aiy = ay * i.y;
    float len2 = aix * aix + aiy * aiy;
    float t = sqrtf_(len2 + i.z * i.z);
    if (i.z >= 0.0f) {
        float a = clampf_(minf_(ax, ay), 0.0f, 1.0f);
        float s = 1.0f + sqrtf_(i.x * i.x + i.y * i.y);
        float a2 = a * a, s2 = s * s;
        float k = bounded ? ((1.0f - a2) * s2) / (s2 + (a2 * i.z) * i.z) : 1.0f;
        return ndf / (2.0f * (k * i.z + t));
    }
    return (ndf * (t - i.z)) / (2.0f * len2);
}

RT_D V3 sample_gtr1(float alpha, uint32_t& rng) {                           // :96-111
    float r1 = rnd(rng);
    float r2 = rnd(rng);
    float a = maxf_(0.001f, alpha);
    float c = sqrtf_((1.0f - expf_(r1 * logf_(a * a))) / (1.0f - a * a));
    float sE = (c > 1.0f || c < -1.0f || c != c) ? __builtin_nanf("") : sqrtf_(1.0f - c * c);
    float hAz = (2.0f * kPi) * r2;
    float s, co;
    sincos_(hAz, &s, &co);
    return v3(sE * co, sE * s, c);
}

RT_D float fd_term(float fd90, V3 n, V3 w) { return 1.0f + (fd90 - 1.0f) * pow5_(1.0f - maxf_(dot(n, w), 0.0f)); }

RT_D V3 eval_diffuse(float roughness, float subsurface, V3 bc, V3 n, V3 wi, V3 wo, V3 h) {   // :124-168
    float hw = maxf_(dot(h, wo), 0.0f);
    float fd90 = 0.5f + ((2.0f * roughness) * hw) * hw;
    float FDin = fd_term(fd90, n, wi), FDout = fd_term(fd90, n, wo);
    V3 base = scale(scale(v3(bc.x / kPi, bc.y / kPi, bc.z / kPi), FDin), FDout);
    V3 k = scale(scale(bc, 1.25f), kInvPi);
    float fss90 = (roughness * hw) * hw;
    float fssIn = 1.0f + (fss90 - 1.0f) * pow5_(1.0f - maxf_(dot(n, wi), 0.0f));
    float fssOut = 1.0f + (fss90 - 1.0f) * pow5_(1.0f - maxf_(dot(n, wo), 0.0f));
    float third = 1.0f / (maxf_(dot(n, wi), 0.0f) + maxf_(dot(n, wo), 0.0f)) - 0.5f;
    V3 subs = scale(k, (fssIn * fssOut) * third + 0.5f);
    return mix(base, subs, subsurface);
}

RT_D float pdf_diffuse(V3 wo, V3 n) {                                       // :170-179
    float c = dot(n, wo);
    if (c <= 0.0f) return 0.0f;
    return c * kInvPi;
}

RT_D V3 sample_diffuse(V3 n, uint32_t& rng) {                               // :181-201
    float xi1 = rnd(rng);
    float xi2 = rnd(rng);
    float r = sqrtf_(xi1);
    float phi = (2.0f * kPi) * xi2;
    float s, c;
    sincos_(phi, &s, &c);
    float x = r * c, y = r * s;
    float z = sqrtf_(maxf_(0.0f, 1.0f - xi1));
    V3 t = fabsf(n.x) < 0.5f ? normalize(cross(n, v3(1.0f, 0.0f, 0.0f))) : normalize(cross(n, v3(0.0f, 1.0f, 0.0f)));
    V3 b = cross(n, t);
    return normalize(add(add(scale(t, x), scale(b, y)), scale(n, z)));
}

RT_D float eval_r0(float ior) { return ((ior - 1.0f) * (ior - 1.0f)) / ((ior + 1.0f) * (ior + 1.0f)); }
RT_D float luminance(V3 c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }

RT_D V3 eval_fm(V3 bc, V3 h, V3 wo, float specular, V3 specularTint, float metallic, float eta) {   // :215-222
    float lum = luminance(bc);
    V3 ctint = lum > 0.0f ? divs(bc, lum) : v3(1.0f, 1.0f, 1.0f);
    V3 ks = add(sub(v3(1.0f, 1.0f, 1.0f), ctint), mul(specularTint, ctint));
    V3 c0 = add(scale(ks, (specular * eval_r0(eta)) * (1.0f - metallic)), scale(bc, metallic));
    float p = pow5_(1.0f - fabsf(dot(h, wo)));
    return add(c0, scale(sub(v3(1.0f, 1.0f, 1.0f), c0), p));
}

RT_D float eval_dm(V3 hl, float ax, float ay) {                             // :224-229
    float constant = (kPi * ax) * ay;
    float q = (pow2_(hl.x) / pow2_(ax) + pow2_(hl.y) / pow2_(ay)) + pow2_(hl.z);
    return 1.0f / (constant * pow2_(q));
}

RT_D float smith_g(V3 wl, float ax, float ay) {                             // :231-239
    float st = sqrtf_(1.0f + (pow2_(wl.x * ax) + pow2_(wl.y * ay)) / pow2_(wl.z));
    return 1.0f / (1.0f + (st - 1.0f) / 2.0f);
}

RT_D void alphas(float anisotropic, float roughness, float& ax, float& ay) {
    float aspect = sqrtf_(1.0f - 0.9f * anisotropic);
    ax = maxf_(0.0001f, (roughness * roughness) / aspect);
    ay = maxf_(0.0001f, (roughness * roughness) * aspect);
}

RT_D V3 eval_metal(const M3& tbn, V3 bc, float anisotropic, float roughness, V3 n, V3 wi, V3 wo, V3 h,
                   float specular, V3 specularTint, float metallic, float eta) {   // :247-267
    float ax, ay;
    alphas(anisotropic, roughness, ax, ay);
    V3 fm = eval_fm(bc, h, wo, specular, specularTint, metallic, eta);
    V3 wiT = normalize(tmul(tbn, wi)), woT = normalize(tmul(tbn, wo)), hT = normalize(tmul(tbn, h));
    float dm = eval_dm(hT, ax, ay);
    float gm = smith_g(wiT, ax, ay) * smith_g(woT, ax, ay);
    float den = (4.0f * fabsf(dot(n, wi))) * fabsf(dot(n, wo));
    return v3(((fm.x * dm) * gm) / den, ((fm.y * dm) * gm) / den, ((fm.z * dm) * gm) / den);
}

RT_D V3 sample_metal(const M3& tbn, float anisotropic, float roughness, V3 wi, uint32_t& rng) {   // :269-285
    float ax, ay;
    alphas(anisotropic, roughness, ax, ay);
    V3 h = sample_ggx_vndf(tmul(tbn, wi), ax, ay, rng);
    h = normalize(mmul(tbn, h));
    return reflect(neg(wi), h);
}

RT_D float pdf_metal(const M3& tbn, V3 wi, V3 wo, float anisotropic, float roughness, bool bounded) {   // :287-299
    float ax, ay;
    alphas(anisotropic, roughness, ax, ay);
    return pdf_ggx_reflection(tmul(tbn, wi), tmul(tbn, wo), ax, ay, bounded);
}

RT_D float g1c(V3 w, float a) {                                             // :313-319
    float a2 = a * a;
    return 2.0f / (1.0f + sqrtf_(a2 + ((1.0f - a2) * w.z) * w.z));
}

RT_D float eval_dc(float ag, V3 hl) {                                       // :328-333
    float num = ag * ag - 1.0f;
    float den = (kPi * logf_(ag * ag)) * (1.0f + (ag * ag - 1.0f) * (hl.z * hl.z));
    return num / den;
}

RT_D V3 eval_clearcoat(const M3& tbn, V3 wi, V3 wo, float cg, V3 h) {      // :335-354
    float ag = (1.0f - cg) * 0.1f + cg * 0.001f;
    V3 hT = tmul(tbn, h), wiT = tmul(tbn, wi), woT = tmul(tbn, wo);
    float r0 = eval_r0(1.5f);
    float fc = r0 + (1.0f - r0) * pow5_(1.0f - dot(h, wo));
    float gc = g1c(wiT, 0.25f) * g1c(woT, 0.25f);
    float v = ((0.25f * fc) * gc) * eval_dc(ag, hT);
    return v3(v, v, v);
}

RT_D V3 sample_clearcoat(const M3& tbn, float cg, V3 wi, uint32_t& rng) {  // :356-366
    float ag = (1.0f - cg) * 0.1f + cg * 0.001f;
    V3 h = normalize(sample_gtr1(ag, rng));
    h = normalize(mmul(tbn, h));
    return normalize(reflect(neg(wi), h));
}

RT_D float pdf_clearcoat(const M3& tbn, V3 wi, V3 wo, V3 h, float cg) {    // :368-383
    float ag = (1.0f - cg) * 0.1f + cg * 0.001f;
    V3 wiT = tmul(tbn, wi), woT = tmul(tbn, wo), hT = tmul(tbn, h);
    if (wiT.z <= 0.0f || woT.z <= 0.0f) return 0.0f;
    float dc = eval_dc(ag, hT) * hT.z;
    return dc / (4.0f * fabsf(dot(woT, hT)));
}

// brdfDisney.h.glsl:389-394 and dielectric.rchit.glsl:7-12 (Schlick)
RT_D float reflectance(float cosine, float ri) {
    float r0 = (1.0f - ri) / (1.0f + ri);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * pow5_(1.0f - cosine);
}

RT_D V3 sample_glass(const M3& tbn, V3 wi, float roughness, float anisotropic, float eta, uint32_t& rng,
                     bool& refracted) {                                     // :396-431
    float ax, ay;
    alphas(anisotropic, roughness, ax, ay);
    V3 wiT = tmul(tbn, wi);
    V3 hT = sample_ggx_vndf(wiT, ax, ay, rng);
    float cosTheta = dot(wiT, hT);
    V3 hW = normalize(mmul(tbn, hT));
    float refl = reflectance(cosTheta, eta);
    float sinTheta = sqrtf_(1.0f - cosTheta * cosTheta);
    bool cannot = eta * sinTheta > 1.0f;
    if (cannot || refl > rnd(rng)) {
        refracted = false;
        return reflect(neg(wi), hW);
    }
    refracted = true;
    return refract(neg(wi), hW, eta);
}

RT_D float smith_g_aniso(float NdV, float VdX, float VdY, float ax, float ay) {   // :433-439
    float a = VdX * ax, b = VdY * ay, c = NdV;
    return (2.0f * NdV) / (NdV + sqrtf_((a * a + b * b) + c * c));
}

RT_D V3 eval_microfacet_refraction(V3 bc, float anisotropic, float roughness, float eta, V3 V, V3 L, V3 H,
                                   float& pdf) {                            // :441-467
    float ax, ay;
    alphas(anisotropic, roughness, ax, ay);
    pdf = 0.0f;
    if (L.z >= 0.0f) return v3(0.0f, 0.0f, 0.0f);
    float LdH = dot(L, H), VdH = dot(V, H);
    float D = eval_dm(H, ax, ay);
    float G1 = smith_g_aniso(fabsf(V.z), V.x, V.y, ax, ay);
    float G2 = G1 * smith_g_aniso(fabsf(L.z), L.x, L.y, ax, ay);
    float denom = LdH + VdH * eta;
    denom = denom * denom;
    float jac = (fabsf(LdH) * (eta * eta)) / denom;
    pdf = (((G1 * maxf_(0.0f, VdH)) * D) * jac) / V.z;
    float F = reflectance(dot(V, H), eta);
    V3 r = v3(powf_(bc.x, 0.5f), powf_(bc.y, 0.5f), powf_(bc.z, 0.5f));
    r = scale(r, 1.0f - F);
    r = scale(r, D);
    r = scale(r, G2);
    r = scale(r, fabsf(VdH));
    r = scale(r, jac);
    return divs(r, fabsf(L.z * V.z));
}

// branch_pdf: the pdf includes the probability of sample_glass's branch, F for
// reflection (1 under total internal reflection) and 1 - F for refraction
// (REINA_RT_TRACE_FIX_DISNEY_MIX); the reference leaves it out.
RT_D V3 eval_glass(const M3& tbn, V3 bc, float anisotropic, float roughness, float eta, V3 n, V3 wi, V3 wo,
                   bool didRefract, bool bounded, bool branch_pdf, float& pdf) {   // :539-581
    V3 h = didRefract ? normalize(add(wo, scale(wi, eta))) : normalize(add(wo, wi));
    if (dot(h, n) < 0.0f) h = neg(h);
    const V3 wiT = tmul(tbn, wi), hT = tmul(tbn, h);
    if (didRefract) {
        float tp;
        V3 f = eval_microfacet_refraction(bc, anisotropic, roughness, eta, wiT, tmul(tbn, wo), hT, tp);
        pdf = branch_pdf ? tp * (1.0f - reflectance(dot(wiT, hT), eta)) : tp;
        return f;
    }
    pdf = pdf_metal(tbn, wi, wo, anisotropic, roughness, bounded);
    if (branch_pdf) {
        const float c = dot(wiT, hT);
        const float s = sqrtf_(1.0f - c * c);
        pdf = pdf * ((eta * s > 1.0f) ? 1.0f : reflectance(c, eta));
    }
    return eval_metal(tbn, bc, anisotropic, roughness, n, wi, wo, h, 0.0f, v3(1.0f, 1.0f, 1.0f), 1.0f, eta);
}

RT_D V3 eval_sheen(V3 bc, V3 wo, V3 h, V3 n, V3 sheenTint) {               // :587-594
    float lum = luminance(bc);
    V3 ctint = lum > 0.0f ? divs(bc, lum) : v3(1.0f, 1.0f, 1.0f);
    V3 cs = v3(mixf(1.0f, ctint.x, sheenTint.x), mixf(1.0f, ctint.y, sheenTint.y), mixf(1.0f, ctint.z, sheenTint.z));
    float p = pow5_(1.0f - maxf_(dot(h, wo), 0.0f));
    return scale(scale(cs, p), maxf_(dot(n, wo), 0.0f));
}

RT_D V3 sample_disney(const M3& tbn, float anisotropic, float roughness, float cg, float eta, float metallic,
                      float clearcoat, float st, V3 n, V3 wi, bool& didRefract, bool& choseGlass,
                      uint32_t& rng) {                                      // :609-654
    float dW = (1.0f - st) * (1.0f - metallic);
    float mW = metallic;
    float cW = 0.25f * clearcoat;
    float gW = (1.0f - metallic) * st;
    float cdf0 = dW, cdf1 = cdf0 + mW, cdf2 = cdf1 + cW, cdf3 = cdf2 + gW;
    didRefract = false;
    float r = rnd(rng) * cdf3;
    choseGlass = false;
    if (r < cdf0) return sample_diffuse(n, rng);
    if (r < cdf1) return sample_metal(tbn, anisotropic, roughness, wi, rng);
    if (r < cdf2) return sample_clearcoat(tbn, cg, wi, rng);
    choseGlass = true;
    return sample_glass(tbn, wi, roughness, anisotropic, eta, rng, didRefract);
}

RT_D V3 eval_disney(const M3& tbn, V3 bc, V3 specularTint, V3 sheenTint, float anisotropic, float roughness,
                    float subsurface, float cg, float eta, float metallic, float clearcoat, float st, float sheen,
                    bool didRefract, uint32_t fixes, V3 n, V3 wi, V3 wo, V3 h, float& pdf) { // :656-711
    // fixes: REINA_RT_TRACE_FIX_METAL_PDF (sampler density for GGX reflection),
    // REINA_RT_TRACE_FIX_DISNEY_MIX (clearcoat inside the normalised mixture,
    // specular = 1 for the metal lobe's dielectric F0, glass branch pdf)
    const bool bounded = (fixes & REINA_RT_TRACE_FIX_METAL_PDF) == 0u;
    const bool mix = (fixes & REINA_RT_TRACE_FIX_DISNEY_MIX) != 0u;
    float dW = (1.0f - st) * (1.0f - metallic);
    float mW = metallic;
    float cW = 0.25f * clearcoat;
    float gW = (1.0f - metallic) * st;
    float ws = (dW + mW) + gW;
    if (mix) ws = ws + cW;
    const float cN = mix ? cW / ws : cW;
    V3 fd = eval_diffuse(roughness, subsurface, bc, n, wi, wo, h);
    float dP = pdf_diffuse(wo, n);
    V3 fs = eval_sheen(bc, wo, h, n, sheenTint);
    V3 fm = eval_metal(tbn, bc, anisotropic, roughness, n, wi, wo, h, mix ? 1.0f : st, specularTint, metallic, eta);
    float mP = pdf_metal(tbn, wi, wo, anisotropic, roughness, bounded);
    V3 fc = eval_clearcoat(tbn, wi, wo, cg, h);
    float cP = pdf_clearcoat(tbn, wi, wo, h, cg);
    float gP;
    V3 fg = eval_glass(tbn, bc, anisotropic, roughness, eta, n, wi, wo, didRefract, bounded, mix, gP);
    pdf = (((dP * dW) / ws + (mP * mW) / ws) + cP * cN) + (gP * gW) / ws;
    V3 r = scale(add(fd, scale(fs, sheen)), dW / ws);
    r = add(r, scale(fm, mW / ws));
    r = add(r, scale(fc, cN));
    r = add(r, scale(fg, gW / ws));
    return r;
}

// ---- next-event estimation at Disney hits ----------------------------------------
// The reference's directLight evaluates evalDisney for materialID 3
// (raytrace.rgen.glsl:57-83) but never runs (skipNEE = true, :146). Here a
// light sample at a Disney hit is weighted so that NEE + MIS estimates the
// same integral as the hit shader's own BSDF sampling: a continuation from a
// diffuse, metal or clearcoat lobe carries f cos / pdf (evalDisney with
// didRefract = false), one from the glass lobe carries weight 1 (the raygen
// skips that bounce's multiply, Q4). The light sample at w therefore carries
// g(w) = q_ng(w) f(w) cos / pdf(w) + q_g(w), with q_ng and q_g the densities
// of sampleDisney's non-glass and glass lobes at w, and the power heuristic
// against q_ng + q_g, the density the continuation was drawn from.

// Density of sampleDisney's directions at wo (brdfDisney.h.glsl:609-654),
// from the samplers themselves rather than the reference's pdf functions
// (which give D > 0 for back-facing microfacets, cut clearcoat at the
// horizon and carry eta^2 in the refraction Jacobian): lobe probabilities
// times
//   diffuse    cos / pi about the shading normal (sampleDiffuse);
//   metal      Heitz's VNDF reflected: G1(i) D(m) / (4 i.z), m = |i + o|,
//              m.z > 0;
//   clearcoat  GTR1 reflected: Dc(h) h.z / (4 |o.h|), h = +-|i + o| with h.z > 0;
//   glass      VNDF reflected times the reflection probability (1 under
//              total internal reflection, else Schlick's F), plus VNDF
//              refracted times 1 - F with Walter et al.'s half-vector
//              Jacobian |o.h| / (eta i.h + o.h)^2, h = +-|o + eta i|.
// png: diffuse, metal and clearcoat; pg: glass. Tangent frame tbn, i.z > 0.
RT_D float vndf_reflection_density(V3 wiT, V3 woT, float ax, float ay) {
    if (!(wiT.z > 0.0f)) return 0.0f;
    const V3 m = normalize(add(wiT, woT));
    if (!(m.z > 0.0f)) return 0.0f;
    return (eval_dm(m, ax, ay) * smith_g(wiT, ax, ay)) / (4.0f * wiT.z);
}

RT_D void disney_density(const M3& tbn, float anisotropic, float roughness, float cg, float eta, float metallic,
                         float clearcoat, float st, V3 n, V3 wi, V3 wo, float& png, float& pg) {
    const float dW = (1.0f - st) * (1.0f - metallic);
    const float mW = metallic;
    const float cW = 0.25f * clearcoat;
    const float gW = (1.0f - metallic) * st;
    const float ws = ((dW + mW) + gW) + cW;
    float ax, ay;
    alphas(anisotropic, roughness, ax, ay);
    const V3 wiT = tmul(tbn, wi), woT = tmul(tbn, wo);
    const float dP = pdf_diffuse(wo, n);
    const float mr = vndf_reflection_density(wiT, woT, ax, ay);
    const float ag = (1.0f - cg) * 0.1f + cg * 0.001f;
    V3 hc = normalize(add(wiT, woT));
    if (hc.z < 0.0f) hc = neg(hc);
    const float ohc = fabsf(dot(woT, hc));
    const float cP = ohc > 0.0f ? (eval_dc(ag, hc) * hc.z) / (4.0f * ohc) : 0.0f;
    float gr = 0.0f, gt = 0.0f;
    if (mr > 0.0f) {
        const V3 m = normalize(add(wiT, woT));
        const float c = dot(wiT, m);
        const float sn = sqrtf_(1.0f - c * c);
        gr = mr * ((eta * sn > 1.0f) ? 1.0f : reflectance(c, eta));
    }
    if (wiT.z > 0.0f) {   // refraction can also leave above the horizon through a tilted microfacet
        V3 h = normalize(add(woT, scale(wiT, eta)));
        if (h.z < 0.0f) h = neg(h);
        const float ih = dot(wiT, h), oh = dot(woT, h);
        const float sn = sqrtf_(1.0f - ih * ih);
        if (h.z > 0.0f && ih > 0.0f && oh < 0.0f && !(eta * sn > 1.0f)) {
            const float den = eta * ih + oh;
            const float dv = ((smith_g(wiT, ax, ay) * ih) * eval_dm(h, ax, ay)) / wiT.z;
            gt = ((dv * fabsf(oh)) / (den * den)) * (1.0f - reflectance(ih, eta));
        }
    }
    png = ((dP * dW) / ws + (mr * mW) / ws) + (cP * cW) / ws;
    pg = ((gr + gt) * gW) / ws;
}

// What a Disney hit leaves for next-event estimation (chs_disney fills it).
struct DisneyNee {
    M3 tbn;
    V3 albedo, wn, wi, pos, ng;
    float eta;
    float p_bsdf;   // q_ng + q_g of the direction the hit shader sampled
};

// g(w) of a light sample along w, and the sampler's density there (ptot).
RT_D V3 disney_nee_weight(const DisneyNee& c, const ReinaInstanceProperties& p, uint32_t fixes, V3 w, float& ptot) {
    const V3 h = normalize(add(w, c.wi));
    float pdf_ev;
    const V3 f = eval_disney(c.tbn, c.albedo, ldv3(p.specularTint), ldv3(p.sheenTint), p.anisotropic, p.roughness,
                             p.subsurface, p.clearcoatGloss, c.eta, p.metallic, p.clearcoat, p.specularTransmission,
                             p.sheen, false, fixes, c.wn, c.wi, w, h, pdf_ev);
    float png, pg;
    disney_density(c.tbn, p.anisotropic, p.roughness, p.clearcoatGloss, c.eta, p.metallic, p.clearcoat,
                   p.specularTransmission, c.wn, c.wi, w, png, pg);
    ptot = png + pg;
    const float cx = maxf_(dot(c.wn, w), 0.0f);
    V3 g = v3(pg, pg, pg);
    if (pdf_ev > 0.0f && cx > 0.0f) g = add(scale(f, (png / pdf_ev) * cx), g);
    return g;
}

// ---- hit shaders ---------------------------------------------------------------
// ---- textures -------------------------------------------------------------
// The hit shaders' texture(textures[id], uv): RGBA8 UNORM, linear filter,
// repeat addressing (sampler of src/tools/vktools.cpp:765-788, one mip level,
// src/graphics/Image.cpp:56). A ray-tracing stage has no derivatives, so
// texture() samples the base level and the sampler's anisotropy has no
// effect: the result is the bilinear filter of the Vulkan specification
// (texel coordinates u * w - 0.5, v * h - 0.5), evaluated here in FP32 as two
// linear interpolations along x and one along y. Hardware filtering rounds its weights to a few
// fraction bits, so texel-edge values of the reference are implementation
// defined; this order is the library's definition.
RT_D float4 unorm8(uint32_t t) {
    return make_float4((float)(t & 0xFFu) / 255.0f, (float)((t >> 8) & 0xFFu) / 255.0f,
                       (float)((t >> 16) & 0xFFu) / 255.0f, (float)(t >> 24) / 255.0f);
}

// floor(x) mod n for a floored, finite x with |x| <= 2^30
RT_D uint32_t repeat_index(float fx, uint32_t n) {
    const int32_t m = (int32_t)fx % (int32_t)n;
    return (uint32_t)(m < 0 ? m + (int32_t)n : m);
}

RT_D float4 tex_sample(const SceneView& sv, int32_t id, float u, float v) {
    const TexInfo t = sv.tex[id];
    float x = u * (float)t.w - 0.5f, y = v * (float)t.h - 0.5f;
    // non-finite coordinates (a NaN uv) read texel 0; huge ones are clamped so
    // the integer conversion stays defined
    x = (x == x) ? clampf_(x, -1073741824.0f, 1073741824.0f) : 0.0f;
    y = (y == y) ? clampf_(y, -1073741824.0f, 1073741824.0f) : 0.0f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const uint32_t i0 = repeat_index(fx, t.w), i1 = repeat_index(fx + 1.0f, t.w);
    const uint32_t j0 = repeat_index(fy, t.h), j1 = repeat_index(fy + 1.0f, t.h);
    const uint32_t* row0 = sv.texels + t.base + (uint64_t)j0 * t.w;
    const uint32_t* row1 = sv.texels + t.base + (uint64_t)j1 * t.w;
    const float4 t00 = unorm8(row0[i0]), t10 = unorm8(row0[i1]), t01 = unorm8(row1[i0]), t11 = unorm8(row1[i1]);
    // two linear interpolations along x, then one along y: exact for equal texels
    const auto lerp2 = [a, b](float p00, float p10, float p01, float p11) {
        const float top = p00 + a * (p10 - p00), bottom = p01 + a * (p11 - p01);
        return top + b * (bottom - top);
    };
    return make_float4(lerp2(t00.x, t10.x, t01.x, t11.x), lerp2(t00.y, t10.y, t01.y, t11.y),
                       lerp2(t00.z, t10.z, t01.z, t11.z), lerp2(t00.w, t10.w, t01.w, t11.w));
}

// texutils.h.glsl:4-41 (parallax occlusion search). The layer loop ends
// within 0.2 * 512 + 1 steps for finite input; the cap only guards NaN.
constexpr int kBumpMaxSteps = 1024;
RT_D void bump_mapping(const SceneView& sv, int32_t id, float& u, float& v, V3 rayIn, const M3& T) {
    const float heightScale = 0.2f, minLayers = 64.0f, maxLayers = 512.0f;
    const V3 V = normalize(tmul(T, neg(rayIn)));
    if (V.z <= 0.0f) return;
    const float numLayers = mixf(maxLayers, minLayers, clampf_(V.z, 0.0f, 1.0f));
    const float layerDepth = 1.0f / numLayers;
    const float Px = V.x / V.z * heightScale, Py = V.y / V.z * heightScale;
    const float du = Px / numLayers, dv = Py / numLayers;
    float cu = u, cv = v, depthSum = 0.0f;
    float h = tex_sample(sv, id, cu, cv).x * heightScale;
    for (int k = 0; k < kBumpMaxSteps && depthSum < h; ++k) {
        cu = cu - du;
        cv = cv - dv;
        depthSum = depthSum + layerDepth;
        h = tex_sample(sv, id, cu, cv).x * heightScale;
    }
    const float pu = cu + du, pv = cv + dv;
    const float hPrev = tex_sample(sv, id, pu, pv).x * heightScale;
    const float sumPrev = depthSum - layerDepth;
    const float after = h - depthSum;
    const float before = hPrev - sumPrev;
    const float weight = after / (after - before);
    u = mixf(pu, cu, weight);
    v = mixf(pv, cv, weight);
}

// tangentNormal = texture(normalMap, uv).rgb * 2 - 1; y flipped;
// normalize(tbn * tangentNormal) (lambertian.rchit.glsl:37-41)
RT_D V3 normal_map(const SceneView& sv, int32_t id, float u, float v, const M3& tbn) {
    const float4 t = tex_sample(sv, id, u, v);
    V3 tn = v3(t.x * 2.0f - 1.0f, t.y * 2.0f - 1.0f, t.z * 2.0f - 1.0f);
    tn.y = -tn.y;
    return normalize(mmul(tbn, tn));
}

RT_D void skip_hit(Payload& pld, const HitInfo& hi, V3 D) {                // closestHitCommon.h.glsl:185-193
    pld.rayOrigin = offset_along(hi.worldPosition, neg(hi.worldNormal));
    pld.rayDirection = D;
    pld.rayHitSky = false;
    pld.skip = true;
}

// lambertian.rchit.glsl:11-78 (metal = false) and metal.rchit.glsl:7-70 (metal = true)
RT_D void chs_lambert_metal(Payload& pld, const SceneView& sv, const HitInfo& hi, const ReinaInstanceProperties& p,
                            V3 D, bool metal) {
    if (p.cullBackface != 0u && !hi.frontFace) { skip_hit(pld, hi, D); return; }
    float ux = hi.uvx - floorf(hi.uvx), uy = hi.uvy - floorf(hi.uvy);   // mod(uv, 1.0)
    if (p.bumpMapTexID >= 0) bump_mapping(sv, p.bumpMapTexID, ux, uy, normalize(D), hi.tbn);
    if (!metal && (ux < 0.0f || ux > 1.0f || uy < 0.0f || uy > 1.0f)) { skip_hit(pld, hi, D); return; }
    V3 wn = hi.worldNormal;
    if (p.normalMapTexID >= 0) wn = normal_map(sv, p.normalMapTexID, ux, uy, hi.tbn);
    pld.color = ldv3(p.albedo);
    if (p.textureID >= 0) {
        const float4 tc = tex_sample(sv, p.textureID, ux, uy);
        if (tc.w < 0.999f && rnd(pld.rng) > tc.w) { skip_hit(pld, hi, D); return; }   // transparency
        pld.color = mul(pld.color, v3(tc.x, tc.y, tc.z));
    }
    pld.albedo = pld.color;
    pld.surfaceNormal = wn;
    pld.emission = ldv3(p.emission);
    pld.rayOrigin = offset_along(hi.worldPosition, hi.worldNormalGeometry);
    if (!metal) pld.rayDirection = diffuse_reflection(wn, pld.rng);
    else pld.rayDirection = fuzzy_reflection(D, wn, p.roughness, pld.rng);
    pld.rayHitSky = false;
    pld.skip = false;
    pld.insideDielectric = false;
    pld.accumulatedDistance = 0.0f;
}

// dielectric.rchit.glsl:40-113
RT_D void chs_dielectric(Payload& pld, const SceneView& sv, const HitInfo& hi, const ReinaInstanceProperties& p, V3 O,
                         V3 D) {
    float ri = hi.frontFace ? 1.0f / p.ior : p.ior;
    V3 unitDir = normalize(D);
    float cosTheta = minf_(dot(neg(unitDir), hi.worldNormal), 1.0f);
    float sinTheta = sqrtf_(1.0f - cosTheta * cosTheta);
    bool cannot = ri * sinTheta > 1.0f;
    float refl = reflectance(cosTheta, ri);
    bool prevInside = pld.insideDielectric;
    float ux = hi.uvx - floorf(hi.uvx), uy = hi.uvy - floorf(hi.uvy);   // mod(uv, 1.0)
    if (p.bumpMapTexID >= 0) bump_mapping(sv, p.bumpMapTexID, ux, uy, normalize(D), hi.tbn);
    V3 albedo = ldv3(p.albedo);
    if (p.textureID >= 0) {
        const float4 tc = tex_sample(sv, p.textureID, ux, uy);
        albedo = mul(albedo, v3(tc.x, tc.y, tc.z));
    }
    V3 wn = hi.worldNormal;
    if (p.normalMapTexID >= 0) wn = normal_map(sv, p.normalMapTexID, ux, uy, hi.tbn);
    if (cannot || refl > rnd(pld.rng)) {
        pld.rayDirection = fuzzy_reflection(unitDir, wn, p.roughness, pld.rng);
        pld.color = v3(1.0f, 1.0f, 1.0f);
        pld.rayOrigin = offset_along(hi.worldPosition, hi.worldNormalGeometry);
    } else {
        pld.insideDielectric = hi.frontFace;
        pld.rayDirection = add(refract(unitDir, wn, ri), scale(random_unit_vec(pld.rng), p.roughness));
        pld.color = albedo;
        pld.rayOrigin = offset_dielectric(hi.worldPosition, hi.worldNormalGeometry, unitDir);
    }
    pld.accumulatedDistance = pld.accumulatedDistance + (prevInside ? length(sub(hi.worldPosition, O)) : 0.0f);
    if (prevInside && !pld.insideDielectric) {
        float e = expf_(-p.absorption * pld.accumulatedDistance);
        pld.color = mul(v3(e, e, e), albedo);
        pld.accumulatedDistance = 0.0f;
    }
    pld.albedo = pld.color;
    pld.surfaceNormal = wn;
    pld.emission = ldv3(p.emission);
    pld.rayHitSky = false;
    pld.skip = false;
}

// disney.rchit.glsl:36-198
// fixes: the REINA_RT_TRACE_FIX_* bits of the trace (eval_disney reads two)
RT_D void chs_disney(Payload& pld, const SceneView& sv, const HitInfo& hi, const ReinaInstanceProperties& p, V3 O,
                     V3 D, uint32_t fixes, DisneyNee* dn = nullptr) {
    if (p.cullBackface != 0u && !hi.frontFace) { skip_hit(pld, hi, D); return; }
    float ux = hi.uvx - floorf(hi.uvx), uy = hi.uvy - floorf(hi.uvy);   // mod(uv, 1.0)
    if (p.bumpMapTexID >= 0) bump_mapping(sv, p.bumpMapTexID, ux, uy, normalize(D), hi.tbn);
    if (ux < 0.0f || ux > 1.0f || uy < 0.0f || uy > 1.0f) { skip_hit(pld, hi, D); return; }
    V3 wn = hi.worldNormal;
    if (p.normalMapTexID >= 0) wn = normal_map(sv, p.normalMapTexID, ux, uy, hi.tbn);
    V3 albedo = ldv3(p.albedo);
    if (p.textureID >= 0) {
        const float4 tc = tex_sample(sv, p.textureID, ux, uy);
        if (tc.w < 0.999f && rnd(pld.rng) > tc.w) { skip_hit(pld, hi, D); return; }   // transparency
        albedo = mul(albedo, v3(tc.x, tc.y, tc.z));
    }
    float eta = hi.frontFace ? 1.0f / p.ior : p.ior;
    bool didRefract = false, choseGlass = false;
    V3 wi = neg(D);
    V3 rayDir = sample_disney(hi.tbn, p.anisotropic, p.roughness, p.clearcoatGloss, eta, p.metallic, p.clearcoat,
                              p.specularTransmission, wn, wi, didRefract, choseGlass, pld.rng);
    V3 h = normalize(add(rayDir, wi));
    float pdf;
    V3 f = eval_disney(hi.tbn, albedo, ldv3(p.specularTint), ldv3(p.sheenTint), p.anisotropic, p.roughness,
                       p.subsurface, p.clearcoatGloss, eta, p.metallic, p.clearcoat, p.specularTransmission, p.sheen,
                       didRefract, fixes, wn, wi, rayDir, h, pdf);
    float cosThetaI = maxf_(dot(wn, rayDir), 0.0f);
    pld.color = divs(scale(f, cosThetaI), pdf);
    pld.albedo = albedo;
    pld.surfaceNormal = wn;
    pld.emission = ldv3(p.emission);
    pld.rayOrigin = offset_dielectric(hi.worldPosition, hi.worldNormalGeometry, rayDir);
    pld.rayDirection = rayDir;
    pld.rayHitSky = false;
    pld.skip = false;
    pld.insideDielectric = choseGlass;
    if (pld.insideDielectric)
        pld.accumulatedDistance = pld.accumulatedDistance + length(sub(hi.worldPosition, O));
    else
        pld.accumulatedDistance = 0.0f;
    if (dn) {
        dn->tbn = hi.tbn;
        dn->albedo = albedo;
        dn->wn = wn;
        dn->wi = wi;
        dn->pos = hi.worldPosition;
        dn->ng = hi.worldNormalGeometry;
        dn->eta = eta;
        float png, pg;
        disney_density(hi.tbn, p.anisotropic, p.roughness, p.clearcoatGloss, eta, p.metallic, p.clearcoat,
                       p.specularTransmission, wn, wi, rayDir, png, pg);
        dn->p_bsdf = png + pg;
    }
}

// ---- raygen camera (raytrace.rgen.glsl:34-41, 194-247) ------------------------
struct Camera {
    float invView[16];
    float invProj[16];
    float focusDist, defocus;
    float W, H;
};

RT_D void m4mul(const float* m, float x, float y, float z, float w, float* o) {
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r] * w;
}

// lens_advances: the lens draws advance rng (REINA_RT_TRACE_FIX_LENS_RNG)
RT_D void starting_ray(const Camera& cam, float px, float py, uint32_t& rng, V3& o, V3& d, bool lens_advances) {
    float u1 = maxf_(1e-5f, rnd(rng));
    float u2 = rnd(rng);
    float r = sqrtf_(-2.0f * logf_(u1));
    float theta = (2.0f * kPi) * u2;
    float s, c;
    sincos_(theta, &s, &c);
    float gx = r * c, gy = r * s;
    float cx = (px + 0.5f) + 0.375f * gx;
    float cy = (py + 0.5f) + 0.375f * gy;
    float ndcx = (cx / cam.W) * 2.0f - 1.0f;
    float ndcy = -((cy / cam.H) * 2.0f - 1.0f);
    float vp[4];
    m4mul(cam.invProj, ndcx, ndcy, -1.0f, 1.0f, vp);
    float w = vp[3];
    V3 viewDir = normalize(v3(vp[0] / w, vp[1] / w, vp[2] / w));
    float wd[4];
    m4mul(cam.invView, viewDir.x, viewDir.y, viewDir.z, 0.0f, wd);
    V3 rayDirection = normalize(v3(wd[0], wd[1], wd[2]));
    V3 origin = v3(cam.invView[12], cam.invView[13], cam.invView[14]);
    V3 focal = add(origin, scale(rayDirection, cam.focusDist));
    // randomInUnitHexagon takes the state by value: the stream does not advance
    uint32_t hr = rng;
    const float sqrt3 = 1.73205080757f;
    float hx, hy;
    do {
        hx = 2.0f * rnd(hr) - 1.0f;
        hy = (rnd(hr) - 0.5f) * sqrt3;
    } while (fabsf(hy) > sqrt3 * 0.5f || (sqrt3 * fabsf(hx) + fabsf(hy)) > sqrt3);
    if (lens_advances) rng = hr;
    float lx = hx * cam.defocus, ly = hy * cam.defocus;
    V3 right = normalize(v3(cam.invView[0], cam.invView[1], cam.invView[2]));
    V3 up = normalize(v3(cam.invView[4], cam.invView[5], cam.invView[6]));
    V3 no = add(origin, add(scale(right, lx), scale(up, ly)));
    o = no;
    d = normalize(sub(focal, no));
}

}  // namespace rt
