/*
 * or_shade.c — ORACLE (test infrastructure only): CPU restatement of the
 * raygen, closest-hit and miss shaders of Reina (shaders/raytrace/ *.glsl),
 * structured like the shaders themselves: one HitPayload per pixel, a
 * sequential spp loop, a bounce loop calling one "hit shader" per segment.
 * NEE is disabled exactly as in the reference (raytrace.rgen.glsl:146,
 * skipNEE = true), so weightBRDF = 1 and the direct term is 0.
 *
 * Reference behaviours reproduced (each tested): one RNG state per pixel
 * shared by all samples and bounces (raytrace.rgen.glsl:259-264), the hexagon
 * lens sample drawn on a copy of the state (:186-204), Disney sets
 * insideDielectric = choseGlass (disney.rchit.glsl:191), the evalDisney lobe
 * weights (brdfDisney.h.glsl:656-711), directClamp on the whole sample
 * (raytrace.rgen.glsl:266), accumulatedDistance not reset per path (:98;
 * initial value 0), random() = float(w) * 2^-32 (shaderCommon.h.glsl:39-45),
 * NaN samples dropped and NaN when every sample is dropped
 * (raytrace.rgen.glsl:268-277), the world normal by mat3(O2W) and the TBN read
 * even without interpNormals (closestHitCommon.h.glsl:113-136; a model without
 * TBN data reads zeros), metal sampled by the unbounded VNDF but priced by the
 * bounded-VNDF pdf (brdfDisney.h.glsl:9-73). or_scene_set_nee switches on this
 * build's correct NEE + MIS instead.
 *
 * evalDisney's mixture keeps the reference's three odd terms (clearcoat
 * outside the normalising sum, specularTransmission passed as `specular`, no
 * glass branch probability). Each is pinned by a closed form built from
 * separately pinned lobes (tests/test_oracle_disney_mixture.py): at normal
 * incidence f and pdf reduce to D = 1/(pi ax ay), the bounded-VNDF constant
 * and GTR1; a pure-glass pdf integrates to the bounded-VNDF reflection mass
 * plus eta^2 times the refracted VNDF mass; with metallic = 0 and no
 * transmission the pdf is (1 + clearcoat/4) times sampleDisney's density.
 * With REINA_RT_TRACE_FIX_DISNEY_MIX the terms are replaced, and the mixture
 * pdf is then the density of sampleDisney. Each lobe beneath is pinned
 * against its mathematics through or_disney_probe.
 *
 * OR_PLANT (a compile-time switch, 0 by default) builds a copy of the oracle
 * with one deliberate mistake, so the tests can show that their checks catch
 * it: 1 clearcoat inside the sum, 2 specular = 1 in the reference mixture,
 * 3 glass pdf times its branch probability in the reference mixture, 4 the
 * refraction Jacobian without eta^2 (or_post.c holds 5-9), 10 a Disney light
 * sample weighted by f cos alone, 11 without its glass term.
 */
#include "oracle.h"

#ifndef OR_PLANT
#define OR_PLANT 0
#endif

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define K_PI 3.14159265f       /* shaderCommon.h.glsl:47 */
#define K_INV_PI 0.31830989f   /* shaderCommon.h.glsl:48 */

void or_closest_hit(const OrScene* s, const ReinaRay* ray, ReinaHit* hit, int use_bvh);

/* shaderCommon.h.glsl:39-45 */
float or_random(uint32_t* st) {
    *st = *st * 747796405u + 1u;
    uint32_t w = ((*st >> ((*st >> 28) + 4u)) ^ *st) * 277803737u;
    w = (w >> 22) ^ w;
    return (float)w / 4294967295.0f;
}

void or_pcg_stream(uint32_t seed, uint32_t n, uint32_t* states, float* values) {
    uint32_t s = seed;
    for (uint32_t i = 0; i < n; ++i) {
        values[i] = or_random(&s);
        states[i] = s;
    }
}

/* ---- payload: shaderCommon.h.glsl:18-36 ---------------------------------- */
typedef struct {
    v3 albedo, color, rayOrigin, rayDirection;
    uint32_t rngState;
    int rayHitSky;
    v3 emission, surfaceNormal;
    uint32_t materialID;
    int skip;
    float pdf;
    float accumulatedDistance;
    int insideDielectric;
    m3 tbn;
    float eta;
    int didRefract;
} Payload;

typedef struct {
    v3 objectPosition, worldPosition, worldNormal, worldNormalGeometry;
    float uvx, uvy;
    int frontFace;
    m3 tbn;
} HitInfo;

typedef struct {
    v3 worldRayOrigin, worldRayDirection;
    const OrInstance* inst;
    const ReinaInstanceProperties* props;
    uint32_t prim;
    float bu, bv;
} HitCtx;

static v3 pv3(const float* f) { return V3(f[0], f[1], f[2]); }

static v3 tbn_col(const OrScene* s, uint32_t idx_pos, int col) {
    if (idx_pos >= s->ntbni) return V3(0.0f, 0.0f, 0.0f);
    uint32_t ti = s->tbn_indices[idx_pos];
    if (ti >= s->ntbn) return V3(0.0f, 0.0f, 0.0f);
    const float* m = &s->tbns[9ull * ti + 3 * col];
    return V3(m[0], m[1], m[2]);
}

/* closestHitCommon.h.glsl:52-148 */
static HitInfo getObjectHitInfo(const OrScene* s, const HitCtx* c) {
    HitInfo r;
    const ReinaInstanceProperties* props = c->props;
    uint32_t prim = c->prim;
    uint32_t io = props->indicesOffset;
    uint32_t i0 = s->indices[3u * prim + io + 0];
    uint32_t i1 = s->indices[3u * prim + io + 1];
    uint32_t i2 = s->indices[3u * prim + io + 2];
    v3 v0 = pv3(&s->vertices[4ull * i0]);
    v3 v1 = pv3(&s->vertices[4ull * i1]);
    v3 v2 = pv3(&s->vertices[4ull * i2]);
    float bx = (1.0f - c->bu) - c->bv, by = c->bu, bz = c->bv;
    r.objectPosition = vadd(vadd(vscale(v0, bx), vscale(v1, by)), vscale(v2, bz));
    const OrInstance* in = c->inst;
    r.worldPosition = vadd(vadd(vadd(vscale(in->c0, r.objectPosition.x), vscale(in->c1, r.objectPosition.y)),
                                vscale(in->c2, r.objectPosition.z)), in->c3);
    v3 ong = vnormalize(vcross(vsub(v1, v0), vsub(v2, v0)));
    v3 on;
    uint32_t to = props->tbnsIndicesOffset;
    if (!props->interpNormals) {
        on = ong;
    } else {
        v3 n0 = vnormalize(tbn_col(s, 3u * prim + to + 0, 2));
        v3 n1 = vnormalize(tbn_col(s, 3u * prim + to + 1, 2));
        v3 n2 = vnormalize(tbn_col(s, 3u * prim + to + 2, 2));
        on = vnormalize(vadd(vadd(vscale(n0, bx), vscale(n1, by)), vscale(n2, bz)));
    }
    if (props->texIndicesOffset == 0xFFFFFFFFu) {
        r.uvx = 0.0f; r.uvy = 0.0f;
    } else {
        uint32_t tio = props->texIndicesOffset;
        float u[3], w[3];
        for (int k = 0; k < 3; ++k) {
            uint32_t ti = s->tex_indices[3u * prim + tio + k];
            int ok = (uint64_t)ti < s->ntc;   /* 0xFFFFFFFF entries read as (0, 0) */
            u[k] = ok ? s->tex_coords[2ull * ti] : 0.0f;
            w[k] = ok ? s->tex_coords[2ull * ti + 1] : 0.0f;
        }
        r.uvx = (u[0] * bx + u[1] * by) + u[2] * bz;
        r.uvy = (w[0] * bx + w[1] * by) + w[2] * bz;
    }
    m3 M = {in->c0, in->c1, in->c2};
    /* mat3(O2W) as the reference; its inverse-transpose with REINA_RT_TRACE_FIX_NORMALS */
    float ndet = vdot(M.c0, vcross(M.c1, M.c2));
    float ninv = 1.0f / ndet;
    m3 NI = {vscale(vcross(M.c1, M.c2), ninv), vscale(vcross(M.c2, M.c0), ninv), vscale(vcross(M.c0, M.c1), ninv)};
    m3 NM = (s->fixes & REINA_RT_TRACE_FIX_NORMALS) ? NI : M;
    r.worldNormal = vnormalize(m3mul(NM, on));
    r.worldNormalGeometry = vnormalize(m3mul(NM, ong));
    v3 D = c->worldRayDirection;
    r.frontFace = vdot(D, r.worldNormalGeometry) < 0.0f;
    r.worldNormal = vfaceforward(r.worldNormal, D, r.worldNormalGeometry);
    r.worldNormalGeometry = vfaceforward(r.worldNormalGeometry, D, r.worldNormalGeometry);
    /* TBN (read even without interpNormals) */
    v3 t1 = tbn_col(s, 3u * prim + to + 0, 0), b1 = tbn_col(s, 3u * prim + to + 0, 1), n1 = tbn_col(s, 3u * prim + to + 0, 2);
    v3 t2 = tbn_col(s, 3u * prim + to + 1, 0), b2 = tbn_col(s, 3u * prim + to + 1, 1), n2 = tbn_col(s, 3u * prim + to + 1, 2);
    v3 t3 = tbn_col(s, 3u * prim + to + 2, 0), b3 = tbn_col(s, 3u * prim + to + 2, 1), n3 = tbn_col(s, 3u * prim + to + 2, 2);
    v3 tangent = vnormalize(vadd(vadd(vscale(t1, bx), vscale(t2, by)), vscale(t3, bz)));
    v3 bitangent = vnormalize(vadd(vadd(vscale(b1, bx), vscale(b2, by)), vscale(b3, bz)));
    v3 normal = vnormalize(vadd(vadd(vscale(n1, bx), vscale(n2, by)), vscale(n3, bz)));
    /* N = transpose(inverse(M)) = [b x c, c x a, a x b] / det for M = [a b c] */
    float det = vdot(M.c0, vcross(M.c1, M.c2));
    float invdet = 1.0f / det;
    m3 N = {vscale(vcross(M.c1, M.c2), invdet), vscale(vcross(M.c2, M.c0), invdet), vscale(vcross(M.c0, M.c1), invdet)};
    v3 wT = vnormalize(m3mul(M, tangent));
    v3 wB = vnormalize(m3mul(M, bitangent));
    v3 wN = vnormalize(m3mul(N, normal));
    wT = vnormalize(vsub(wT, vscale(wN, vdot(wN, wT))));
    wB = vnormalize(vsub(wB, vscale(wN, vdot(wN, wB))));
    wB = vscale(wB, -1.0f);
    wN = vscale(wN, r.frontFace ? 1.0f : -1.0f);
    r.tbn.c0 = wT; r.tbn.c1 = wB; r.tbn.c2 = wN;
    return r;
}

/* closestHitCommon.h.glsl:156-177 (dielectric: dielectric.rchit.glsl:14-38) */
static float offc(float p, float n) {
    int32_t of = or_f2i(256.0f * n);
    uint32_t bits = or_f2u(p);
    uint32_t nb = bits + (uint32_t)((p < 0.0f) ? -of : of);
    float pi = or_u2f(nb);
    return (fabsf(p) < (1.0f / 32.0f)) ? p + (1.0f / 65536.0f) * n : pi;
}
static v3 offsetPositionAlongNormal(v3 p, v3 n) { return V3(offc(p.x, n.x), offc(p.y, n.y), offc(p.z, n.z)); }
static v3 offsetPositionForDielectric(v3 p, v3 n, v3 dir) {
    v3 on = (vdot(n, dir) < 0.0f) ? vneg(n) : n;
    return offsetPositionAlongNormal(p, on);
}

void or_offset_position(const float* p, const float* n, float* out, int dielectric, const float* dir) {
    v3 r = dielectric ? offsetPositionForDielectric(pv3(p), pv3(n), pv3(dir))
                      : offsetPositionAlongNormal(pv3(p), pv3(n));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* closestHitCommon.h.glsl:185-193 */
static void skip_hit(Payload* pld, const HitInfo* hi, const HitCtx* c) {
    pld->rayOrigin = offsetPositionAlongNormal(hi->worldPosition, vneg(hi->worldNormal));
    pld->rayDirection = c->worldRayDirection;
    pld->rayHitSky = 0;
    pld->skip = 1;
}

/* closestHitCommon.h.glsl:195-205 */
static v3 randomUnitVec(uint32_t* rng) {
    for (;;) {
        float a = or_random(rng);
        float b = or_random(rng);
        float c = or_random(rng);
        v3 v = vsub(vscale(V3(a, b, c), 2.0f), V3(1.0f, 1.0f, 1.0f));
        float l2 = vdot(v, v);
        if (0.0001f < l2 && l2 < 1.0f) return vnormalize(v);
    }
}
/* closestHitCommon.h.glsl:207-213 */
static v3 addFuzz(v3 ray, float fuzz, uint32_t* rng) { return vadd(ray, vscale(randomUnitVec(rng), fuzz)); }
static v3 fuzzyReflection(v3 rayIn, v3 n, float fuzz, uint32_t* rng) {
    return vadd(vreflect(vnormalize(rayIn), vnormalize(n)), vscale(randomUnitVec(rng), fuzz));
}
/* closestHitCommon.h.glsl:215-222 */
static v3 diffuseReflection(v3 n, uint32_t* rng) {
    float theta = (2.0f * K_PI) * or_random(rng);
    float u = 2.0f * or_random(rng) - 1.0f;
    float r = sqrtf(1.0f - u * u);
    v3 d = vadd(n, V3(r * or_cosf(theta), r * or_sinf(theta), u));
    return vnormalize(d);
}

void or_diffuse_dir(const float* n, uint32_t* rng, float* out) {
    v3 d = diffuseReflection(pv3(n), rng);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

static float fmodf1(float x) { return x - floorf(x); } /* GLSL mod(x, 1.0) */

/* ====================== Disney BSDF: brdfDisney.h.glsl ==================== */
static v3 sampleGGXVNDF(v3 V, float ax, float ay, uint32_t* rng) {           /* :9-40 */
    int flip = V.z < 0.0f;
    if (flip) V.z = V.z * -1.0f;
    float r1 = or_random(rng);
    float r2 = or_random(rng);
    v3 Vh = vnormalize(V3(ax * V.x, ay * V.y, V.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    v3 T1 = lensq > 0.0f ? vscale(V3(-Vh.y, Vh.x, 0.0f), 1.0f / sqrtf(lensq)) : V3(1.0f, 0.0f, 0.0f);
    v3 T2 = vcross(Vh, T1);
    float r = sqrtf(r1);
    float phi = (2.0f * K_PI) * r2;
    float t1 = r * or_cosf(phi);
    float t2 = r * or_sinf(phi);
    float s = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
    v3 Nh = vadd(vadd(vscale(T1, t1), vscale(T2, t2)), vscale(Vh, sqrtf(or_maxf(0.0f, (1.0f - t1 * t1) - t2 * t2))));
    if (flip) Nh.z = Nh.z * -1.0f;
    return vnormalize(V3(ax * Nh.x, ay * Nh.y, or_maxf(0.0f, Nh.z)));
}

static float D_GGX_Aniso(v3 m, float ax, float ay) {                          /* :42-48 */
    float NoM = or_maxf(m.z, 0.0f);
    float Tm = m.x, Bm = m.y;
    float inv = 1.0f / (((Tm / ax) * (Tm / ax) + (Bm / ay) * (Bm / ay)) + NoM * NoM);
    return (inv * inv) / ((K_PI * ax) * ay);
}

/* bounded = 0: k = 1, the density of sampleGGXVNDF (REINA_RT_TRACE_FIX_METAL_PDF) */
static float pdfGGXReflection(v3 i, v3 o, float ax, float ay, int bounded) {  /* :55-73 */
    v3 m = vnormalize(vadd(i, o));
    float ndf = D_GGX_Aniso(m, ax, ay);
    float aix = ax * i.x, aiy = ay * i.y;
    float len2 = aix * aix + aiy * aiy;
    float t = sqrtf(len2 + i.z * i.z);
    if (i.z >= 0.0f) {
        float a = or_clampf(or_minf(ax, ay), 0.0f, 1.0f);
        float s = 1.0f + sqrtf(i.x * i.x + i.y * i.y);
        float a2 = a * a, s2 = s * s;
        float k = bounded ? ((1.0f - a2) * s2) / (s2 + (a2 * i.z) * i.z) : 1.0f;
        return ndf / (2.0f * (k * i.z + t));
    }
    return (ndf * (t - i.z)) / (2.0f * len2);
}

static v3 sampleGTR1_attempt2(float alpha, uint32_t* rng) {                  /* :96-111 */
    float r1 = or_random(rng);
    float r2 = or_random(rng);
    float a = or_maxf(0.001f, alpha);
    float c = sqrtf((1.0f - or_expf(r1 * or_logf(a * a))) / (1.0f - a * a));
    /* sin(acos(c)) for c in [-1, 1]; acos is NaN outside */
    float sE = (c > 1.0f || c < -1.0f || c != c) ? NAN : sqrtf(1.0f - c * c);
    float hAz = (2.0f * K_PI) * r2;
    return V3(sE * or_cosf(hAz), sE * or_sinf(hAz), c);
}

static float FD(float fd90, v3 n, v3 w) {                                     /* :128-131 */
    return 1.0f + (fd90 - 1.0f) * or_pow5(1.0f - or_maxf(vdot(n, w), 0.0f));
}

static v3 evalDiffuse(float roughness, float subsurface, v3 bc, v3 n, v3 wi, v3 wo, v3 h) { /* :124-168 */
    float hw = or_maxf(vdot(h, wo), 0.0f);
    float fd90 = 0.5f + ((2.0f * roughness) * hw) * hw;
    float FDin = FD(fd90, n, wi), FDout = FD(fd90, n, wo);
    v3 base = vscale(vscale(V3(bc.x / K_PI, bc.y / K_PI, bc.z / K_PI), FDin), FDout);
    v3 k = vscale(vscale(bc, 1.25f), K_INV_PI);
    float hw2 = or_maxf(vdot(h, wo), 0.0f);
    float fss90 = (roughness * hw2) * hw2;
    float fssIn = 1.0f + (fss90 - 1.0f) * or_pow5(1.0f - or_maxf(vdot(n, wi), 0.0f));
    float fssOut = 1.0f + (fss90 - 1.0f) * or_pow5(1.0f - or_maxf(vdot(n, wo), 0.0f));
    float third = 1.0f / (or_maxf(vdot(n, wi), 0.0f) + or_maxf(vdot(n, wo), 0.0f)) - 0.5f;
    v3 sub = vscale(k, (fssIn * fssOut) * third + 0.5f);
    return vmix(base, sub, subsurface);
}

static float pdfDiffuse(v3 wo, v3 n) {                                        /* :170-179 */
    float c = vdot(n, wo);
    if (c <= 0.0f) return 0.0f;
    return c * K_INV_PI;
}

static v3 sampleDiffuse(v3 n, uint32_t* rng) {                                /* :181-201 */
    float xi1 = or_random(rng);
    float xi2 = or_random(rng);
    float r = sqrtf(xi1);
    float phi = (2.0f * K_PI) * xi2;
    float x = r * or_cosf(phi);
    float y = r * or_sinf(phi);
    float z = sqrtf(or_maxf(0.0f, 1.0f - xi1));
    v3 t = fabsf(n.x) < 0.5f ? vnormalize(vcross(n, V3(1.0f, 0.0f, 0.0f))) : vnormalize(vcross(n, V3(0.0f, 1.0f, 0.0f)));
    v3 b = vcross(n, t);
    return vnormalize(vadd(vadd(vscale(t, x), vscale(b, y)), vscale(n, z)));
}

static float evalR0(float ior) { return ((ior - 1.0f) * (ior - 1.0f)) / ((ior + 1.0f) * (ior + 1.0f)); } /* :207-209 */
static float luminance(v3 c) { return vdot(c, V3(0.2126f, 0.7152f, 0.0722f)); }                        /* :211-213 */

static v3 evalFm(v3 bc, v3 h, v3 wo, float specular, v3 specularTint, float metallic, float eta) {     /* :215-222 */
    float lum = luminance(bc);
    v3 ctint = lum > 0.0f ? V3(bc.x / lum, bc.y / lum, bc.z / lum) : V3(1.0f, 1.0f, 1.0f);
    v3 ks = vadd(vsub(V3(1.0f, 1.0f, 1.0f), ctint), vmul(specularTint, ctint));
    v3 c0 = vadd(vscale(ks, (specular * evalR0(eta)) * (1.0f - metallic)), vscale(bc, metallic));
    float p = or_pow5(1.0f - fabsf(vdot(h, wo)));
    return vadd(c0, vscale(vsub(V3(1.0f, 1.0f, 1.0f), c0), p));
}

static float evalDm(v3 hl, float ax, float ay) {                              /* :224-229 */
    float constant = (K_PI * ax) * ay;
    float q = (or_pow2(hl.x) / or_pow2(ax) + or_pow2(hl.y) / or_pow2(ay)) + or_pow2(hl.z);
    float other = or_pow2(q);
    return 1.0f / (constant * other);
}

static float lambdaG(v3 wl, float ax, float ay) {                             /* :231-234 */
    float st = sqrtf(1.0f + (or_pow2(wl.x * ax) + or_pow2(wl.y * ay)) / or_pow2(wl.z));
    return (st - 1.0f) / 2.0f;
}
static float smithG(v3 wl, float ax, float ay) { return 1.0f / (1.0f + lambdaG(wl, ax, ay)); } /* :236-239 */
static float evalGm(v3 wi, v3 wo, float ax, float ay) { return smithG(wi, ax, ay) * smithG(wo, ax, ay); } /* :241-245 */

static void alphas(float anisotropic, float roughness, float* ax, float* ay) {
    float aspect = sqrtf(1.0f - 0.9f * anisotropic);
    *ax = or_maxf(0.0001f, (roughness * roughness) / aspect);
    *ay = or_maxf(0.0001f, (roughness * roughness) * aspect);
}

static v3 evalMetal(m3 tbn, v3 bc, float anisotropic, float roughness, v3 n, v3 wi, v3 wo, v3 h,
                    float specular, v3 specularTint, float metallic, float eta) {   /* :247-267 */
    float ax, ay;
    alphas(anisotropic, roughness, &ax, &ay);
    v3 fm = evalFm(bc, h, wo, specular, specularTint, metallic, eta);
    v3 wiT = vnormalize(m3tmul(tbn, wi));
    v3 woT = vnormalize(m3tmul(tbn, wo));
    v3 hT = vnormalize(m3tmul(tbn, h));
    float dm = evalDm(hT, ax, ay);
    float gm = evalGm(wiT, woT, ax, ay);
    float a = fabsf(vdot(n, wi)), b = fabsf(vdot(n, wo));
    float den = (4.0f * a) * b;
    return V3(((fm.x * dm) * gm) / den, ((fm.y * dm) * gm) / den, ((fm.z * dm) * gm) / den);
}

static v3 sampleMetal(m3 tbn, float anisotropic, float roughness, v3 wi, uint32_t* rng) {   /* :269-285 */
    float ax, ay;
    alphas(anisotropic, roughness, &ax, &ay);
    v3 wiT = m3tmul(tbn, wi);
    v3 h = sampleGGXVNDF(wiT, ax, ay, rng);
    h = vnormalize(m3mul(tbn, h));
    return vreflect(vneg(wi), h);
}

static float pdfMetal(m3 tbn, v3 wi, v3 wo, float anisotropic, float roughness, int bounded) { /* :287-299 */
    float ax, ay;
    alphas(anisotropic, roughness, &ax, &ay);
    return pdfGGXReflection(m3tmul(tbn, wi), m3tmul(tbn, wo), ax, ay, bounded);
}

static float G1c(v3 w, float a) {                                             /* :313-319 */
    float a2 = a * a;
    float z = w.z;
    return 2.0f / (1.0f + sqrtf(a2 + ((1.0f - a2) * z) * z));
}

static float evalDc(float ag, v3 hl) {                                        /* :328-333 */
    float num = ag * ag - 1.0f;
    float den = (K_PI * or_logf(ag * ag)) * (1.0f + (ag * ag - 1.0f) * (hl.z * hl.z));
    return num / den;
}

static v3 evalClearcoat(m3 tbn, v3 wi, v3 wo, float cg, v3 h) {               /* :335-354 */
    float ag = (1.0f - cg) * 0.1f + cg * 0.001f;
    v3 hT = m3tmul(tbn, h), wiT = m3tmul(tbn, wi), woT = m3tmul(tbn, wo);
    float r0 = evalR0(1.5f);
    float fc = r0 + (1.0f - r0) * or_pow5(1.0f - vdot(h, wo));
    float gc = G1c(wiT, 0.25f) * G1c(woT, 0.25f);
    float dc = evalDc(ag, hT);
    float v = ((0.25f * fc) * gc) * dc;
    return V3(v, v, v);
}

static v3 sampleClearcoat(m3 tbn, float cg, v3 wi, uint32_t* rng) {          /* :356-366 */
    float ag = (1.0f - cg) * 0.1f + cg * 0.001f;
    v3 h = vnormalize(sampleGTR1_attempt2(ag, rng));
    h = vnormalize(m3mul(tbn, h));
    return vnormalize(vreflect(vneg(wi), h));
}

static float pdfClearcoat(m3 tbn, v3 wi, v3 wo, v3 h, float cg) {            /* :368-383 */
    float ag = (1.0f - cg) * 0.1f + cg * 0.001f;
    v3 wiT = m3tmul(tbn, wi), woT = m3tmul(tbn, wo), hT = m3tmul(tbn, h);
    if (wiT.z <= 0.0f || woT.z <= 0.0f) return 0.0f;
    float dc = evalDc(ag, hT) * hT.z;
    return dc / (4.0f * fabsf(vdot(woT, hT)));
}

static float reflectance(float cosine, float ri) {                            /* :389-394, dielectric.rchit.glsl:7-12 */
    float r0 = (1.0f - ri) / (1.0f + ri);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * or_pow5(1.0f - cosine);
}

static v3 sampleGlass(m3 tbn, v3 wi, float roughness, float anisotropic, float eta, uint32_t* rng,
                      int* refracted) {                                       /* :396-431 */
    float ax, ay;
    alphas(anisotropic, roughness, &ax, &ay);
    v3 wiT = m3tmul(tbn, wi);
    v3 hT = sampleGGXVNDF(wiT, ax, ay, rng);
    float cosTheta = vdot(wiT, hT);
    v3 hW = vnormalize(m3mul(tbn, hT));
    float refl = reflectance(cosTheta, eta);
    float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
    int cannot = eta * sinTheta > 1.0f;
    if (cannot || refl > or_random(rng)) {
        *refracted = 0;
        return vreflect(vneg(wi), hW);
    }
    *refracted = 1;
    return vrefract(vneg(wi), hW, eta);
}

static float SmithGAniso(float NdV, float VdX, float VdY, float ax, float ay) {   /* :433-439 */
    float a = VdX * ax, b = VdY * ay, c = NdV;
    return (2.0f * NdV) / (NdV + sqrtf((a * a + b * b) + c * c));
}

static v3 evalMicrofacetRefraction(v3 bc, float anisotropic, float roughness, float eta, v3 V, v3 L,
                                   v3 H, float* pdf) {                        /* :441-467 */
    float ax, ay;
    alphas(anisotropic, roughness, &ax, &ay);
    *pdf = 0.0f;
    if (L.z >= 0.0f) return V3(0.0f, 0.0f, 0.0f);
    float LdH = vdot(L, H), VdH = vdot(V, H);
    float D = evalDm(H, ax, ay);
    float G1 = SmithGAniso(fabsf(V.z), V.x, V.y, ax, ay);
    float G2 = G1 * SmithGAniso(fabsf(L.z), L.x, L.y, ax, ay);
    float denom = LdH + VdH * eta;
    denom = denom * denom;
    float eta2 = eta * eta;
    float jac = (fabsf(LdH) * eta2) / denom;
#if OR_PLANT == 4
    jac = fabsf(LdH) / denom;
#endif
    *pdf = (((G1 * or_maxf(0.0f, VdH)) * D) * jac) / V.z;
    float F = reflectance(vdot(V, H), eta);
    v3 sq = V3(or_powf(bc.x, 0.5f), or_powf(bc.y, 0.5f), or_powf(bc.z, 0.5f));
    v3 r = vscale(sq, 1.0f - F);
    r = vscale(r, D);
    r = vscale(r, G2);
    r = vscale(r, fabsf(VdH));
    r = vscale(r, jac);
    float den2 = fabsf(L.z * V.z);
    return V3(r.x / den2, r.y / den2, r.z / den2);
}

/* branch_pdf: times the probability of sampleGlass's branch, F for reflection
 * (1 under total internal reflection), 1 - F for refraction
 * (REINA_RT_TRACE_FIX_DISNEY_MIX); the reference leaves it out */
static v3 glassEval(m3 tbn, v3 bc, float anisotropic, float roughness, float eta, v3 n, v3 wi, v3 wo,
                    int didRefract, int bounded, int branch_pdf, float* pdf) {              /* :539-581 */
#if OR_PLANT == 3
    branch_pdf = 1;
#endif
    v3 h = didRefract ? vnormalize(vadd(wo, vscale(wi, eta))) : vnormalize(vadd(wo, wi));
    if (vdot(h, n) < 0.0f) h = vneg(h);
    v3 wiT = m3tmul(tbn, wi), hT = m3tmul(tbn, h);
    if (didRefract) {
        float tp;
        v3 f = evalMicrofacetRefraction(bc, anisotropic, roughness, eta, wiT, m3tmul(tbn, wo), hT, &tp);
        *pdf = branch_pdf ? tp * (1.0f - reflectance(vdot(wiT, hT), eta)) : tp;
        return f;
    }
    *pdf = pdfMetal(tbn, wi, wo, anisotropic, roughness, bounded);
    if (branch_pdf) {
        float c = vdot(wiT, hT);
        float sn = sqrtf(1.0f - c * c);
        *pdf = *pdf * ((eta * sn > 1.0f) ? 1.0f : reflectance(c, eta));
    }
    return evalMetal(tbn, bc, anisotropic, roughness, n, wi, wo, h, 0.0f, V3(1.0f, 1.0f, 1.0f), 1.0f, eta);
}

static v3 evalSheen(v3 bc, v3 wo, v3 h, v3 n, v3 sheenTint) {                 /* :587-594 */
    float lum = luminance(bc);
    v3 ctint = lum > 0.0f ? V3(bc.x / lum, bc.y / lum, bc.z / lum) : V3(1.0f, 1.0f, 1.0f);
    v3 cs = V3(mixf(1.0f, ctint.x, sheenTint.x), mixf(1.0f, ctint.y, sheenTint.y), mixf(1.0f, ctint.z, sheenTint.z));
    float p = or_pow5(1.0f - or_maxf(vdot(h, wo), 0.0f));
    return vscale(vscale(cs, p), or_maxf(vdot(n, wo), 0.0f));
}

static v3 sampleDisney(m3 tbn, v3 bc, float anisotropic, float roughness, float cg, float eta,
                       float metallic, float clearcoat, float st, v3 n, v3 wi, int* didRefract,
                       int* choseGlass, uint32_t* rng) {                       /* :609-654 */
    float dW = (1.0f - st) * (1.0f - metallic);
    float mW = metallic;
    float cW = 0.25f * clearcoat;
    float gW = (1.0f - metallic) * st;
    float cdf0 = dW, cdf1 = cdf0 + mW, cdf2 = cdf1 + cW, cdf3 = cdf2 + gW;
    (void)bc;
    *didRefract = 0;
    float rnd = or_random(rng) * cdf3;
    *choseGlass = 0;
    if (rnd < cdf0) return sampleDiffuse(n, rng);
    if (rnd < cdf1) return sampleMetal(tbn, anisotropic, roughness, wi, rng);
    if (rnd < cdf2) return sampleClearcoat(tbn, cg, wi, rng);
    *choseGlass = 1;
    return sampleGlass(tbn, wi, roughness, anisotropic, eta, rng, didRefract);
}

static v3 evalDisney(m3 tbn, v3 bc, v3 specularTint, v3 sheenTint, float anisotropic, float roughness,
                     float subsurface, float cg, float eta, float metallic, float clearcoat, float st,
                     float sheen, int didRefract, uint32_t fixes, v3 n, v3 wi, v3 wo, v3 h,
                     float* pdf) {   /* :656-711 */
    /* REINA_RT_TRACE_FIX_METAL_PDF: the sampler's density for GGX reflection;
     * REINA_RT_TRACE_FIX_DISNEY_MIX: clearcoat inside the normalised mixture,
     * specular = 1 for the metal lobe's dielectric F0, glass branch pdf */
    int bounded = (fixes & REINA_RT_TRACE_FIX_METAL_PDF) == 0;
    int mix = (fixes & REINA_RT_TRACE_FIX_DISNEY_MIX) != 0;
    float dW = (1.0f - st) * (1.0f - metallic);
    float mW = metallic;
    float cW = 0.25f * clearcoat;
    float gW = (1.0f - metallic) * st;
    float ws = (dW + mW) + gW;
    if (mix || OR_PLANT == 1) ws = ws + cW;
    float cN = (mix || OR_PLANT == 1) ? cW / ws : cW;
    v3 fd = evalDiffuse(roughness, subsurface, bc, n, wi, wo, h);
    float dP = pdfDiffuse(wo, n);
    v3 fs = evalSheen(bc, wo, h, n, sheenTint);
    v3 fm = evalMetal(tbn, bc, anisotropic, roughness, n, wi, wo, h, (mix || OR_PLANT == 2) ? 1.0f : st, specularTint,
                      metallic, eta);
    float mP = pdfMetal(tbn, wi, wo, anisotropic, roughness, bounded);
    v3 fc = evalClearcoat(tbn, wi, wo, cg, h);
    float cP = pdfClearcoat(tbn, wi, wo, h, cg);
    float gP;
    v3 fg = glassEval(tbn, bc, anisotropic, roughness, eta, n, wi, wo, didRefract, bounded, mix, &gP);
    *pdf = (((dP * dW) / ws + (mP * mW) / ws) + cP * cN) + (gP * gW) / ws;
    v3 r = vscale(vadd(fd, vscale(fs, sheen)), dW / ws);
    r = vadd(r, vscale(fm, mW / ws));
    r = vadd(r, vscale(fc, cN));
    r = vadd(r, vscale(fg, gW / ws));
    return r;
}

/* ---- next-event estimation at Disney hits (the library's, shade.h) --------
 * The reference's directLight evaluates evalDisney for materialID 3
 * (raytrace.rgen.glsl:57-83) but never runs. A light sample at a Disney hit
 * carries g(w) = q_ng f cos / pdf + q_g: the expected weight of the hit
 * shader's own continuation (f cos / pdf from a diffuse, metal or clearcoat
 * lobe, 1 from the glass lobe, Q4), with q_ng, q_g the densities of
 * sampleDisney's non-glass and glass lobes at w; MIS uses q_ng + q_g. */

/* sampleDisney's density at wo (brdfDisney.h.glsl:609-654), from the
 * samplers themselves: lobe probabilities times the cosine density (diffuse),
 * Heitz's VNDF reflected, G1(i) D(m) / (4 i.z) with m.z > 0 (metal), GTR1
 * reflected, Dc(h) h.z / (4 |o.h|) with h oriented up (clearcoat), and for
 * glass the VNDF reflected times the reflection probability plus the VNDF
 * refracted times 1 - F with Walter et al.'s half-vector Jacobian
 * |o.h| / (eta i.h + o.h)^2. png: the non-glass lobes, pg: glass. */
static float vndf_reflection_density(v3 wiT, v3 woT, float ax, float ay) {
    if (!(wiT.z > 0.0f)) return 0.0f;
    const v3 m = vnormalize(vadd(wiT, woT));
    if (!(m.z > 0.0f)) return 0.0f;
    return (evalDm(m, ax, ay) * smithG(wiT, ax, ay)) / (4.0f * wiT.z);
}

static void disney_density(m3 tbn, float anisotropic, float roughness, float cg, float eta, float metallic,
                           float clearcoat, float st, v3 n, v3 wi, v3 wo, float* png, float* pg) {
    const float dW = (1.0f - st) * (1.0f - metallic);
    const float mW = metallic;
    const float cW = 0.25f * clearcoat;
    const float gW = (1.0f - metallic) * st;
    const float ws = ((dW + mW) + gW) + cW;
    float ax, ay;
    alphas(anisotropic, roughness, &ax, &ay);
    const v3 wiT = m3tmul(tbn, wi), woT = m3tmul(tbn, wo);
    const float dP = pdfDiffuse(wo, n);
    const float mr = vndf_reflection_density(wiT, woT, ax, ay);
    const float ag = (1.0f - cg) * 0.1f + cg * 0.001f;
    v3 hc = vnormalize(vadd(wiT, woT));
    if (hc.z < 0.0f) hc = vneg(hc);
    const float ohc = fabsf(vdot(woT, hc));
    const float cP = ohc > 0.0f ? (evalDc(ag, hc) * hc.z) / (4.0f * ohc) : 0.0f;
    float gr = 0.0f, gt = 0.0f;
    if (mr > 0.0f) {
        const v3 m = vnormalize(vadd(wiT, woT));
        const float c = vdot(wiT, m);
        const float sn = sqrtf(1.0f - c * c);
        gr = mr * ((eta * sn > 1.0f) ? 1.0f : reflectance(c, eta));
    }
    if (wiT.z > 0.0f) {   /* refraction can also leave above the horizon through a tilted microfacet */
        v3 h = vnormalize(vadd(woT, vscale(wiT, eta)));
        if (h.z < 0.0f) h = vneg(h);
        const float ih = vdot(wiT, h), oh = vdot(woT, h);
        const float sn = sqrtf(1.0f - ih * ih);
        if (h.z > 0.0f && ih > 0.0f && oh < 0.0f && !(eta * sn > 1.0f)) {
            const float den = eta * ih + oh;
            const float dv = ((smithG(wiT, ax, ay) * ih) * evalDm(h, ax, ay)) / wiT.z;
            gt = ((dv * fabsf(oh)) / (den * den)) * (1.0f - reflectance(ih, eta));
        }
    }
    *png = ((dP * dW) / ws + (mr * mW) / ws) + (cP * cW) / ws;
    *pg = ((gr + gt) * gW) / ws;
}

typedef struct {
    m3 tbn;
    v3 albedo, wn, wi, pos, ng;
    float eta;
    float p_bsdf;   /* density of the direction the hit shader sampled */
} OrDisneyNee;

static v3 disney_nee_weight(const OrDisneyNee* c, const ReinaInstanceProperties* p, uint32_t fixes, v3 w,
                            float* ptot) {
    const v3 h = vnormalize(vadd(w, c->wi));
    float pdf_ev;
    const v3 f = evalDisney(c->tbn, c->albedo, pv3(p->specularTint), pv3(p->sheenTint), p->anisotropic,
                            p->roughness, p->subsurface, p->clearcoatGloss, c->eta, p->metallic, p->clearcoat,
                            p->specularTransmission, p->sheen, 0, fixes, c->wn, c->wi, w, h, &pdf_ev);
    float png, pg;
    disney_density(c->tbn, p->anisotropic, p->roughness, p->clearcoatGloss, c->eta, p->metallic, p->clearcoat,
                   p->specularTransmission, c->wn, c->wi, w, &png, &pg);
    *ptot = png + pg;
    const float cx = or_maxf(vdot(c->wn, w), 0.0f);
#if OR_PLANT == 10
    return vscale(f, cx);   /* the textbook f cos, ignoring how the hit shader weights its continuations */
#endif
#if OR_PLANT == 11
    pg = 0.0f;              /* the glass lobe's weight-1 continuations left out */
#endif
    v3 g = V3(pg, pg, pg);
    if (pdf_ev > 0.0f && cx > 0.0f) g = vadd(vscale(f, (png / pdf_ev) * cx), g);
    return g;
}

/* Batch probes of the Disney pieces above for the tests, in the tangent frame
 * (tbn = identity, n = +z). prm holds the piece's scalar parameters; rng is one
 * stream carried across the n samples of a sampling probe.
 *   0: D_GGX_Aniso(a), evalDm(a)            -> out[2k], out[2k+1]  (prm ax, ay)
 *   1: smithG(a), SmithGAniso(|a.z|, a.x, a.y) -> out[2k], out[2k+1]  (ax, ay)
 *   2: sampleGGXVNDF(a)                     -> out[3k..]            (ax, ay)
 *   3: pdfGGXReflection(a, b)               -> out[k]   (ax, ay, 1 for k = 1)
 *   4: sampleGTR1_attempt2                  -> out[3k..]            (alpha)
 *   5: evalDc(alpha, a)                     -> out[k]               (alpha)
 *   6: evalDiffuse(wi a, wo b, h = |a + b|) -> out[3k..]  (roughness, subsurface, bc[3])
 *   7: evalSheen(wo b, h = |a + b|)         -> out[3k..]  (-, -, bc[3], sheenTint[3])
 *   8: pdfClearcoat(wi a, wo b, h = |a + b|) -> out[k]              (clearcoatGloss)
 *   9: sampleClearcoat(wi a)                -> out[3k..]            (clearcoatGloss)
 *  10: sampleDisney(wi a)                   -> out[4k..] (wo, didRefract)
 *  11: evalDisney pdf(wi a, wo b, h = |a + b|, didRefract = b.z < 0) -> out[k]
 *  12: evalDisney f and pdf (as 11)         -> out[4k..] (f.xyz, pdf)
 *  13: the Disney hit shader's continuation from wi a: sampleDisney, then
 *      its weight f cos / pdf, or 1 when the glass lobe was chosen (the
 *      raygen skips that multiply, Q4)     -> out[7k..] (wo, weight, choseGlass)
 *  14: next-event estimation at wi a toward wo b: disney_nee_weight's g and
 *      the sampler's density there          -> out[4k..] (g.xyz, density)
 *      10-14 read the material from prm: bc[3] 0-2, specularTint[3] 3-5,
 *      sheenTint[3] 6-8, anisotropic 9, roughness 10, subsurface 11,
 *      clearcoatGloss 12, eta 13, metallic 14, clearcoat 15,
 *      specularTransmission 16, sheen 17, fixes 18 (as a float) */
void or_disney_probe(int fn, uint64_t n, const float* a, const float* b, const float* prm, uint32_t* rng,
                     float* out) {
    const m3 I = {V3(1.0f, 0.0f, 0.0f), V3(0.0f, 1.0f, 0.0f), V3(0.0f, 0.0f, 1.0f)};
    const v3 nz = V3(0.0f, 0.0f, 1.0f);
    for (uint64_t k = 0; k < n; ++k) {
        v3 x = a ? V3(a[3 * k], a[3 * k + 1], a[3 * k + 2]) : nz;
        v3 y = b ? V3(b[3 * k], b[3 * k + 1], b[3 * k + 2]) : nz;
        v3 r = V3(0.0f, 0.0f, 0.0f);
        switch (fn) {
            case 0: out[2 * k] = D_GGX_Aniso(x, prm[0], prm[1]); out[2 * k + 1] = evalDm(x, prm[0], prm[1]); continue;
            case 1:
                out[2 * k] = smithG(x, prm[0], prm[1]);
                out[2 * k + 1] = SmithGAniso(fabsf(x.z), x.x, x.y, prm[0], prm[1]);
                continue;
            case 2: r = sampleGGXVNDF(x, prm[0], prm[1], rng); break;
            case 3: out[k] = pdfGGXReflection(x, y, prm[0], prm[1], prm[2] == 0.0f); continue;
            case 4: r = sampleGTR1_attempt2(prm[0], rng); break;
            case 5: out[k] = evalDc(prm[0], x); continue;
            case 6: r = evalDiffuse(prm[0], prm[1], V3(prm[2], prm[3], prm[4]), nz, x, y, vnormalize(vadd(x, y))); break;
            case 7:
                r = evalSheen(V3(prm[2], prm[3], prm[4]), y, vnormalize(vadd(x, y)), nz, V3(prm[5], prm[6], prm[7]));
                break;
            case 8: out[k] = pdfClearcoat(I, x, y, vnormalize(vadd(x, y)), prm[0]); continue;
            case 9: r = sampleClearcoat(I, prm[0], x, rng); break;
            case 10: {
                int dr, cg;
                r = sampleDisney(I, V3(prm[0], prm[1], prm[2]), prm[9], prm[10], prm[12], prm[13], prm[14], prm[15],
                                 prm[16], nz, x, &dr, &cg, rng);
                out[4 * k] = r.x; out[4 * k + 1] = r.y; out[4 * k + 2] = r.z; out[4 * k + 3] = (float)dr;
                continue;
            }
            case 11: {
                float pdf;
                (void)evalDisney(I, V3(prm[0], prm[1], prm[2]), V3(prm[3], prm[4], prm[5]), V3(prm[6], prm[7], prm[8]),
                                 prm[9], prm[10], prm[11], prm[12], prm[13], prm[14], prm[15], prm[16], prm[17],
                                 y.z < 0.0f, (uint32_t)prm[18], nz, x, y, vnormalize(vadd(x, y)), &pdf);
                out[k] = pdf;
                continue;
            }
            case 13: {
                int dr, cg;
                v3 wo = sampleDisney(I, V3(prm[0], prm[1], prm[2]), prm[9], prm[10], prm[12], prm[13], prm[14],
                                     prm[15], prm[16], nz, x, &dr, &cg, rng);
                float pdf;
                v3 f = evalDisney(I, V3(prm[0], prm[1], prm[2]), V3(prm[3], prm[4], prm[5]), V3(prm[6], prm[7], prm[8]),
                                  prm[9], prm[10], prm[11], prm[12], prm[13], prm[14], prm[15], prm[16], prm[17],
                                  dr, (uint32_t)prm[18], nz, x, wo, vnormalize(vadd(wo, x)), &pdf);
                v3 wgt = vscale(f, or_maxf(vdot(nz, wo), 0.0f));
                wgt = cg ? V3(1.0f, 1.0f, 1.0f) : V3(wgt.x / pdf, wgt.y / pdf, wgt.z / pdf);
                out[7 * k] = wo.x; out[7 * k + 1] = wo.y; out[7 * k + 2] = wo.z;
                out[7 * k + 3] = wgt.x; out[7 * k + 4] = wgt.y; out[7 * k + 5] = wgt.z; out[7 * k + 6] = (float)cg;
                continue;
            }
            case 14: {
                ReinaInstanceProperties pp;
                memset(&pp, 0, sizeof pp);
                for (int q = 0; q < 3; ++q) { pp.specularTint[q] = prm[3 + q]; pp.sheenTint[q] = prm[6 + q]; }
                pp.anisotropic = prm[9]; pp.roughness = prm[10]; pp.subsurface = prm[11]; pp.clearcoatGloss = prm[12];
                pp.metallic = prm[14]; pp.clearcoat = prm[15]; pp.specularTransmission = prm[16]; pp.sheen = prm[17];
                OrDisneyNee dn;
                dn.tbn = I; dn.albedo = V3(prm[0], prm[1], prm[2]); dn.wn = nz; dn.wi = x;
                dn.pos = V3(0.0f, 0.0f, 0.0f); dn.ng = nz; dn.eta = prm[13]; dn.p_bsdf = 0.0f;
                float pt;
                v3 g = disney_nee_weight(&dn, &pp, (uint32_t)prm[18], y, &pt);
                out[4 * k] = g.x; out[4 * k + 1] = g.y; out[4 * k + 2] = g.z; out[4 * k + 3] = pt;
                continue;
            }
            case 12: {
                float pdf;
                v3 f = evalDisney(I, V3(prm[0], prm[1], prm[2]), V3(prm[3], prm[4], prm[5]), V3(prm[6], prm[7], prm[8]),
                                  prm[9], prm[10], prm[11], prm[12], prm[13], prm[14], prm[15], prm[16], prm[17],
                                  y.z < 0.0f, (uint32_t)prm[18], nz, x, y, vnormalize(vadd(x, y)), &pdf);
                out[4 * k] = f.x; out[4 * k + 1] = f.y; out[4 * k + 2] = f.z; out[4 * k + 3] = pdf;
                continue;
            }
            default: return;
        }
        out[3 * k] = r.x; out[3 * k + 1] = r.y; out[3 * k + 2] = r.z;
    }
}

/* ============================ textures ==================================== */
/* texture(textures[id], uv) of the hit shaders: RGBA8 UNORM, bilinear, repeat,
 * base level (a ray-tracing stage has no derivatives; Image.cpp:56,
 * vktools.cpp:765-788): two linear interpolations along x, one along y. */
typedef struct { float x, y, z, w; } f4;

static f4 unorm8(const uint8_t* p) {
    f4 r = {(float)p[0] / 255.0f, (float)p[1] / 255.0f, (float)p[2] / 255.0f, (float)p[3] / 255.0f};
    return r;
}

/* two linear interpolations along x, then one along y */
static float lerp2(float a, float b, float p00, float p10, float p01, float p11) {
    float top = p00 + a * (p10 - p00), bottom = p01 + a * (p11 - p01);
    return top + b * (bottom - top);
}

static uint32_t repeat_index(float fx, uint32_t n) {
    int32_t m = (int32_t)fx % (int32_t)n;
    return (uint32_t)(m < 0 ? m + (int32_t)n : m);
}

static f4 tex_sample(const OrScene* s, int32_t id, float u, float v) {
    const struct OrTexture* t = &s->tex[id];
    float x = u * (float)t->w - 0.5f, y = v * (float)t->h - 0.5f;
    x = (x == x) ? or_clampf(x, -1073741824.0f, 1073741824.0f) : 0.0f;
    y = (y == y) ? or_clampf(y, -1073741824.0f, 1073741824.0f) : 0.0f;
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    uint32_t i0 = repeat_index(fx, t->w), i1 = repeat_index(fx + 1.0f, t->w);
    uint32_t j0 = repeat_index(fy, t->h), j1 = repeat_index(fy + 1.0f, t->h);
    f4 t00 = unorm8(t->rgba8 + 4ull * ((uint64_t)j0 * t->w + i0)), t10 = unorm8(t->rgba8 + 4ull * ((uint64_t)j0 * t->w + i1));
    f4 t01 = unorm8(t->rgba8 + 4ull * ((uint64_t)j1 * t->w + i0)), t11 = unorm8(t->rgba8 + 4ull * ((uint64_t)j1 * t->w + i1));
    f4 r = {lerp2(a, b, t00.x, t10.x, t01.x, t11.x), lerp2(a, b, t00.y, t10.y, t01.y, t11.y),
            lerp2(a, b, t00.z, t10.z, t01.z, t11.z), lerp2(a, b, t00.w, t10.w, t01.w, t11.w)};
    return r;
}

/* texutils.h.glsl:4-41 */
static void bumpMapping(const OrScene* s, int32_t id, float* u, float* v, v3 rayIn, m3 T) {
    const float heightScale = 0.2f, minLayers = 64.0f, maxLayers = 512.0f;
    v3 V = vnormalize(m3tmul(T, vneg(rayIn)));
    if (V.z <= 0.0f) return;
    float numLayers = mixf(maxLayers, minLayers, or_clampf(V.z, 0.0f, 1.0f));
    float layerDepth = 1.0f / numLayers;
    float Px = V.x / V.z * heightScale, Py = V.y / V.z * heightScale;
    float du = Px / numLayers, dv = Py / numLayers;
    float cu = *u, cv = *v, depthSum = 0.0f;
    float h = tex_sample(s, id, cu, cv).x * heightScale;
    for (int k = 0; k < 1024 && depthSum < h; ++k) {   /* ends within 104 steps for finite input */
        cu = cu - du;
        cv = cv - dv;
        depthSum = depthSum + layerDepth;
        h = tex_sample(s, id, cu, cv).x * heightScale;
    }
    float pu = cu + du, pv = cv + dv;
    float hPrev = tex_sample(s, id, pu, pv).x * heightScale;
    float sumPrev = depthSum - layerDepth;
    float after = h - depthSum;
    float before = hPrev - sumPrev;
    float weight = after / (after - before);
    *u = mixf(pu, cu, weight);
    *v = mixf(pv, cv, weight);
}

/* lambertian.rchit.glsl:37-41: texture(normalMap).rgb * 2 - 1, y flipped, tbn * n */
static v3 normalMap(const OrScene* s, int32_t id, float u, float v, m3 tbn) {
    f4 t = tex_sample(s, id, u, v);
    v3 tn = V3(t.x * 2.0f - 1.0f, t.y * 2.0f - 1.0f, t.z * 2.0f - 1.0f);
    tn.y = -tn.y;
    return vnormalize(m3mul(tbn, tn));
}

/* entry points for the tests: the sampler and the parallax search alone */
void or_tex_sample_at(const OrScene* s, int32_t id, float u, float v, float* out4) {
    f4 r = tex_sample(s, id, u, v);
    out4[0] = r.x; out4[1] = r.y; out4[2] = r.z; out4[3] = r.w;
}

void or_bump_at(const OrScene* s, int32_t id, float u, float v, const float* dir3, const float* tbn9,
                float* out2) {
    m3 T;
    T.c0 = V3(tbn9[0], tbn9[1], tbn9[2]);
    T.c1 = V3(tbn9[3], tbn9[4], tbn9[5]);
    T.c2 = V3(tbn9[6], tbn9[7], tbn9[8]);
    bumpMapping(s, id, &u, &v, V3(dir3[0], dir3[1], dir3[2]), T);
    out2[0] = u;
    out2[1] = v;
}

/* ============================ hit shaders ================================= */
/* lambertian.rchit.glsl:11-78 and metal.rchit.glsl:7-70 */
static void chs_lambert_metal(const OrScene* s, Payload* pld, const HitCtx* c, int metal) {
    HitInfo hi = getObjectHitInfo(s, c);
    const ReinaInstanceProperties* p = c->props;
    if (p->cullBackface != 0u && !hi.frontFace) { skip_hit(pld, &hi, c); return; }
    float ux = fmodf1(hi.uvx), uy = fmodf1(hi.uvy);
    if (p->bumpMapTexID >= 0) bumpMapping(s, p->bumpMapTexID, &ux, &uy, vnormalize(c->worldRayDirection), hi.tbn);
    if (!metal && (ux < 0.0f || ux > 1.0f || uy < 0.0f || uy > 1.0f)) { skip_hit(pld, &hi, c); return; }
    v3 wn = hi.worldNormal;
    if (p->normalMapTexID >= 0) wn = normalMap(s, p->normalMapTexID, ux, uy, hi.tbn);
    pld->color = pv3(p->albedo);
    if (p->textureID >= 0) {
        f4 tc = tex_sample(s, p->textureID, ux, uy);
        if (tc.w < 0.999f && or_random(&pld->rngState) > tc.w) { skip_hit(pld, &hi, c); return; }
        pld->color = vmul(pld->color, V3(tc.x, tc.y, tc.z));
    }
    pld->albedo = pld->color;
    pld->emission = pv3(p->emission);
    pld->rayOrigin = offsetPositionAlongNormal(hi.worldPosition, hi.worldNormalGeometry);
    if (!metal) pld->rayDirection = diffuseReflection(wn, &pld->rngState);
    else pld->rayDirection = fuzzyReflection(c->worldRayDirection, wn, p->roughness, &pld->rngState);
    pld->rayHitSky = 0;
    pld->skip = 0;
    pld->insideDielectric = 0;
    pld->materialID = metal ? 1u : 0u;
    pld->surfaceNormal = wn;
    pld->pdf = metal ? 0.0f : or_maxf(vdot(pld->surfaceNormal, pld->rayDirection), 0.0f) / K_PI;
    pld->tbn = hi.tbn;
    pld->didRefract = 0;
    pld->eta = 0.0f;
    pld->accumulatedDistance = 0.0f; /* insideDielectric was just set to false */
}

/* dielectric.rchit.glsl:40-113 */
static void chs_dielectric(const OrScene* s, Payload* pld, const HitCtx* c) {
    HitInfo hi = getObjectHitInfo(s, c);
    const ReinaInstanceProperties* p = c->props;
    float ri = hi.frontFace ? 1.0f / p->ior : p->ior;
    v3 unitDir = vnormalize(c->worldRayDirection);
    float cosTheta = or_minf(vdot(vneg(unitDir), hi.worldNormal), 1.0f);
    float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
    int cannot = ri * sinTheta > 1.0f;
    float refl = reflectance(cosTheta, ri);
    int prevInside = pld->insideDielectric;
    float ux = fmodf1(hi.uvx), uy = fmodf1(hi.uvy);
    if (p->bumpMapTexID >= 0) bumpMapping(s, p->bumpMapTexID, &ux, &uy, vnormalize(c->worldRayDirection), hi.tbn);
    v3 albedo = pv3(p->albedo);
    if (p->textureID >= 0) {
        f4 tc = tex_sample(s, p->textureID, ux, uy);
        albedo = vmul(albedo, V3(tc.x, tc.y, tc.z));
    }
    v3 wn = hi.worldNormal;
    if (p->normalMapTexID >= 0) wn = normalMap(s, p->normalMapTexID, ux, uy, hi.tbn);
    if (cannot || refl > or_random(&pld->rngState)) {
        pld->rayDirection = fuzzyReflection(unitDir, wn, p->roughness, &pld->rngState);
        pld->color = V3(1.0f, 1.0f, 1.0f);
        pld->rayOrigin = offsetPositionAlongNormal(hi.worldPosition, hi.worldNormalGeometry);
    } else {
        pld->insideDielectric = hi.frontFace;
        pld->rayDirection = addFuzz(vrefract(unitDir, wn, ri), p->roughness, &pld->rngState);
        pld->color = albedo;
        pld->rayOrigin = offsetPositionForDielectric(hi.worldPosition, hi.worldNormalGeometry, unitDir);
    }
    pld->accumulatedDistance = pld->accumulatedDistance +
        (prevInside ? vlength(vsub(hi.worldPosition, c->worldRayOrigin)) : 0.0f);
    if (prevInside && !pld->insideDielectric) {
        float e = or_expf(-p->absorption * pld->accumulatedDistance);
        pld->color = vmul(V3(e, e, e), albedo);
        pld->accumulatedDistance = 0.0f;
    }
    pld->albedo = pld->color;
    pld->emission = pv3(p->emission);
    pld->rayHitSky = 0;
    pld->skip = 0;
    pld->materialID = 2;
    pld->surfaceNormal = wn;
    pld->pdf = 0.0f;
    pld->tbn = hi.tbn;
    pld->didRefract = 0;
    pld->eta = 0.0f;
}

/* disney.rchit.glsl:36-198; dn (optional): what next-event estimation needs */
static void chs_disney(const OrScene* s, Payload* pld, const HitCtx* c, OrDisneyNee* dn) {
    HitInfo hi = getObjectHitInfo(s, c);
    const ReinaInstanceProperties* p = c->props;
    if (p->cullBackface != 0u && !hi.frontFace) { skip_hit(pld, &hi, c); return; }
    float ux = fmodf1(hi.uvx), uy = fmodf1(hi.uvy);
    if (p->bumpMapTexID >= 0) bumpMapping(s, p->bumpMapTexID, &ux, &uy, vnormalize(c->worldRayDirection), hi.tbn);
    if (ux < 0.0f || ux > 1.0f || uy < 0.0f || uy > 1.0f) { skip_hit(pld, &hi, c); return; }
    v3 wn = hi.worldNormal;
    if (p->normalMapTexID >= 0) wn = normalMap(s, p->normalMapTexID, ux, uy, hi.tbn);
    v3 albedo = pv3(p->albedo);
    if (p->textureID >= 0) {
        f4 tc = tex_sample(s, p->textureID, ux, uy);
        if (tc.w < 0.999f && or_random(&pld->rngState) > tc.w) { skip_hit(pld, &hi, c); return; }
        albedo = vmul(albedo, V3(tc.x, tc.y, tc.z));
    }
    float eta = hi.frontFace ? 1.0f / p->ior : p->ior;
    int didRefract = 0, choseGlass = 0;
    v3 wi = vneg(c->worldRayDirection);
    v3 rayDir = sampleDisney(hi.tbn, albedo, p->anisotropic, p->roughness, p->clearcoatGloss, eta,
                             p->metallic, p->clearcoat, p->specularTransmission, wn, wi, &didRefract,
                             &choseGlass, &pld->rngState);
    v3 h = vnormalize(vadd(rayDir, wi));
    float pdf;
    v3 f = evalDisney(hi.tbn, albedo, pv3(p->specularTint), pv3(p->sheenTint), p->anisotropic,
                      p->roughness, p->subsurface, p->clearcoatGloss, eta, p->metallic, p->clearcoat,
                      p->specularTransmission, p->sheen, didRefract, s->fixes, wn, wi, rayDir, h, &pdf);
    float cosThetaI = or_maxf(vdot(wn, rayDir), 0.0f);
    v3 fc = vscale(f, cosThetaI);
    pld->color = V3(fc.x / pdf, fc.y / pdf, fc.z / pdf);
    pld->albedo = albedo;
    pld->pdf = pdf;
    pld->emission = pv3(p->emission);
    pld->rayOrigin = offsetPositionForDielectric(hi.worldPosition, hi.worldNormalGeometry, rayDir);
    pld->rayDirection = rayDir;
    pld->rayHitSky = 0;
    pld->skip = 0;
    pld->materialID = 3;
    pld->surfaceNormal = wn;
    pld->tbn = hi.tbn;
    pld->didRefract = 0;
    pld->eta = 0.0f;
    pld->insideDielectric = choseGlass;
    if (pld->insideDielectric)
        pld->accumulatedDistance = pld->accumulatedDistance + vlength(vsub(hi.worldPosition, c->worldRayOrigin));
    else
        pld->accumulatedDistance = 0.0f;
    if (dn) {
        dn->tbn = hi.tbn;
        dn->albedo = albedo;
        dn->wn = wn;
        dn->wi = wi;
        dn->pos = hi.worldPosition;
        dn->ng = hi.worldNormalGeometry;
        dn->eta = eta;
        float png, pg;
        disney_density(hi.tbn, p->anisotropic, p->roughness, p->clearcoatGloss, eta, p->metallic, p->clearcoat,
                       p->specularTransmission, wn, wi, rayDir, &png, &pg);
        dn->p_bsdf = png + pg;
    }
}

/* raytrace.rmiss.glsl:10-19 */
static v3 sky(v3 dir) {
    float y = vnormalize(dir).y;
    float t = 0.5f * (y + 1.0f);
    return vscale(vmix(V3(0.1f, 0.1f, 0.1f), V3(0.4f, 1.7f, 2.0f), t), 0.07f);
}
void or_sky(const float* dir, float* out) {
    v3 c = sky(pv3(dir));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

/* ================== next-event estimation (the library's) ================ */
/* The reference never runs its NEE (skipNEE = true, raytrace.rgen.glsl:146).
 * This restates the library's shade_ray<NEE> (shade_step.h): a light sample
 * at Lambertian and Disney hits whose continuation can still be traced,
 * power-heuristic MIS against the density with which the hit's own sampler
 * draws the same direction (cosine; disney_density), and the matching weight
 * on emitters reached by the continuation. A Disney hit samples a light also
 * when its continuation took the glass lobe. */
static float power_heuristic(float a, float b) { return (a * a) / (a * a + b * b); }

static float light_pdf(float pdf_area, float d, v3 ng, v3 w, int cull) {
    float cl = vdot(ng, vneg(w));
    cl = cull ? or_maxf(cl, 0.0f) : fabsf(cl);
    return cl > 0.0f ? (pdf_area * (d * d)) / cl : 0.0f;
}

static void world_triangle(const OrScene* s, const OrInstance* in, const ReinaInstanceProperties* p, uint32_t prim,
                           v3* w) {
    for (int k = 0; k < 3; ++k) {
        const float* v = &s->vertices[4ull * s->indices[3u * prim + p->indicesOffset + k]];
        w[k] = vadd(vadd(vadd(vscale(in->c0, v[0]), vscale(in->c1, v[1])), vscale(in->c2, v[2])), in->c3);
    }
}

static uint32_t cdf_pick_emi(const OrScene* s, float u) {
    uint32_t lo = 0, hi = s->nemi - 1u;
    while (lo < hi) { uint32_t mid = (lo + hi) / 2u; if (u <= s->emi[mid].cdf) hi = mid; else lo = mid + 1u; }
    return lo;
}

static uint32_t cdf_pick_tri(const OrScene* s, uint32_t lo, uint32_t hi, float u) {
    while (lo < hi) { uint32_t mid = (lo + hi) / 2u; if (u <= s->tri_cdf[mid]) hi = mid; else lo = mid + 1u; }
    return lo;
}

/* ============================= raygen ===================================== */
static uint64_t trace_segments(const OrScene* s, Payload* pld, v3 o, v3 d, uint32_t maxBounces,
                               int use_bvh, v3* Lout) {                      /* raytrace.rgen.glsl:97-184 */
    uint64_t nrays = 0;
    pld->insideDielectric = 0;
    v3 acc = V3(1.0f, 1.0f, 1.0f);
    v3 L = V3(0.0f, 0.0f, 0.0f);
    const int nee = s->nee;
    v3 mis_x = V3(0.0f, 0.0f, 0.0f);   /* origin of the last Lambertian continuation */
    float mis_pb = 0.0f;               /* its cosine pdf, 0: no MIS                   */
    for (uint32_t seg = 0; seg < maxBounces; ++seg) {
        int lambert_hit = 0, disney_hit = 0, emissive_mis = 0;
        OrDisneyNee dn;
        const ReinaInstanceProperties* hp = NULL;
        v3 nee_n = V3(0.0f, 0.0f, 0.0f), nee_albedo = nee_n;
        float emis_w = 1.0f;
        ReinaRay ray = {o.x, o.y, o.z, 0.0f, d.x, d.y, d.z, 10000.0f};
        ReinaHit hit;
        or_closest_hit(s, &ray, &hit, use_bvh);
        nrays++;
        if (hit.prim == 0xFFFFFFFFu) {
            pld->color = sky(d);
            pld->albedo = pld->color;
            pld->rayHitSky = 1;
            pld->skip = 0;
        } else {
            HitCtx c;
            c.worldRayOrigin = o;
            c.worldRayDirection = d;
            c.inst = &s->inst[s->tri_inst[hit.prim]];
            c.props = &s->props[c.inst->props_id];
            c.prim = s->tri_prim[hit.prim];
            c.bu = hit.u;
            c.bv = hit.v;
            switch (c.inst->material) {
                case 0: chs_lambert_metal(s, pld, &c, 0); break;
                case 1: chs_lambert_metal(s, pld, &c, 1); break;
                case 2: chs_dielectric(s, pld, &c); break;
                default: chs_disney(s, pld, &c, nee ? &dn : NULL); break;
            }
            if (nee && !pld->skip) {
                HitInfo hi = getObjectHitInfo(s, &c);
                lambert_hit = c.inst->material == 0u;
                disney_hit = c.inst->material == 3u;
                hp = c.props;
                nee_n = pld->surfaceNormal;   /* normal-mapped shading normal */
                nee_albedo = pld->albedo;     /* textured colour */
                const int32_t ek = s->inst_emi[s->tri_inst[hit.prim]];
                if (mis_pb > 0.0f && ek >= 0) {
                    v3 w[3];
                    world_triangle(s, c.inst, c.props, c.prim, w);
                    v3 ng = vnormalize(vcross(vsub(w[1], w[0]), vsub(w[2], w[0])));
                    v3 dx = vsub(hi.worldPosition, mis_x);
                    float pl = light_pdf(s->emi[ek].pdf_area, vlength(dx), ng, vnormalize(d),
                                         c.props->cullBackface != 0u);
                    emis_w = power_heuristic(mis_pb, pl);
                    emissive_mis = 1;
                }
            }
        }
        o = pld->rayOrigin;
        d = pld->rayDirection;
        if (pld->skip) continue;
        if (pld->rayHitSky) {
            L = vadd(L, vmul(pld->color, acc));
            break;
        }
        {
            const int inside = pld->insideDielectric;
            if (!inside) {
                v3 contrib = vadd(V3(0.0f, 0.0f, 0.0f), (nee && emissive_mis) ? vscale(pld->emission, emis_w)
                                                                            : pld->emission);
                L = vadd(L, vmul(contrib, acc));
            }
            int shadow = 0;
            v3 scontrib = V3(0.0f, 0.0f, 0.0f);
            ReinaRay sr;
            if (nee && (lambert_hit || disney_hit) && seg + 1u < maxBounces && s->nemi > 0u) {
                const float u1 = or_random(&pld->rngState), u2 = or_random(&pld->rngState);
                const float u3 = or_random(&pld->rngState), u4 = or_random(&pld->rngState);
                const struct OrEmissive* ei = &s->emi[cdf_pick_emi(s, u1)];
                const uint32_t t = cdf_pick_tri(s, ei->cdf_base, ei->cdf_base + ei->ntri - 1u, u2) - ei->cdf_base;
                const OrInstance* li = &s->inst[ei->inst];
                const ReinaInstanceProperties* lp = &s->props[li->props_id];
                v3 w[3];
                world_triangle(s, li, lp, t, w);
                const float beta = 1.0f - sqrtf(u3);
                const float gamma = (1.0f - beta) * u4;
                const float alpha = (1.0f - beta) - gamma;
                v3 y = vadd(vadd(vscale(w[0], alpha), vscale(w[1], beta)), vscale(w[2], gamma));
                v3 ng = vnormalize(vcross(vsub(w[1], w[0]), vsub(w[2], w[0])));
                if (lambert_hit) {
                    v3 dv = vsub(y, pld->rayOrigin);
                    float dist = vlength(dv);
                    v3 wdir = vnormalize(dv);
                    float pl = light_pdf(ei->pdf_area, dist, ng, wdir, lp->cullBackface != 0u);
                    float cx = or_maxf(vdot(nee_n, wdir), 0.0f);
                    if (pl > 0.0f && cx > 0.0f && dist > 0.001f) {
                        float pb = cx / K_PI;
                        float wn = power_heuristic(pl, pb);
                        v3 f = V3(nee_albedo.x / K_PI, nee_albedo.y / K_PI, nee_albedo.z / K_PI);
                        scontrib = vmul(vscale(vmul(pv3(lp->emission), f), (cx / pl) * wn), acc);
                        shadow = 1;
                        ReinaRay r0 = {pld->rayOrigin.x, pld->rayOrigin.y, pld->rayOrigin.z, 0.0f,
                                       wdir.x, wdir.y, wdir.z, dist - 0.001f};
                        sr = r0;
                    }
                    mis_x = pld->rayOrigin;
                    mis_pb = or_maxf(vdot(nee_n, pld->rayDirection), 0.0f) / K_PI;
                } else {
                    /* the shadow ray leaves on the light's side of the surface */
                    v3 so = offsetPositionForDielectric(dn.pos, dn.ng, vnormalize(vsub(y, dn.pos)));
                    v3 dv = vsub(y, so);
                    float dist = vlength(dv);
                    v3 wdir = vnormalize(dv);
                    float pl = light_pdf(ei->pdf_area, dist, ng, wdir, lp->cullBackface != 0u);
                    if (pl > 0.0f && dist > 0.001f) {
                        float pt;
                        v3 g = disney_nee_weight(&dn, hp, s->fixes, wdir, &pt);
                        if (g.x > 0.0f || g.y > 0.0f || g.z > 0.0f) {
                            float wn = power_heuristic(pl, pt);
                            scontrib = vmul(vscale(vmul(pv3(lp->emission), g), wn / pl), acc);
                            shadow = 1;
                            ReinaRay r0 = {so.x, so.y, so.z, 0.0f, wdir.x, wdir.y, wdir.z, dist - 0.001f};
                            sr = r0;
                        }
                    }
                    mis_x = pld->rayOrigin;
                    mis_pb = dn.p_bsdf;
                }
            } else if (nee) {
                mis_pb = 0.0f;
            }
            if (!inside) acc = vmul(acc, pld->color);
            if (shadow && !or_occluded(s, &sr)) L = vadd(L, scontrib);
        }
    }
    *Lout = L;
    return nrays;
}

/* raytrace.rgen.glsl:34-41 */
static void randomGaussian(uint32_t* rng, float* gx, float* gy) {
    float u1 = or_maxf(1e-5f, or_random(rng));
    float u2 = or_random(rng);
    float r = sqrtf(-2.0f * or_logf(u1));
    float theta = (2.0f * K_PI) * u2;
    *gx = r * or_cosf(theta);
    *gy = r * or_sinf(theta);
}

/* raytrace.rgen.glsl:194-204; advances *rng (the caller passes a copy, as the
 * reference passes the state by value) */
static void randomInUnitHexagon(uint32_t* rng, float* px, float* py) {
    const float sqrt3 = 1.73205080757f;
    float x, y;
    do {
        x = 2.0f * or_random(rng) - 1.0f;
        y = (or_random(rng) - 0.5f) * sqrt3;
    } while (fabsf(y) > sqrt3 * 0.5f || (sqrt3 * fabsf(x) + fabsf(y)) > sqrt3);
    *px = x;
    *py = y;
}

/* mat4 (column-major) * vec4 = ((c0*x + c1*y) + c2*z) + c3*w */
static void m4mul(const float* m, const float* v, float* out) {
    for (int r = 0; r < 4; ++r)
        out[r] = ((m[0 + r] * v[0] + m[4 + r] * v[1]) + m[8 + r] * v[2]) + m[12 + r] * v[3];
}

/* raytrace.rgen.glsl:206-247 */
/* lens_advances: the lens draws advance *rng (REINA_RT_TRACE_FIX_LENS_RNG) */
static void getStartingRay(const RtPushConsts* pc, float px, float py, float W, float H, uint32_t* rng,
                           v3* o, v3* d, int lens_advances) {
    float gx, gy;
    randomGaussian(rng, &gx, &gy);
    float cx = (px + 0.5f) + 0.375f * gx;
    float cy = (py + 0.5f) + 0.375f * gy;
    float ndcx = (cx / W) * 2.0f - 1.0f;
    float ndcy = -((cy / H) * 2.0f - 1.0f);
    float clip[4] = {ndcx, ndcy, -1.0f, 1.0f};
    float vp[4];
    m4mul(pc->invProjection, clip, vp);
    float w = vp[3];
    vp[0] = vp[0] / w; vp[1] = vp[1] / w; vp[2] = vp[2] / w;
    v3 viewDir = vnormalize(V3(vp[0], vp[1], vp[2]));
    float vd4[4] = {viewDir.x, viewDir.y, viewDir.z, 0.0f};
    float wd[4];
    m4mul(pc->invView, vd4, wd);
    v3 rayDirection = vnormalize(V3(wd[0], wd[1], wd[2]));
    const float* iv = pc->invView;
    v3 origin = V3(iv[12], iv[13], iv[14]);
    v3 focal = vadd(origin, vscale(rayDirection, pc->focusDist));
    float lx, ly;
    uint32_t hr = *rng;   /* by value in the reference: the stream does not advance */
    randomInUnitHexagon(&hr, &lx, &ly);
    if (lens_advances) *rng = hr;
    lx = lx * pc->defocusMultiplier;
    ly = ly * pc->defocusMultiplier;
    v3 right = vnormalize(V3(iv[0], iv[1], iv[2]));
    v3 up = vnormalize(V3(iv[4], iv[5], iv[6]));
    v3 off = vadd(vscale(right, lx), vscale(up, ly));
    v3 no = vadd(origin, off);
    *o = no;
    *d = vnormalize(vsub(focal, no));
}

void or_camera_ray(const RtPushConsts* pc, uint32_t px, uint32_t py, uint32_t W, uint32_t H,
                   uint32_t* rng, float* origin, float* dir, int lens_advances) {
    v3 o, d;
    getStartingRay(pc, (float)px, (float)py, (float)W, (float)H, rng, &o, &d, lens_advances);
    origin[0] = o.x; origin[1] = o.y; origin[2] = o.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}

/* raytrace.rgen.glsl:259-276 — the spp loop of one dispatch for one pixel,
 * seeded with `batch`; returns the number of non-NaN samples and their mean
 * (sum / float(actual): NaN when every sample was dropped). */
/* With REINA_RT_TRACE_SAMPLE_GROUPS(n) in the scene's options the samples split
 * into G = 2^n groups of spp / G, group g on its own stream seeded for batch
 * batch * G + g; the group sums are added in group order. G = 1: the
 * reference's single stream per pixel. */
static uint32_t pixel_mean(const OrScene* s, const RtPushConsts* pc, uint32_t batch, uint32_t x, uint32_t y,
                           uint32_t W, uint32_t H, int use_bvh, v3* mean, uint64_t* rays) {
    const uint32_t G = 1u << ((s->fixes & REINA_RT_TRACE_SAMPLE_GROUPS_MASK) >> 8);
    uint32_t actual = 0;
    v3 sum = V3(0.0f, 0.0f, 0.0f);
    for (uint32_t g = 0; g < G; ++g) {
        Payload pld;
        memset(&pld, 0, sizeof pld);
        pld.rngState = ((batch * G + g) * H + y) * W + x;
        uint32_t gact = 0;
        v3 gsum = V3(0.0f, 0.0f, 0.0f);
        for (uint32_t si = 0; si < pc->samplesPerPixel / G; ++si) {
            v3 o, d, L;
            getStartingRay(pc, (float)x, (float)y, (float)W, (float)H, &pld.rngState, &o, &d,
                           (s->fixes & REINA_RT_TRACE_FIX_LENS_RNG) != 0);
            if (s->fixes & REINA_RT_TRACE_FIX_MEDIUM_RESET) pld.accumulatedDistance = 0.0f;
            *rays += trace_segments(s, &pld, o, d, pc->maxBounces, use_bvh, &L);
            v3 c = V3(or_clampf(L.x, 0.0f, pc->directClamp), or_clampf(L.y, 0.0f, pc->directClamp),
                      or_clampf(L.z, 0.0f, pc->directClamp));
            if (visnan(c)) continue;
            gact++;
            gsum = vadd(gsum, c);
        }
        sum = g == 0 ? gsum : vadd(sum, gsum);
        actual += gact;
    }
    float fa = (float)actual;
    *mean = V3(sum.x / fa, sum.y / fa, sum.z / fa);
    return actual;
}

/* raytrace.rgen.glsl:250-285 for one pixel; returns the number of rays */
static uint64_t render_pixel(const OrScene* s, const RtPushConsts* pc, uint32_t x, uint32_t y,
                             uint32_t W, uint32_t H, float* px, int use_bvh) {
    uint64_t rays = 0;
    v3 fin;
    pixel_mean(s, pc, pc->sampleBatch, x, y, W, H, use_bvh, &fin, &rays);
    if (pc->sampleBatch > 0) {
        float b = (float)pc->sampleBatch;
        float b1 = (float)(pc->sampleBatch + 1u);
        fin = V3((px[0] * b + fin.x) / b1, (px[1] * b + fin.y) / b1, (px[2] * b + fin.z) / b1);
    }
    px[0] = fin.x; px[1] = fin.y; px[2] = fin.z; px[3] = 1.0f;
    return rays;
}

typedef struct {
    const OrScene* s;
    const RtPushConsts* pc;
    uint32_t W, H, x0, y0, x1, y1;
    float* img;
    int use_bvh;
    uint64_t rays;
    uint64_t* pnext;      /* shared pixel counter (atomic) */
    /* sum64 mode */
    double *sum, *sumsq;
    uint32_t* cnt;
    uint32_t first, nb;
} RJob;

static void* rjob_run(void* arg) {
    RJob* j = (RJob*)arg;
    uint32_t w = j->x1 - j->x0;
    uint64_t npx = (uint64_t)w * (j->y1 - j->y0);
    for (;;) {
        uint64_t i = __atomic_fetch_add(j->pnext, 1, __ATOMIC_RELAXED);
        if (i >= npx) break;
        uint32_t x = j->x0 + (uint32_t)(i % w), y = j->y0 + (uint32_t)(i / w);
        uint64_t pi = (uint64_t)y * j->W + x;
        if (!j->sum) {
            j->rays += render_pixel(j->s, j->pc, x, y, j->W, j->H, &j->img[4ull * pi], j->use_bvh);
            continue;
        }
        /* one dispatch per batch, each contributing its per-pixel mean */
        for (uint32_t b = 0; b < j->nb; ++b) {
            v3 m;
            if (pixel_mean(j->s, j->pc, j->first + b, x, y, j->W, j->H, j->use_bvh, &m, &j->rays) == 0) continue;
            double v[3] = {m.x, m.y, m.z};
            for (int k = 0; k < 3; ++k) {
                j->sum[3 * pi + k] += v[k];
                j->sumsq[3 * pi + k] += v[k] * v[k];
            }
            j->cnt[pi] += 1;
        }
    }
    return NULL;
}

static void run_render(RJob* proto, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    RJob jobs[256];
    uint64_t next = 0;
    for (int k = 0; k < nthreads; ++k) {
        jobs[k] = *proto;
        jobs[k].pnext = &next;
        jobs[k].rays = 0;
        pthread_create(&th[k], NULL, rjob_run, &jobs[k]);
    }
    proto->rays = 0;
    for (int k = 0; k < nthreads; ++k) {
        pthread_join(th[k], NULL);
        proto->rays += jobs[k].rays;
    }
}

void or_render(const OrScene* s, const RtPushConsts* pc, uint32_t W, uint32_t H, float* img, uint32_t x0,
               uint32_t y0, uint32_t x1, uint32_t y1, int use_bvh, int nthreads, uint64_t* rays_out) {
    RJob j;
    memset(&j, 0, sizeof j);
    j.s = s; j.pc = pc; j.W = W; j.H = H; j.x0 = x0; j.y0 = y0; j.x1 = x1; j.y1 = y1;
    j.img = img; j.use_bvh = use_bvh;
    run_render(&j, nthreads);
    if (rays_out) *rays_out = j.rays;
}

void or_render_sum64(const OrScene* s, const RtPushConsts* pc, uint32_t W, uint32_t H, uint32_t first_batch,
                     uint32_t nbatches, double* sum64, double* sumsq64, uint32_t* cnt, uint32_t x0,
                     uint32_t y0, uint32_t x1, uint32_t y1, int use_bvh, int nthreads) {
    RJob j;
    memset(&j, 0, sizeof j);
    j.s = s; j.pc = pc; j.W = W; j.H = H; j.x0 = x0; j.y0 = y0; j.x1 = x1; j.y1 = y1;
    j.use_bvh = use_bvh; j.sum = sum64; j.sumsq = sumsq64; j.cnt = cnt;
    j.first = first_batch; j.nb = nbatches;
    run_render(&j, nthreads);
}

/* ---- math exports -------------------------------------------------------- */
void or_math_eval(int fn, const float* in, const float* in2, float* out, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        float x = in[i];
        switch (fn) {
            case 0: out[i] = or_sinf(x); break;
            case 1: out[i] = or_cosf(x); break;
            case 2: out[i] = or_logf(x); break;
            case 3: out[i] = or_expf(x); break;
            case 4: out[i] = sqrtf(x); break;
            case 5: out[i] = or_powf(x, in2[i]); break;
            case 6: out[i] = or_exp2f(x); break;
            default: out[i] = NAN;
        }
    }
}
