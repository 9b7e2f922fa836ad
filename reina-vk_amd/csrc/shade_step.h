// shade_step.h — shading of one queued ray: the closest-hit / miss shader of
// its hit and the raygen bounce logic (shaders/raytrace/raytrace.rgen.glsl:
// 97-184, 265-273), shared by the single-scene and the multi-request
// wavefront kernels; optionally with next-event estimation.
#pragma once

#include "shade.h"
#include "wavefront.h"

namespace rt {

struct ShadeResult {
    bool push;      // a next ray (no, nd) continues this pixel's path or starts its next sample
    bool finished;  // the pixel has finished all its samples
    V3 no, nd;
    // next-event estimation: a shadow ray (so, sd, [0, stmax]) whose light,
    // if unoccluded, adds scontrib to the pixel's radiance
    bool shadow;
    V3 so, sd, scontrib;
    float stmax;
};

// ---- next-event estimation (NEE) with multiple importance sampling --------
//
// The reference compiles NEE but never runs it (skipNEE = true,
// shaders/raytrace/raytrace.rgen.glsl:146); when enabled it pairs the light
// pdf of one direction with the BSDF pdf of another (:153-170). This is the
// standard pairing: at a Lambertian or Disney hit x whose continuation can
// still be traced (segment + 1 < maxBounces) one point y is sampled on the
// emitters (nee.h.glsl:52-119: instance by weight, triangle by area, uniform
// point), connected by a shadow ray and weighted by the power heuristic of
// its solid-angle light pdf p_l against the density p_b with which x's own
// sampler draws that same direction (cosine for Lambertian; disney_density
// for Disney, whose light samples carry disney_nee_weight); an emitter that
// the continuation from x reaches is weighted by p_b^2 / (p_b^2 + p_l^2) with
// p_l computed for that hit. Emission reached from the camera or from a
// metal or dielectric bounce keeps weight 1. A Disney hit samples a light
// also when its continuation took the glass lobe (the raygen then skips the
// bounce's emission and multiply, Q4): the light sample is independent of
// the lobe the continuation drew.

// Power heuristic, beta = 2 (shaders/raytrace/pdf.h.glsl:4-7).
RT_D float power_heuristic(float a, float b) { return (a * a) / (a * a + b * b); }

// Lowest index in [lo, hi] with u <= cdf(index) (nee.h.glsl:52-86).
template <typename F>
RT_D uint32_t cdf_pick(uint32_t lo, uint32_t hi, float u, F cdf) {
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (u <= cdf(mid)) hi = mid;
        else lo = mid + 1u;
    }
    return lo;
}

// World-space vertices of triangle `prim` of instance `in`, bit-identical to
// the flattened triangles of the BVH (((c0 x + c1 y) + c2 z) + c3).
RT_D void world_triangle(const SceneView& sv, const InstanceRec& in, const ReinaInstanceProperties& p, uint32_t prim,
                         V3& w0, V3& w1, V3& w2) {
    const V3 c0 = ldv3(in.c0), c1 = ldv3(in.c1), c2 = ldv3(in.c2), c3 = ldv3(in.c3);
    V3 w[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float4 v = sv.vertices[sv.indices[3u * prim + p.indicesOffset + k]];
        w[k] = add(add(add(scale(c0, v.x), scale(c1, v.y)), scale(c2, v.z)), c3);
    }
    w0 = w[0];
    w1 = w[1];
    w2 = w[2];
}

// Light pdf in solid angle of reaching a point (geometric normal ng, back
// face culled or not) along the unit direction w at distance d.
RT_D float light_pdf(float pdf_area, float d, V3 ng, V3 w, bool cull) {
    float cl = dot(ng, neg(w));
    cl = cull ? maxf_(cl, 0.0f) : fabsf(cl);
    return cl > 0.0f ? (pdf_area * (d * d)) / cl : 0.0f;
}

// pp: index of the pixel's path state; px: pixel index in its image (y * W + x).
// NEE: next-event estimation with MIS; S.st4 holds the last Lambertian bounce
// (its continuation ray's origin and the cosine pdf of that direction; 0
// when the continuation is not MIS-weighted).
template <bool NEE>
RT_D ShadeResult shade_ray(const SceneView& sv, const TraceParams& P, const PathState& S, uint32_t pp,
                           uint32_t px, float4 ra, float4 rb, float4 hh) {
    ShadeResult r;
    r.push = false;
    r.shadow = false;
    r.no = v3(0.0f, 0.0f, 0.0f);
    r.nd = r.no;
    r.so = r.no;
    r.sd = r.no;
    r.scontrib = r.no;
    r.stmax = 0.0f;
    // NEE code is compiled in with NEE and runs for requests that ask for it
    const bool nee = NEE && P.nee != 0u;
    float4 s4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (nee) s4 = S.st4[pp];
    bool lambert_hit = false, disney_hit = false, emissive_mis = false;
    DisneyNee dn;
    const ReinaInstanceProperties* hp = nullptr;
    V3 nee_n = r.no, nee_albedo = r.no;
    float emis_w = 1.0f;
    uint4 s0 = S.st0[pp];
    float4 s1 = S.st1[pp], s2 = S.st2[pp], s3 = S.st3[pp];
    V3 O = v3(ra.x, ra.y, ra.z), D = v3(rb.x, rb.y, rb.z);
    Payload pld;
    pld.rng = s0.x;
    pld.insideDielectric = s0.w != 0u;
    pld.accumulatedDistance = s1.w;
    pld.skip = false;
    pld.rayHitSky = false;
    pld.rayOrigin = O;
    pld.rayDirection = D;
    pld.emission = v3(0.0f, 0.0f, 0.0f);
    pld.color = v3(0.0f, 0.0f, 0.0f);
    uint32_t prim = __float_as_uint(hh.y);
    if (prim == 0xFFFFFFFFu) {
        pld.color = sky_color(D);            // raytrace.rmiss.glsl:10-19
        pld.rayHitSky = true;
        pld.skip = false;
    } else {
        const float4* rec = sv.srec + (size_t)kShadeRecF4 * prim;
        const uint32_t ii = __float_as_uint(rec[0].w), pm = __float_as_uint(rec[1].w);
        const InstanceRec in = sv.inst[ii];
        const ReinaInstanceProperties& pr = sv.props[__float_as_uint(rec[2].w)];
        HitInfo hi = hit_info(sv, in, pr, rec, pm, hh.z, hh.w, D, (P.fixes & REINA_RT_TRACE_FIX_NORMALS) != 0u);
        switch (in.material) {
            case 0: chs_lambert_metal(pld, sv, hi, pr, D, false); break;
            case 1: chs_lambert_metal(pld, sv, hi, pr, D, true); break;
            case 2: chs_dielectric(pld, sv, hi, pr, O, D); break;
            default: chs_disney(pld, sv, hi, pr, O, D, P.fixes, nee ? &dn : nullptr); break;
        }
        if (nee && !pld.skip) {
            lambert_hit = in.material == 0u;
            disney_hit = in.material == 3u;
            hp = &pr;
            nee_n = pld.surfaceNormal;     // the shading normal (normal-mapped) and
            nee_albedo = pld.albedo;       // textured colour of the Lambertian BSDF
            // an emitter reached by the continuation of a Lambertian bounce: MIS weight
            const int32_t ek = sv.inst_emi[ii];
            if (s4.w > 0.0f && ek >= 0) {
                V3 w0, w1, w2;
                world_triangle(sv, in, pr, pm, w0, w1, w2);
                const V3 ng = normalize(cross(sub(w1, w0), sub(w2, w0)));
                const V3 dx = sub(hi.worldPosition, v3(s4.x, s4.y, s4.z));
                const float pl = light_pdf(sv.emi[ek].pdf_area, length(dx), ng, normalize(D), pr.cullBackface != 0u);
                emis_w = power_heuristic(s4.w, pl);
                emissive_mis = true;
            }
        }
    }
    // raytrace.rgen.glsl:124-180 (parity mode: skipNEE = true, weightBRDF = 1, direct = 0)
    V3 thr = v3(s1.x, s1.y, s1.z), L = v3(s2.x, s2.y, s2.z);
    uint32_t seg = s0.z;
    bool end = false;
    if (!pld.skip) {
        if (pld.rayHitSky) {
            L = add(L, mul(pld.color, thr));
            end = true;
        } else {
            const bool inside = pld.insideDielectric;
            if (!inside) {
                V3 contrib = add(v3(0.0f, 0.0f, 0.0f), (nee && emissive_mis) ? scale(pld.emission, emis_w) : pld.emission);
                L = add(L, mul(contrib, thr));
            }
            if (nee && (lambert_hit || disney_hit) && seg + 1u < P.maxBounces && sv.num_emi > 0u) {
                // light sample (4 draws after the BSDF's), connected by the shadow kernel
                const float u1 = rnd(pld.rng), u2 = rnd(pld.rng), u3 = rnd(pld.rng), u4 = rnd(pld.rng);
                const uint32_t k = cdf_pick(0u, sv.num_emi - 1u, u1, [&](uint32_t i) { return sv.emi[i].cdf; });
                const EmissiveInst ei = sv.emi[k];
                const uint32_t t = cdf_pick(ei.cdf_base, ei.cdf_base + ei.ntri - 1u, u2,
                                            [&](uint32_t i) { return sv.tri_cdf[i]; }) - ei.cdf_base;
                const InstanceRec li = sv.inst[ei.inst];
                const ReinaInstanceProperties& lp = sv.props[li.props_id];
                V3 w0, w1, w2;
                world_triangle(sv, li, lp, t, w0, w1, w2);
                const float beta = 1.0f - sqrtf_(u3);           // Shirley 2019, nee.h.glsl:88-95
                const float gamma = (1.0f - beta) * u4;
                const float alpha = (1.0f - beta) - gamma;
                const V3 y = add(add(scale(w0, alpha), scale(w1, beta)), scale(w2, gamma));
                const V3 ng = normalize(cross(sub(w1, w0), sub(w2, w0)));
                if (lambert_hit) {
                    const V3 dv = sub(y, pld.rayOrigin);
                    const float dist = length(dv);
                    const V3 w = normalize(dv);
                    const float pl = light_pdf(ei.pdf_area, dist, ng, w, lp.cullBackface != 0u);
                    const float cx = maxf_(dot(nee_n, w), 0.0f);
                    if (pl > 0.0f && cx > 0.0f && dist > 0.001f) {
                        const float pb = cx / kPi;
                        const float wn = power_heuristic(pl, pb);
                        const V3 f = divs(nee_albedo, kPi);
                        r.scontrib = mul(scale(mul(ldv3(lp.emission), f), (cx / pl) * wn), thr);
                        r.shadow = true;
                        r.so = pld.rayOrigin;
                        r.sd = w;
                        r.stmax = dist - 0.001f;
                    }
                    s4 = make_float4(pld.rayOrigin.x, pld.rayOrigin.y, pld.rayOrigin.z,
                                     maxf_(dot(nee_n, pld.rayDirection), 0.0f) / kPi);
                } else {
                    // the shadow ray leaves on the light's side of the surface
                    const V3 so = offset_dielectric(dn.pos, dn.ng, normalize(sub(y, dn.pos)));
                    const V3 dv = sub(y, so);
                    const float dist = length(dv);
                    const V3 w = normalize(dv);
                    const float pl = light_pdf(ei.pdf_area, dist, ng, w, lp.cullBackface != 0u);
                    if (pl > 0.0f && dist > 0.001f) {
                        float pt;
                        const V3 g = disney_nee_weight(dn, *hp, P.fixes, w, pt);
                        if (g.x > 0.0f || g.y > 0.0f || g.z > 0.0f) {
                            const float wn = power_heuristic(pl, pt);
                            r.scontrib = mul(scale(mul(ldv3(lp.emission), g), wn / pl), thr);
                            r.shadow = true;
                            r.so = so;
                            r.sd = w;
                            r.stmax = dist - 0.001f;
                        }
                    }
                    s4 = make_float4(pld.rayOrigin.x, pld.rayOrigin.y, pld.rayOrigin.z, dn.p_bsdf);
                }
            } else if (nee) {
                s4.w = 0.0f;
            }
            if (!inside) thr = mul(thr, pld.color);
        }
    }
    seg += 1u;
    if (seg >= P.maxBounces) end = true;
    uint32_t sample = s0.y;
    V3 sum = v3(s3.x, s3.y, s3.z);
    uint32_t actual = __float_as_uint(s3.w);
    if (end) {
        // raytrace.rgen.glsl:265-273
        V3 c = v3(clampf_(L.x, 0.0f, P.directClamp), clampf_(L.y, 0.0f, P.directClamp),
                  clampf_(L.z, 0.0f, P.directClamp));
        if (!anynan(c)) { actual += 1u; sum = add(sum, c); }
        sample += 1u;
        if (sample < P.spp) {
            uint32_t x = px % P.W, y = (px / P.W) % P.H;   // px: slot of a sample group
            starting_ray(P.cam, (float)x, (float)y, pld.rng, r.no, r.nd, (P.fixes & REINA_RT_TRACE_FIX_LENS_RNG) != 0u);
            thr = v3(1.0f, 1.0f, 1.0f);
            L = v3(0.0f, 0.0f, 0.0f);
            seg = 0;
            pld.insideDielectric = false;   // raytrace.rgen.glsl:98
            if (P.fixes & REINA_RT_TRACE_FIX_MEDIUM_RESET) pld.accumulatedDistance = 0.0f;
            s4.w = 0.0f;
            r.push = true;
        }
        S.st3[pp] = make_float4(sum.x, sum.y, sum.z, __uint_as_float(actual));
    } else {
        r.no = pld.rayOrigin;
        r.nd = pld.rayDirection;
        r.push = true;
    }
    S.st0[pp] = make_uint4(pld.rng, sample, seg, pld.insideDielectric ? 1u : 0u);
    S.st1[pp] = make_float4(thr.x, thr.y, thr.z, pld.accumulatedDistance);
    S.st2[pp] = make_float4(L.x, L.y, L.z, 0.0f);
    if (nee) S.st4[pp] = s4;
    r.finished = !r.push;
    return r;
}

}  // namespace rt
