// shade_step.h — shading of one queued ray: the closest-hit / miss shader of
// its hit and the raygen bounce logic (shaders/raytrace/raytrace.rgen.glsl:
// 97-184, 265-273), shared by the single-scene and the multi-request
// wavefront kernels.
#pragma once

#include "shade.h"
#include "wavefront.h"

namespace rt {

struct ShadeResult {
    bool push;      // a next ray (no, nd) continues this pixel's path or starts its next sample
    bool finished;  // the pixel has finished all its samples
    V3 no, nd;
};

// pp: index of the pixel's path state; px: pixel index in its image (y * W + x).
RT_D ShadeResult shade_ray(const SceneView& sv, const TraceParams& P, const PathState& S, uint32_t pp,
                           uint32_t px, float4 ra, float4 rb, float4 hh) {
    ShadeResult r;
    r.push = false;
    r.no = v3(0.0f, 0.0f, 0.0f);
    r.nd = r.no;
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
        const InstanceRec in = sv.inst[sv.tri_inst[prim]];
        const ReinaInstanceProperties& pr = sv.props[in.props_id];
        HitInfo hi = hit_info(sv, in, pr, sv.tri_prim[prim], hh.z, hh.w, D);
        switch (in.material) {
            case 0: chs_lambert_metal(pld, hi, pr, D, false); break;
            case 1: chs_lambert_metal(pld, hi, pr, D, true); break;
            case 2: chs_dielectric(pld, hi, pr, O, D); break;
            default: chs_disney(pld, hi, pr, O, D); break;
        }
    }
    // raytrace.rgen.glsl:124-180 (skipNEE = true: weightBRDF = 1, direct = 0)
    V3 thr = v3(s1.x, s1.y, s1.z), L = v3(s2.x, s2.y, s2.z);
    uint32_t seg = s0.z;
    bool end = false;
    if (!pld.skip) {
        if (pld.rayHitSky) {
            L = add(L, mul(pld.color, thr));
            end = true;
        } else if (!pld.insideDielectric) {
            V3 contrib = add(v3(0.0f, 0.0f, 0.0f), pld.emission);
            L = add(L, mul(contrib, thr));
            thr = mul(thr, pld.color);
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
            uint32_t x = px % P.W, y = px / P.W;
            starting_ray(P.cam, (float)x, (float)y, pld.rng, r.no, r.nd);
            thr = v3(1.0f, 1.0f, 1.0f);
            L = v3(0.0f, 0.0f, 0.0f);
            seg = 0;
            pld.insideDielectric = false;   // raytrace.rgen.glsl:98
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
    r.finished = !r.push;
    return r;
}

}  // namespace rt
