// rt_math.h — device-side FP32 arithmetic for the path-tracing kernels.
//
// Every elementary function here is defined by an explicit operation sequence
// (the file is compiled with -ffp-contract=off, so no multiply-add is fused),
// which makes the kernels' results reproducible bit for bit on any IEEE-754
// binary32 machine that evaluates the same sequence. GLSL leaves the precision
// of sin/cos/log/exp/pow implementation-defined, so the reference shaders
// (shaders/raytrace/brdfDisney.h.glsl, closestHitCommon.h.glsl,
// raytrace.rgen.glsl) do not pin their rounding; this file is the library's
// definition of them:
//   sin/cos: Cody-Waite reduction by pi/2 in three parts + Taylor series of
//            degree 9 / 10 on |r| <= pi/4;
//   log:     x = m 2^e, m in [sqrt(1/2), sqrt(2)), atanh series in s = f/(2+f);
//   exp:     x = k ln2 + r, Taylor series of degree 8, exact 2^k scaling;
//   pow:     exp(y log x) for x > 0; the shaders' integer exponents 2 and 5 are
//            products (GLSL leaves pow(x < 0, 2) undefined).
//   min/max: GLSL's comparison definitions (NaN in the first operand propagates).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RT_D __device__ __forceinline__

namespace rt {

RT_D float minf_(float x, float y) { return (y < x) ? y : x; }
RT_D float maxf_(float x, float y) { return (x < y) ? y : x; }
RT_D float clampf_(float x, float lo, float hi) { return minf_(maxf_(x, lo), hi); }

RT_D float sin_poly(float r) {
    float r2 = r * r;
    float p = 2.7557319e-6f;
    p = p * r2 + -1.984127e-4f;
    p = p * r2 + 8.333333e-3f;
    p = p * r2 + -0.16666667f;
    return r + r * (r2 * p);
}

RT_D float cos_poly(float r) {
    float r2 = r * r;
    float p = -2.7557319e-7f;
    p = p * r2 + 2.4801587e-5f;
    p = p * r2 + -1.3888889e-3f;
    p = p * r2 + 4.1666668e-2f;
    float hr2 = 0.5f * r2;
    return (1.0f - hr2) + (r2 * r2) * p;
}

// Returns sin(x) in *s and cos(x) in *c with one shared reduction.
RT_D void sincos_(float x, float* s, float* c) {
    float kf = floorf(x * 0.63661975f + 0.5f);
    float r = ((x - kf * 1.5703125f) - kf * 4.837512969970703125e-4f) - kf * 7.549789415861596e-8f;
    float km = kf - 4.0f * floorf(kf * 0.25f);
    int q = (km == km) ? ((int)km & 3) : 0;
    float sp = sin_poly(r), cp = cos_poly(r);
    float ss, cc;
    if (q == 0) { ss = sp; cc = cp; }
    else if (q == 1) { ss = cp; cc = -sp; }
    else if (q == 2) { ss = -sp; cc = -cp; }
    else { ss = -cp; cc = sp; }
    *s = ss;
    *c = cc;
}

RT_D float logf_(float x) {
    if (x != x || x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_inff();
    if (x == __builtin_inff()) return __builtin_inff();
    int e = 0;
    uint32_t b = __float_as_uint(x);
    if (b < 0x00800000u) { x = x * 8388608.0f; b = __float_as_uint(x); e = -23; }
    e += (int)(b >> 23) - 127;
    float m = __uint_as_float((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = 0.22222222f;
    R = R * z + 0.2857143f;
    R = R * z + 0.4f;
    R = R * z + 0.6666667f;
    R = R * z;
    float hfsq = 0.5f * (f * f);
    float dk = (float)e;
    return dk * 0.693145751953125f - ((hfsq - (s * (hfsq + R) + dk * 1.4286068e-06f)) - f);
}

RT_D float pow2i_(int j) { return __uint_as_float((uint32_t)(j + 127) << 23); }

RT_D float exp_tail(float r, float kf) {
    float p = 2.4801587e-5f;
    p = p * r + 1.9841270e-4f;
    p = p * r + 1.3888889e-3f;
    p = p * r + 8.3333333e-3f;
    p = p * r + 4.1666668e-2f;
    p = p * r + 0.16666667f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int k = (int)kf;
    int k1 = k / 2;
    int k2 = k - k1;
    return (p * pow2i_(k1)) * pow2i_(k2);
}

RT_D float expf_(float x) {
    if (x != x) return x;
    if (x > 88.72283f) return __builtin_inff();
    if (x < -103.97208f) return 0.0f;
    float kf = floorf(x * 1.442695f + 0.5f);
    float r = (x - kf * 0.693145751953125f) - kf * 1.4286068e-06f;
    return exp_tail(r, kf);
}

RT_D float exp2f_(float x) {
    if (x != x) return x;
    if (x > 128.0f) return __builtin_inff();
    if (x < -150.0f) return 0.0f;
    float kf = floorf(x + 0.5f);
    float r = (x - kf) * 0.6931472f;
    return exp_tail(r, kf);
}

RT_D float powf_(float x, float y) {
    if (x != x || y != y) return __builtin_nanf("");
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : ((y == 0.0f) ? 1.0f : __builtin_inff());
    return expf_(y * logf_(x));
}
RT_D float pow2_(float x) { return x * x; }
RT_D float pow5_(float x) { float x2 = x * x; return (x2 * x2) * x; }

// IEEE square root and division (the build keeps correctly-rounded fp32 div/sqrt).
RT_D float sqrtf_(float x) { return __builtin_sqrtf(x); }

RT_D int32_t f2i_(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)x;
}

// ---- 3-vectors -----------------------------------------------------------
struct V3 { float x, y, z; };
RT_D V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
RT_D V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
RT_D V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
RT_D V3 mul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
RT_D V3 scale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
RT_D V3 divs(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
RT_D V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
RT_D float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
RT_D V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
RT_D float length(V3 a) { return sqrtf_(dot(a, a)); }
RT_D V3 normalize(V3 a) { float s = 1.0f / sqrtf_(dot(a, a)); return scale(a, s); }
RT_D V3 reflect(V3 i, V3 n) { return sub(i, scale(n, 2.0f * dot(n, i))); }
RT_D V3 refract(V3 i, V3 n, float eta) {
    float d = dot(n, i);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return v3(0.0f, 0.0f, 0.0f);
    return sub(scale(i, eta), scale(n, eta * d + sqrtf_(k)));
}
RT_D V3 faceforward(V3 n, V3 i, V3 nref) { return (dot(nref, i) < 0.0f) ? n : neg(n); }
RT_D V3 mix(V3 x, V3 y, float a) { return add(scale(x, 1.0f - a), scale(y, a)); }
RT_D float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
RT_D bool anynan(V3 a) { return (a.x != a.x) || (a.y != a.y) || (a.z != a.z); }

// column-major 3x3
struct M3 { V3 c0, c1, c2; };
RT_D V3 mmul(const M3& m, V3 v) { return add(add(scale(m.c0, v.x), scale(m.c1, v.y)), scale(m.c2, v.z)); }
RT_D V3 tmul(const M3& m, V3 v) { return v3(dot(m.c0, v), dot(m.c1, v), dot(m.c2, v)); }

}  // namespace rt
