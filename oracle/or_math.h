/*
 * or_math.h — ORACLE (test infrastructure only; never linked into the product).
 *
 * FP32 elementary functions with an explicit, fixed operation sequence, so
 * that the oracle (gcc, x86-64 SSE, -ffp-contract=off) and the HIP kernels
 * (hipcc, -ffp-contract=off) produce identical bits. GLSL leaves the
 * precision of sin/cos/log/exp/pow implementation-defined (the Vulkan
 * environment only bounds their error), so the reference shaders
 * (shaders/raytrace/ *.glsl) pin no particular rounding; these definitions are
 * the build's choice and are checked against libm in double precision
 * (tests/test_oracle_math.py: <= 2 ulp on the domains the path uses).
 *
 *   sin/cos : Cody-Waite reduction by pi/2 (3-part split), Taylor series of
 *             degree 9 (sin) / 10 (cos) on |r| <= pi/4.
 *   log     : x = m * 2^e, m in [sqrt(1/2), sqrt(2)); log(1+f) via the atanh
 *             series s = f/(2+f), evaluated as f - (hfsq - s*(hfsq + R)).
 *   exp     : x = k*ln2 + r, |r| <= ln2/2, Taylor series of degree 8, scaled
 *             by 2^k in two exact steps.
 *   pow     : exp(y*log(x)) for x > 0; integer-constant exponents used by the
 *             shaders (2, 5) are products (pow2, pow5) — pow(x<0, 2) is
 *             undefined in GLSL and the reference's renders require x*x.
 *   min/max : GLSL's definitions, min(x,y) = y < x ? y : x and
 *             max(x,y) = x < y ? y : x (a NaN in x propagates).
 */
#ifndef OR_MATH_H
#define OR_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

static inline uint32_t or_f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float or_u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

#define OR_PIO2_1 1.5703125f                 /* 0x1.92p+0           */
#define OR_PIO2_2 4.837512969970703125e-4f   /* 0x1.fb4p-12         */
#define OR_PIO2_3 7.549789415861596e-8f      /* 0x1.4442d2p-24      */
#define OR_2_OVER_PI 0.63661975f
#define OR_LN2_HI 0.693145751953125f         /* 0x1.62e4p-1         */
#define OR_LN2_LO 1.4286068e-06f             /* 0x1.7f7d1cp-20      */
#define OR_INV_LN2 1.442695f

static inline float or_minf(float x, float y) { return (y < x) ? y : x; }
static inline float or_maxf(float x, float y) { return (x < y) ? y : x; }
static inline float or_clampf(float x, float lo, float hi) { return or_minf(or_maxf(x, lo), hi); }

static inline float or_sin_poly(float r) {
    float r2 = r * r;
    float p = 2.7557319e-6f;            /*  1/9! */
    p = p * r2 + -1.984127e-4f;         /* -1/7! */
    p = p * r2 + 8.333333e-3f;          /*  1/5! */
    p = p * r2 + -0.16666667f;          /* -1/3! */
    return r + r * (r2 * p);
}

static inline float or_cos_poly(float r) {
    float r2 = r * r;
    float p = -2.7557319e-7f;           /* -1/10! */
    p = p * r2 + 2.4801587e-5f;         /*  1/8!  */
    p = p * r2 + -1.3888889e-3f;        /* -1/6!  */
    p = p * r2 + 4.1666668e-2f;         /*  1/4!  */
    float hr2 = 0.5f * r2;
    return (1.0f - hr2) + (r2 * r2) * p;
}

/* quadrant offset: 0 for sin, 1 for cos */
static inline float or_sincos_q(float x, int off) {
    float kf = floorf(x * OR_2_OVER_PI + 0.5f);
    float r = ((x - kf * OR_PIO2_1) - kf * OR_PIO2_2) - kf * OR_PIO2_3;
    float km = kf - 4.0f * floorf(kf * 0.25f);          /* kf mod 4, exact */
    int q = (km == km) ? (((int)km + off) & 3) : 0;     /* NaN-safe */
    switch (q) {
        case 0: return or_sin_poly(r);
        case 1: return or_cos_poly(r);
        case 2: return -or_sin_poly(r);
        default: return -or_cos_poly(r);
    }
}
static inline float or_sinf(float x) { return or_sincos_q(x, 0); }
static inline float or_cosf(float x) { return or_sincos_q(x, 1); }

static inline float or_logf(float x) {
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    int e = 0;
    uint32_t b = or_f2u(x);
    if (b < 0x00800000u) { x = x * 8388608.0f; b = or_f2u(x); e = -23; }
    e += (int)(b >> 23) - 127;
    float m = or_u2f((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = 0.22222222f;              /* 2/9 */
    R = R * z + 0.2857143f;             /* 2/7 */
    R = R * z + 0.4f;                   /* 2/5 */
    R = R * z + 0.6666667f;             /* 2/3 */
    R = R * z;
    float hfsq = 0.5f * (f * f);
    float dk = (float)e;
    return dk * OR_LN2_HI - ((hfsq - (s * (hfsq + R) + dk * OR_LN2_LO)) - f);
}

static inline float or_pow2i(int j) { /* 2^j for j in [-126, 127] */
    return or_u2f((uint32_t)(j + 127) << 23);
}

static inline float or_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283f) return INFINITY;
    if (x < -103.97208f) return 0.0f;
    float kf = floorf(x * OR_INV_LN2 + 0.5f);
    float r = (x - kf * OR_LN2_HI) - kf * OR_LN2_LO;
    float p = 2.4801587e-5f;            /* 1/8! */
    p = p * r + 1.9841270e-4f;          /* 1/7! */
    p = p * r + 1.3888889e-3f;          /* 1/6! */
    p = p * r + 8.3333333e-3f;          /* 1/5! */
    p = p * r + 4.1666668e-2f;          /* 1/4! */
    p = p * r + 0.16666667f;            /* 1/3! */
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int k = (int)kf;
    int k1 = k / 2;
    int k2 = k - k1;
    return (p * or_pow2i(k1)) * or_pow2i(k2);
}

static inline float or_exp2f(float x) {
    if (x != x) return x;
    if (x > 128.0f) return INFINITY;
    if (x < -150.0f) return 0.0f;
    float kf = floorf(x + 0.5f);
    float r = (x - kf) * 0.6931472f;    /* (x - k) is exact */
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
    return (p * or_pow2i(k1)) * or_pow2i(k2);
}

static inline float or_powf(float x, float y) {
    if (x != x || y != y) return NAN;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : ((y == 0.0f) ? 1.0f : INFINITY);
    return or_expf(y * or_logf(x));
}
static inline float or_pow2(float x) { return x * x; }
static inline float or_pow5(float x) { float x2 = x * x; return (x2 * x2) * x; }

/* float -> int as the GPU converts (truncate, NaN -> 0, saturate). */
static inline int32_t or_f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)x;
}

/* ---- vec3 ---------------------------------------------------------------- */
typedef struct { float x, y, z; } v3;
static inline v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vdiv(v3 a, v3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 vscale(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 vneg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline float vdot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float vlength(v3 a) { return sqrtf(vdot(a, a)); }
/* normalize(x) = x * (1 / sqrt(dot(x, x))) */
static inline v3 vnormalize(v3 a) { float s = 1.0f / sqrtf(vdot(a, a)); return vscale(a, s); }
/* reflect(I, N) = I - 2 * dot(N, I) * N */
static inline v3 vreflect(v3 i, v3 n) { return vsub(i, vscale(n, 2.0f * vdot(n, i))); }
/* refract(I, N, eta): GLSL definition */
static inline v3 vrefract(v3 i, v3 n, float eta) {
    float d = vdot(n, i);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return V3(0.0f, 0.0f, 0.0f);
    return vsub(vscale(i, eta), vscale(n, eta * d + sqrtf(k)));
}
/* faceforward(N, I, Nref) = dot(Nref, I) < 0 ? N : -N */
static inline v3 vfaceforward(v3 n, v3 i, v3 nref) { return (vdot(nref, i) < 0.0f) ? n : vneg(n); }
/* mix(x, y, a) = x * (1 - a) + y * a */
static inline v3 vmix(v3 x, v3 y, float a) { return vadd(vscale(x, 1.0f - a), vscale(y, a)); }
static inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
static inline int visnan(v3 a) { return (a.x != a.x) || (a.y != a.y) || (a.z != a.z); }

/* column-major 3x3: c0, c1, c2 are the columns. m * v = c0*v.x + c1*v.y + c2*v.z */
typedef struct { v3 c0, c1, c2; } m3;
static inline v3 m3mul(m3 m, v3 v) {
    return vadd(vadd(vscale(m.c0, v.x), vscale(m.c1, v.y)), vscale(m.c2, v.z));
}
/* transpose(m) * v = (dot(c0, v), dot(c1, v), dot(c2, v)) */
static inline v3 m3tmul(m3 m, v3 v) { return V3(vdot(m.c0, v), vdot(m.c1, v), vdot(m.c2, v)); }

#endif
