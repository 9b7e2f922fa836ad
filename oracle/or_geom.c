/*
 * or_geom.c — ORACLE (test infrastructure only): scene flattening and the
 * closest-hit query that the driver performs for traceRayEXT
 * (shaders/raytrace/raytrace.rgen.glsl:110-122: opaque, no culling, mask 0xFF,
 * tmin 0, tmax 10000).
 *
 *  - or_scene_create flattens every TLAS instance (src/tools/vktools.cpp:460-488:
 *    3x4 object-to-world rows, customIndex = instancePropertiesID, SBT offset =
 *    materialIdx) over its BLAS triangles (src/graphics/Blas.cpp:17-39: u32
 *    non-offset indices + firstVertex; the offset index buffer of
 *    src/scene/Models.cpp:78-81 addresses the same vertices).
 *  - or_intersect_brute32: the watertight FP32 ray/triangle test (Woop, Benthin,
 *    Wald, JCGT 2(1) 2013) over all triangles; closest hit = argmin over the
 *    pair (t, flattened id), which makes the answer independent of visit order.
 *  - or_intersect_bvh: the same test driven by a BVH8 traversal; must return the
 *    brute-force answer bit for bit, and defines the node-visit / triangle-test
 *    counts of the library's extend kernel.
 *  - or_intersect_fp64: textbook Moller-Trumbore in double precision over all
 *    triangles, with an ambiguity flag for edge/vertex/near-equal-t ties.
 */
#include "oracle.h"

#include <float.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* world = ((c0*x + c1*y) + c2*z) + c3 : gl_ObjectToWorldEXT * vec4(p, 1) */
static v3 xform_point(const OrInstance* in, v3 p) {
    return vadd(vadd(vadd(vscale(in->c0, p.x), vscale(in->c1, p.y)), vscale(in->c2, p.z)), in->c3);
}

#define DUP(dst, src, count, type)                                             \
    do {                                                                       \
        (dst) = (type*)malloc((((count) != 0) ? (count) : 1) * sizeof(type));         \
        if (((count) != 0) && (src)) memcpy((dst), (src), (count) * sizeof(type));    \
    } while (0)

/* Light-sampling tables for next-event estimation (the library's
 * build_emissive, api.cpp): an instance is emissive when |e|^2 > 1e-10
 * (src/scene/Instance.cpp:75-77); weight = world-space area x luminance
 * (Instance.cpp:10-11, 15-53, with the transform applied as a column-major
 * matrix); instance CDF by weight, triangle CDF by area; pdf_area =
 * (weight / total) / area. Per-triangle flags: 1 = back faces culled,
 * 2 = mirroring transform (det < 0). */
static void build_emissive(OrScene* s) {
    s->inst_emi = (int32_t*)malloc((s->ninst ? s->ninst : 1) * sizeof(int32_t));
    s->tri_flags = (uint8_t*)calloc(s->ntri ? s->ntri : 1, 1);
    s->emi = (struct OrEmissive*)calloc(s->ninst ? s->ninst : 1, sizeof(struct OrEmissive));
    s->tri_cdf = (float*)malloc((s->ntri ? s->ntri : 1) * sizeof(float));
    double* weight = (double*)malloc((s->ninst ? s->ninst : 1) * sizeof(double));
    double* cum = (double*)malloc((s->ntri ? s->ntri : 1) * sizeof(double));
    double total = 0.0;
    s->nemi = 0;
    s->ntri_cdf = 0;
    for (uint32_t i = 0; i < s->ninst; ++i) {
        const OrInstance* r = &s->inst[i];
        const ReinaInstanceProperties* p = &s->props[r->props_id];
        const uint32_t ntri = s->models[r->model].indexCount;
        s->inst_emi[i] = -1;
        const double det = (double)r->c0.x * ((double)r->c1.y * r->c2.z - (double)r->c2.y * r->c1.z) -
                           (double)r->c1.x * ((double)r->c0.y * r->c2.z - (double)r->c2.y * r->c0.z) +
                           (double)r->c2.x * ((double)r->c0.y * r->c1.z - (double)r->c1.y * r->c0.z);
        const uint8_t fl = (uint8_t)((p->cullBackface ? 1 : 0) | (det < 0.0 ? 2 : 0));
        for (uint32_t t = 0; t < ntri; ++t) s->tri_flags[r->tri_base + t] = fl;
        const float* e = p->emission;
        if (!((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2] > 0.00001f * 0.00001f)) continue;
        const double lum = 0.2126 * e[0] + 0.7152 * e[1] + 0.0722 * e[2];
        double area = 0.0;
        for (uint32_t t = 0; t < ntri; ++t) {
            const float* w = &s->wtri[9ull * (r->tri_base + t)];
            double ax = (double)w[3] - w[0], ay = (double)w[4] - w[1], az = (double)w[5] - w[2];
            double bx = (double)w[6] - w[0], by = (double)w[7] - w[1], bz = (double)w[8] - w[2];
            double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
            area += 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
            cum[t] = area;
        }
        if (!(area > 0.0) || !(lum > 0.0)) continue;
        struct OrEmissive* ei = &s->emi[s->nemi];
        ei->inst = i;
        ei->ntri = ntri;
        ei->cdf_base = s->ntri_cdf;
        for (uint32_t t = 0; t < ntri; ++t) s->tri_cdf[s->ntri_cdf + t] = (float)(cum[t] / area);
        s->ntri_cdf += ntri;
        s->tri_cdf[s->ntri_cdf - 1] = 1.0f;
        weight[s->nemi] = area * lum;
        total += area * lum;
        ei->pdf_area = (float)area;
        s->inst_emi[i] = (int32_t)s->nemi;
        s->nemi++;
    }
    double c = 0.0;
    for (uint32_t k = 0; k < s->nemi; ++k) {
        c += weight[k];
        s->emi[k].cdf = (float)(c / total);
        s->emi[k].pdf_area = (float)((weight[k] / total) / (double)s->emi[k].pdf_area);
    }
    if (s->nemi) s->emi[s->nemi - 1].cdf = 1.0f;
    free(weight);
    free(cum);
}

uint64_t or_export_emissive(const OrScene* s, void* buf) {
    const uint64_t need = 16 + 32ull * s->nemi + 4ull * s->ntri_cdf + 4ull * s->ninst + s->ntri;
    if (!buf) return need;
    uint8_t* p = (uint8_t*)buf;
    memset(p, 0, need);
    uint32_t hdr[4] = {s->nemi, s->ntri_cdf, s->ninst, (uint32_t)s->ntri};
    memcpy(p, hdr, 16); p += 16;
    for (uint32_t k = 0; k < s->nemi; ++k) {
        memcpy(p, &s->emi[k].inst, 4);
        memcpy(p + 4, &s->emi[k].ntri, 4);
        memcpy(p + 8, &s->emi[k].cdf_base, 4);
        memcpy(p + 12, &s->emi[k].cdf, 4);
        memcpy(p + 16, &s->emi[k].pdf_area, 4);
        p += 32;
    }
    memcpy(p, s->tri_cdf, 4ull * s->ntri_cdf); p += 4ull * s->ntri_cdf;
    memcpy(p, s->inst_emi, 4ull * s->ninst); p += 4ull * s->ninst;
    memcpy(p, s->tri_flags, s->ntri);
    return need;
}

void or_scene_set_nee(OrScene* s, int on) { s->nee = on; }
void or_scene_set_fixes(OrScene* s, uint32_t fixes) { s->fixes = fixes; }

OrScene* or_scene_create(const ReinaSceneDesc* d) {
    OrScene* s = (OrScene*)calloc(1, sizeof(OrScene));
    s->nv = d->num_vertices;     DUP(s->vertices, d->vertices, s->nv * 4, float);
    s->ni = d->num_indices;      DUP(s->indices, d->indices, s->ni, uint32_t);
    s->ntbn = d->num_tbns;       DUP(s->tbns, d->tbns, s->ntbn * 9, float);
    s->ntbni = d->num_tbn_indices; DUP(s->tbn_indices, d->tbn_indices, s->ntbni, uint32_t);
    s->ntc = d->num_tex_coords;  DUP(s->tex_coords, d->tex_coords, s->ntc * 2, float);
    s->ntci = d->num_tex_indices; DUP(s->tex_indices, d->tex_indices, s->ntci, uint32_t);
    s->nprops = d->num_props;    DUP(s->props, d->props, s->nprops, ReinaInstanceProperties);
    s->nmodels = d->num_models;  DUP(s->models, d->models, s->nmodels, ReinaModelRange);
    s->ntex = d->num_textures;
    s->tex = (struct OrTexture*)calloc(s->ntex ? s->ntex : 1, sizeof(struct OrTexture));
    for (uint32_t t = 0; t < s->ntex; ++t) {
        s->tex[t].w = d->textures[t].width;
        s->tex[t].h = d->textures[t].height;
        DUP(s->tex[t].rgba8, d->textures[t].rgba8, 4ull * s->tex[t].w * s->tex[t].h, uint8_t);
    }
    s->ninst = d->num_instances;
    s->inst = (OrInstance*)calloc(s->ninst ? s->ninst : 1, sizeof(OrInstance));
    uint64_t ntri = 0;
    for (uint32_t i = 0; i < s->ninst; ++i) {
        const ReinaInstanceDesc* id = &d->instances[i];
        OrInstance* in = &s->inst[i];
        const float* m = id->transform; /* column-major mat4 */
        in->c0 = V3(m[0], m[1], m[2]);
        in->c1 = V3(m[4], m[5], m[6]);
        in->c2 = V3(m[8], m[9], m[10]);
        in->c3 = V3(m[12], m[13], m[14]);
        in->props_id = id->instancePropertiesID;
        in->material = id->materialIdx;
        in->model = id->modelID;
        in->tri_base = (uint32_t)ntri;
        ntri += s->models[id->modelID].indexCount;
    }
    s->ntri = ntri;
    s->wtri = (float*)malloc((ntri ? ntri : 1) * 9 * sizeof(float));
    s->tri_inst = (uint32_t*)malloc((ntri ? ntri : 1) * sizeof(uint32_t));
    s->tri_prim = (uint32_t*)malloc((ntri ? ntri : 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < s->ninst; ++i) {
        const OrInstance* in = &s->inst[i];
        const ReinaModelRange* mr = &s->models[in->model];
        for (uint32_t p = 0; p < mr->indexCount; ++p) {
            uint64_t g = (uint64_t)in->tri_base + p;
            for (int k = 0; k < 3; ++k) {
                uint32_t vi = s->indices[3ull * p + mr->indexOffset + k];
                v3 ov = V3(s->vertices[4ull * vi], s->vertices[4ull * vi + 1], s->vertices[4ull * vi + 2]);
                v3 w = xform_point(in, ov);
                s->wtri[9 * g + 3 * k + 0] = w.x;
                s->wtri[9 * g + 3 * k + 1] = w.y;
                s->wtri[9 * g + 3 * k + 2] = w.z;
            }
            s->tri_inst[g] = i;
            s->tri_prim[g] = p;
        }
    }
    build_emissive(s);
    return s;
}

void or_scene_destroy(OrScene* s) {
    if (!s) return;
    free(s->vertices); free(s->indices); free(s->tbns); free(s->tbn_indices);
    free(s->tex_coords); free(s->tex_indices); free(s->props); free(s->models);
    free(s->inst); free(s->wtri); free(s->tri_inst); free(s->tri_prim);
    free(s->emi); free(s->tri_cdf); free(s->inst_emi); free(s->tri_flags);
    for (uint32_t t = 0; t < s->ntex; ++t) free(s->tex[t].rgba8);
    free(s->tex);
    free(s);
}

int or_scene_set_bvh(OrScene* s, const void* nodes, uint64_t nodes_bytes, const void* tris,
                     uint64_t tris_bytes, uint32_t node_bytes) {
    if (node_bytes != 256 && node_bytes != 80) return -1;
    s->nodes = (const uint8_t*)nodes;
    s->node_bytes = node_bytes;
    s->nnodes = nodes_bytes / node_bytes;
    s->btris = (const float*)tris;
    s->nbtris = tris_bytes / 48;
    return 0;
}

/* ---- watertight FP32 test ------------------------------------------------ */
typedef struct {
    v3 o, d;
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float tmin, tmax;
    int valid;   /* finite origin and direction, direction != 0; invalid rays miss */
} WRay;

static int or_finite(float x) { return (x == x) && (x - x == 0.0f); }

static float comp(v3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

static void wray_init(WRay* r, const ReinaRay* in) {
    r->o = V3(in->ox, in->oy, in->oz);
    r->d = V3(in->dx, in->dy, in->dz);
    r->tmin = in->tmin;
    r->tmax = in->tmax;
    r->valid = or_finite(r->o.x) && or_finite(r->o.y) && or_finite(r->o.z) && or_finite(r->d.x) && or_finite(r->d.y) &&
               or_finite(r->d.z) && !(r->d.x == 0.0f && r->d.y == 0.0f && r->d.z == 0.0f);
    float ax = fabsf(r->d.x), ay = fabsf(r->d.y), az = fabsf(r->d.z);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (comp(r->d, kz) < 0.0f) { int t = kx; kx = ky; ky = t; }
    r->kx = kx; r->ky = ky; r->kz = kz;
    float dz = comp(r->d, kz);
    r->Sx = comp(r->d, kx) / dz;
    r->Sy = comp(r->d, ky) / dz;
    r->Sz = 1.0f / dz;
}

/* returns 1 and writes t,u,v if the triangle is hit with t in [tmin, tbest] */
static int wtest(const WRay* r, const float* v0p, const float* v1p, const float* v2p, float tbest,
                 float* to, float* uo, float* vo) {
    v3 A = vsub(V3(v0p[0], v0p[1], v0p[2]), r->o);
    v3 B = vsub(V3(v1p[0], v1p[1], v1p[2]), r->o);
    v3 C = vsub(V3(v2p[0], v2p[1], v2p[2]), r->o);
    float Akz = comp(A, r->kz), Bkz = comp(B, r->kz), Ckz = comp(C, r->kz);
    float Ax = comp(A, r->kx) - r->Sx * Akz;
    float Ay = comp(A, r->ky) - r->Sy * Akz;
    float Bx = comp(B, r->kx) - r->Sx * Bkz;
    float By = comp(B, r->ky) - r->Sy * Bkz;
    float Cx = comp(C, r->kx) - r->Sx * Ckz;
    float Cy = comp(C, r->ky) - r->Sy * Ckz;
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
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    float det = (U + V) + W;
    if (det == 0.0f) return 0;
    float Az = r->Sz * Akz, Bz = r->Sz * Bkz, Cz = r->Sz * Ckz;
    float T = (U * Az + V * Bz) + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= r->tmin && t <= tbest)) return 0;
    *to = t;
    *uo = V * rcp;
    *vo = W * rcp;
    return 1;
}

/* closest = argmin (t, id) */
static inline int better(float t, uint32_t id, float tb, uint32_t idb) {
    return (t < tb) || (t == tb && id < idb);
}

static void brute32_one(const OrScene* s, const ReinaRay* ray, ReinaHit* hit) {
    WRay r;
    wray_init(&r, ray);
    float tb = r.tmax;
    uint32_t ib = 0xFFFFFFFFu;
    float ub = 0.0f, vb = 0.0f;
    for (uint64_t g = 0; r.valid && g < s->ntri; ++g) {
        const float* v = &s->wtri[9 * g];
        float t, u, w;
        if (wtest(&r, v, v + 3, v + 6, tb, &t, &u, &w) && better(t, (uint32_t)g, tb, ib)) {
            tb = t; ib = (uint32_t)g; ub = u; vb = w;
        }
    }
    hit->t = tb; hit->prim = ib; hit->u = ub; hit->v = vb;
}

/* ---- BVH8 traversal restatement ------------------------------------------
 * Node layout (256 bytes): float lo_x[8], hi_x[8], lo_y[8], hi_y[8], lo_z[8],
 * hi_z[8]; u32 child_base | imask << 24; u32 tri_base | leafmask << 24;
 * u8 meta[8]; 48 bytes padding (imask / leafmask: slots holding internal
 * children / leaves).
 * meta 0 = empty slot; 0x80|rank = internal child child_base+rank; otherwise a
 * leaf of (meta>>5)&3 triangles starting at tri_base + (meta & 31).
 * Triangles: float4 x3 (v0.xyz|id, v1.xyz|-, v2.xyz|-).
 * Visit order: leaves of a node first, in slot order i ^ oct (oct = ray
 * direction sign bits); then the node's internal hits become a group
 * {child_base, internal-slot mask, ordered hit mask} on a LIFO stack, popped
 * lowest ordered bit first. Box test: slabs with the near/far plane picked
 * by the direction sign; a plane distance is fmaf(plane, 1/d, -o/d -+ E)
 * with E = 2^-16 |o/d| when every |1/d| <= 2^40 and o/d is finite, else
 * (plane - o) * (1/d) with +-inf interval ends clamped to +-3e38; then the
 * interval is widened by 2^-16 |t| (conservative: never rejects a box whose
 * triangle passes). Rays with a non-finite component or a zero direction
 * miss everything. */
#define BOX_EPS 1.52587890625e-05f /* 2^-16 */
#define SLAB_BIG 3.0e38f
#define FMA_INV_MAX 1099511627776.0f /* 2^40 */

static int finitef_(float x) { return (x == x) && (x - x == 0.0f); }

typedef struct { uint32_t base, imask, om; } Group;

static void bvh_one(const OrScene* s, const ReinaRay* ray, ReinaHit* hit, uint32_t* cnt) {
    WRay r;
    wray_init(&r, ray);
    float tb = r.tmax;
    uint32_t ib = 0xFFFFFFFFu;
    float ub = 0.0f, vb = 0.0f;
    uint32_t nvis = 0, ntst = 0;
    /* inverse direction; a zero component gives +-inf */
    float ix = 1.0f / r.d.x, iy = 1.0f / r.d.y, iz = 1.0f / r.d.z;
    /* slab constants -o/d -+ E, E = 2^-16 |o/d| */
    float oix = -(r.o.x * ix), oiy = -(r.o.y * iy), oiz = -(r.o.z * iz);
    float ex = fabsf(oix) * BOX_EPS, ey = fabsf(oiy) * BOX_EPS, ez = fabsf(oiz) * BOX_EPS;
    float cnx = oix - ex, cny = oiy - ey, cnz = oiz - ez;
    float cfx = oix + ex, cfy = oiy + ey, cfz = oiz + ez;
    int fma_slabs = finitef_(oix) && finitef_(oiy) && finitef_(oiz) && fabsf(ix) <= FMA_INV_MAX &&
                    fabsf(iy) <= FMA_INV_MAX && fabsf(iz) <= FMA_INV_MAX;
    uint32_t oct = (r.d.x < 0.0f ? 1u : 0u) | (r.d.y < 0.0f ? 2u : 0u) | (r.d.z < 0.0f ? 4u : 0u);
    Group stack[64];
    int sp = 0;
    uint32_t cur = 0;
    int have = (s->nnodes > 0) && r.valid;
    while (have) {
        const uint8_t* nb = s->nodes + (uint64_t)cur * s->node_bytes;
        float fq[48];
        const float* f;
        uint32_t w48, w49;
        const uint8_t* meta;
        if (s->node_bytes == 256) {
            f = (const float*)nb;
            memcpy(&w48, nb + 192, 4);
            memcpy(&w49, nb + 196, 4);
            meta = nb + 200;
        } else {
            /* quantized node: p.xyz, exponent bytes, header, meta, then
             * qlo_x qhi_x qlo_y qhi_y qlo_z qhi_z (8 bytes each); plane =
             * fmaf(q, 2^(e-127), p) */
            float p[3];
            uint32_t ex;
            memcpy(p, nb, 12);
            memcpy(&ex, nb + 12, 4);
            for (int a = 0; a < 3; ++a) {
                uint32_t eb = ((ex >> (8 * a)) & 0xFFu) << 23;
                float sc;
                memcpy(&sc, &eb, 4);
                for (int k = 0; k < 8; ++k) {
                    fq[16 * a + k] = fmaf((float)nb[32 + 16 * a + k], sc, p[a]);
                    fq[16 * a + 8 + k] = fmaf((float)nb[32 + 16 * a + 8 + k], sc, p[a]);
                }
            }
            f = fq;
            memcpy(&w48, nb + 16, 4);
            memcpy(&w49, nb + 20, 4);
            meta = nb + 24;
        }
        uint32_t child_base = w48 & 0xFFFFFFu, tri_base = w49 & 0xFFFFFFu;
        uint32_t imask = w48 >> 24, leafmask = w49 >> 24;
        nvis++;
        uint32_t hitm = 0;
        /* near/far plane per axis by direction sign (lo <= hi, so this is min/max);
         * the interval is widened by 2^-16 |t| with one fused multiply-add */
        int sx = ix >= 0.0f, sy = iy >= 0.0f, sz = iz >= 0.0f;
        for (int k = 0; k < 8; ++k) {
            float nx = sx ? f[0 + k] : f[8 + k], fx = sx ? f[8 + k] : f[0 + k];
            float ny = sy ? f[16 + k] : f[24 + k], fy = sy ? f[24 + k] : f[16 + k];
            float nz = sz ? f[32 + k] : f[40 + k], fz = sz ? f[40 + k] : f[32 + k];
            float tn, tf;
            if (fma_slabs) {
                tn = fmaxf(fmaxf(fmaf(nx, ix, cnx), fmaf(ny, iy, cny)), fmaf(nz, iz, cnz));
                tf = fminf(fminf(fmaf(fx, ix, cfx), fmaf(fy, iy, cfy)), fmaf(fz, iz, cfz));
            } else {
                /* d = 0 on an axis: its distances are +-inf, or NaN (0 * inf)
                 * for a ray inside a plane, which fmaxf/fminf ignore (closed
                 * slab); +inf / -inf interval ends are clamped before the
                 * widening */
                tn = fmaxf(fmaxf((nx - r.o.x) * ix, (ny - r.o.y) * iy), (nz - r.o.z) * iz);
                tf = fminf(fminf((fx - r.o.x) * ix, (fy - r.o.y) * iy), (fz - r.o.z) * iz);
                tn = fminf(tn, SLAB_BIG);
                tf = fmaxf(tf, -SLAB_BIG);
            }
            tn = fmaxf(fmaf(-fabsf(tn), BOX_EPS, tn), r.tmin);
            tf = fminf(fmaf(fabsf(tf), BOX_EPS, tf), tb);
            if (tn <= tf) hitm |= 1u << k;
        }
        uint32_t om = 0;
        for (uint32_t i = 0; i < 8; ++i) {
            uint32_t k = i ^ oct;
            if (!(hitm & (1u << k))) continue;
            if (imask & (1u << k)) { om |= 1u << i; continue; }
            if (!(leafmask & (1u << k))) continue;
            uint8_t m = meta[k];
            uint32_t cntl = (m >> 5) & 3u, off = m & 31u;
            for (uint32_t j = 0; j < cntl; ++j) {
                const float* tv = s->btris + 12ull * (tri_base + off + j);
                uint32_t id;
                memcpy(&id, &tv[3], 4);
                float t, u, w;
                ntst++;
                if (wtest(&r, tv, tv + 4, tv + 8, tb, &t, &u, &w) && better(t, id, tb, ib)) {
                    tb = t; ib = id; ub = u; vb = w;
                }
            }
        }
        if (om) { stack[sp].base = child_base; stack[sp].imask = imask; stack[sp].om = om; sp++; }
        have = 0;
        if (sp > 0) {
            Group* g = &stack[sp - 1];
            uint32_t i = (uint32_t)__builtin_ctz(g->om);
            g->om &= g->om - 1;
            uint32_t slot = i ^ oct;
            cur = g->base + (uint32_t)__builtin_popcount(g->imask & ((1u << slot) - 1u));
            if (g->om == 0) sp--;
            have = 1;
        }
    }
    hit->t = tb; hit->prim = ib; hit->u = ub; hit->v = vb;
    if (cnt) { cnt[0] = nvis; cnt[1] = ntst; }
}

/* ---- FP64 Moller-Trumbore ------------------------------------------------ */
typedef struct { double x, y, z; } d3;
static d3 D3(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static d3 dsub(d3 a, d3 b) { return D3(a.x - b.x, a.y - b.y, a.z - b.z); }
static double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 dcross(d3 a, d3 b) { return D3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

/* Per-ray FP64 closest-hit accumulator. Order independent: the best hit is
 * the argmin of (t, id); t2 is the smallest lower t bound of the other hits;
 * nm_t the smallest lower t bound of a near miss.
 *
 * Ties ("ambiguous" rays) are where FP32 rounding can decide the answer.
 * The FP32 test translates the vertices to the ray origin and evaluates 2D
 * edge functions, so its rounding moves an edge by up to about
 * delta = TIE_ULPS * 2^-24 * max_i |v_i - o| in the plane perpendicular to
 * the ray. For edge k that is a barycentric margin delta * |edge k| * |d| / |det|
 * (|det| / |d| is twice the area projected along the ray), and a t margin
 * delta / |d|; both are at least tie_rel (relative). A ray is ambiguous if
 * the hit lies within its margin of an edge, if another triangle's hit lies
 * within the t margins, if a triangle misses only by its margins at a t
 * within the margins of the hit, or if the hit is within the margin of tmin. */
#define F32_EPS 5.9604644775390625e-08 /* 2^-24 */
#define TIE_ULPS 16.0

typedef struct {
    double tb, t2, ub, vb, nm_t;
    double mu, mv, mw, mt;   /* margins of the best hit */
    uint32_t ib;
} F64Acc;

static void f64_init(F64Acc* a, const ReinaRay* ray) {
    a->tb = ray->tmax; a->t2 = INFINITY; a->ub = 0; a->vb = 0; a->nm_t = INFINITY;
    a->mu = a->mv = a->mw = 0.0; a->mt = 0.0;
    a->ib = 0xFFFFFFFFu;
}

static double dlen(d3 a) { return sqrt(ddot(a, a)); }

static void f64_tri(F64Acc* a, d3 o, d3 d, const ReinaRay* ray, const float* v, uint32_t g, double tie_rel) {
    d3 v0 = D3(v[0], v[1], v[2]), v1 = D3(v[3], v[4], v[5]), v2 = D3(v[6], v[7], v[8]);
    d3 e1 = dsub(v1, v0), e2 = dsub(v2, v0);
    d3 p = dcross(d, e2);
    double det = ddot(e1, p);
    if (det == 0.0) return;
    double inv = 1.0 / det;
    d3 sv = dsub(o, v0);
    double uu = ddot(sv, p) * inv;
    d3 q = dcross(sv, e1);
    double vv = ddot(d, q) * inv;
    double tt = ddot(e2, q) * inv;
    double dn = dlen(d);
    double dist = fmax(fmax(dlen(dsub(v0, o)), dlen(dsub(v1, o))), dlen(dsub(v2, o)));
    double delta = TIE_ULPS * F32_EPS * dist;
    double k = delta * dn / fabs(det);
    double mu = fmax(tie_rel, k * dlen(e2)), mv = fmax(tie_rel, k * dlen(e1));
    double mw = fmax(tie_rel, k * dlen(dsub(e2, e1)));
    double mt = fmax(tie_rel * fabs(tt), delta / dn);
    int inside = !(uu < 0.0 || uu > 1.0) && !(vv < 0.0 || uu + vv > 1.0);
    if (inside && tt >= ray->tmin && tt <= ray->tmax) {
        if (tt < a->tb || (tt == a->tb && g < a->ib)) {
            if (a->ib != 0xFFFFFFFFu && a->tb - a->mt < a->t2) a->t2 = a->tb - a->mt;
            a->tb = tt; a->ib = g; a->ub = uu; a->vb = vv;
            a->mu = mu; a->mv = mv; a->mw = mw; a->mt = mt;
        } else if (tt - mt < a->t2) {
            a->t2 = tt - mt;
        }
        return;
    }
    /* near miss: inside the triangle enlarged by its margins? */
    if (uu >= -mu && vv >= -mv && 1.0 - uu - vv >= -mw && tt >= ray->tmin - mt && tt <= ray->tmax + mt)
        if (tt - mt < a->nm_t) a->nm_t = tt - mt;
}

static void f64_finish(const F64Acc* a, const ReinaRay* ray, double tie_rel, double* to, uint32_t* po,
                       double* uo, double* vo, uint8_t* amb) {
    *to = a->tb; *po = a->ib; *uo = a->ub; *vo = a->vb;
    int am = a->nm_t <= a->tb + a->mt;
    if (a->ib != 0xFFFFFFFFu) {
        double wb = 1.0 - a->ub - a->vb;
        if (a->ub < a->mu || a->vb < a->mv || wb < a->mw) am = 1;
        if (a->t2 <= a->tb + a->mt) am = 1;
        if (a->tb <= ray->tmin + a->mt + tie_rel) am = 1;
    }
    (void)tie_rel;
    *amb = (uint8_t)am;
}

static void fp64_one(const OrScene* s, const ReinaRay* ray, double* to, uint32_t* po, double* uo,
                     double* vo, uint8_t* amb, double tie_rel) {
    d3 o = D3(ray->ox, ray->oy, ray->oz), d = D3(ray->dx, ray->dy, ray->dz);
    F64Acc a;
    f64_init(&a, ray);
    for (uint64_t g = 0; g < s->ntri; ++g) f64_tri(&a, o, d, ray, &s->wtri[9 * g], (uint32_t)g, tie_rel);
    f64_finish(&a, ray, tie_rel, to, po, uo, vo, amb);
}

/* The same FP64 query with the FP32 BVH8 (256-byte nodes) as a culling
 * structure only: every child box is padded to cover the margin-enlarged
 * triangles of its subtree and tested in double; boxes entered beyond the
 * current best t plus the margins are skipped.
 * Result identical to fp64_one (the accumulator is order independent). */
static void fp64_bvh_one(const OrScene* s, const ReinaRay* ray, double* to, uint32_t* po, double* uo,
                         double* vo, uint8_t* amb, double tie_rel) {
    d3 o = D3(ray->ox, ray->oy, ray->oz), d = D3(ray->dx, ray->dy, ray->dz);
    F64Acc a;
    f64_init(&a, ray);
    double dir[3] = {d.x, d.y, d.z}, org[3] = {o.x, o.y, o.z};
    double inv[3] = {1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
    double dnorm = dlen(d);
    uint32_t stack[64 * 8];
    int sp = 0;
    if (s->nnodes && s->node_bytes == 256) stack[sp++] = 0;
    while (sp > 0) {
        const uint8_t* nb = s->nodes + (uint64_t)stack[--sp] * 256;
        const float* f = (const float*)nb;
        uint32_t w48, w49;
        memcpy(&w48, nb + 192, 4);
        memcpy(&w49, nb + 196, 4);
        uint32_t child_base = w48 & 0xFFFFFFu, tri_base = w49 & 0xFFFFFFu;
        const uint8_t* meta = nb + 200;
        for (int k = 0; k < 8; ++k) {
            uint8_t mk = meta[k];
            if (!mk) continue;
            double lo[3] = {f[0 + k], f[16 + k], f[32 + k]}, hi[3] = {f[8 + k], f[24 + k], f[40 + k]};
            /* pad: the barycentric tie_rel margin, plus twice the largest
             * FP32 edge shift of a triangle in this box (see F64Acc) */
            double ext = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
            double far = 0.0;
            for (int ax = 0; ax < 3; ++ax) far = fmax(far, fmax(fabs(lo[ax] - org[ax]), fabs(hi[ax] - org[ax])));
            double shift = 2.0 * TIE_ULPS * F32_EPS * 1.7320508075688772 * far;
            double pad = 2.0 * tie_rel * ext + shift + 1e-12;
            double tn = -INFINITY, tf = INFINITY;
            for (int ax = 0; ax < 3; ++ax) {
                double l = lo[ax] - pad, h = hi[ax] + pad;
                if (dir[ax] == 0.0) {   /* runs parallel to the slab: inside or never */
                    if (org[ax] < l || org[ax] > h) tn = INFINITY;
                    continue;
                }
                double t0 = (l - org[ax]) * inv[ax], t1 = (h - org[ax]) * inv[ax];
                tn = fmax(tn, fmin(t0, t1));
                tf = fmin(tf, fmax(t0, t1));
            }
            double tpad = shift / dnorm;
            tn = fmax(tn - 1e-9 * fabs(tn), ray->tmin - tie_rel * fabs(ray->tmin) - tpad);
            tf = fmin(tf + 1e-9 * fabs(tf), a.tb * (1.0 + tie_rel) + tpad + 1e-300);
            if (!(tn <= tf)) continue;
            if (mk & 0x80) {
                if (sp < 64 * 8) stack[sp++] = child_base + (mk & 7u);
            } else {
                uint32_t cntl = (mk >> 5) & 3u, off = mk & 31u;
                for (uint32_t j = 0; j < cntl; ++j) {
                    const float* tv = s->btris + 12ull * (tri_base + off + j);
                    uint32_t id;
                    memcpy(&id, &tv[3], 4);
                    f64_tri(&a, o, d, ray, &s->wtri[9ull * id], id, tie_rel);
                }
            }
        }
    }
    f64_finish(&a, ray, tie_rel, to, po, uo, vo, amb);
}

/* ---- threading ----------------------------------------------------------- */
typedef struct {
    int kind;
    const OrScene* s;
    const ReinaRay* rays;
    ReinaHit* hits;
    uint32_t* counters;
    double *t, *u, *v;
    uint32_t* prim;
    uint8_t* amb;
    double tie;
    uint64_t begin, end;
} IJob;

static void* ijob_run(void* arg) {
    IJob* j = (IJob*)arg;
    for (uint64_t i = j->begin; i < j->end; ++i) {
        if (j->kind == 0) brute32_one(j->s, &j->rays[i], &j->hits[i]);
        else if (j->kind == 1) bvh_one(j->s, &j->rays[i], &j->hits[i], j->counters ? &j->counters[2 * i] : NULL);
        else if (j->kind == 2) fp64_one(j->s, &j->rays[i], &j->t[i], &j->prim[i], &j->u[i], &j->v[i], &j->amb[i], j->tie);
        else fp64_bvh_one(j->s, &j->rays[i], &j->t[i], &j->prim[i], &j->u[i], &j->v[i], &j->amb[i], j->tie);
    }
    return NULL;
}

static void run_jobs(IJob* proto, uint64_t n, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if ((uint64_t)nthreads > n) nthreads = n ? (int)n : 1;
    pthread_t th[256];
    IJob jobs[256];
    if (nthreads > 256) nthreads = 256;
    for (int k = 0; k < nthreads; ++k) {
        jobs[k] = *proto;
        jobs[k].begin = n * (uint64_t)k / (uint64_t)nthreads;
        jobs[k].end = n * (uint64_t)(k + 1) / (uint64_t)nthreads;
        pthread_create(&th[k], NULL, ijob_run, &jobs[k]);
    }
    for (int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
}

void or_intersect_brute32(const OrScene* s, const ReinaRay* rays, ReinaHit* hits, uint64_t n, int nthreads) {
    IJob j; memset(&j, 0, sizeof j);
    j.kind = 0; j.s = s; j.rays = rays; j.hits = hits;
    run_jobs(&j, n, nthreads);
}

void or_intersect_bvh(const OrScene* s, const ReinaRay* rays, ReinaHit* hits, uint32_t* counters,
                      uint64_t n, int nthreads) {
    IJob j; memset(&j, 0, sizeof j);
    j.kind = 1; j.s = s; j.rays = rays; j.hits = hits; j.counters = counters;
    run_jobs(&j, n, nthreads);
}

void or_intersect_fp64(const OrScene* s, const ReinaRay* rays, uint64_t n, double* t_out,
                       uint32_t* prim_out, double* u_out, double* v_out, uint8_t* ambiguous,
                       double tie_rel, int use_bvh, int nthreads) {
    IJob j; memset(&j, 0, sizeof j);
    j.kind = (use_bvh && s->nnodes && s->node_bytes == 256) ? 3 : 2; j.s = s; j.rays = rays; j.t = t_out; j.prim = prim_out; j.u = u_out; j.v = v_out;
    j.amb = ambiguous; j.tie = tie_rel;
    run_jobs(&j, n, nthreads);
}

/* used by or_shade.c */
/* Shadow ray (any hit): the library's k_trace ANY mode restated as a
 * brute-force scan; back faces (world-space winding, flipped for mirroring
 * transforms) of culling instances are passed through. */
int or_occluded(const OrScene* s, const ReinaRay* ray) {
    WRay r;
    wray_init(&r, ray);
    if (!r.valid) return 0;
    for (uint64_t g = 0; g < s->ntri; ++g) {
        const float* v = &s->wtri[9 * g];
        float t, u, w;
        if (!wtest(&r, v, v + 3, v + 6, r.tmax, &t, &u, &w)) continue;
        const uint8_t fl = s->tri_flags[g];
        if (fl & 1) {
            const float ex = v[3] - v[0], ey = v[4] - v[1], ez = v[5] - v[2];
            const float fx = v[6] - v[0], fy = v[7] - v[1], fz = v[8] - v[2];
            float gg = (r.d.x * (ey * fz - ez * fy) + r.d.y * (ez * fx - ex * fz)) + r.d.z * (ex * fy - ey * fx);
            if (fl & 2) gg = -gg;
            if (!(gg < 0.0f)) continue;
        }
        return 1;
    }
    return 0;
}

void or_closest_hit(const OrScene* s, const ReinaRay* ray, ReinaHit* hit, int use_bvh) {
    if (use_bvh && s->nnodes) bvh_one(s, ray, hit, NULL);
    else brute32_one(s, ray, hit);
}
