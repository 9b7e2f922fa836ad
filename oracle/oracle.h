/*
 * oracle.h — ORACLE: CPU restatement of Reina's ray-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (reina-vk_amd/) includes,
 * links or loads anything under oracle/. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, as the checker / CPU baseline.
 *
 * Each function cites the reference file:line it restates. The reference
 * publishes no tests or golden vectors; every function here is pinned by the
 * checks listed in tests/test_oracle_*.py (closed forms, invariants, brute
 * force on small inputs, libm in double precision).
 */
#ifndef REINA_ORACLE_H
#define REINA_ORACLE_H

#include <stdint.h>
#include "../include/reina_rt.h"
#include "or_math.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- flattened scene (oracle-owned copy of the reference buffers) -------- */
typedef struct OrInstance {
    v3 c0, c1, c2, c3;       /* gl_ObjectToWorldEXT columns (mat4x3)          */
    uint32_t props_id;       /* gl_InstanceCustomIndexEXT                     */
    uint32_t material;       /* SBT offset = hit group                        */
    uint32_t model;
    uint32_t tri_base;       /* first flattened triangle of this instance     */
} OrInstance;

typedef struct OrScene {
    /* reference buffers (copied) */
    float* vertices; uint64_t nv;
    uint32_t* indices; uint64_t ni;
    float* tbns; uint64_t ntbn;
    uint32_t* tbn_indices; uint64_t ntbni;
    float* tex_coords; uint64_t ntc;
    uint32_t* tex_indices; uint64_t ntci;
    ReinaInstanceProperties* props; uint32_t nprops;
    ReinaModelRange* models; uint32_t nmodels;
    OrInstance* inst; uint32_t ninst;
    /* textures (RGBA8, 4 bytes per texel, row-major; copied) */
    struct OrTexture { uint32_t w, h; uint8_t* rgba8; } *tex;
    uint32_t ntex;
    /* flattened world-space triangles, instance-major */
    uint64_t ntri;
    float* wtri;             /* 9 floats per triangle: v0, v1, v2              */
    uint32_t* tri_inst;      /* instance index                                */
    uint32_t* tri_prim;      /* gl_PrimitiveID (local to the model)           */
    /* next-event estimation (restatement of the library's light tables) */
    struct OrEmissive { uint32_t inst, ntri, cdf_base; float cdf, pdf_area; } *emi;
    uint32_t nemi;
    float* tri_cdf; uint32_t ntri_cdf;
    int32_t* inst_emi;
    uint8_t* tri_flags;
    int nee;                 /* next-event estimation on in or_render        */
    uint32_t fixes;          /* REINA_RT_TRACE_FIX_* corrections (0: as the reference) */
    /* optional BVH8 (exported from the library) */
    const uint8_t* nodes; uint64_t nnodes;
    uint32_t node_bytes;     /* 256: FP32 node, 80: 8-bit quantized node        */
    const float* btris; uint64_t nbtris;  /* float4 x3 per triangle            */
} OrScene;

OrScene* or_scene_create(const ReinaSceneDesc* desc);
/* texture(textures[id], (u, v)) and bumpMapping of the hit shaders (tests) */
void or_tex_sample_at(const OrScene* s, int32_t id, float u, float v, float* out4);
void or_bump_at(const OrScene* s, int32_t id, float u, float v, const float* dir3, const float* tbn9, float* out2);
void or_scene_destroy(OrScene* s);
int or_scene_set_bvh(OrScene* s, const void* nodes, uint64_t nodes_bytes,
                     const void* tris, uint64_t tris_bytes, uint32_t node_bytes);

/* Light tables in the layout of reina_rt_scene_export_emissive; returns the
 * byte size (buf may be NULL). */
uint64_t or_export_emissive(const OrScene* s, void* buf);
void or_scene_set_nee(OrScene* s, int on);
/* Shadow-ray visibility: 1 if a triangle other than a back face of a
 * culling instance is hit with t in [tmin, tmax]. */
int or_occluded(const OrScene* s, const ReinaRay* ray);

/* ---- intersection -------------------------------------------------------- */
void or_intersect_brute32(const OrScene* s, const ReinaRay* rays, ReinaHit* hits, uint64_t n,
                          int nthreads);
void or_intersect_bvh(const OrScene* s, const ReinaRay* rays, ReinaHit* hits,
                      uint32_t* counters, uint64_t n, int nthreads);
void or_intersect_fp64(const OrScene* s, const ReinaRay* rays, uint64_t n, double* t_out,
                       uint32_t* prim_out, double* u_out, double* v_out, uint8_t* ambiguous,
                       double tie_rel, int use_bvh, int nthreads);

/* ---- integrator ---------------------------------------------------------- */
float or_random(uint32_t* state);
void or_pcg_stream(uint32_t seed, uint32_t n, uint32_t* states, float* values);
void or_offset_position(const float* p, const float* n, float* out, int dielectric, const float* dir);
void or_sky(const float* dir, float* out);
void or_camera_ray(const RtPushConsts* pc, uint32_t px, uint32_t py, uint32_t W, uint32_t H,
                   uint32_t* rng, float* origin, float* dir, int lens_advances);
void or_scene_set_fixes(OrScene* s, uint32_t fixes);
void or_diffuse_dir(const float* n, uint32_t* rng, float* out);

/* Full raygen (raytrace.rgen.glsl:250-285) for the pixel rectangle
 * [x0,x1) x [y0,y1) of a W x H image; img is float4[W*H] (in/out). */
void or_render(const OrScene* s, const RtPushConsts* pc, uint32_t W, uint32_t H, float* img,
               uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, int use_bvh, int nthreads,
               uint64_t* rays_out);

/* Converged reference: nbatches dispatches with sampleBatch = first..first+n-1,
 * each pixel's per-dispatch mean accumulated in double (img64: double3[W*H]
 * holding the sum of means and cnt the number of non-NaN dispatches). */
void or_render_sum64(const OrScene* s, const RtPushConsts* pc, uint32_t W, uint32_t H,
                     uint32_t first_batch, uint32_t nbatches, double* sum64, double* sumsq64,
                     uint32_t* cnt, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
                     int use_bvh, int nthreads);

/* ---- post-process -------------------------------------------------------- */
void or_postprocess(const float* in, uint32_t W, uint32_t H, const BloomPushConsts* bloom,
                    const TonemappingPushConsts* tm, float* ping, float* pong, uint8_t* out);

/* ---- math exports (tests) ------------------------------------------------ */
void or_math_eval(int fn, const float* in, const float* in2, float* out, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
