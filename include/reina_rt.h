/*
 * reina_rt.h — C ABI of the MI355X path-tracing library (libreina_rt.so).
 *
 * The library replaces Reina's Vulkan ray-tracing pipeline: the driver-built
 * BLAS/TLAS (src/graphics/Blas.cpp:8-124, src/tools/vktools.cpp:460-596), one
 * vkCmdTraceRaysKHR over the raygen / closest-hit / miss shaders
 * (src/Reina.cpp:425-470, shaders/raytrace/ *.glsl) and, as the first
 * downstream consumer, the bloom + tone-map compute passes
 * (src/Reina.cpp:472-577).
 *
 * Conventions
 *   - Every function returns an int status: 0 = REINA_RT_OK, negative = error.
 *     reina_rt_last_error() returns a thread-local message for the last error.
 *     (The reference throws std::runtime_error everywhere, src/main.cpp:11-14.)
 *   - Host pointers in descriptors are copied during the call; the caller may
 *     free them afterwards.
 *   - Image and ray/hit pointers passed to trace/intersect/postprocess are
 *     DEVICE pointers (hipMalloc'd or torch tensors' data_ptr()).
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream). All
 *     device work is enqueued on it; functions documented as "synchronous"
 *     block until that work completes.
 *   - A scene handle is not thread-safe; distinct handles may be used from
 *     distinct threads.
 */
#ifndef REINA_RT_H
#define REINA_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REINA_RT_ABI_VERSION 1u

/* ---- status codes ------------------------------------------------------- */
#define REINA_RT_OK                0
#define REINA_RT_ERR_INVALID_ARG  -1  /* null pointer, zero size, bad index   */
#define REINA_RT_ERR_ABI          -2  /* struct_size / ABI version mismatch   */
#define REINA_RT_ERR_OUT_OF_MEMORY -3 /* device allocation failed             */
#define REINA_RT_ERR_OVER_BUDGET  -4  /* request exceeds the HBM budget       */
#define REINA_RT_ERR_DEVICE       -5  /* HIP runtime / kernel launch error    */
#define REINA_RT_ERR_UNSUPPORTED  -6  /* feature not built (e.g. textures)    */
#define REINA_RT_ERR_NO_EMISSIVE  -7  /* NEE requested on a scene w/o lights  */
#define REINA_RT_ERR_BUSY         -8  /* server queue full (admission reject) */

/* ---- byte-identical copies of polyglot/raytrace.h ----------------------- */

/* polyglot/raytrace.h:18-41 (120 bytes, GLSL scalar layout).
 * interpNormals is a C++ bool followed by 3 padding bytes in the reference;
 * the device reads only that byte (padding is ignored). */
typedef struct ReinaInstanceProperties {
    uint32_t indicesOffset;        /*   0 */
    float    albedo[3];            /*   4 */
    float    emission[3];          /*  16 */
    uint32_t tbnsIndicesOffset;    /*  28 */
    uint32_t texIndicesOffset;     /*  32 (0xFFFFFFFF: model has no UVs) */
    float    roughness;            /*  36 */
    float    ior;                  /*  40 */
    uint8_t  interpNormals;        /*  44 (C++ bool) */
    uint8_t  _pad0[3];             /*  45 */
    float    absorption;           /*  48 */
    int32_t  textureID;            /*  52 (-1: none) */
    int32_t  normalMapTexID;       /*  56 */
    int32_t  bumpMapTexID;         /*  60 */
    uint32_t cullBackface;         /*  64 */
    float    anisotropic;          /*  68 */
    float    subsurface;           /*  72 */
    float    clearcoatGloss;       /*  76 */
    float    sheenTint[3];         /*  80 */
    float    specularTint[3];      /*  92 */
    float    metallic;             /* 104 */
    float    clearcoat;            /* 108 */
    float    specularTransmission; /* 112 */
    float    sheen;                /* 116 */
} ReinaInstanceProperties;

/* polyglot/raytrace.h:43-54 (160 bytes). Matrices are glm::mat4, column-major. */
typedef struct RtPushConsts {
    float    invView[16];          /*   0 */
    float    invProjection[16];    /*  64 */
    uint32_t sampleBatch;          /* 128 */
    float    totalEmissiveWeight;  /* 132 */
    float    focusDist;            /* 136 */
    float    defocusMultiplier;    /* 140 (already divided by 100, Reina.cpp:150) */
    float    directClamp;          /* 144 */
    float    indirectClamp;        /* 148 (unused by the active path) */
    uint32_t samplesPerPixel;      /* 152 */
    uint32_t maxBounces;           /* 156 */
} RtPushConsts;

/* src/scene/Instances.h:15-27 (alignas(16), 112 bytes): emissive metadata for
 * next-event estimation. Produced on the device from the scene description;
 * exposed for inspection through reina_rt_scene_emissive(). */
typedef struct ReinaInstanceData {
    float    transform[16];        /*   0 */
    uint32_t materialOffset;       /*  64 */
    uint32_t cdfRangeStart;        /*  68 */
    uint32_t cdfRangeEnd;          /*  72 */
    uint32_t indexOffset;          /*  76 */
    float    emission[3];          /*  80 */
    float    weight;               /*  92 */
    float    area;                 /*  96 */
    uint8_t  cullBackface;         /* 100 (C++ bool) */
    uint8_t  _pad0[3];             /* 101 */
    float    padding[2];           /* 104 */
} ReinaInstanceData;

/* polyglot/bloom.h:4-8 and polyglot/tonemapping.h:14-16 */
typedef struct BloomPushConsts { float radius, threshold, intensity; } BloomPushConsts;
typedef struct TonemappingPushConsts { float exposure; } TonemappingPushConsts;

/* ---- scene description (replaces Scene::build, src/scene/Scene.cpp:56-125) */

/* src/scene/Models.h:12-21 */
typedef struct ReinaModelRange {
    uint32_t firstVertex;
    uint32_t firstNormal;
    uint32_t indexOffset;
    uint32_t tbnsIndexOffset;
    uint32_t texIndexOffset;
    uint32_t indexCount;      /* triangles */
    uint32_t tbnsIndexCount;  /* triangles */
    uint32_t texIndexCount;   /* triangles */
} ReinaModelRange;

/* One TLAS instance (src/tools/vktools.cpp:460-488, Scene.cpp:99-111):
 * object-to-world transform (glm::mat4, column-major; the bottom row is
 * ignored exactly as VkTransformMatrixKHR drops it), the custom index
 * (gl_InstanceCustomIndexEXT) and the hit group (SBT record offset). */
typedef struct ReinaInstanceDesc {
    float    transform[16];
    uint32_t instancePropertiesID;
    uint32_t materialIdx;      /* 0 lambertian, 1 metal, 2 dielectric, 3 disney (Reina.cpp:159-167) */
    uint32_t modelID;
    uint32_t _pad0;
} ReinaInstanceDesc;

/* BVH / build options */
#define REINA_RT_BUILD_DEVICE      0x1u   /* binned-SAH build on the GPU: the default (flags 0 too) */
#define REINA_RT_BUILD_HOST        0x4u   /* binned-SAH build on the host CPU (threads) */
#define REINA_RT_NODES_QUANTIZED   0x2u   /* 8-bit child boxes (80-byte nodes)      */

/* The packed buffers exactly as Models::addModel produces them
 * (src/scene/Models.cpp:24-92) plus the instance and properties lists. */
/* One texture of the hit shaders' bindless `sampler2D textures[]`
 * (src/graphics/Image.cpp:10-86: RGBA8 UNORM, one mip level): texel (x, y)
 * at rgba8[4 * (y * width + x)], y = 0 the first row (texture v = 0). */
typedef struct ReinaTextureDesc {
    uint32_t width, height;
    const uint8_t* rgba8;
} ReinaTextureDesc;

typedef struct ReinaSceneDesc {
    uint32_t struct_size;                    /* sizeof(ReinaSceneDesc)            */
    uint32_t flags;                          /* REINA_RT_BUILD_* | REINA_RT_NODES_* */
    const float*    vertices;     uint64_t num_vertices;     /* vec4 each (binding 2)        */
    const uint32_t* indices;      uint64_t num_indices;      /* offset indices (binding 3)   */
    const float*    tbns;         uint64_t num_tbns;         /* mat3 col-major (binding 5)   */
    const uint32_t* tbn_indices;  uint64_t num_tbn_indices;  /* offset (binding 6)           */
    const float*    tex_coords;   uint64_t num_tex_coords;   /* vec2 each (binding 10)       */
    const uint32_t* tex_indices;  uint64_t num_tex_indices;  /* offset (binding 11)          */
    const ReinaInstanceProperties* props;     uint32_t num_props;     /* binding 4 */
    const ReinaModelRange*         models;    uint32_t num_models;
    const ReinaInstanceDesc*       instances; uint32_t num_instances;
    uint32_t device;                         /* HIP device ordinal               */
    /* InstanceProperties.textureID / normalMapTexID / bumpMapTexID index this
     * array (-1: none); copied at scene creation */
    const ReinaTextureDesc*        textures;  uint32_t num_textures;
    uint32_t _pad0;
} ReinaSceneDesc;

typedef struct ReinaScene ReinaScene;

typedef struct ReinaSceneInfo {
    uint64_t num_triangles;   /* flattened world-space triangles          */
    uint64_t num_nodes;       /* BVH8 nodes                               */
    uint64_t node_bytes;      /* bytes per node as stored (256 or 80)     */
    uint64_t device_bytes;    /* total HBM held by the scene              */
    uint32_t max_depth;       /* BVH8 depth (root = 1)                    */
    uint32_t quantized;       /* 1 if 8-bit child boxes                   */
    float    build_ms;        /* BVH build wall time                      */
    float    sah_cost;        /* SAH cost of the BVH8 (ci=1, ct=1)        */
    uint32_t num_emissive;    /* instances sampled by next-event estimation */
    uint32_t _pad0;
} ReinaSceneInfo;

/* ---- rays and hits for the closest-hit query ---------------------------- */

/* One ray: traceRayEXT(tlas, gl_RayFlagsOpaqueEXT, 0xFF, 0, 0, 0, o, tmin, d, tmax)
 * (shaders/raytrace/raytrace.rgen.glsl:110-122). 32 bytes. */
typedef struct ReinaRay {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} ReinaRay;

/* Closest hit: t, flattened triangle id (instance-major order: all triangles
 * of instance 0, then instance 1, ...), barycentrics (u, v) of vertices 1 and
 * 2 as in gl hitAttributeEXT. prim = 0xFFFFFFFF and t = tmax on a miss. */
typedef struct ReinaHit {
    float    t;
    uint32_t prim;
    float    u, v;
} ReinaHit;

/* ---- API ---------------------------------------------------------------- */

/* ABI version and the struct sizes the library was built with
 * (120 / 160 / 112 bytes for the three polyglot structs). */
int reina_rt_abi_version(uint32_t* version, uint32_t* sizeof_instance_properties,
                         uint32_t* sizeof_push_consts, uint32_t* sizeof_instance_data);

/* Thread-local message describing the last failed call ("" if none). */
const char* reina_rt_last_error(void);

/* Replaces Scene::build's BLAS + TLAS construction (src/scene/Scene.cpp:93-116,
 * src/graphics/Blas.cpp:8-124, src/tools/vktools.cpp:460-596): flattens every
 * instance to world-space triangles and builds a BVH8 in HBM. Synchronous. */
int reina_rt_scene_create(const ReinaSceneDesc* desc, ReinaScene** out_scene);
int reina_rt_scene_destroy(ReinaScene* scene);
int reina_rt_scene_info(const ReinaScene* scene, ReinaSceneInfo* out_info);

/* Copies the BVH8 and triangle arrays back to host memory (for the traversal
 * oracle and for inspection). Pass NULL buffers to query the sizes first. */
int reina_rt_scene_export_bvh(const ReinaScene* scene, void* nodes, uint64_t* nodes_bytes,
                              void* tris, uint64_t* tris_bytes);

/* Copies the light-sampling tables back to host memory (for the oracle):
 * uint32 {num_emissive, num_tri_cdf, num_instances, num_triangles}, then
 * num_emissive 32-byte records {uint32 instance, ntri, cdf_base; float cdf,
 * pdf_area; pad}, float tri_cdf[num_tri_cdf], int32 inst_emissive[num_instances]
 * (-1: not sampled), uint8 tri_flags[num_triangles] (bit 0: back faces
 * culled, bit 1: mirroring transform). Pass buf = NULL to query the size. */
int reina_rt_scene_export_emissive(const ReinaScene* scene, void* buf, uint64_t* bytes);

/* Host-only BVH8 build of a world-space triangle soup (9 floats per triangle),
 * the same builder reina_rt_scene_create uses; needs no GPU. Pass NULL
 * buffers to query the sizes first. Layout of the outputs: see
 * reina_rt_scene_export_bvh. */
int reina_rt_bvh_build_host(const float* triangles, uint64_t num_triangles, void* nodes,
                            uint64_t* nodes_bytes, void* tris, uint64_t* tris_bytes,
                            uint32_t* max_depth);

/* 8-bit quantized copy (80-byte nodes) of a BVH8 from reina_rt_bvh_build_host
 * or reina_rt_scene_export_bvh: the node layout scenes created with
 * REINA_RT_NODES_QUANTIZED hold. Every decoded child box contains the FP32
 * child box. Host only; pass qnodes = NULL to query the size first. */
int reina_rt_bvh_quantize_host(const void* nodes, uint64_t nodes_bytes, void* qnodes,
                               uint64_t* qnodes_bytes);

/* Replaces one vkCmdTraceRaysKHR(W, H, 1) of the raygen/closest-hit/miss
 * pipeline (src/Reina.cpp:425-470): renders samplesPerPixel samples per pixel
 * and writes the (running-mean) radiance into image_rgba32f (device pointer,
 * float4[W*H], row-major, y = 0 at the top, w = 1). When pc->sampleBatch > 0
 * the previous image content is blended as (prev*b + new)/(b+1). The caller
 * increments sampleBatch after the call, as Reina::traceRays does. Enqueued
 * on `stream`; returns after enqueueing all bounce iterations. */
int reina_rt_trace(ReinaScene* scene, const RtPushConsts* pc, float* image_rgba32f,
                   uint32_t width, uint32_t height, void* stream);

/* Statistics of the last reina_rt_trace on this scene (synchronises it). */
typedef struct ReinaTraceStats {
    uint64_t extension_rays;  /* every traced extension ray (primary + secondary) */
    uint64_t shadow_rays;     /* NEE shadow/connect rays (0 unless NEE enabled)   */
    uint64_t node_visits;     /* BVH8 nodes visited (REINA_RT_TRACE_COUNT only)    */
    uint64_t tri_tests;       /* triangles tested   (REINA_RT_TRACE_COUNT only)    */
    uint32_t iterations;      /* wavefront bounce iterations launched             */
    uint32_t extend_launches; /* extend kernel launches                           */
    float    trace_ms;        /* GPU time of the whole trace (events)             */
    float    extend_ms;       /* summed extend-kernel time (REINA_RT_TRACE_TIME_EXTEND) */
    uint64_t tail_rays;       /* extension rays traced by the tail kernel (in extension_rays) */
    float    tail_ms;         /* GPU time of the tail kernel (0 if not used)      */
    uint32_t tail_launches;   /* 1 if the frame ended in the tail kernel          */
} ReinaTraceStats;

/* Trace options (sticky per scene): record a HIP event pair around every
 * extend launch on the trace's stream, use the counting build of the
 * extend kernel (per-ray node visits / triangle tests summed on the device),
 * and/or next-event estimation. */
#define REINA_RT_TRACE_TIME_EXTEND 0x1u
#define REINA_RT_TRACE_COUNT       0x2u
/* Next-event estimation with multiple importance sampling at Lambertian and
 * Disney hits (one light sample and shadow ray per bounce, power heuristic),
 * as the reference's directLight intends for materialIDs 0 and 3
 * (shaders/raytrace/raytrace.rgen.glsl:43-95). The reference compiles this
 * but never runs it (skipNEE = true, :146), so it is off by default; the
 * default renders match the reference's active path. A Disney light sample
 * carries the expected weight of the hit shader's own continuation, so NEE
 * changes the variance and not the expected image. Needs an emissive
 * instance (REINA_RT_ERR_NO_EMISSIVE otherwise). */
#define REINA_RT_TRACE_NEE         0x4u
/* Corrections of reference behaviours that the default renders reproduce
 * (no counterpart in the reference: each flag replaces one behaviour of its
 * shaders by the textbook one, so frames then differ from the reference's).
 *   FIX_LENS_RNG      the thin-lens sample advances the pixel's RNG stream;
 *                     randomInUnitHexagon takes the state by value
 *                     (shaders/raytrace/raytrace.rgen.glsl:194, 234), so the
 *                     reference reuses the lens draws for the first bounce.
 *   FIX_MEDIUM_RESET  accumulatedDistance is cleared when a sample starts;
 *                     the reference clears only insideDielectric (:98), so
 *                     distance travelled in glass by a path that ends there
 *                     is charged to the next sample's first exit.
 *   FIX_NORMALS       world normals by the inverse-transpose of the
 *                     object-to-world matrix; the reference uses mat3(O2W)
 *                     (closestHitCommon.h.glsl:113-114), which tilts normals
 *                     under non-uniform scale.
 *   FIX_METAL_PDF     the GGX reflection pdf is the density of the sampler in
 *                     use (Heitz's VNDF, brdfDisney.h.glsl:9-40); the
 *                     reference prices it with the bounded-VNDF pdf of Eto and
 *                     Tokuyoshi (:55-73), larger by (i.z + t) / (k i.z + t).
 *   FIX_DISNEY_MIX    evalDisney's mixture as sampleDisney draws it: the
 *                     clearcoat weight inside the normalising sum (the
 *                     reference leaves it out, :680-709), the metal lobe's
 *                     dielectric F0 from the IOR alone (the reference passes
 *                     specularTransmission as `specular`, :692, so an opaque
 *                     dielectric has F0 = 0), and the glass pdf times the
 *                     probability of the branch taken (F or 1 - F, :571-580).
 *                     With FIX_METAL_PDF the pdf is then the sampler's density. */
#define REINA_RT_TRACE_FIX_LENS_RNG     0x8u
#define REINA_RT_TRACE_FIX_MEDIUM_RESET 0x10u
#define REINA_RT_TRACE_FIX_NORMALS      0x20u
#define REINA_RT_TRACE_FIX_METAL_PDF    0x40u
#define REINA_RT_TRACE_FIX_DISNEY_MIX   0x80u
#define REINA_RT_TRACE_FIXES            0xF8u
/* Independent sample streams (no counterpart in the reference, where one RNG
 * stream per pixel runs all of a pixel's samples in sequence,
 * shaders/raytrace/raytrace.rgen.glsl:259-277). Bits 8-11 hold log2(G),
 * G in 1..16: a pixel's samples split into G groups of spp / G (spp must be
 * a multiple of G) that are traced in parallel, group g seeded as the
 * reference seeds sample batch sampleBatch * G + g:
 * ((sampleBatch * G + g) * H + y) * W + x. The pixel's mean is the sum of
 * the groups' sample sums (added in group order) over the sum of their
 * counts. G = 1 is the reference. Server requests take G from their scene. */
#define REINA_RT_TRACE_SAMPLE_GROUPS(log2g) ((uint32_t)(log2g) << 8)
#define REINA_RT_TRACE_SAMPLE_GROUPS_MASK   0xF00u
/* reina_rt_intersect on FP32 nodes only (REINA_RT_ERR_UNSUPPORTED on 8-bit
 * nodes; reina_rt_trace ignores it): ray-direction inverses rounded to FP16,
 * slab intervals widened by 2^-10 of |t| so the box test stays conservative.
 * Hits are identical to the FP32 inverses'; only node visits differ. A
 * precision experiment (no counterpart in the reference). */
#define REINA_RT_TRACE_F16_INV              0x1000u
int reina_rt_set_trace_options(ReinaScene* scene, uint32_t options);
int reina_rt_trace_stats(ReinaScene* scene, ReinaTraceStats* out);

/* Launch parameters of the traversal / shading kernels. Defaults are set at
 * scene creation (stack depth from the BVH depth, refill 32, 8 traversal
 * blocks per CU at stack depth 8 and 4 above, 16 shade blocks per CU, the
 * compiler's register budget, tail launch at 2^20 live paths, claims of 128
 * rays); the autotuner in the Python package searches them per scene. */
typedef struct ReinaTuning {
    uint32_t stack_depth;   /* LDS stack entries per lane: 8, 16, 24 or 32 (>= BVH depth - 1) */
    uint32_t refill_min;    /* idle lanes per wave that trigger a refill, 1..64                */
    uint32_t grid_blocks;   /* persistent traversal blocks (256 threads)                       */
    uint32_t shade_blocks;  /* shade blocks (256 threads), a multiple of 8                      */
    uint32_t waves_per_simd;/* traversal register budget: 0 (compiler's choice), 4, 5, 6 or 8      */
    uint32_t tail_paths;    /* when at most this many paths are live, the frame ends in one
                               tail launch (per-wave extend/shade steps); 0 = never            */
    uint32_t claim_rays;    /* rays a traversal wave claims from its queue partition with one
                               atomic: 64, 128, 256 or 512 (0 = 128)                           */
} ReinaTuning;
int reina_rt_get_tuning(const ReinaScene* scene, ReinaTuning* out);
int reina_rt_set_tuning(ReinaScene* scene, const ReinaTuning* tuning);

/* Closest-hit query for a batch of rays (device pointers): the traceRayEXT
 * of shaders/raytrace/raytrace.rgen.glsl:110-122 as a stand-alone kernel.
 * If counters != NULL (device uint32[2*n]) the per-ray node visits and
 * triangle tests are written there (counting build of the same kernel). */
int reina_rt_intersect(ReinaScene* scene, const ReinaRay* rays, ReinaHit* hits,
                       uint32_t n, uint32_t* counters, void* stream);

/* reina_rt_intersect without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as code
 * under HIP graph capture and tight serving loops. The traversal kernel's
 * claim counters and used-up flag (1152 bytes) live in a static device
 * array of the library, one per GPU, held until the process ends, and are zeroed by a fill kernel on
 * `stream` at every call. At most 2^30 rays per call (REINA_RT_ERR_INVALID_ARG
 * beyond). Calls that overlap on one GPU (different streams or threads) race
 * on the shared counters; serialise them. Hits equal reina_rt_intersect's. */
int reina_rt_intersect_noalloc(ReinaScene* scene, const ReinaRay* rays, ReinaHit* hits, uint32_t n, void* stream);

/* Bloom (blurX with threshold, blurY, combine) then ACES-fitted tone map to
 * RGBA8 UNORM: shaders/postprocessing/bloom/{blurX,blurY,combine}.comp.glsl, tonemap/tonemapping.comp.glsl,
 * dispatch order src/Reina.cpp:472-577. in_rgba32f, scratch (2 * W*H float4)
 * and out_rgba8 (W*H*4 bytes) are device pointers. */
int reina_rt_postprocess(const float* in_rgba32f, float* scratch_rgba32f, uint8_t* out_rgba8,
                         uint32_t width, uint32_t height, const BloomPushConsts* bloom,
                         const TonemappingPushConsts* tonemap, void* stream);

/* ---- batching and serving: many render requests in one wavefront --------
 *
 * No counterpart call in the reference, which renders one scene per
 * vkCmdTraceRaysKHR (src/Reina.cpp:425-470); each request here is exactly
 * such a call (scene, RtPushConsts, W x H image, same semantics as
 * reina_rt_trace), and a request's image is bit-identical to its
 * reina_rt_trace image whatever else runs beside it.
 *
 * A server owns a worker thread, a stream, and a pool of path-state slots
 * sized by path_budget_bytes (REINA_RT_SERVER_BYTES_PER_PIXEL per pixel in
 * flight). Submitted requests wait in a bounded queue; the worker admits them
 * (in admit_order) as soon as enough contiguous slots and a request index are
 * free (and, with admit_paths, while fewer paths than that are in flight:
 * under overload this keeps fewer requests sharing the GPU, so each
 * finishes sooner), generates their camera rays into the shared queues, and runs bounce
 * iterations over every admitted request at once (captured in a HIP graph of
 * graph_iters iterations). When at most tuning.tail_paths paths are live after a graph,
 * one tail launch (per-wave extend/shade steps, as reina_rt_trace's) finishes
 * every running request; it runs beside the loop, which keeps admitting and
 * running new requests (REINA_RT_SERVER_TAIL_INLINE: requests arriving
 * meanwhile wait for it). A request
 * completes when all its pixels have finished all
 * samples; its image is then written. Scenes must stay alive until their
 * requests complete. submit / poll / wait may be called from any thread. */
#define REINA_RT_SERVER_BYTES_PER_PIXEL 152u
/* with REINA_RT_SERVER_NEE: + 16 B of light-sampling state and a 48 B shadow
 * ray per pixel in flight */
#define REINA_RT_SERVER_BYTES_PER_PIXEL_NEE 216u
/* ReinaServerDesc.flags: hold next-event-estimation state, so requests on
 * scenes whose trace options include REINA_RT_TRACE_NEE render with NEE + MIS
 * (as reina_rt_trace does on that scene); without it such requests are
 * refused with REINA_RT_ERR_UNSUPPORTED. */
#define REINA_RT_SERVER_NEE 0x1u
/* ReinaServerDesc.flags: run the tail launch on the worker's stream as one
 * step of the request loop, so requests arriving meanwhile wait for it.
 * Without it the tail launch runs on a second stream with a grid of
 * tail_cu_percent of the GPU's resident blocks, while the loop keeps admitting requests and
 * running graphs for them on the shared queues; the server then holds a tail
 * queue of 288 * tuning.tail_paths bytes. */
#define REINA_RT_SERVER_TAIL_INLINE 0x2u

typedef struct ReinaServer ReinaServer;

typedef struct ReinaServerDesc {
    uint32_t struct_size;        /* sizeof(ReinaServerDesc)                              */
    uint32_t device;             /* HIP device ordinal (scenes must live on it)          */
    uint64_t path_budget_bytes;  /* HBM for in-flight pixels; at least 8 slot tiles      */
    uint32_t max_pending;        /* bounded FIFO; submit beyond it -> REINA_RT_ERR_BUSY  */
    uint32_t quantized;          /* node format of every scene served: 0 FP32, 1 8-bit   */
    ReinaTuning tuning;          /* shared-kernel launch parameters; zeros = defaults
                                    (stack_depth bounds the BVH depth of served scenes;
                                    tail_paths 0 = 2^20, refill_min 0 = 32,
                                    REINA_RT_TAIL_OFF = never)                           */
    uint32_t flags;              /* REINA_RT_SERVER_NEE or 0                             */
    uint32_t admit_paths;        /* admission stops while this many paths are in flight
                                    in the shared queues (live paths of the last graph +
                                    paths admitted since; a tail launch's are not counted);
                                    a request is always admitted when nothing runs there.
                                    0 or REINA_RT_ADMIT_ALL: admit whenever slots are free */
    uint32_t slot_tile;          /* pixels per slot tile (the unit of path-slot allocation
                                    and of the queue-partition interleave): 256, 512, 1024,
                                    2048 or 4096; 0 = 1024                                */
    uint32_t graph_iters;        /* bounce iterations per graph replay, between two
                                    admission / completion points: 6, 12, 18 or 24; 0 = 6 */
    uint32_t admit_order;        /* REINA_RT_ADMIT_FIFO, _SHORTEST or _LONGEST              */
    uint32_t max_wait_ms;        /* SHORTEST / LONGEST: a request waiting this long goes
                                    before the others (oldest first); 0 = 500             */
    uint32_t max_requests;       /* submit -> REINA_RT_ERR_BUSY while this many requests are
                                    queued or running (a capacity bound like a plain
                                    renderer's 1 running + max_pending queued); 0 = none  */
    uint32_t tail_cu_percent;    /* share of the GPU the concurrent tail launch takes, in
                                    percent of the tail kernel's resident blocks, 1..100;
                                    0 = 50 when requests are waiting at its launch, else
                                    the whole GPU                                          */
} ReinaServerDesc;

/* ReinaServerDesc.admit_paths: no limit but the path pool. */
#define REINA_RT_ADMIT_ALL 0xFFFFFFFFu

/* ReinaServerDesc.admit_order. FIFO: submission order. SHORTEST: least
 * estimated work first (path slots x samples per pixel x max bounces).
 * LONGEST: longest per-pixel sample chain first (samples per pixel x max
 * bounces, then more path slots), which shortens the makespan of a batch,
 * since a pixel's samples run in sequence. With SHORTEST and LONGEST,
 * requests waiting max_wait_ms or longer go first, oldest first, and
 * admission stops at the first of those that does not fit: no request waits
 * more than max_wait_ms plus the time for the requests ahead of it to free
 * their slots. */
#define REINA_RT_ADMIT_FIFO     0u
#define REINA_RT_ADMIT_SHORTEST 1u
#define REINA_RT_ADMIT_LONGEST  2u

/* ReinaServerDesc.tuning.tail_paths: never end with a tail launch. */
#define REINA_RT_TAIL_OFF 0xFFFFFFFFu

#define REINA_RT_REQ_QUEUED    0u
#define REINA_RT_REQ_RUNNING   1u
#define REINA_RT_REQ_DONE      2u
#define REINA_RT_REQ_FAILED    3u   /* a device error stopped the server (reina_rt_last_error) */

typedef struct ReinaRequestStatus {
    uint32_t state;          /* REINA_RT_REQ_*                                   */
    uint32_t npx;            /* W * H                                            */
    double   submit_s;       /* seconds since server creation                    */
    double   admit_s;        /* rays generated (0 while queued)                  */
    double   done_s;         /* image written (0 until done)                     */
} ReinaRequestStatus;

typedef struct ReinaServerStats {
    uint64_t submitted, completed, rejected_busy, rejected_budget;
    uint64_t extension_rays;     /* traced by the shared kernels                   */
    uint64_t iterations;         /* bounce iterations run                          */
    uint64_t graph_replays;      /* captured graphs (graph_iters iterations) launched */
    uint32_t pending, running;   /* now                                            */
    uint32_t slot_tiles_total, slot_tiles_used;   /* tiles of slot_tile pixels     */
    double   busy_s;             /* worker time with admitted requests             */
    uint64_t tail_launches;      /* tail launches that finished every running request */
    uint64_t tail_rays;          /* extension rays traced by those launches        */
} ReinaServerStats;

int reina_rt_server_create(const ReinaServerDesc* desc, ReinaServer** out_server);
/* Finishes every admitted and queued request, then stops the worker. */
int reina_rt_server_destroy(ReinaServer* server);
/* Enqueues one render request (image: device float4[W*H]; its previous
 * content is blended when pc->sampleBatch > 0). Returns immediately:
 * REINA_RT_ERR_BUSY when the FIFO is full, REINA_RT_ERR_OVER_BUDGET when the
 * request alone needs more slots than the server has, REINA_RT_ERR_UNSUPPORTED
 * when the scene's node format or BVH depth does not fit the server or the
 * scene's trace options ask for NEE on a server created without
 * REINA_RT_SERVER_NEE, REINA_RT_ERR_NO_EMISSIVE when they ask for NEE and the
 * scene has no emissive instance. */
int reina_rt_submit(ReinaServer* server, ReinaScene* scene, const RtPushConsts* pc, float* image_rgba32f,
                    uint32_t width, uint32_t height, uint64_t* ticket);
/* Status of a request; the record is released by the first call that
 * reports REINA_RT_REQ_DONE (later calls return REINA_RT_ERR_INVALID_ARG). */
int reina_rt_poll(ReinaServer* server, uint64_t ticket, ReinaRequestStatus* out);
/* Blocks until the request is done (or timeout_ms passes: REINA_RT_ERR_BUSY);
 * fills *out like reina_rt_poll and releases the record when done. */
int reina_rt_wait(ReinaServer* server, uint64_t ticket, uint32_t timeout_ms, ReinaRequestStatus* out);
int reina_rt_server_stats(ReinaServer* server, ReinaServerStats* out);

/* ---- scene ingest (host only: no device, no stream) ---------------------- */

/* One model's buffers as src/scene/Models.h:23-30 (ModelData) holds them and
 * Models::addModel (src/scene/Models.cpp:24-92) consumes them: the arrays of
 * one ReinaSceneDesc model before the per-model offsets are added. The index
 * arrays are model-local. num_indices need not be a multiple of 3 (line and
 * point records of an OBJ file add 2- and 1-index faces, as in the
 * reference); a model has num_indices / 3 triangles. */
typedef struct ReinaModelData {
    float*    vertices;     uint64_t num_vertices;     /* x, y, z, 1 per vertex            */
    uint32_t* indices;      uint64_t num_indices;
    float*    tbns;         uint64_t num_tbns;         /* 9 floats each: columns T, B, N  */
    uint32_t* tbn_indices;  uint64_t num_tbn_indices;
    float*    tex_coords;   uint64_t num_tex_coords;   /* u, v per coordinate (0 = none) */
    uint32_t* tex_indices;  uint64_t num_tex_indices;
} ReinaModelData;

/* Models::getObjData (src/scene/Models.cpp:117-175): the first mesh of a
 * Wavefront OBJ file as Assimp imports it with aiProcess_Triangulate |
 * GenSmoothNormals | JoinIdenticalVertices | CalcTangentSpace | FlipUVs.
 * `text` holds the file's bytes (need not be NUL-terminated). The result is
 * released with reina_rt_model_free. REINA_RT_ERR_INVALID_ARG on a malformed
 * file or one without faces (the reference throws there). */
int reina_rt_model_load_obj(const char* path, ReinaModelData** out);
int reina_rt_model_load_obj_memory(const char* text, uint64_t len, ReinaModelData** out);
void reina_rt_model_free(ReinaModelData* model);

/* An RGBA8 image as stb_image returns it with 4 channels requested
 * (src/graphics/Image.cpp:10-41). From a file, rows are flipped vertically
 * (stbi_set_flip_vertically_on_load(true), Image.cpp:14: row 0 = the file's
 * last row); from memory (glTF-embedded images, Image.cpp:24-41) they are
 * not. PNG (all bit depths, colour types, tRNS and Adam7 interlacing) and
 * baseline and progressive JPEG (grey, YCbCr, RGB, CMYK or YCCK, any
 * subsampling, restart intervals, stb's integer IDCT, upsampling and colour
 * conversion); 12-bit, lossless and arithmetic-coded JPEGs (which stb refuses
 * too) and other formats return REINA_RT_ERR_UNSUPPORTED. Released with
 * reina_rt_image_free. */
typedef struct ReinaImage {
    uint32_t width, height;
    uint8_t* rgba8;              /* width * height * 4 bytes, row-major */
} ReinaImage;

int reina_rt_image_load(const char* path, ReinaImage** out);
int reina_rt_image_load_memory(const uint8_t* bytes, uint64_t len, ReinaImage** out);
void reina_rt_image_free(ReinaImage* image);

/* src/scene/Scene.h:21-46 (Material), field for field in that order; the bools
 * are uint32 (0 / 1). What Scene::addInstance turns into
 * ReinaInstanceProperties. */
typedef struct ReinaMaterial {
    uint32_t materialIdx;             /* 0 Lambertian, 1 metal, 2 dielectric, 3 Disney */
    int32_t  textureID, normalMapID, bumpMapID;
    float    albedo[3], emission[3];
    float    roughness, ior;
    uint32_t interpNormals;
    float    absorption;
    uint32_t cullBackface;
    float    anisotropic, subsurface, clearcoatGloss;
    float    sheenTint[3], specularTint[3];
    float    metallic, clearcoat, specularTransmission, sheen;
} ReinaMaterial;

typedef struct ReinaGltfInstance {
    uint32_t model;                   /* index into ReinaGltfScene.models */
    float transform[16];              /* column-major object-to-world (glm::make_mat4 order) */
    ReinaMaterial material;
} ReinaGltfInstance;

typedef struct ReinaGltfScene {
    ReinaModelData* models;       uint32_t num_models;      /* one per mesh primitive */
    ReinaImage* images;           uint32_t num_images;      /* Material texture IDs index these */
    ReinaGltfInstance* instances; uint32_t num_instances;   /* one per node and primitive */
} ReinaGltfScene;

/* Images the decoders cannot read (12-bit JPEG, other formats) become texture ID -1 instead of
 * failing the load (the reference's fallback for an unmapped texture). */
#define REINA_RT_GLTF_SKIP_UNDECODABLE_IMAGES 0x1u

/* reina::scene::gltf::loadScene (src/scene/gltf/gltfloader.cpp:432-456): the
 * default scene of a .gltf (external buffers and images next to it, or data:
 * URIs) or .glb, as models, decoded images and per-node instances with the
 * reference's material mapping. The memory form takes a .glb or a
 * self-contained .gltf. REINA_RT_ERR_INVALID_ARG on a malformed asset,
 * REINA_RT_ERR_UNSUPPORTED for a required extension outside the reference's
 * list or an undecodable image. Released with reina_rt_gltf_free. */
int reina_rt_gltf_load(const char* path, uint32_t flags, ReinaGltfScene** out);
int reina_rt_gltf_load_memory(const uint8_t* bytes, uint64_t len, uint32_t flags, ReinaGltfScene** out);
void reina_rt_gltf_free(ReinaGltfScene* scene);

/* Diagnostic: evaluates the library's FP32 elementary functions on n inputs
 * (device pointers) so that tests can compare them with the host definitions.
 * fn: 0 sin, 1 cos, 2 log, 3 exp, 4 sqrt, 5 pow(x, y = in2). */
int reina_rt_debug_math(int fn, const float* in, const float* in2, float* out,
                        uint32_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* REINA_RT_H */
