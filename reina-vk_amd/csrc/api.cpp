// api.cpp — the C ABI of libreina_rt (include/reina_rt.h).
//
// reina_rt_scene_create  replaces Scene::build's BLAS/TLAS construction
//                        (src/scene/Scene.cpp:93-116): validate the packed
//                        Models/Instances buffers, flatten every instance to
//                        world space, build the BVH8, upload one HBM arena.
// reina_rt_trace         replaces Reina::traceRays' vkCmdTraceRaysKHR
//                        (src/Reina.cpp:425-470): init, then extend/shade
//                        bounce iterations until every pixel has finished its
//                        samples, then finish (running mean into the image).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "bvh_host.h"
#include "rt_layout.h"

#include "wavefront.h"

using namespace rt;

#include "api_internal.h"

static thread_local std::string g_err;

int rt::fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

extern "C" {

int reina_rt_abi_version(uint32_t* version, uint32_t* s_ip, uint32_t* s_pc, uint32_t* s_id) {
    if (version) *version = REINA_RT_ABI_VERSION;
    if (s_ip) *s_ip = (uint32_t)sizeof(ReinaInstanceProperties);
    if (s_pc) *s_pc = (uint32_t)sizeof(RtPushConsts);
    if (s_id) *s_id = (uint32_t)sizeof(ReinaInstanceData);
    return REINA_RT_OK;
}

const char* reina_rt_last_error(void) { return g_err.c_str(); }

static int validate(const ReinaSceneDesc* d) {
    if (!d) return fail(REINA_RT_ERR_INVALID_ARG, "scene description is null");
    if (d->struct_size != sizeof(ReinaSceneDesc))
        return fail(REINA_RT_ERR_ABI, "ReinaSceneDesc.struct_size mismatch (expected " +
                                          std::to_string(sizeof(ReinaSceneDesc)) + ")");
    if (d->flags & ~(REINA_RT_BUILD_DEVICE | REINA_RT_BUILD_HOST | REINA_RT_NODES_QUANTIZED))
        return fail(REINA_RT_ERR_INVALID_ARG, "unknown ReinaSceneDesc.flags bits");
    if ((d->flags & REINA_RT_BUILD_DEVICE) && (d->flags & REINA_RT_BUILD_HOST))
        return fail(REINA_RT_ERR_INVALID_ARG, "REINA_RT_BUILD_DEVICE and REINA_RT_BUILD_HOST are exclusive");
    if (d->num_instances && (!d->instances || !d->models || !d->props))
        return fail(REINA_RT_ERR_INVALID_ARG, "instances given without models/props");
    for (uint32_t m = 0; m < d->num_models; ++m) {
        const ReinaModelRange& r = d->models[m];
        if ((uint64_t)r.indexOffset + 3ull * r.indexCount > d->num_indices)
            return fail(REINA_RT_ERR_INVALID_ARG, "model " + std::to_string(m) + " index range exceeds indices");
    }
    for (uint64_t i = 0; i < d->num_indices; ++i)
        if (d->indices[i] >= d->num_vertices)
            return fail(REINA_RT_ERR_INVALID_ARG, "index " + std::to_string(i) + " out of vertex range");
    for (uint64_t i = 0; i < d->num_tbn_indices; ++i)
        if (d->tbn_indices[i] >= d->num_tbns)
            return fail(REINA_RT_ERR_INVALID_ARG, "tbn index " + std::to_string(i) + " out of range");
    // texture-coordinate indices are read only through an instance whose
    // texIndicesOffset is set: a model without UVs (texIndexOffset = -1,
    // Models.cpp:50) keeps index values with no coordinates behind them
    std::vector<uint8_t> tex_reached(d->num_tex_indices, 0);
    for (uint32_t i = 0; i < d->num_instances; ++i) {
        const ReinaInstanceDesc& in = d->instances[i];
        if (in.modelID >= d->num_models)
            return fail(REINA_RT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": bad modelID");
        if (in.instancePropertiesID >= d->num_props)
            return fail(REINA_RT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": bad instancePropertiesID");
        if (in.materialIdx > 3)
            return fail(REINA_RT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": materialIdx > 3");
        const ReinaInstanceProperties& p = d->props[in.instancePropertiesID];
        const ReinaModelRange& r = d->models[in.modelID];
        if ((uint64_t)p.indicesOffset + 3ull * r.indexCount > d->num_indices)
            return fail(REINA_RT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": indicesOffset out of range");
        if (p.texIndicesOffset != 0xFFFFFFFFu && (uint64_t)p.texIndicesOffset + 3ull * r.indexCount > d->num_tex_indices)
            return fail(REINA_RT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": texIndicesOffset out of range");
        if (p.texIndicesOffset != 0xFFFFFFFFu && r.indexCount)
            std::memset(tex_reached.data() + p.texIndicesOffset, 1, 3ull * r.indexCount);
        const int32_t ids[3] = {p.textureID, p.normalMapTexID, p.bumpMapTexID};
        for (int32_t id : ids)
            if (id >= 0 && (uint32_t)id >= d->num_textures)
                return fail(REINA_RT_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": texture index " +
                                                          std::to_string(id) + " out of range");
    }
    for (uint64_t i = 0; i < d->num_tex_indices; ++i)
        if (tex_reached[i] && d->tex_indices[i] != 0xFFFFFFFFu && d->tex_indices[i] >= d->num_tex_coords)
            return fail(REINA_RT_ERR_INVALID_ARG, "tex index " + std::to_string(i) + " out of range");
    if (d->num_textures && !d->textures) return fail(REINA_RT_ERR_INVALID_ARG, "num_textures without textures");
    for (uint32_t t = 0; t < d->num_textures; ++t) {
        const ReinaTextureDesc& x = d->textures[t];
        if (x.width == 0 || x.height == 0 || x.width > (1u << 16) || x.height > (1u << 16) || !x.rgba8)
            return fail(REINA_RT_ERR_INVALID_ARG, "texture " + std::to_string(t) +
                                                      ": width and height must be 1..65536 with data");
    }
    return REINA_RT_OK;
}

// Light-sampling tables (EmissiveInst, rt_layout.h) from the flattened
// world-space triangles, in double precision.
static void build_emissive(ReinaScene* s, const ReinaSceneDesc* d, const std::vector<InstanceRec>& inst,
                           const std::vector<float>& wtri) {
    s->emi.clear();
    s->tri_cdf.clear();
    s->inst_emi.assign(inst.size(), -1);
    s->tri_flags.assign(wtri.size() / 9, 0);
    std::vector<double> weight;
    double total = 0.0;
    for (uint32_t i = 0; i < inst.size(); ++i) {
        const InstanceRec& r = inst[i];
        const ReinaInstanceProperties& p = d->props[r.props_id];
        const uint32_t ntri = d->models[r.model].indexCount;
        const double det = (double)r.c0[0] * ((double)r.c1[1] * r.c2[2] - (double)r.c2[1] * r.c1[2]) -
                           (double)r.c1[0] * ((double)r.c0[1] * r.c2[2] - (double)r.c2[1] * r.c0[2]) +
                           (double)r.c2[0] * ((double)r.c0[1] * r.c1[2] - (double)r.c1[1] * r.c0[2]);
        const uint8_t fl = (uint8_t)((p.cullBackface ? kTriCull : 0) | (det < 0.0 ? kTriMirror : 0));
        for (uint32_t t = 0; t < ntri; ++t) s->tri_flags[r.tri_base + t] = fl;
        const float* e = p.emission;
        if (!((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2] > 0.00001f * 0.00001f)) continue;
        const double lum = 0.2126 * e[0] + 0.7152 * e[1] + 0.0722 * e[2];
        std::vector<double> cum(ntri);
        double area = 0.0;
        for (uint32_t t = 0; t < ntri; ++t) {
            const float* w = &wtri[9ull * (r.tri_base + t)];
            double ax = (double)w[3] - w[0], ay = (double)w[4] - w[1], az = (double)w[5] - w[2];
            double bx = (double)w[6] - w[0], by = (double)w[7] - w[1], bz = (double)w[8] - w[2];
            double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
            area += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
            cum[t] = area;
        }
        if (!(area > 0.0) || !(lum > 0.0)) continue;   // not sampled; its emission is found by BSDF rays only
        EmissiveInst ei{};
        ei.inst = i;
        ei.ntri = ntri;
        ei.cdf_base = (uint32_t)s->tri_cdf.size();
        for (uint32_t t = 0; t < ntri; ++t) s->tri_cdf.push_back((float)(cum[t] / area));
        s->tri_cdf.back() = 1.0f;
        weight.push_back(area * lum);
        total += area * lum;
        ei.pdf_area = (float)area;   // area for now; pdf below
        s->inst_emi[i] = (int32_t)s->emi.size();
        s->emi.push_back(ei);
    }
    double c = 0.0;
    for (size_t k = 0; k < s->emi.size(); ++k) {
        c += weight[k];
        s->emi[k].cdf = (float)(c / total);
        s->emi[k].pdf_area = (float)((weight[k] / total) / (double)s->emi[k].pdf_area);
    }
    if (!s->emi.empty()) s->emi.back().cdf = 1.0f;
}

static int scene_upload(ReinaScene* s, const ReinaSceneDesc* d, const std::vector<float>& wtri,
                        const std::vector<uint32_t>& tri_inst, const std::vector<uint32_t>& tri_prim,
                        const std::vector<InstanceRec>& inst) {
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    struct Part { const void* src; uint64_t bytes; uint64_t off; };
    std::vector<Part> parts;
    uint64_t off = 0;
    auto add = [&](const void* src, uint64_t bytes) {
        parts.push_back({src, bytes, off});
        off += al(bytes ? bytes : 4);
        return parts.size() - 1;
    };
    size_t p_nodes = s->quantized ? add(s->host_qnodes.data(), s->host_qnodes.size() * sizeof(QBvh8Node))
                                  : add(s->host_nodes.data(), s->host_nodes.size() * sizeof(Bvh8Node));
    size_t p_tris = add(s->host_tris.data(), s->host_tris.size() * sizeof(float));
    size_t p_ti = add(tri_inst.data(), tri_inst.size() * 4);
    size_t p_tp = add(tri_prim.data(), tri_prim.size() * 4);
    size_t p_in = add(inst.data(), inst.size() * sizeof(InstanceRec));
    size_t p_pr = add(d->props, (uint64_t)d->num_props * sizeof(ReinaInstanceProperties));
    size_t p_v = add(d->vertices, d->num_vertices * 16);
    size_t p_i = add(d->indices, d->num_indices * 4);
    size_t p_t = add(d->tbns, d->num_tbns * 36);
    size_t p_tix = add(d->tbn_indices, d->num_tbn_indices * 4);
    size_t p_tc = add(d->tex_coords, d->num_tex_coords * 8);
    size_t p_tci = add(d->tex_indices, d->num_tex_indices * 4);
    size_t p_emi = add(s->emi.data(), s->emi.size() * sizeof(EmissiveInst));
    size_t p_cdf = add(s->tri_cdf.data(), s->tri_cdf.size() * 4);
    size_t p_ie = add(s->inst_emi.data(), s->inst_emi.size() * 4);
    size_t p_tf = add(s->tri_flags.data(), s->tri_flags.size());
    size_t p_rec = add(nullptr, tri_inst.size() * (uint64_t)kShadeRecF4 * 16);   // filled on the device
    // textures: a TexInfo table, then every texture's texels (RGBA8 packed in 32 bits)
    std::vector<TexInfo> tinfo(d->num_textures);
    uint64_t ntexel = 0;
    for (uint32_t t = 0; t < d->num_textures; ++t) {
        tinfo[t] = TexInfo{d->textures[t].width, d->textures[t].height, ntexel};
        ntexel += (uint64_t)d->textures[t].width * d->textures[t].height;
    }
    size_t p_tinfo = add(tinfo.data(), tinfo.size() * sizeof(TexInfo));
    size_t p_texel = add(nullptr, ntexel * 4);
    (void)wtri;
    HIP_TRY(hipSetDevice(s->device));
    hipError_t e = hipMalloc(&s->arena, off);
    if (e != hipSuccess) return fail(REINA_RT_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(off) + " bytes failed");
    s->arena_bytes = off;
    for (const Part& p : parts)
        if (p.bytes && p.src) HIP_TRY(hipMemcpy((char*)s->arena + p.off, p.src, p.bytes, hipMemcpyHostToDevice));
    char* a = (char*)s->arena;
    for (uint32_t t = 0; t < d->num_textures; ++t)   // byte order r, g, b, a = the packed word's low to high byte
        HIP_TRY(hipMemcpy(a + parts[p_texel].off + 4 * tinfo[t].base, d->textures[t].rgba8,
                          4ull * tinfo[t].w * tinfo[t].h, hipMemcpyHostToDevice));
    s->sv.nodes = s->quantized ? nullptr : (const Bvh8Node RT_GLOBAL*)(a + parts[p_nodes].off);
    s->sv.qnodes = s->quantized ? (const QBvh8Node RT_GLOBAL*)(a + parts[p_nodes].off) : nullptr;
    s->sv.tris = (const float4 RT_GLOBAL*)(a + parts[p_tris].off);
    s->sv.tri_inst = (const uint32_t RT_GLOBAL*)(a + parts[p_ti].off);
    s->sv.tri_prim = (const uint32_t RT_GLOBAL*)(a + parts[p_tp].off);
    s->sv.inst = (const InstanceRec RT_GLOBAL*)(a + parts[p_in].off);
    s->sv.props = (const ReinaInstanceProperties RT_GLOBAL*)(a + parts[p_pr].off);
    s->sv.vertices = (const float4 RT_GLOBAL*)(a + parts[p_v].off);
    s->sv.indices = (const uint32_t RT_GLOBAL*)(a + parts[p_i].off);
    s->sv.tbns = (const float RT_GLOBAL*)(a + parts[p_t].off);
    s->sv.tbn_indices = (const uint32_t RT_GLOBAL*)(a + parts[p_tix].off);
    s->sv.tex_coords = (const float RT_GLOBAL*)(a + parts[p_tc].off);
    s->sv.tex_indices = (const uint32_t RT_GLOBAL*)(a + parts[p_tci].off);
    s->sv.num_tbns = d->num_tbns;
    s->sv.num_tbn_indices = d->num_tbn_indices;
    s->sv.num_tex_coords = d->num_tex_coords;
    s->sv.num_nodes = s->host_nodes.size();
    s->sv.num_tris = tri_inst.size();
    s->sv.emi = (const EmissiveInst RT_GLOBAL*)(a + parts[p_emi].off);
    s->sv.tri_cdf = (const float RT_GLOBAL*)(a + parts[p_cdf].off);
    s->sv.inst_emi = (const int32_t RT_GLOBAL*)(a + parts[p_ie].off);
    s->sv.tri_flags = (const uint8_t RT_GLOBAL*)(a + parts[p_tf].off);
    s->sv.num_emi = (uint32_t)s->emi.size();
    s->sv.num_inst = (uint32_t)inst.size();
    s->sv.tex = (const TexInfo RT_GLOBAL*)(a + parts[p_tinfo].off);
    s->sv.texels = (const uint32_t RT_GLOBAL*)(a + parts[p_texel].off);
    s->sv.num_tex = d->num_textures;
    float4* rec = (float4*)(a + parts[p_rec].off);
    HIP_TRY(launch_shade_records(s->sv, rec, (uint32_t)tri_inst.size(), nullptr));
    HIP_TRY(hipStreamSynchronize(nullptr));
    s->sv.srec = (const float4 RT_GLOBAL*)rec;
    return REINA_RT_OK;
}

int reina_rt_scene_create(const ReinaSceneDesc* d, ReinaScene** out) {
    if (!out) return fail(REINA_RT_ERR_INVALID_ARG, "out_scene is null");
    *out = nullptr;
    int rc = validate(d);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(REINA_RT_ERR_DEVICE, "no HIP device available");
    if ((int)d->device >= ndev) return fail(REINA_RT_ERR_INVALID_ARG, "device ordinal out of range");
    auto t0 = std::chrono::steady_clock::now();
    // flatten: world = ((c0*x + c1*y) + c2*z) + c3, instance-major order
    uint64_t ntri = 0;
    std::vector<InstanceRec> inst(d->num_instances);
    for (uint32_t i = 0; i < d->num_instances; ++i) {
        const ReinaInstanceDesc& id = d->instances[i];
        InstanceRec& r = inst[i];
        const float* m = id.transform;
        for (int k = 0; k < 3; ++k) { r.c0[k] = m[k]; r.c1[k] = m[4 + k]; r.c2[k] = m[8 + k]; r.c3[k] = m[12 + k]; }
        r.props_id = id.instancePropertiesID;
        r.material = id.materialIdx;
        r.model = id.modelID;
        r.tri_base = (uint32_t)ntri;
        ntri += d->models[id.modelID].indexCount;
    }
    if (ntri >= (1ull << 24)) return fail(REINA_RT_ERR_INVALID_ARG, "more than 2^24 triangles in one scene");
    std::vector<float> wtri(9 * ntri);
    std::vector<uint32_t> tri_inst(ntri), tri_prim(ntri);
    for (uint32_t i = 0; i < d->num_instances; ++i) {
        const InstanceRec& r = inst[i];
        const ReinaModelRange& mr = d->models[r.model];
        for (uint32_t p = 0; p < mr.indexCount; ++p) {
            uint64_t g = (uint64_t)r.tri_base + p;
            for (int k = 0; k < 3; ++k) {
                uint32_t vi = d->indices[3ull * p + mr.indexOffset + k];
                float x = d->vertices[4ull * vi], y = d->vertices[4ull * vi + 1], z = d->vertices[4ull * vi + 2];
                for (int c = 0; c < 3; ++c)
                    wtri[9 * g + 3 * k + c] = ((r.c0[c] * x + r.c1[c] * y) + r.c2[c] * z) + r.c3[c];
            }
            tri_inst[g] = i;
            tri_prim[g] = p;
        }
    }
    ReinaScene* s = new ReinaScene();
    s->device = (int)d->device;
    Bvh8Result bvh;
    std::string err;
    if (!(d->flags & REINA_RT_BUILD_HOST)) {   // the device build is the default
        if (hipSetDevice(s->device) != hipSuccess) {
            delete s;
            return fail(REINA_RT_ERR_DEVICE, "hipSetDevice failed");
        }
        if (!build_bvh8_device(wtri.data(), ntri, bvh, err, &s->device_build_ms)) {
            delete s;
            return fail(REINA_RT_ERR_DEVICE, "device BVH build: " + err);
        }
    } else if (!build_bvh8_host(wtri.data(), ntri, bvh, err)) {
        delete s;
        return fail(REINA_RT_ERR_INVALID_ARG, err);
    }
    if (bvh.max_depth > kMaxStackDepth + 1) {
        delete s;
        return fail(REINA_RT_ERR_UNSUPPORTED, "BVH8 depth " + std::to_string(bvh.max_depth) +
                                                  " exceeds the traversal stack (" + std::to_string(kMaxStackDepth) + ")");
    }
    s->host_nodes = std::move(bvh.nodes);
    s->host_tris = std::move(bvh.tris);
    if (d->flags & REINA_RT_NODES_QUANTIZED) {
        if (!quantize_bvh8(s->host_nodes, s->host_qnodes, err)) {
            delete s;
            return fail(REINA_RT_ERR_INVALID_ARG, err);
        }
        s->quantized = true;
    }
    build_emissive(s, d, inst, wtri);
    rc = scene_upload(s, d, wtri, tri_inst, tri_prim, inst);
    if (rc) {
        if (s->arena) hipFree(s->arena);
        delete s;
        return rc;
    }
    auto t1 = std::chrono::steady_clock::now();
    s->info.num_triangles = ntri;
    s->info.num_nodes = s->host_nodes.size();
    s->info.node_bytes = s->quantized ? sizeof(QBvh8Node) : sizeof(Bvh8Node);
    s->info.device_bytes = s->arena_bytes;
    s->info.max_depth = bvh.max_depth;
    s->info.quantized = s->quantized ? 1u : 0u;
    s->info.num_emissive = (uint32_t)s->emi.size();
    s->info.build_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
    s->info.sah_cost = bvh.sah_cost;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, s->device));
    uint32_t cus = (uint32_t)prop.multiProcessorCount;
    uint32_t need = bvh.max_depth > 1 ? bvh.max_depth - 1 : 1;
    s->tune.stack_depth = need <= 8 ? 8 : (need <= 16 ? 16 : (need <= 24 ? 24 : 32));
    s->tune.refill_min = 32;
    s->tune.grid = cus * (s->tune.stack_depth <= 8 ? 8 : 4);
    s->tune.shade_grid = cus * 16;   // fills every CU to its register-limited occupancy (latency hiding)
    s->tune.tail_paths = kDefaultTailPaths;
    s->min_stack = s->tune.stack_depth;
    *out = s;
    return REINA_RT_OK;
}

static void free_trace_buffers(ReinaScene* s) {
    if (s->tbuf) hipFree(s->tbuf);
    s->tbuf = nullptr;
    s->cap_px = 0;
}

int reina_rt_scene_destroy(ReinaScene* s) {
    if (!s) return fail(REINA_RT_ERR_INVALID_ARG, "scene is null");
    hipSetDevice(s->device);
    hipDeviceSynchronize();
    free_trace_buffers(s);
    if (s->ctrl) hipFree(s->ctrl);
    if (s->h_count) hipHostFree(s->h_count);
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ev1) hipEventDestroy(s->ev1);
    if (s->tev0) hipEventDestroy(s->tev0);
    if (s->tev1) hipEventDestroy(s->tev1);
    for (int r = 0; r < 2; ++r)
        if (s->cev[r]) hipEventDestroy(s->cev[r]);
    for (hipEvent_t e : s->xev) hipEventDestroy(e);
    if (s->arena) hipFree(s->arena);
    delete s;
    return REINA_RT_OK;
}

int reina_rt_scene_info(const ReinaScene* s, ReinaSceneInfo* o) {
    if (!s || !o) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    *o = s->info;
    return REINA_RT_OK;
}

int reina_rt_scene_export_bvh(const ReinaScene* s, void* nodes, uint64_t* nodes_bytes, void* tris,
                              uint64_t* tris_bytes) {
    if (!s || !nodes_bytes || !tris_bytes) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    const void* src = s->quantized ? (const void*)s->host_qnodes.data() : (const void*)s->host_nodes.data();
    uint64_t nb = s->quantized ? s->host_qnodes.size() * sizeof(QBvh8Node) : s->host_nodes.size() * sizeof(Bvh8Node);
    uint64_t tb = s->host_tris.size() * sizeof(float);
    if (nodes) {
        if (*nodes_bytes < nb) return fail(REINA_RT_ERR_INVALID_ARG, "nodes buffer too small");
        std::memcpy(nodes, src, nb);
    }
    if (tris) {
        if (*tris_bytes < tb) return fail(REINA_RT_ERR_INVALID_ARG, "tris buffer too small");
        std::memcpy(tris, s->host_tris.data(), tb);
    }
    *nodes_bytes = nb;
    *tris_bytes = tb;
    return REINA_RT_OK;
}

int reina_rt_scene_export_emissive(const ReinaScene* s, void* buf, uint64_t* bytes) {
    if (!s || !bytes) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    const uint32_t hdr[4] = {(uint32_t)s->emi.size(), (uint32_t)s->tri_cdf.size(), (uint32_t)s->inst_emi.size(),
                             (uint32_t)s->tri_flags.size()};
    const uint64_t need = 16 + s->emi.size() * sizeof(EmissiveInst) + s->tri_cdf.size() * 4 + s->inst_emi.size() * 4 +
                          s->tri_flags.size();
    if (buf) {
        if (*bytes < need) return fail(REINA_RT_ERR_INVALID_ARG, "buffer too small");
        char* p = (char*)buf;
        std::memcpy(p, hdr, 16); p += 16;
        std::memcpy(p, s->emi.data(), s->emi.size() * sizeof(EmissiveInst)); p += s->emi.size() * sizeof(EmissiveInst);
        std::memcpy(p, s->tri_cdf.data(), s->tri_cdf.size() * 4); p += s->tri_cdf.size() * 4;
        std::memcpy(p, s->inst_emi.data(), s->inst_emi.size() * 4); p += s->inst_emi.size() * 4;
        std::memcpy(p, s->tri_flags.data(), s->tri_flags.size());
    }
    *bytes = need;
    return REINA_RT_OK;
}

int reina_rt_bvh_build_host(const float* tri, uint64_t ntri, void* nodes, uint64_t* nodes_bytes, void* tris,
                            uint64_t* tris_bytes, uint32_t* max_depth) {
    if ((ntri && !tri) || !nodes_bytes || !tris_bytes) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    Bvh8Result r;
    std::string err;
    if (!build_bvh8_host(tri, ntri, r, err)) return fail(REINA_RT_ERR_INVALID_ARG, err);
    uint64_t nb = r.nodes.size() * sizeof(Bvh8Node), tb = r.tris.size() * sizeof(float);
    if (nodes) {
        if (*nodes_bytes < nb) return fail(REINA_RT_ERR_INVALID_ARG, "nodes buffer too small");
        std::memcpy(nodes, r.nodes.data(), nb);
    }
    if (tris) {
        if (*tris_bytes < tb) return fail(REINA_RT_ERR_INVALID_ARG, "tris buffer too small");
        std::memcpy(tris, r.tris.data(), tb);
    }
    *nodes_bytes = nb;
    *tris_bytes = tb;
    if (max_depth) *max_depth = r.max_depth;
    return REINA_RT_OK;
}

int reina_rt_bvh_quantize_host(const void* nodes, uint64_t nodes_bytes, void* qnodes, uint64_t* qnodes_bytes) {
    if ((nodes_bytes && !nodes) || !qnodes_bytes) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    if (nodes_bytes % sizeof(Bvh8Node)) return fail(REINA_RT_ERR_INVALID_ARG, "nodes_bytes is not a multiple of 256");
    std::vector<Bvh8Node> in(nodes_bytes / sizeof(Bvh8Node));
    if (nodes_bytes) std::memcpy(in.data(), nodes, nodes_bytes);
    std::vector<QBvh8Node> q;
    std::string err;
    if (!quantize_bvh8(in, q, err)) return fail(REINA_RT_ERR_INVALID_ARG, err);
    uint64_t qb = q.size() * sizeof(QBvh8Node);
    if (qnodes) {
        if (*qnodes_bytes < qb) return fail(REINA_RT_ERR_INVALID_ARG, "qnodes buffer too small");
        std::memcpy(qnodes, q.data(), qb);
    }
    *qnodes_bytes = qb;
    return REINA_RT_OK;
}

static int ensure_trace_buffers(ReinaScene* s, uint64_t npx) {
    if (!s->trace_ready) {
        // 3 iteration slots, the stats words, and the claim counters and used-up
        // flag of reina_rt_intersect
        if (!s->ctrl) HIP_TRY(hipMalloc(&s->ctrl, 3 * sizeof(Ctrl) + 256 + (kParts + 1) * kLine * 4));
        s->stats = (unsigned long long*)((char*)s->ctrl + 3 * sizeof(Ctrl));
        s->fetch = (uint32_t*)((char*)s->ctrl + 3 * sizeof(Ctrl) + 256);
        if (!s->h_count) HIP_TRY(hipHostMalloc((void**)&s->h_count, 2 * sizeof(Ctrl::count), hipHostMallocDefault));
        if (!s->ev0) HIP_TRY(hipEventCreate(&s->ev0));
        if (!s->ev1) HIP_TRY(hipEventCreate(&s->ev1));
        if (!s->tev0) HIP_TRY(hipEventCreate(&s->tev0));
        if (!s->tev1) HIP_TRY(hipEventCreate(&s->tev1));
        for (int r = 0; r < 2; ++r)
            if (!s->cev[r]) HIP_TRY(hipEventCreateWithFlags(&s->cev[r], hipEventDisableTiming));
        s->trace_ready = true;
    }
    if (npx <= s->cap_px) return REINA_RT_OK;
    free_trace_buffers(s);
    uint64_t qn = (uint64_t)kParts * part_capacity((uint32_t)npx);   // queue entries (partitioned)
    uint64_t bytes = npx * 80 /*state*/ + qn * (64 /*2 ray queues*/ + 8 /*2 pix queues*/ + 16 /*hits*/ +
                                                48 /*shadow rays + contributions*/) + 16 * 256;
    hipError_t e = hipMalloc(&s->tbuf, bytes);
    if (e != hipSuccess) return fail(REINA_RT_ERR_OUT_OF_MEMORY, "trace buffers: hipMalloc failed");
    char* p = (char*)s->tbuf;
    auto take = [&](uint64_t b) { char* r = p; p += (b + 255) & ~255ull; return r; };
    s->S.st0 = (uint4*)take(npx * 16);
    s->S.st1 = (float4*)take(npx * 16);
    s->S.st2 = (float4*)take(npx * 16);
    s->S.st3 = (float4*)take(npx * 16);
    s->S.st4 = (float4*)take(npx * 16);
    s->Q.ray[0] = (float4*)take(qn * 32);
    s->Q.ray[1] = (float4*)take(qn * 32);
    s->Q.pix[0] = (uint32_t*)take(qn * 4);
    s->Q.pix[1] = (uint32_t*)take(qn * 4);
    s->Q.hit = (float4*)take(qn * 16);
    s->Q.sray = (float4*)take(qn * 32);
    s->Q.scon = (float4*)take(qn * 16);
    s->cap_px = npx;
    return REINA_RT_OK;
}

int reina_rt_trace(ReinaScene* s, const RtPushConsts* pc, float* image, uint32_t W, uint32_t H, void* stream) {
    if (!s || !pc || !image) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    if (W == 0 || H == 0) return fail(REINA_RT_ERR_INVALID_ARG, "zero image size");
    if ((uint64_t)W * H >= (1ull << 31)) return fail(REINA_RT_ERR_INVALID_ARG, "image too large");
    hipStream_t st = (hipStream_t)stream;
    const uint32_t groups = 1u << ((s->options & REINA_RT_TRACE_SAMPLE_GROUPS_MASK) >> 8);
    if (pc->samplesPerPixel % groups != 0u)
        return fail(REINA_RT_ERR_INVALID_ARG, "samplesPerPixel " + std::to_string(pc->samplesPerPixel) +
                                                  " is not a multiple of the " + std::to_string(groups) +
                                                  " sample groups");
    if ((uint64_t)W * H * groups >= (1ull << 31)) return fail(REINA_RT_ERR_INVALID_ARG, "image x sample groups too large");
    HIP_TRY(hipSetDevice(s->device));
    int rc = ensure_trace_buffers(s, (uint64_t)W * H * groups);
    if (rc) return rc;
    TraceParams P = make_trace_params(pc, W, H);
    P.groups = groups;
    P.spp = pc->samplesPerPixel / groups;
    const bool nee = (s->options & REINA_RT_TRACE_NEE) != 0;
    if (nee && s->emi.empty())
        return fail(REINA_RT_ERR_NO_EMISSIVE, "next-event estimation needs at least one emissive instance");
    P.nee = nee ? 1u : 0u;
    P.fixes = s->options & REINA_RT_TRACE_FIXES;
    const bool timex = (s->options & REINA_RT_TRACE_TIME_EXTEND) != 0;
    const bool count = (s->options & REINA_RT_TRACE_COUNT) != 0;
    const uint32_t npx = W * H * groups;   // path slots: one per pixel and sample group
    s->Q.cap = part_capacity(npx);
    HIP_TRY(hipEventRecord(s->ev0, st));
    HIP_TRY(launch_init(P, s->S, s->Q, s->ctrl, s->stats, 0, npx, st));
    const uint64_t max_iters = (uint64_t)P.spp * P.maxBounces;
    uint64_t live = max_iters ? npx : 0;
    bool pending[2] = {false, false};
    uint32_t it = 0, rb = 0;
    const uint32_t K = 4;   // iterations between queue-length readbacks
    s->tail_count = 0;
    s->xev_used = 0;
    while (it < max_iters && live != 0) {
        const uint32_t slot = it % 3, in_q = it & 1u;
        if (live <= s->tune.tail_paths) {
            // few paths left: finish them in one launch
            HIP_TRY(hipEventRecord(s->tev0, st));
            HIP_TRY(launch_tail(s->sv, P, s->S, s->Q, in_q, s->ctrl, slot, s->stats, s->tune, (uint32_t)live, st));
            HIP_TRY(hipEventRecord(s->tev1, st));
            s->tail_count = 1;
            break;
        }
        if (timex) {
            if (s->xev.size() < 2 * (size_t)(s->xev_used + 1)) {
                hipEvent_t e0, e1;
                HIP_TRY(hipEventCreate(&e0));
                HIP_TRY(hipEventCreate(&e1));
                s->xev.push_back(e0);
                s->xev.push_back(e1);
            }
            HIP_TRY(hipEventRecord(s->xev[2 * s->xev_used], st));
        }
        TraceInput tin{&s->ctrl[slot].count[0][0], s->Q.cap, 0, &s->ctrl[slot].fetch[0][0], &s->ctrl[slot].done[0]};
        HIP_TRY(launch_trace(s->sv, s->Q.ray[in_q], s->Q.hit, tin, s->stats, nullptr, count, s->tune, st));
        if (timex) HIP_TRY(hipEventRecord(s->xev[2 * s->xev_used++ + 1], st));
        HIP_TRY(launch_shade(s->sv, P, s->S, s->Q, in_q, s->ctrl, slot, s->tune.shade_grid, nee, st));
        if (nee) HIP_TRY(launch_connect(s->sv, s->Q, s->S, s->ctrl, slot, s->stats, s->tune, st));
        ++it;
        if (it % K == 0) {
            // The queue length after this iteration is copied back in stream
            // order; the host reads the previous readback, so it waits only on
            // work enqueued K iterations ago and the GPU keeps K iterations
            // queued (no idle gap per readback). Decisions lag by K iterations.
            const uint32_t r = rb & 1u;
            HIP_TRY(hipMemcpyAsync(s->h_count + r * kParts * kLine, &s->ctrl[it % 3].count[0][0],
                                   sizeof(Ctrl::count), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipEventRecord(s->cev[r], st));
            pending[r] = true;
            if (pending[r ^ 1u]) {
                pending[r ^ 1u] = false;
                HIP_TRY(hipEventSynchronize(s->cev[r ^ 1u]));
                const uint32_t* h = s->h_count + (r ^ 1u) * kParts * kLine;
                live = 0;
                for (uint32_t p = 0; p < kParts; ++p) live += h[p * kLine];
            }
            rb++;
        }
    }
    HIP_TRY(launch_finish(P, s->S, (float4*)image, st));
    HIP_TRY(hipEventRecord(s->ev1, st));
    s->traced = true;
    s->last_iters = it;
    s->last_stream = st;
    return REINA_RT_OK;
}

int reina_rt_trace_stats(ReinaScene* s, ReinaTraceStats* o) {
    if (!s || !o) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    std::memset(o, 0, sizeof *o);
    if (!s->traced) return REINA_RT_OK;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipEventSynchronize(s->ev1));
    unsigned long long st3[5] = {0, 0, 0, 0, 0};
    HIP_TRY(hipMemcpy(st3, s->stats, sizeof st3, hipMemcpyDeviceToHost));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    o->extension_rays = st3[0];
    o->shadow_rays = st3[3];
    o->node_visits = st3[1];
    o->tri_tests = st3[2];
    o->iterations = s->last_iters;
    o->extend_launches = s->last_iters;
    o->trace_ms = ms;
    float xms = 0.0f;
    for (uint32_t k = 0; k < s->xev_used; ++k) {
        float e = 0.0f;
        HIP_TRY(hipEventElapsedTime(&e, s->xev[2 * k], s->xev[2 * k + 1]));
        xms += e;
    }
    o->extend_ms = xms;
    o->tail_rays = st3[4];
    o->tail_launches = s->tail_count;
    if (s->tail_count) HIP_TRY(hipEventElapsedTime(&o->tail_ms, s->tev0, s->tev1));
    return REINA_RT_OK;
}

int reina_rt_get_tuning(const ReinaScene* s, ReinaTuning* o) {
    if (!s || !o) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    o->stack_depth = s->tune.stack_depth;
    o->refill_min = s->tune.refill_min;
    o->grid_blocks = s->tune.grid;
    o->shade_blocks = s->tune.shade_grid;
    o->waves_per_simd = s->tune.waves_per_simd;
    o->tail_paths = s->tune.tail_paths;
    o->claim_rays = s->tune.claim;
    return REINA_RT_OK;
}

int reina_rt_set_tuning(ReinaScene* s, const ReinaTuning* t) {
    if (!s || !t) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    uint32_t sd = t->stack_depth;
    if (sd != 8 && sd != 16 && sd != 24 && sd != 32)
        return fail(REINA_RT_ERR_INVALID_ARG, "stack_depth must be 8, 16, 24 or 32");
    if (sd < s->min_stack)
        return fail(REINA_RT_ERR_INVALID_ARG, "stack_depth below what this BVH needs (" + std::to_string(s->min_stack) + ")");
    if (t->refill_min < 1 || t->refill_min > 64) return fail(REINA_RT_ERR_INVALID_ARG, "refill_min must be in 1..64");
    if (t->grid_blocks < 1 || t->grid_blocks > 65536 || t->shade_blocks < kParts || t->shade_blocks > 65536 ||
        t->shade_blocks % kParts)
        return fail(REINA_RT_ERR_INVALID_ARG, "grid_blocks must be in 1..65536 and shade_blocks a multiple of 8");
    if (t->waves_per_simd != 0 && t->waves_per_simd != 4 && t->waves_per_simd != 5 && t->waves_per_simd != 6 && t->waves_per_simd != 8)
        return fail(REINA_RT_ERR_INVALID_ARG, "waves_per_simd must be 0, 4, 5, 6 or 8");
    const uint32_t claim = t->claim_rays ? t->claim_rays : 128u;
    if (claim != 64 && claim != 128 && claim != 256 && claim != 512)
        return fail(REINA_RT_ERR_INVALID_ARG, "claim_rays must be 0, 64, 128, 256 or 512");
    s->tune.stack_depth = sd;
    s->tune.refill_min = t->refill_min;
    s->tune.grid = t->grid_blocks;
    s->tune.shade_grid = t->shade_blocks;
    s->tune.waves_per_simd = t->waves_per_simd;
    s->tune.tail_paths = t->tail_paths;
    s->tune.claim = claim;
    return REINA_RT_OK;
}

int reina_rt_set_trace_options(ReinaScene* s, uint32_t options) {
    if (!s) return fail(REINA_RT_ERR_INVALID_ARG, "scene is null");
    if (options & ~(REINA_RT_TRACE_TIME_EXTEND | REINA_RT_TRACE_COUNT | REINA_RT_TRACE_NEE | REINA_RT_TRACE_FIXES |
                    REINA_RT_TRACE_SAMPLE_GROUPS_MASK | REINA_RT_TRACE_F16_INV))
        return fail(REINA_RT_ERR_INVALID_ARG, "unknown trace option bits");
    s->options = options;
    return REINA_RT_OK;
}

int reina_rt_intersect(ReinaScene* s, const ReinaRay* rays, ReinaHit* hits, uint32_t n, uint32_t* counters,
                       void* stream) {
    if (!s || (n && (!rays || !hits))) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(s->device));
    int rc = ensure_trace_buffers(s, 0);
    if (rc) return rc;
    if (n == 0) return REINA_RT_OK;
    const bool h16 = (s->options & REINA_RT_TRACE_F16_INV) != 0;
    if (h16 && s->sv.qnodes) return fail(REINA_RT_ERR_UNSUPPORTED, "FP16 inverse directions need FP32 nodes");
    HIP_TRY(hipMemsetAsync(s->fetch, 0, (kParts + 1) * kLine * 4, st));
    TraceInput tin{nullptr, 0, n, s->fetch, s->fetch + kParts * kLine};
    if (h16)
        HIP_TRY(launch_trace_f32_h16(s->sv, (const float4*)rays, (float4*)hits, tin, nullptr, counters,
                                     counters != nullptr, s->tune, st));
    else
        HIP_TRY(launch_trace(s->sv, (const float4*)rays, (float4*)hits, tin, nullptr, counters, counters != nullptr,
                             s->tune, st));
    return REINA_RT_OK;
}

}  // extern "C"

// ---- reina_rt_intersect_noalloc -----------------------------------------
// The claim counters of the persistent traversal kernel (one 128-byte line
// per XCD partition) live in this file-scope device array instead of the
// scene's library-allocated control block. Each device holds its own copy;
// its address on a device comes from hipGetSymbolAddress, taken once per
// device with that device current. Host code never passes `scratch` itself
// to a kernel: in host code the name is a stand-in, not a device address.
namespace {
constexpr uint32_t kNoAllocScratchWords = (kParts + 1) * kLine;   // claim counters and used-up flag, 1152 bytes
constexpr uint32_t kNoAllocMaxRays = 1u << 30;   // keeps every claim counter far below 2^32
constexpr int kNoAllocMaxDevices = 64;
}  // namespace

static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

namespace {
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kNoAllocMaxDevices] = {};

__global__ void k_zero_words(uint32_t* p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// The device address of `scratch` on `device` (which must be current).
int scratch_address(int device, uint32_t** out) {
    if (device < 0 || device >= kNoAllocMaxDevices)
        return fail(REINA_RT_ERR_INVALID_ARG, "device ordinal " + std::to_string(device) + " out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(REINA_RT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = (uint32_t*)p;
    }
    *out = g_scratch_ptr[device];
    return REINA_RT_OK;
}
}  // namespace

extern "C" {

int reina_rt_intersect_noalloc(ReinaScene* s, const ReinaRay* rays, ReinaHit* hits, uint32_t n, void* stream) {
    if (!s || (n && (!rays || !hits))) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    if (n > kNoAllocMaxRays)
        return fail(REINA_RT_ERR_INVALID_ARG, "reina_rt_intersect_noalloc takes at most " +
                                                  std::to_string(kNoAllocMaxRays) + " rays per call, got " +
                                                  std::to_string(n));
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(s->device));
    uint32_t* fetch = nullptr;
    int rc = scratch_address(s->device, &fetch);
    if (rc) return rc;
    if (n == 0) return REINA_RT_OK;
    // the claim counters start at zero: a fill kernel, not hipMemsetAsync, so
    // the call can be captured in a HIP graph
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(256), 0, st, fetch, kNoAllocScratchWords);
    HIP_TRY(hipGetLastError());
    TraceInput tin{nullptr, 0, n, fetch, fetch + kParts * kLine};
    HIP_TRY(launch_trace(s->sv, (const float4*)rays, (float4*)hits, tin, nullptr, nullptr, false, s->tune, st));
    return REINA_RT_OK;
}

int reina_rt_postprocess(const float* in, float* scratch, uint8_t* out, uint32_t W, uint32_t H,
                         const BloomPushConsts* bloom, const TonemappingPushConsts* tm, void* stream) {
    if (!in || !scratch || !out || !bloom || !tm) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    if (W == 0 || H == 0) return fail(REINA_RT_ERR_INVALID_ARG, "zero image size");
    float4* ping = (float4*)scratch;
    float4* pong = ping + (uint64_t)W * H;
    HIP_TRY(launch_postprocess((const float4*)in, ping, pong, (uchar4*)out, W, H, *bloom, *tm, (hipStream_t)stream));
    return REINA_RT_OK;
}

int reina_rt_debug_math(int fn, const float* in, const float* in2, float* out, uint32_t n, void* stream) {
    if (n && (!in || !out || (fn == 5 && !in2))) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    HIP_TRY(launch_debug_math(fn, in, in2, out, n, (hipStream_t)stream));
    return REINA_RT_OK;
}

}  // extern "C"
