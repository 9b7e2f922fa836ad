"""numpy / ctypes mirrors of the C ABI in include/reina_rt.h.

The three polyglot structs keep the reference's byte layout:
InstanceProperties 120 B (polyglot/raytrace.h:18-41), RtPushConsts 160 B
(polyglot/raytrace.h:43-54), InstanceData 112 B (src/scene/Instances.h:15-27).
"""
import ctypes

import numpy as np

INSTANCE_PROPERTIES = np.dtype({
    "names": ["indicesOffset", "albedo", "emission", "tbnsIndicesOffset", "texIndicesOffset",
              "roughness", "ior", "interpNormals", "absorption", "textureID", "normalMapTexID",
              "bumpMapTexID", "cullBackface", "anisotropic", "subsurface", "clearcoatGloss",
              "sheenTint", "specularTint", "metallic", "clearcoat", "specularTransmission", "sheen"],
    "formats": ["<u4", ("<f4", 3), ("<f4", 3), "<u4", "<u4", "<f4", "<f4", "u1", "<f4", "<i4", "<i4",
                "<i4", "<u4", "<f4", "<f4", "<f4", ("<f4", 3), ("<f4", 3), "<f4", "<f4", "<f4", "<f4"],
    "offsets": [0, 4, 16, 28, 32, 36, 40, 44, 48, 52, 56, 60, 64, 68, 72, 76, 80, 92, 104, 108, 112, 116],
    "itemsize": 120,
})

PUSH_CONSTS = np.dtype({
    "names": ["invView", "invProjection", "sampleBatch", "totalEmissiveWeight", "focusDist",
              "defocusMultiplier", "directClamp", "indirectClamp", "samplesPerPixel", "maxBounces"],
    "formats": [("<f4", 16), ("<f4", 16), "<u4", "<f4", "<f4", "<f4", "<f4", "<f4", "<u4", "<u4"],
    "offsets": [0, 64, 128, 132, 136, 140, 144, 148, 152, 156],
    "itemsize": 160,
})

INSTANCE_DATA = np.dtype({
    "names": ["transform", "materialOffset", "cdfRangeStart", "cdfRangeEnd", "indexOffset", "emission",
              "weight", "area", "cullBackface", "padding"],
    "formats": [("<f4", 16), "<u4", "<u4", "<u4", "<u4", ("<f4", 3), "<f4", "<f4", "u1", ("<f4", 2)],
    "offsets": [0, 64, 68, 72, 76, 80, 92, 96, 100, 104],
    "itemsize": 112,
})

MODEL_RANGE = np.dtype([("firstVertex", "<u4"), ("firstNormal", "<u4"), ("indexOffset", "<u4"),
                        ("tbnsIndexOffset", "<u4"), ("texIndexOffset", "<u4"), ("indexCount", "<u4"),
                        ("tbnsIndexCount", "<u4"), ("texIndexCount", "<u4")])

INSTANCE_DESC = np.dtype([("transform", "<f4", 16), ("instancePropertiesID", "<u4"), ("materialIdx", "<u4"),
                          ("modelID", "<u4"), ("_pad0", "<u4")])

RAY = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("tmin", "<f4"),
                ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("tmax", "<f4")])
HIT = np.dtype([("t", "<f4"), ("prim", "<u4"), ("u", "<f4"), ("v", "<f4")])

BLOOM = np.dtype([("radius", "<f4"), ("threshold", "<f4"), ("intensity", "<f4")])
TONEMAP = np.dtype([("exposure", "<f4")])

assert INSTANCE_PROPERTIES.itemsize == 120
assert PUSH_CONSTS.itemsize == 160
assert INSTANCE_DATA.itemsize == 112
assert MODEL_RANGE.itemsize == 32
assert INSTANCE_DESC.itemsize == 80


class ReinaTextureDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("rgba8", ctypes.c_void_p)]


class ReinaModelData(ctypes.Structure):
    _fields_ = [
        ("vertices", ctypes.POINTER(ctypes.c_float)), ("num_vertices", ctypes.c_uint64),
        ("indices", ctypes.POINTER(ctypes.c_uint32)), ("num_indices", ctypes.c_uint64),
        ("tbns", ctypes.POINTER(ctypes.c_float)), ("num_tbns", ctypes.c_uint64),
        ("tbn_indices", ctypes.POINTER(ctypes.c_uint32)), ("num_tbn_indices", ctypes.c_uint64),
        ("tex_coords", ctypes.POINTER(ctypes.c_float)), ("num_tex_coords", ctypes.c_uint64),
        ("tex_indices", ctypes.POINTER(ctypes.c_uint32)), ("num_tex_indices", ctypes.c_uint64),
    ]


class ReinaImage(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("rgba8", ctypes.POINTER(ctypes.c_uint8))]


class ReinaMaterial(ctypes.Structure):
    _fields_ = [
        ("materialIdx", ctypes.c_uint32), ("textureID", ctypes.c_int32), ("normalMapID", ctypes.c_int32),
        ("bumpMapID", ctypes.c_int32), ("albedo", ctypes.c_float * 3), ("emission", ctypes.c_float * 3),
        ("roughness", ctypes.c_float), ("ior", ctypes.c_float), ("interpNormals", ctypes.c_uint32),
        ("absorption", ctypes.c_float), ("cullBackface", ctypes.c_uint32), ("anisotropic", ctypes.c_float),
        ("subsurface", ctypes.c_float), ("clearcoatGloss", ctypes.c_float), ("sheenTint", ctypes.c_float * 3),
        ("specularTint", ctypes.c_float * 3), ("metallic", ctypes.c_float), ("clearcoat", ctypes.c_float),
        ("specularTransmission", ctypes.c_float), ("sheen", ctypes.c_float),
    ]


class ReinaGltfInstance(ctypes.Structure):
    _fields_ = [("model", ctypes.c_uint32), ("transform", ctypes.c_float * 16), ("material", ReinaMaterial)]


class ReinaGltfScene(ctypes.Structure):
    _fields_ = [
        ("models", ctypes.POINTER(ReinaModelData)), ("num_models", ctypes.c_uint32),
        ("images", ctypes.POINTER(ReinaImage)), ("num_images", ctypes.c_uint32),
        ("instances", ctypes.POINTER(ReinaGltfInstance)), ("num_instances", ctypes.c_uint32),
    ]


GLTF_SKIP_UNDECODABLE_IMAGES = 0x1


class ReinaSceneDesc(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("flags", ctypes.c_uint32),
        ("vertices", ctypes.c_void_p), ("num_vertices", ctypes.c_uint64),
        ("indices", ctypes.c_void_p), ("num_indices", ctypes.c_uint64),
        ("tbns", ctypes.c_void_p), ("num_tbns", ctypes.c_uint64),
        ("tbn_indices", ctypes.c_void_p), ("num_tbn_indices", ctypes.c_uint64),
        ("tex_coords", ctypes.c_void_p), ("num_tex_coords", ctypes.c_uint64),
        ("tex_indices", ctypes.c_void_p), ("num_tex_indices", ctypes.c_uint64),
        ("props", ctypes.c_void_p), ("num_props", ctypes.c_uint32),
        ("models", ctypes.c_void_p), ("num_models", ctypes.c_uint32),
        ("instances", ctypes.c_void_p), ("num_instances", ctypes.c_uint32),
        ("device", ctypes.c_uint32),
        ("textures", ctypes.c_void_p), ("num_textures", ctypes.c_uint32), ("_pad0", ctypes.c_uint32),
    ]


class ReinaSceneInfo(ctypes.Structure):
    _fields_ = [("num_triangles", ctypes.c_uint64), ("num_nodes", ctypes.c_uint64),
                ("node_bytes", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64),
                ("max_depth", ctypes.c_uint32), ("quantized", ctypes.c_uint32),
                ("build_ms", ctypes.c_float), ("sah_cost", ctypes.c_float),
                ("num_emissive", ctypes.c_uint32), ("_pad0", ctypes.c_uint32)]


class ReinaTraceStats(ctypes.Structure):
    _fields_ = [("extension_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
                ("node_visits", ctypes.c_uint64), ("tri_tests", ctypes.c_uint64),
                ("iterations", ctypes.c_uint32), ("extend_launches", ctypes.c_uint32),
                ("trace_ms", ctypes.c_float), ("extend_ms", ctypes.c_float),
                ("tail_rays", ctypes.c_uint64), ("tail_ms", ctypes.c_float), ("tail_launches", ctypes.c_uint32)]


class ReinaTuning(ctypes.Structure):
    _fields_ = [("stack_depth", ctypes.c_uint32), ("refill_min", ctypes.c_uint32),
                ("grid_blocks", ctypes.c_uint32), ("shade_blocks", ctypes.c_uint32),
                ("waves_per_simd", ctypes.c_uint32), ("tail_paths", ctypes.c_uint32),
                ("claim_rays", ctypes.c_uint32)]


class ReinaServerDesc(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("device", ctypes.c_uint32),
                ("path_budget_bytes", ctypes.c_uint64), ("max_pending", ctypes.c_uint32),
                ("quantized", ctypes.c_uint32), ("tuning", ReinaTuning), ("flags", ctypes.c_uint32),
                ("admit_paths", ctypes.c_uint32), ("slot_tile", ctypes.c_uint32), ("graph_iters", ctypes.c_uint32),
                ("admit_order", ctypes.c_uint32), ("max_wait_ms", ctypes.c_uint32),
                ("max_requests", ctypes.c_uint32), ("tail_cu_percent", ctypes.c_uint32)]


class ReinaRequestStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint32), ("npx", ctypes.c_uint32), ("submit_s", ctypes.c_double),
                ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double)]


class ReinaServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("rejected_busy", ctypes.c_uint64), ("rejected_budget", ctypes.c_uint64),
                ("extension_rays", ctypes.c_uint64), ("iterations", ctypes.c_uint64),
                ("graph_replays", ctypes.c_uint64), ("pending", ctypes.c_uint32), ("running", ctypes.c_uint32),
                ("slot_tiles_total", ctypes.c_uint32), ("slot_tiles_used", ctypes.c_uint32),
                ("busy_s", ctypes.c_double), ("tail_launches", ctypes.c_uint64), ("tail_rays", ctypes.c_uint64)]


SERVER_BYTES_PER_PIXEL = 152
SERVER_BYTES_PER_PIXEL_NEE = 216
SERVER_NEE = 0x1
TAIL_OFF = 0xFFFFFFFF
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3

TRACE_TIME_EXTEND = 0x1
TRACE_COUNT = 0x2
TRACE_NEE = 0x4
TRACE_FIX_LENS_RNG = 0x8
TRACE_FIX_MEDIUM_RESET = 0x10
TRACE_FIX_NORMALS = 0x20
TRACE_FIX_METAL_PDF = 0x40
TRACE_FIX_DISNEY_MIX = 0x80
TRACE_FIXES = 0xF8
TRACE_SAMPLE_GROUPS_MASK = 0xF00
TRACE_F16_INV = 0x1000


def TRACE_SAMPLE_GROUPS(log2g):
    """Trace-option bits for 2**log2g independent sample streams per pixel."""
    return (int(log2g) & 0xF) << 8


# status codes
OK = 0
ERR_INVALID_ARG = -1
ERR_ABI = -2
ERR_OUT_OF_MEMORY = -3
ERR_OVER_BUDGET = -4
ERR_DEVICE = -5
ERR_UNSUPPORTED = -6
ERR_NO_EMISSIVE = -7
ERR_BUSY = -8

BUILD_DEVICE = 0x1   # the default (flags 0 builds on the device too)
BUILD_HOST = 0x4
NODES_QUANTIZED = 0x2
ADMIT_ALL = 0xFFFFFFFF
SERVER_TAIL_INLINE = 0x2
ADMIT_FIFO = 0
ADMIT_SHORTEST = 1
ADMIT_LONGEST = 2
