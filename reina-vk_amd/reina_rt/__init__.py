"""Python bindings (ctypes) of libreina_rt — the MI355X path-tracing hot path.

The product path is the HIP library libreina_rt.so built in this directory's
parent by build_lib.py. Loading fails loudly if it is missing: there is no
CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import abi
from .scene import Material, ModelData, Scene, load_image, load_obj, make_desc, push_consts

_HERE = os.path.dirname(os.path.abspath(__file__))
# REINA_RT_LIBRARY selects another build of the library (default: the in-tree one)
LIB_PATH = os.environ.get("REINA_RT_LIBRARY") or os.path.join(os.path.dirname(_HERE), "libreina_rt.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "reina_rt.h")

_lib = None


class ReinaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("reina_rt error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load libreina_rt.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libreina_rt.so not built: run `python reina-vk_amd/build_lib.py` "
                           "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    c_p, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    sig = {
        "reina_rt_abi_version": [ctypes.POINTER(u32)] * 4,
        "reina_rt_scene_create": [ctypes.POINTER(abi.ReinaSceneDesc), ctypes.POINTER(c_p)],
        "reina_rt_scene_destroy": [c_p],
        "reina_rt_scene_info": [c_p, ctypes.POINTER(abi.ReinaSceneInfo)],
        "reina_rt_scene_export_bvh": [c_p, c_p, ctypes.POINTER(u64), c_p, ctypes.POINTER(u64)],
        "reina_rt_trace": [c_p, c_p, c_p, u32, u32, c_p],
        "reina_rt_trace_stats": [c_p, ctypes.POINTER(abi.ReinaTraceStats)],
        "reina_rt_set_trace_options": [c_p, u32],
        "reina_rt_get_tuning": [c_p, ctypes.POINTER(abi.ReinaTuning)],
        "reina_rt_set_tuning": [c_p, ctypes.POINTER(abi.ReinaTuning)],
        "reina_rt_intersect": [c_p, c_p, c_p, u32, c_p, c_p],
        "reina_rt_intersect_noalloc": [c_p, c_p, c_p, u32, c_p],
        "reina_rt_postprocess": [c_p, c_p, c_p, u32, u32, c_p, c_p, c_p],
        "reina_rt_debug_math": [i32, c_p, c_p, c_p, u32, c_p],
        "reina_rt_bvh_build_host": [c_p, u64, c_p, ctypes.POINTER(u64), c_p, ctypes.POINTER(u64),
                                    ctypes.POINTER(u32)],
        "reina_rt_bvh_quantize_host": [c_p, u64, c_p, ctypes.POINTER(u64)],
        "reina_rt_scene_export_emissive": [c_p, c_p, ctypes.POINTER(u64)],
        "reina_rt_server_create": [c_p, c_p],
        "reina_rt_server_destroy": [c_p],
        "reina_rt_submit": [c_p, c_p, c_p, c_p, u32, u32, ctypes.POINTER(u64)],
        "reina_rt_poll": [c_p, u64, c_p],
        "reina_rt_wait": [c_p, u64, u32, c_p],
        "reina_rt_server_stats": [c_p, c_p],
        "reina_rt_model_load_obj": [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(abi.ReinaModelData))],
        "reina_rt_model_load_obj_memory": [ctypes.c_char_p, u64, ctypes.POINTER(ctypes.POINTER(abi.ReinaModelData))],
        "reina_rt_image_load": [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(abi.ReinaImage))],
        "reina_rt_image_load_memory": [ctypes.c_char_p, u64, ctypes.POINTER(ctypes.POINTER(abi.ReinaImage))],
        "reina_rt_gltf_load": [ctypes.c_char_p, u32, ctypes.POINTER(ctypes.POINTER(abi.ReinaGltfScene))],
        "reina_rt_gltf_load_memory": [ctypes.c_char_p, u64, u32, ctypes.POINTER(ctypes.POINTER(abi.ReinaGltfScene))],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = i32
    L.reina_rt_model_free.argtypes = [ctypes.POINTER(abi.ReinaModelData)]
    L.reina_rt_model_free.restype = None
    L.reina_rt_image_free.argtypes = [ctypes.POINTER(abi.ReinaImage)]
    L.reina_rt_image_free.restype = None
    L.reina_rt_gltf_free.argtypes = [ctypes.POINTER(abi.ReinaGltfScene)]
    L.reina_rt_gltf_free.restype = None
    L.reina_rt_last_error.argtypes = []
    L.reina_rt_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise ReinaError(rc, lib().reina_rt_last_error().decode())
    return rc


def exported_symbols():
    return ["reina_rt_abi_version", "reina_rt_last_error", "reina_rt_scene_create", "reina_rt_scene_destroy",
            "reina_rt_scene_info", "reina_rt_scene_export_bvh", "reina_rt_trace", "reina_rt_trace_stats",
            "reina_rt_intersect", "reina_rt_postprocess", "reina_rt_debug_math", "reina_rt_bvh_build_host",
            "reina_rt_set_trace_options", "reina_rt_get_tuning", "reina_rt_set_tuning",
            "reina_rt_bvh_quantize_host", "reina_rt_server_create", "reina_rt_server_destroy", "reina_rt_submit",
            "reina_rt_poll", "reina_rt_wait", "reina_rt_server_stats", "reina_rt_scene_export_emissive",
            "reina_rt_model_load_obj", "reina_rt_model_load_obj_memory", "reina_rt_model_free",
            "reina_rt_image_load", "reina_rt_image_load_memory", "reina_rt_image_free",
            "reina_rt_gltf_load", "reina_rt_gltf_load_memory", "reina_rt_gltf_free", "reina_rt_intersect_noalloc"]


class DeviceScene:
    """A scene resident in HBM (reina_rt_scene_create / _destroy)."""

    def __init__(self, scene_or_arrays, flags=0, device=0):
        arrays = scene_or_arrays.build_arrays() if isinstance(scene_or_arrays, Scene) else scene_or_arrays
        self.arrays = arrays
        desc = make_desc(arrays, flags=flags, device=device)
        h = ctypes.c_void_p()
        check(lib().reina_rt_scene_create(ctypes.byref(desc), ctypes.byref(h)))
        self.handle = h

    def info(self):
        inf = abi.ReinaSceneInfo()
        check(lib().reina_rt_scene_info(self.handle, ctypes.byref(inf)))
        return inf

    def export_bvh(self):
        nb, tb = ctypes.c_uint64(0), ctypes.c_uint64(0)
        check(lib().reina_rt_scene_export_bvh(self.handle, None, ctypes.byref(nb), None, ctypes.byref(tb)))
        nodes = np.zeros(nb.value, np.uint8)
        tris = np.zeros(tb.value // 4, np.float32)
        check(lib().reina_rt_scene_export_bvh(self.handle, nodes.ctypes.data, ctypes.byref(nb),
                                              tris.ctypes.data, ctypes.byref(tb)))
        return nodes, tris

    def export_emissive(self):
        """Light-sampling tables as a byte blob (see reina_rt_scene_export_emissive)."""
        nb = ctypes.c_uint64(0)
        check(lib().reina_rt_scene_export_emissive(self.handle, None, ctypes.byref(nb)))
        buf = np.zeros(nb.value, np.uint8)
        check(lib().reina_rt_scene_export_emissive(self.handle, buf.ctypes.data, ctypes.byref(nb)))
        return buf

    def trace(self, pc, image_ptr, width, height, stream=None):
        pcb = np.ascontiguousarray(np.asarray(pc, dtype=abi.PUSH_CONSTS).reshape(()))
        check(lib().reina_rt_trace(self.handle, pcb.ctypes.data, ctypes.c_void_p(image_ptr), width, height,
                                   ctypes.c_void_p(stream) if stream else None))

    def get_tuning(self):
        t = abi.ReinaTuning()
        check(lib().reina_rt_get_tuning(self.handle, ctypes.byref(t)))
        return dict(stack_depth=t.stack_depth, refill_min=t.refill_min, grid_blocks=t.grid_blocks,
                    shade_blocks=t.shade_blocks, waves_per_simd=t.waves_per_simd, tail_paths=t.tail_paths,
                    claim_rays=t.claim_rays)

    def set_tuning(self, **kw):
        cur = self.get_tuning()
        cur.update(kw)
        t = abi.ReinaTuning(cur["stack_depth"], cur["refill_min"], cur["grid_blocks"], cur["shade_blocks"],
                            cur["waves_per_simd"], cur["tail_paths"], cur["claim_rays"])
        check(lib().reina_rt_set_tuning(self.handle, ctypes.byref(t)))

    def set_trace_options(self, options):
        check(lib().reina_rt_set_trace_options(self.handle, options))

    def stats(self):
        st = abi.ReinaTraceStats()
        check(lib().reina_rt_trace_stats(self.handle, ctypes.byref(st)))
        return st

    def intersect(self, rays_ptr, hits_ptr, n, counters_ptr=None, stream=None):
        check(lib().reina_rt_intersect(self.handle, ctypes.c_void_p(rays_ptr), ctypes.c_void_p(hits_ptr), n,
                                       ctypes.c_void_p(counters_ptr) if counters_ptr else None,
                                       ctypes.c_void_p(stream) if stream else None))

    def close(self):
        if self.handle:
            check(lib().reina_rt_scene_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Batching / serving request loop (reina_rt_server_*): render requests
    on resident DeviceScenes share one wavefront; submit() returns a ticket
    at once, wait()/poll() report completion with timestamps.

    slot_tile, graph_iters, admit_order (abi.ADMIT_FIFO / ADMIT_SHORTEST),
    max_wait_ms, max_requests and tail_cu_percent are ReinaServerDesc's
    fields (0 = default); tail_inline sets REINA_RT_SERVER_TAIL_INLINE; the
    remaining keyword arguments are ReinaTuning fields."""

    def __init__(self, path_budget_bytes=1 << 30, max_pending=1024, quantized=False, device=0, nee=False,
                 admit_paths=0, slot_tile=0, graph_iters=0, admit_order=0, max_wait_ms=0, max_requests=0,
                 tail_inline=False, tail_cu_percent=0, **tuning):
        d = abi.ReinaServerDesc()
        d.struct_size = ctypes.sizeof(abi.ReinaServerDesc)
        d.device = device
        d.path_budget_bytes = int(path_budget_bytes)
        d.max_pending = max_pending
        d.quantized = int(bool(quantized))
        d.flags = (abi.SERVER_NEE if nee else 0) | (abi.SERVER_TAIL_INLINE if tail_inline else 0)
        d.admit_paths = int(admit_paths)
        d.slot_tile = int(slot_tile)
        d.graph_iters = int(graph_iters)
        d.admit_order = int(admit_order)
        d.max_wait_ms = int(max_wait_ms)
        d.max_requests = int(max_requests)
        d.tail_cu_percent = int(tail_cu_percent)
        for k, v in tuning.items():
            setattr(d.tuning, k, v)
        h = ctypes.c_void_p()
        check(lib().reina_rt_server_create(ctypes.byref(d), ctypes.byref(h)))
        self.handle = h
        self._scenes = {}   # keep scenes alive while their requests run

    def submit(self, dev, pc, image_ptr, width, height):
        pcb = np.ascontiguousarray(np.asarray(pc, dtype=abi.PUSH_CONSTS).reshape(()))
        t = ctypes.c_uint64(0)
        check(lib().reina_rt_submit(self.handle, dev.handle, pcb.ctypes.data, ctypes.c_void_p(image_ptr), width, height,
                                    ctypes.byref(t)))
        self._scenes[t.value] = dev
        return t.value

    def wait(self, ticket, timeout_ms=600000):
        st = abi.ReinaRequestStatus()
        check(lib().reina_rt_wait(self.handle, ticket, timeout_ms, ctypes.byref(st)))
        self._scenes.pop(ticket, None)
        return st

    def poll(self, ticket):
        st = abi.ReinaRequestStatus()
        check(lib().reina_rt_poll(self.handle, ticket, ctypes.byref(st)))
        if st.state in (abi.REQ_DONE, abi.REQ_FAILED):
            self._scenes.pop(ticket, None)
        return st

    def stats(self):
        st = abi.ReinaServerStats()
        check(lib().reina_rt_server_stats(self.handle, ctypes.byref(st)))
        return st

    def close(self):
        if self.handle:
            check(lib().reina_rt_server_destroy(self.handle))
            self.handle = None
            self._scenes.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bvh_build_host(world_tris):
    """Host-only BVH8 build of a (n, 9) float32 triangle array -> (nodes u8, tris f32, depth)."""
    t = np.ascontiguousarray(world_tris, np.float32).reshape(-1, 9)
    nb, tb, dep = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint32(0)
    check(lib().reina_rt_bvh_build_host(t.ctypes.data if len(t) else None, len(t), None, ctypes.byref(nb), None,
                                        ctypes.byref(tb), ctypes.byref(dep)))
    nodes = np.zeros(nb.value, np.uint8)
    tris = np.zeros(tb.value // 4, np.float32)
    check(lib().reina_rt_bvh_build_host(t.ctypes.data if len(t) else None, len(t), nodes.ctypes.data,
                                        ctypes.byref(nb), tris.ctypes.data, ctypes.byref(tb), ctypes.byref(dep)))
    return nodes, tris, dep.value


def bvh_quantize_host(nodes):
    """8-bit quantized copy (80-byte nodes, u8 array) of FP32 BVH8 nodes (u8 array)."""
    n = np.ascontiguousarray(nodes, np.uint8)
    qb = ctypes.c_uint64(0)
    check(lib().reina_rt_bvh_quantize_host(n.ctypes.data if n.size else None, n.nbytes, None, ctypes.byref(qb)))
    q = np.zeros(qb.value, np.uint8)
    check(lib().reina_rt_bvh_quantize_host(n.ctypes.data if n.size else None, n.nbytes, q.ctypes.data,
                                           ctypes.byref(qb)))
    return q


def postprocess(in_ptr, scratch_ptr, out_ptr, width, height, radius=5.0, threshold=1.0, intensity=0.05,
                exposure=1.0, stream=None):
    b = np.array([(radius, threshold, intensity)], abi.BLOOM)
    t = np.array([(exposure,)], abi.TONEMAP)
    check(lib().reina_rt_postprocess(ctypes.c_void_p(in_ptr), ctypes.c_void_p(scratch_ptr), ctypes.c_void_p(out_ptr),
                                     width, height, b.ctypes.data, t.ctypes.data,
                                     ctypes.c_void_p(stream) if stream else None))


def debug_math(fn, in_ptr, in2_ptr, out_ptr, n, stream=None):
    check(lib().reina_rt_debug_math(fn, ctypes.c_void_p(in_ptr), ctypes.c_void_p(in2_ptr) if in2_ptr else None,
                                    ctypes.c_void_p(out_ptr), n, ctypes.c_void_p(stream) if stream else None))
