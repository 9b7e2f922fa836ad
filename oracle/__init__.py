"""ORACLE — ctypes binding of liboracle.so, the CPU restatement of Reina's
ray-tracing hot path (see oracle.h). Test infrastructure only: used by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product path.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    r = subprocess.run(["make", "-C", HERE, "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


_planted = {}


@contextlib.contextmanager
def planted(k):
    """Inside the block every function of this module runs a copy of the
    oracle built with deliberate mistake k (OR_PLANT in or_shade.c /
    or_post.c), so a test can show that its check fails on that mistake."""
    global _lib
    if k not in _planted:
        name = "_plant/liboracle_plant%d.so" % k
        r = subprocess.run(["make", "-C", HERE, "-s", name], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
        _planted[k] = _bind(ctypes.CDLL(os.path.join(HERE, name)))
    saved = lib()
    _lib = _planted[k]
    try:
        yield
    finally:
        _lib = saved


def _bind(L):
    c_p, u32, u64, i32, f64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    L.or_scene_create.argtypes = [c_p]
    L.or_scene_create.restype = c_p
    L.or_scene_destroy.argtypes = [c_p]
    L.or_scene_destroy.restype = None
    L.or_export_emissive.argtypes = [c_p, c_p]
    L.or_export_emissive.restype = u64
    L.or_scene_set_nee.argtypes = [c_p, i32]
    L.or_scene_set_nee.restype = None
    L.or_scene_set_fixes.argtypes = [c_p, u32]
    L.or_scene_set_fixes.restype = None
    L.or_occluded.argtypes = [c_p, c_p]
    L.or_occluded.restype = i32
    L.or_scene_set_bvh.argtypes = [c_p, c_p, u64, c_p, u64, u32]
    L.or_scene_set_bvh.restype = i32
    L.or_intersect_brute32.argtypes = [c_p, c_p, c_p, u64, i32]
    L.or_intersect_brute32.restype = None
    L.or_intersect_bvh.argtypes = [c_p, c_p, c_p, c_p, u64, i32]
    L.or_intersect_bvh.restype = None
    L.or_intersect_fp64.argtypes = [c_p, c_p, u64, c_p, c_p, c_p, c_p, c_p, f64, i32, i32]
    L.or_intersect_fp64.restype = None
    L.or_random.argtypes = [ctypes.POINTER(u32)]
    L.or_random.restype = ctypes.c_float
    L.or_pcg_stream.argtypes = [u32, u32, c_p, c_p]
    L.or_pcg_stream.restype = None
    L.or_offset_position.argtypes = [c_p, c_p, c_p, i32, c_p]
    L.or_offset_position.restype = None
    L.or_sky.argtypes = [c_p, c_p]
    L.or_sky.restype = None
    L.or_camera_ray.argtypes = [c_p, u32, u32, u32, u32, ctypes.POINTER(u32), c_p, c_p, i32]
    L.or_camera_ray.restype = None
    L.or_diffuse_dir.argtypes = [c_p, ctypes.POINTER(u32), c_p]
    L.or_diffuse_dir.restype = None
    L.or_render.argtypes = [c_p, c_p, u32, u32, c_p, u32, u32, u32, u32, i32, i32, ctypes.POINTER(u64)]
    L.or_render.restype = None
    L.or_render_sum64.argtypes = [c_p, c_p, u32, u32, u32, u32, c_p, c_p, c_p, u32, u32, u32, u32, i32, i32]
    L.or_render_sum64.restype = None
    L.or_postprocess.argtypes = [c_p, u32, u32, c_p, c_p, c_p, c_p, c_p]
    L.or_postprocess.restype = None
    L.or_math_eval.argtypes = [i32, c_p, c_p, c_p, u64]
    L.or_math_eval.restype = None
    L.or_tex_sample_at.argtypes = [c_p, i32, ctypes.c_float, ctypes.c_float, c_p]
    L.or_tex_sample_at.restype = None
    L.or_bump_at.argtypes = [c_p, i32, ctypes.c_float, ctypes.c_float, c_p, c_p, c_p]
    L.or_bump_at.restype = None
    L.or_disney_probe.argtypes = [i32, u64, c_p, c_p, c_p, ctypes.POINTER(u32), c_p]
    L.or_disney_probe.restype = None
    return L


def _p(a):
    return a.ctypes.data if a is not None else None


class OracleScene:
    """Flattened copy of a packed scene (+ optionally the library's BVH8)."""

    def __init__(self, arrays, make_desc):
        self.arrays = arrays
        d = make_desc(arrays)
        self.h = lib().or_scene_create(ctypes.byref(d))
        self._bvh = None

    def set_bvh(self, nodes, tris, node_bytes=256):
        """node_bytes: 256 for FP32 BVH8 nodes, 80 for 8-bit quantized nodes."""
        self._bvh = (np.ascontiguousarray(nodes), np.ascontiguousarray(tris))
        rc = lib().or_scene_set_bvh(self.h, _p(self._bvh[0]), self._bvh[0].nbytes, _p(self._bvh[1]),
                                    self._bvh[1].nbytes, node_bytes)
        if rc:
            raise ValueError("node_bytes must be 256 or 80")

    def set_nee(self, on):
        """Next-event estimation with MIS in render / render_sum64 (the
        library's REINA_RT_TRACE_NEE)."""
        lib().or_scene_set_nee(self.h, int(bool(on)))

    def set_fixes(self, fixes):
        """The library's REINA_RT_TRACE_FIX_* corrections of reference
        behaviours (a bit mask; 0 renders as the reference)."""
        lib().or_scene_set_fixes(self.h, int(fixes))

    def export_emissive(self):
        """Light tables in the byte layout of reina_rt_scene_export_emissive."""
        n = lib().or_export_emissive(self.h, None)
        buf = np.zeros(n, np.uint8)
        lib().or_export_emissive(self.h, buf.ctypes.data)
        return buf

    def tex_sample(self, tex_id, u, v):
        """texture(textures[tex_id], (u, v)) as the hit shaders sample it
        (RGBA, float32)."""
        out = np.zeros(4, np.float32)
        lib().or_tex_sample_at(self.h, tex_id, float(u), float(v), out.ctypes.data)
        return out

    def bump(self, tex_id, u, v, ray_dir, tbn):
        """bumpMapping(uv, rayIn, T, textures[tex_id]) (texutils.h.glsl:4-41);
        tbn: 3x3 with columns T, B, N."""
        d = np.ascontiguousarray(ray_dir, np.float32).reshape(3)
        t = np.ascontiguousarray(np.asarray(tbn, np.float32).T.reshape(9))   # column-major
        out = np.zeros(2, np.float32)
        lib().or_bump_at(self.h, tex_id, float(u), float(v), d.ctypes.data, t.ctypes.data, out.ctypes.data)
        return out

    def occluded(self, ray):
        r = np.ascontiguousarray(np.asarray(ray).reshape(()))
        return bool(lib().or_occluded(self.h, r.ctypes.data))

    def brute32(self, rays, nthreads=8):
        hits = np.zeros(len(rays), _hit_dtype())
        lib().or_intersect_brute32(self.h, _p(rays), _p(hits), len(rays), nthreads)
        return hits

    def bvh(self, rays, nthreads=8):
        hits = np.zeros(len(rays), _hit_dtype())
        cnt = np.zeros((len(rays), 2), np.uint32)
        lib().or_intersect_bvh(self.h, _p(rays), _p(hits), _p(cnt), len(rays), nthreads)
        return hits, cnt

    def fp64(self, rays, tie_rel=1e-5, use_bvh=False, nthreads=8):
        """FP64 Moller-Trumbore closest hit (argmin of (t, id)) and an
        ambiguity flag (edge/vertex/equal-t ties and near misses within
        tie_rel). use_bvh: cull with the FP32 BVH8 given to set_bvh (boxes
        padded by tie_rel; same result as the full scan)."""
        n = len(rays)
        t = np.zeros(n)
        prim = np.zeros(n, np.uint32)
        u = np.zeros(n)
        v = np.zeros(n)
        amb = np.zeros(n, np.uint8)
        lib().or_intersect_fp64(self.h, _p(rays), n, _p(t), _p(prim), _p(u), _p(v), _p(amb), tie_rel, int(use_bvh), nthreads)
        return t, prim, u, v, amb.astype(bool)

    def render(self, pc, W, H, img=None, rect=None, use_bvh=True, nthreads=8):
        if img is None:
            img = np.zeros((H, W, 4), np.float32)
        x0, y0, x1, y1 = rect if rect else (0, 0, W, H)
        pcb = np.ascontiguousarray(np.asarray(pc).reshape(()))
        rays = ctypes.c_uint64(0)
        lib().or_render(self.h, pcb.ctypes.data, W, H, _p(img), x0, y0, x1, y1, int(use_bvh), nthreads,
                        ctypes.byref(rays))
        return img, rays.value

    def render_sum64(self, pc, W, H, first_batch, nbatches, rect=None, use_bvh=True, nthreads=8):
        s = np.zeros((H, W, 3))
        sq = np.zeros((H, W, 3))
        cnt = np.zeros((H, W), np.uint32)
        x0, y0, x1, y1 = rect if rect else (0, 0, W, H)
        pcb = np.ascontiguousarray(np.asarray(pc).reshape(()))
        lib().or_render_sum64(self.h, pcb.ctypes.data, W, H, first_batch, nbatches, _p(s), _p(sq), _p(cnt),
                              x0, y0, x1, y1, int(use_bvh), nthreads)
        return s, sq, cnt

    def __del__(self):
        try:
            if self.h:
                lib().or_scene_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _hit_dtype():
    return np.dtype([("t", "<f4"), ("prim", "<u4"), ("u", "<f4"), ("v", "<f4")])


def pcg_stream(seed, n):
    st = np.zeros(n, np.uint32)
    val = np.zeros(n, np.float32)
    lib().or_pcg_stream(seed, n, _p(st), _p(val))
    return st, val


def offset_position(p, n, dielectric=False, d=None):
    p = np.asarray(p, np.float32)
    n = np.asarray(n, np.float32)
    d = np.zeros(3, np.float32) if d is None else np.asarray(d, np.float32)
    out = np.zeros(3, np.float32)
    lib().or_offset_position(_p(p), _p(n), _p(out), int(dielectric), _p(d))
    return out


def sky(d):
    d = np.asarray(d, np.float32)
    out = np.zeros(3, np.float32)
    lib().or_sky(_p(d), _p(out))
    return out


def camera_ray(pc, px, py, W, H, rng, lens_advances=False):
    pcb = np.ascontiguousarray(np.asarray(pc).reshape(()))
    st = ctypes.c_uint32(rng)
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib().or_camera_ray(pcb.ctypes.data, px, py, W, H, ctypes.byref(st), _p(o), _p(d), int(lens_advances))
    return o, d, st.value


DISNEY_PROBES = {"ggx_d": (0, 2), "smith_g1": (1, 2), "vndf": (2, 3), "pdf_refl": (3, 1), "gtr1": (4, 3),
                 "dc": (5, 1), "diffuse": (6, 3), "sheen": (7, 3), "pdf_clearcoat": (8, 1), "clearcoat": (9, 3),
                 "disney_sample": (10, 4), "disney_pdf": (11, 1), "disney_eval": (12, 4),
                 "disney_continue": (13, 7), "disney_nee": (14, 4)}


def disney_probe(name, n, a=None, b=None, prm=(), seed=1):
    """Batch evaluation of one piece of the oracle's Disney BSDF in the
    tangent frame (normal +z); see or_disney_probe. Returns (n, width) float32
    and the final RNG state of a sampling probe."""
    fn, width = DISNEY_PROBES[name]
    a = None if a is None else np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float32), (n, 3)))
    b = None if b is None else np.ascontiguousarray(np.broadcast_to(np.asarray(b, np.float32), (n, 3)))
    p = np.zeros(32, np.float32)
    p[:len(prm)] = prm
    st = ctypes.c_uint32(seed)
    out = np.zeros((n, width), np.float32)
    lib().or_disney_probe(fn, n, _p(a), _p(b), _p(p), ctypes.byref(st), _p(out))
    return out, st.value


def diffuse_dir(n, rng):
    n = np.asarray(n, np.float32)
    st = ctypes.c_uint32(rng)
    out = np.zeros(3, np.float32)
    lib().or_diffuse_dir(_p(n), ctypes.byref(st), _p(out))
    return out, st.value


def math_eval(fn, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, np.float32)
    out = np.zeros_like(x)
    lib().or_math_eval(fn, _p(x), _p(y), _p(out), len(x))
    return out


def postprocess(img, radius=5.0, threshold=1.0, intensity=0.05, exposure=1.0, return_bloom=False):
    """bloom + tonemap: (RGBA8 output, combined image[, blurred bloom image])."""
    H, W = img.shape[:2]
    img = np.ascontiguousarray(img, np.float32)
    ping = np.zeros_like(img)
    pong = np.zeros_like(img)
    out = np.zeros((H, W, 4), np.uint8)
    b = np.array([radius, threshold, intensity], np.float32)
    t = np.array([exposure], np.float32)
    lib().or_postprocess(_p(img), W, H, _p(b), _p(t), _p(ping), _p(pong), _p(out))
    return (out, ping, pong) if return_bloom else (out, ping)
