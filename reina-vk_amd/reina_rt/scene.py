"""Host-side scene assembly mirroring Reina's Scene / Models API.

`Scene.define_object` / `add_instance` / `add_object` / `build_desc` pack the
buffers exactly as src/scene/Models.cpp:24-92 (Models::addModel) and
src/scene/Scene.cpp:27-54 (Scene::addInstance) do, and hand them to
reina_rt_scene_create through a ReinaSceneDesc. Camera matrices follow
src/graphics/Camera.cpp:10-11 (glm::perspective / glm::lookAt, OpenGL
right-handed clip space with depth in [-1, 1]).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class ModelData:
    """src/scene/Models.h:23-30. vertices: (n, 4) float32 (w = 1); indices:
    (m,) uint32 per-model; tbns: (k, 3, 3) columns (T, B, N); texCoords (j, 2)."""
    vertices: np.ndarray
    indices: np.ndarray
    tbns: np.ndarray = None
    tbnsIndices: np.ndarray = None
    texCoords: np.ndarray = None
    texIndices: np.ndarray = None


def _copy(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def load_obj(path_or_text):
    """Models::getObjData (src/scene/Models.cpp:117-175) through the library's
    native OBJ loader (reina_rt_model_load_obj): the file's first mesh,
    triangulated, with smooth normals where the file has none, tangent space,
    joined vertices and flipped v. `path_or_text` is a path, or the file's
    bytes. Returns ModelData (texCoords None when the mesh has none)."""
    from . import check, lib
    L = lib()
    out = ctypes.POINTER(abi.ReinaModelData)()
    if isinstance(path_or_text, (bytes, bytearray)):
        data = bytes(path_or_text)
        check(L.reina_rt_model_load_obj_memory(data, len(data), ctypes.byref(out)))
    else:
        check(L.reina_rt_model_load_obj(str(path_or_text).encode(), ctypes.byref(out)))
    try:
        m = out.contents
        v = _copy(m.vertices, 4 * m.num_vertices, np.float32).reshape(-1, 4)
        idx = _copy(m.indices, m.num_indices, np.uint32)
        tb = _copy(m.tbns, 9 * m.num_tbns, np.float32).reshape(-1, 3, 3)
        tbi = _copy(m.tbn_indices, m.num_tbn_indices, np.uint32)
        tc = _copy(m.tex_coords, 2 * m.num_tex_coords, np.float32).reshape(-1, 2) if m.num_tex_coords else None
        ti = _copy(m.tex_indices, m.num_tex_indices, np.uint32)
    finally:
        L.reina_rt_model_free(out)
    return ModelData(v, idx, tb, tbi, tc, ti)


def load_image(path_or_bytes):
    """A PNG as the RGBA8 (height, width, 4) array stb_image returns with 4
    channels (src/graphics/Image.cpp:10-41): a path loads with rows flipped
    (row 0 = the file's last row), bytes load unflipped."""
    from . import check, lib
    L = lib()
    out = ctypes.POINTER(abi.ReinaImage)()
    if isinstance(path_or_bytes, (bytes, bytearray)):
        data = bytes(path_or_bytes)
        check(L.reina_rt_image_load_memory(data, len(data), ctypes.byref(out)))
    else:
        check(L.reina_rt_image_load(str(path_or_bytes).encode(), ctypes.byref(out)))
    try:
        im = out.contents
        img = _copy(im.rgba8, 4 * im.width * im.height, np.uint8).reshape(im.height, im.width, 4)
    finally:
        L.reina_rt_image_free(out)
    return img


@dataclass
class Material:
    """src/scene/Scene.h:21-46 (field order kept)."""
    materialIdx: int = 0
    textureID: int = -1
    normalMapID: int = -1
    bumpMapID: int = -1
    albedo: tuple = (0.73, 0.73, 0.73)
    emission: tuple = (0.0, 0.0, 0.0)
    roughness: float = 0.0
    ior: float = 1.5
    interpNormals: bool = False
    absorption: float = 0.0
    cullBackface: bool = False
    anisotropic: float = 0.0
    subsurface: float = 0.0
    clearcoatGloss: float = 0.0
    sheenTint: tuple = (0.0, 0.0, 0.0)
    specularTint: tuple = (1.0, 1.0, 1.0)
    metallic: float = 0.0
    clearcoat: float = 0.0
    specularTransmission: float = 0.0
    sheen: float = 0.0


class Scene:
    def __init__(self):
        self._models = []          # ModelData
        self._ranges = []          # ModelRange dicts
        self._props = []
        self._instances = []       # (propsID, materialIdx, objectID, transform)
        self._vert_count = 0       # in floats (vec4 per vertex)
        self._tbn_count = 0        # matrices
        self._tex_count = 0        # floats (vec2 per coord)
        self._idx_count = 0
        self._tbni_count = 0
        self._texi_count = 0
        self._textures = []        # (H, W, 4) uint8 RGBA images (Scene::textures)
        self._built = False

    def add_texture(self, rgba8) -> int:
        """An RGBA8 image, row 0 = texture v = 0 (the order stb loads a file
        in, src/graphics/Image.cpp:10-86); returns the texture index that
        Material.textureID / normalMapID / bumpMapID refer to."""
        img = np.ascontiguousarray(rgba8, dtype=np.uint8)
        if img.ndim != 3 or img.shape[2] != 4:
            raise ValueError("texture must be (height, width, 4) uint8")
        self._textures.append(img)
        return len(self._textures) - 1

    def define_texture(self, path_or_bytes) -> int:
        """Scene::defineTexture (src/scene/Scene.cpp:15-25): a PNG file
        (rows flipped as stbi_load with flip-on-load) or PNG bytes (unflipped)."""
        return self.add_texture(load_image(path_or_bytes))

    # Models::addModel (src/scene/Models.cpp:24-92); a path goes through
    # Models::addModel(filepath) = getObjData first
    def define_object(self, md) -> int:
        if not isinstance(md, ModelData):
            md = load_obj(md)
        if self._built:
            raise RuntimeError("Could not add model; buffers are already built")
        v = np.ascontiguousarray(md.vertices, dtype=np.float32).reshape(-1, 4)
        ind = np.ascontiguousarray(md.indices, dtype=np.uint32).reshape(-1)
        tb = np.zeros((0, 3, 3), np.float32) if md.tbns is None else np.ascontiguousarray(md.tbns, np.float32).reshape(-1, 3, 3)
        tbi = np.zeros(0, np.uint32) if md.tbnsIndices is None else np.ascontiguousarray(md.tbnsIndices, np.uint32).reshape(-1)
        tc = np.zeros((0, 2), np.float32) if md.texCoords is None else np.ascontiguousarray(md.texCoords, np.float32).reshape(-1, 2)
        ti = np.zeros(0, np.uint32) if md.texIndices is None else np.ascontiguousarray(md.texIndices, np.uint32).reshape(-1)
        md2 = ModelData(v, ind, tb, tbi, tc, ti)
        rng = dict(
            firstVertex=self._vert_count // 4,
            firstNormal=self._tbn_count,
            indexOffset=self._idx_count,
            tbnsIndexOffset=self._tbni_count,
            texIndexOffset=0xFFFFFFFF if len(tc) == 0 else self._texi_count,
            indexCount=len(ind) // 3,
            tbnsIndexCount=len(ind) // 3,
            texIndexCount=len(ti) // 3,
        )
        md2._offsets = (self._vert_count // 4, self._tbn_count, self._tex_count // 2)
        self._ranges.append(rng)
        self._models.append(md2)
        self._vert_count += v.size
        self._tbn_count += len(tb)
        self._tex_count += tc.size
        self._idx_count += len(ind)
        self._tbni_count += len(tbi)
        self._texi_count += len(ti)
        return len(self._ranges) - 1

    # Scene::addInstance (src/scene/Scene.cpp:27-54)
    def add_instance(self, object_id: int, transform, mat: Material):
        r = self._ranges[object_id]
        p = np.zeros((), abi.INSTANCE_PROPERTIES)
        p["indicesOffset"] = r["indexOffset"]
        p["albedo"] = mat.albedo
        p["emission"] = mat.emission
        p["tbnsIndicesOffset"] = r["tbnsIndexOffset"]
        p["texIndicesOffset"] = r["texIndexOffset"]
        p["roughness"] = mat.roughness
        p["ior"] = mat.ior
        p["interpNormals"] = 1 if mat.interpNormals else 0
        p["absorption"] = mat.absorption
        p["textureID"] = mat.textureID
        p["normalMapTexID"] = mat.normalMapID
        p["bumpMapTexID"] = mat.bumpMapID
        p["cullBackface"] = 1 if mat.cullBackface else 0
        p["anisotropic"] = mat.anisotropic
        p["subsurface"] = mat.subsurface
        p["clearcoatGloss"] = mat.clearcoatGloss
        p["sheenTint"] = mat.sheenTint
        p["specularTint"] = mat.specularTint
        p["metallic"] = mat.metallic
        p["clearcoat"] = mat.clearcoat
        p["specularTransmission"] = mat.specularTransmission
        p["sheen"] = mat.sheen
        self._props.append(p)
        t = np.asarray(transform, dtype=np.float32).reshape(4, 4)
        self._instances.append((len(self._props) - 1, int(mat.materialIdx), int(object_id), t))

    def add_object(self, md, transform, mat: Material) -> int:
        oid = self.define_object(md)
        self.add_instance(oid, transform, mat)
        return oid

    def build_arrays(self):
        """Packed buffers (the same arrays Models::buildBuffers uploads)."""
        nv = self._vert_count // 4
        verts = np.zeros((nv, 4), np.float32)
        tbns = np.zeros((self._tbn_count, 3, 3), np.float32)
        texc = np.zeros((self._tex_count // 2, 2), np.float32)
        idx = np.zeros(self._idx_count, np.uint32)
        tbi = np.zeros(self._tbni_count, np.uint32)
        tei = np.zeros(self._texi_count, np.uint32)
        for md, r in zip(self._models, self._ranges):
            vo, to, xo = md._offsets
            verts[vo:vo + len(md.vertices)] = md.vertices
            tbns[to:to + len(md.tbns)] = md.tbns
            texc[xo:xo + len(md.texCoords)] = md.texCoords
            io = r["indexOffset"]
            idx[io:io + len(md.indices)] = md.indices + np.uint32(vo)                 # Models.cpp:78-81
            ko = r["tbnsIndexOffset"]
            tbi[ko:ko + len(md.tbnsIndices)] = md.tbnsIndices + np.uint32(to)         # Models.cpp:83-85
            if len(md.texIndices):
                so = r["texIndexOffset"]
                ti = md.texIndices.copy()
                m = ti != 0xFFFFFFFF
                ti[m] = ti[m] + np.uint32(xo)                                         # Models.cpp:87-89
                tei[so:so + len(ti)] = ti
        models = np.zeros(len(self._ranges), abi.MODEL_RANGE)
        for k, r in enumerate(self._ranges):
            for name in abi.MODEL_RANGE.names:
                models[k][name] = r[name]
        props = np.array(self._props, dtype=abi.INSTANCE_PROPERTIES).reshape(-1)
        inst = np.zeros(len(self._instances), abi.INSTANCE_DESC)
        for k, (pid, mi, oid, t) in enumerate(self._instances):
            # glm::mat4 is column-major: column c = t[:, c]
            inst[k]["transform"] = t.T.reshape(-1)
            inst[k]["instancePropertiesID"] = pid
            inst[k]["materialIdx"] = mi
            inst[k]["modelID"] = oid
        return dict(vertices=verts, indices=idx, tbns=tbns, tbn_indices=tbi, tex_coords=texc,
                    tex_indices=tei, props=props, models=models, instances=inst, textures=list(self._textures))


def make_desc(arrays, flags=0, device=0):
    """ReinaSceneDesc pointing into `arrays` (keep `arrays` alive while used)."""
    d = abi.ReinaSceneDesc()
    d.struct_size = ctypes.sizeof(abi.ReinaSceneDesc)
    d.flags = flags
    d.device = device

    def ptr(a):
        return a.ctypes.data if a.size else None

    a = arrays
    d.vertices, d.num_vertices = ptr(a["vertices"]), len(a["vertices"])
    d.indices, d.num_indices = ptr(a["indices"]), len(a["indices"])
    d.tbns, d.num_tbns = ptr(a["tbns"]), len(a["tbns"])
    d.tbn_indices, d.num_tbn_indices = ptr(a["tbn_indices"]), len(a["tbn_indices"])
    d.tex_coords, d.num_tex_coords = ptr(a["tex_coords"]), len(a["tex_coords"])
    d.tex_indices, d.num_tex_indices = ptr(a["tex_indices"]), len(a["tex_indices"])
    d.props, d.num_props = ptr(a["props"]), len(a["props"])
    d.models, d.num_models = ptr(a["models"]), len(a["models"])
    d.instances, d.num_instances = ptr(a["instances"]), len(a["instances"])
    tex = a.get("textures") or []
    if tex:
        td = (abi.ReinaTextureDesc * len(tex))()
        for k, img in enumerate(tex):
            td[k].width, td[k].height, td[k].rgba8 = img.shape[1], img.shape[0], img.ctypes.data
        d._texture_descs = td         # kept alive with the descriptor
        d.textures, d.num_textures = ctypes.addressof(td), len(tex)
    return d


# ---- camera (src/graphics/Camera.cpp:10-11, Reina.cpp:138-156) --------------------
def perspective(fovy, aspect, znear, zfar):
    f = np.float32(1.0) / np.tan(np.float32(fovy) / np.float32(2.0), dtype=np.float32)
    m = np.zeros((4, 4), np.float32)
    m[0, 0] = f / np.float32(aspect)
    m[1, 1] = f
    m[2, 2] = -(np.float32(zfar) + np.float32(znear)) / (np.float32(zfar) - np.float32(znear))
    m[3, 2] = np.float32(-1.0)
    m[2, 3] = -(np.float32(2.0) * np.float32(zfar) * np.float32(znear)) / (np.float32(zfar) - np.float32(znear))
    return m


def look_at(eye, center, up):
    eye = np.asarray(eye, np.float32)
    f = np.asarray(center, np.float32) - eye
    f = f / np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, np.float32))
    s = s / np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4, dtype=np.float32)
    m[0, :3] = s
    m[1, :3] = u
    m[2, :3] = -f
    m[0, 3] = -np.dot(s, eye)
    m[1, 3] = -np.dot(u, eye)
    m[2, 3] = np.dot(f, eye)
    return m.astype(np.float32)


def push_consts(eye, center, fov_deg, width, height, spp, max_bounces, sample_batch=0,
                focus_dist=2.2, defocus=0.0, direct_clamp=100.0, indirect_clamp=10.0, up=(0.0, 1.0, 0.0)):
    """RtPushConsts as Reina fills them (Reina.cpp:144-155); matrices column-major."""
    aspect = float(width) / float(height)
    inv_proj = np.linalg.inv(perspective(np.radians(fov_deg), aspect, 0.1, 100.0).astype(np.float64)).astype(np.float32)
    inv_view = np.linalg.inv(look_at(eye, center, up).astype(np.float64)).astype(np.float32)
    pc = np.zeros((), abi.PUSH_CONSTS)
    pc["invView"] = inv_view.T.reshape(-1)
    pc["invProjection"] = inv_proj.T.reshape(-1)
    pc["sampleBatch"] = sample_batch
    pc["totalEmissiveWeight"] = 0.0
    pc["focusDist"] = focus_dist
    pc["defocusMultiplier"] = defocus
    pc["directClamp"] = direct_clamp
    pc["indirectClamp"] = indirect_clamp
    pc["samplesPerPixel"] = spp
    pc["maxBounces"] = max_bounces
    return pc


# ---- packed scene file (input of the C++ headless host) ----------------------------
SCENE_FILE_MAGIC = b"RSCN"
SCENE_FILE_ARRAYS = ("vertices", "indices", "tbns", "tbn_indices", "tex_coords", "tex_indices", "props", "models",
                     "instances")


def save_scene(arrays, path):
    """Writes the packed buffers of build_arrays() (the ReinaSceneDesc
    arrays, in its field order) as: "RSCN", uint32 version 2, then per array
    uint64 element count and the raw little-endian elements (vertices 16 B,
    indices 4 B, tbns 36 B, tbn / tex indices 4 B, tex coords 8 B, props 120 B,
    model ranges 32 B, instances 80 B), then uint64 texture count and per
    texture uint32 width, uint32 height and width * height RGBA8 texels (row 0
    first). Version 1 files (no texture section) are still read."""
    import struct
    with open(path, "wb") as f:
        f.write(SCENE_FILE_MAGIC + struct.pack("<I", 2))
        for name in SCENE_FILE_ARRAYS:
            a = np.ascontiguousarray(arrays[name])
            f.write(struct.pack("<Q", len(a)))
            f.write(a.tobytes())
        tex = arrays.get("textures") or []
        f.write(struct.pack("<Q", len(tex)))
        for t in tex:
            t = np.ascontiguousarray(t, np.uint8)
            f.write(struct.pack("<II", t.shape[1], t.shape[0]))
            f.write(t.tobytes())
