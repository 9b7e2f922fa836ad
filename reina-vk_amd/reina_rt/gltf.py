"""glTF scenes through the library's native loader (reina_rt_gltf_load):
reina::scene::gltf::loadScene (src/scene/gltf/gltfloader.cpp:432-456) as a
Scene -- one model per mesh primitive, the decoded images as textures, one
instance per node and primitive with the reference's Disney material mapping.
"""
import ctypes

import numpy as np

from . import abi
from .scene import Material, ModelData, Scene, _copy


def _material(m):
    return Material(materialIdx=int(m.materialIdx), textureID=int(m.textureID), normalMapID=int(m.normalMapID),
                    bumpMapID=int(m.bumpMapID), albedo=tuple(m.albedo), emission=tuple(m.emission),
                    roughness=float(m.roughness), ior=float(m.ior), interpNormals=bool(m.interpNormals),
                    absorption=float(m.absorption), cullBackface=bool(m.cullBackface),
                    anisotropic=float(m.anisotropic), subsurface=float(m.subsurface),
                    clearcoatGloss=float(m.clearcoatGloss), sheenTint=tuple(m.sheenTint),
                    specularTint=tuple(m.specularTint), metallic=float(m.metallic), clearcoat=float(m.clearcoat),
                    specularTransmission=float(m.specularTransmission), sheen=float(m.sheen))


def load(path_or_bytes, skip_undecodable_images=False):
    """The loader's raw result: (models [ModelData], images [(h, w, 4) uint8],
    instances [(model, 4x4 row-major float32 transform, Material)])."""
    from . import check, lib
    L = lib()
    flags = abi.GLTF_SKIP_UNDECODABLE_IMAGES if skip_undecodable_images else 0
    out = ctypes.POINTER(abi.ReinaGltfScene)()
    if isinstance(path_or_bytes, (bytes, bytearray)):
        data = bytes(path_or_bytes)
        check(L.reina_rt_gltf_load_memory(data, len(data), flags, ctypes.byref(out)))
    else:
        check(L.reina_rt_gltf_load(str(path_or_bytes).encode(), flags, ctypes.byref(out)))
    try:
        s = out.contents
        models = []
        for i in range(s.num_models):
            m = s.models[i]
            models.append(ModelData(
                _copy(m.vertices, 4 * m.num_vertices, np.float32).reshape(-1, 4),
                _copy(m.indices, m.num_indices, np.uint32),
                _copy(m.tbns, 9 * m.num_tbns, np.float32).reshape(-1, 3, 3),
                _copy(m.tbn_indices, m.num_tbn_indices, np.uint32),
                _copy(m.tex_coords, 2 * m.num_tex_coords, np.float32).reshape(-1, 2),
                _copy(m.tex_indices, m.num_tex_indices, np.uint32)))
        images = []
        for i in range(s.num_images):
            im = s.images[i]
            images.append(_copy(im.rgba8, 4 * im.width * im.height, np.uint8).reshape(im.height, im.width, 4))
        instances = []
        for i in range(s.num_instances):
            inst = s.instances[i]
            t = np.array(inst.transform, np.float32).reshape(4, 4).T   # column-major -> row-major
            instances.append((int(inst.model), t, _material(inst.material)))
    finally:
        L.reina_rt_gltf_free(out)
    return models, images, instances


def load_scene(path_or_bytes, skip_undecodable_images=False, scene=None):
    """gltf::loadScene: defines the primitives' models, the images as textures
    and the instances in a Scene (a new one, or `scene`). Returns the Scene."""
    models, images, instances = load(path_or_bytes, skip_undecodable_images)
    s = scene if scene is not None else Scene()
    first_tex = len(s._textures)
    ids = [s.define_object(m) for m in models]
    for img in images:
        s.add_texture(img)
    for model, t, mat in instances:
        if first_tex:
            if mat.textureID >= 0:
                mat.textureID += first_tex
            if mat.normalMapID >= 0:
                mat.normalMapID += first_tex
        s.add_instance(ids[model], t, mat)
    return s
