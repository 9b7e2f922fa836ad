"""Scenes built from files through the library's loaders render on the GPU
exactly as the oracle renders the same arrays: the OBJ scene of
src/Reina.cpp:103-115 (models/cornell_box.obj with the palette PNG
textures/cornell_texture.png, models/cornell_light.obj and a Disney UV sphere),
with and without next-event estimation."""
import os

import numpy as np
import pytest

import oracle
import reina_rt
from reina_rt import abi, scenes
from reina_rt.scene import make_desc

pytestmark = pytest.mark.gpu

ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "assets")


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _same(a, b):
    return bool(((_bits(a) == _bits(b)) | (np.isnan(a) & np.isnan(b))).all())


@pytest.fixture(scope="module")
def obj_scene():
    s, cam = scenes.obj_cornell(ASSETS, subject="uv_sphere.obj")
    return s.build_arrays(), cam


@pytest.mark.parametrize("nee", [False, True], ids=["plain", "nee"])
def test_obj_scene_equals_oracle(cuda, obj_scene, nee):
    torch = cuda
    arrays, cam = obj_scene
    W = H = 40
    pc = reina_rt.push_consts(cam["eye"], cam["center"], cam["fov"], W, H, spp=4, max_bounces=8)
    dev = reina_rt.DeviceScene(arrays)
    orc = oracle.OracleScene(arrays, make_desc)
    if nee:
        dev.set_trace_options(abi.TRACE_NEE)
        orc.set_nee(True)
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    dev.trace(pc, img.data_ptr(), W, H, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = img.cpu().numpy()
    ref, nrays = orc.render(pc, W, H, use_bvh=False)
    assert _same(got, ref), "pixels differing: %d" % (_bits(got) != _bits(ref)).any(-1).sum()
    assert dev.stats().extension_rays == nrays
    assert np.isfinite(ref).all() and ref[..., :3].mean() > 0.01
    dev.close()


@pytest.mark.parametrize("nee", [False, True], ids=["plain", "nee"])
def test_gltf_scene_equals_oracle(cuda, nee):
    """The glTF asset of test_gltf.py (Disney materials with transmission, a
    base-colour texture, a normal map, emissive quads as lights, quantized and
    sparse accessors, a node hierarchy) through reina_rt.gltf.load_scene."""
    import test_gltf
    from reina_rt import gltf
    torch = cuda
    arrays = gltf.load_scene(test_gltf.build_asset().glb()).build_arrays()
    W = H = 40
    pc = reina_rt.push_consts((2.0, 4.0, 9.0), (1.0, 0.0, 0.0), 45.0, W, H, spp=4, max_bounces=8)
    dev = reina_rt.DeviceScene(arrays)
    orc = oracle.OracleScene(arrays, make_desc)
    if nee:
        dev.set_trace_options(abi.TRACE_NEE)
        orc.set_nee(True)
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    dev.trace(pc, img.data_ptr(), W, H, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = img.cpu().numpy()
    ref, nrays = orc.render(pc, W, H, use_bvh=False)
    assert _same(got, ref), "pixels differing: %d" % (_bits(got) != _bits(ref)).any(-1).sum()
    assert dev.stats().extension_rays == nrays
    # mesh 2 faces +x without TANGENT: its frame T = (1, 0, 0) is parallel to
    # N, the hit shader's re-orthogonalised tangent is NaN and every sample
    # there is dropped (the reference's result for such a primitive)
    assert np.nanmean(ref[..., :3]) > 0.001
    dev.close()
