"""The C-ABI boundary without a GPU: struct layouts (polyglot/raytrace.h byte
offsets), every declared entry point exported, argument validation and
error codes, and the Models::addModel buffer packing."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import reina_rt
from reina_rt import abi, scenes
from reina_rt.scene import Material, ModelData, Scene, make_desc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "reina_rt.h")

# Appendix of byte offsets: polyglot/raytrace.h:18-54, src/scene/Instances.h:15-27
OFFSETS = {
    "ReinaInstanceProperties": dict(indicesOffset=0, albedo=4, emission=16, tbnsIndicesOffset=28,
                                    texIndicesOffset=32, roughness=36, ior=40, interpNormals=44, absorption=48,
                                    textureID=52, normalMapTexID=56, bumpMapTexID=60, cullBackface=64,
                                    anisotropic=68, subsurface=72, clearcoatGloss=76, sheenTint=80,
                                    specularTint=92, metallic=104, clearcoat=108, specularTransmission=112,
                                    sheen=116),
    "RtPushConsts": dict(invView=0, invProjection=64, sampleBatch=128, totalEmissiveWeight=132, focusDist=136,
                         defocusMultiplier=140, directClamp=144, indirectClamp=148, samplesPerPixel=152,
                         maxBounces=156),
    "ReinaInstanceData": dict(transform=0, materialOffset=64, cdfRangeStart=68, cdfRangeEnd=72, indexOffset=76,
                              emission=80, weight=92, area=96, cullBackface=100, padding=104),
}
SIZES = {"ReinaInstanceProperties": 120, "RtPushConsts": 160, "ReinaInstanceData": 112}


def test_header_layout_with_gcc():
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void){"]
    for st, fields in OFFSETS.items():
        src.append('printf("%s %%zu\\n", sizeof(%s));' % (st, st))
        for f in fields:
            src.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (st, f, st, f))
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-std=c11", "-o", exe, c], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(line.split() for line in out if line)
    for st, fields in OFFSETS.items():
        assert int(got[st]) == SIZES[st]
        for f, off in fields.items():
            assert int(got[st + "." + f]) == off, (st, f)


def test_ctypes_structs_match_header():
    """Every ctypes mirror of a header struct has the C layout (gcc)."""
    structs = [abi.ReinaSceneDesc, abi.ReinaSceneInfo, abi.ReinaTraceStats, abi.ReinaTuning, abi.ReinaServerDesc,
               abi.ReinaRequestStatus, abi.ReinaServerStats, abi.ReinaTextureDesc, abi.ReinaModelData,
               abi.ReinaImage]
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void){"]
    for st in structs:
        n = st.__name__
        src.append('printf("%s %%zu\\n", sizeof(%s));' % (n, n))
        for f, _ in st._fields_:
            src.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (n, f, n, f))
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-std=c11", "-o", exe, c], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(line.split() for line in out if line)
    for st in structs:
        n = st.__name__
        assert int(got[n]) == ctypes.sizeof(st), n
        for f, _ in st._fields_:
            assert int(got[n + "." + f]) == getattr(st, f).offset, (n, f)


def test_numpy_dtypes_match():
    for st, dt in (("ReinaInstanceProperties", abi.INSTANCE_PROPERTIES), ("RtPushConsts", abi.PUSH_CONSTS),
                   ("ReinaInstanceData", abi.INSTANCE_DATA)):
        assert dt.itemsize == SIZES[st]
        for f, off in OFFSETS[st].items():
            assert dt.fields[f][1] == off


def test_library_sizes_and_version():
    L = reina_rt.lib()
    v, a, b, c = (ctypes.c_uint32() for _ in range(4))
    assert L.reina_rt_abi_version(ctypes.byref(v), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == 0
    assert (v.value, a.value, b.value, c.value) == (1, 120, 160, 112)


def test_every_declared_symbol_is_exported():
    text = open(HEADER).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char\*)\s+(reina_rt_\w+)\s*\(", text, re.M))
    assert len(declared) >= 10
    L = reina_rt.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert declared == set(reina_rt.exported_symbols())
    nm = subprocess.run(["nm", "-D", "--defined-only", reina_rt.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (reina_rt_\w+)", nm))
    assert declared <= exported


def _create(arrays, **overrides):
    d = make_desc(arrays)
    for k, v in overrides.items():
        setattr(d, k, v)
    h = ctypes.c_void_p()
    rc = reina_rt.lib().reina_rt_scene_create(ctypes.byref(d), ctypes.byref(h))
    msg = reina_rt.lib().reina_rt_last_error().decode()
    if rc == 0:
        reina_rt.lib().reina_rt_scene_destroy(h)
    return rc, msg


def test_create_rejects_bad_descriptions(cornell_arrays):
    """Validation happens before any device work (C-ABI status codes replace the
    reference's std::runtime_error throws, src/main.cpp:11-14)."""
    arrays, _ = cornell_arrays
    rc, msg = _create(arrays, struct_size=8)
    assert rc == abi.ERR_ABI and "struct_size" in msg
    bad = {k: v.copy() for k, v in arrays.items()}
    bad["instances"][0]["modelID"] = 99
    assert _create(bad)[0] == abi.ERR_INVALID_ARG
    bad = {k: v.copy() for k, v in arrays.items()}
    bad["instances"][1]["materialIdx"] = 4
    assert _create(bad)[0] == abi.ERR_INVALID_ARG
    bad = {k: v.copy() for k, v in arrays.items()}
    bad["indices"][5] = len(bad["vertices"]) + 3
    assert _create(bad)[0] == abi.ERR_INVALID_ARG
    bad = {k: v.copy() for k, v in arrays.items()}
    bad["props"][2]["textureID"] = 0        # the scene has no textures
    rc, msg = _create(bad)
    assert rc == abi.ERR_INVALID_ARG and "texture index" in msg
    bad = {k: v.copy() for k, v in arrays.items()}
    bad["instances"][0]["instancePropertiesID"] = 1000
    assert _create(bad)[0] == abi.ERR_INVALID_ARG
    L = reina_rt.lib()
    assert L.reina_rt_scene_create(None, ctypes.byref(ctypes.c_void_p())) == abi.ERR_INVALID_ARG
    assert L.reina_rt_trace(None, None, None, 0, 0, None) == abi.ERR_INVALID_ARG
    assert L.reina_rt_scene_destroy(None) == abi.ERR_INVALID_ARG


def test_create_without_device_reports_device_error(cornell_arrays):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    rc, msg = _create(cornell_arrays[0])
    assert rc == abi.ERR_DEVICE


def test_models_packing_matches_addModel():
    """Offsets exactly as Models::addModel (src/scene/Models.cpp:24-92)."""
    s = Scene()
    v1 = np.array([[0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1]], np.float32)
    t1 = np.tile(np.eye(3, dtype=np.float32), (3, 1, 1))
    m1 = ModelData(v1, np.array([0, 1, 2], np.uint32), t1, np.array([0, 1, 2], np.uint32))
    v2 = np.concatenate([v1, [[1, 1, 0, 1]]]).astype(np.float32)
    t2 = np.tile(np.eye(3, dtype=np.float32), (4, 1, 1))
    m2 = ModelData(v2, np.array([0, 1, 2, 1, 3, 2], np.uint32), t2, np.array([0, 1, 2, 1, 3, 2], np.uint32),
                   np.array([[0, 0], [1, 0], [0, 1], [1, 1]], np.float32),
                   np.array([0, 1, 2, 1, 0xFFFFFFFF, 2], np.uint32))
    a = s.define_object(m1)
    b = s.define_object(m2)
    s.add_instance(b, np.eye(4), Material())
    s.add_instance(a, np.eye(4), Material())
    ar = s.build_arrays()
    r = ar["models"]
    assert (r[0]["firstVertex"], r[0]["indexOffset"], r[0]["texIndexOffset"]) == (0, 0, 0xFFFFFFFF)
    assert (r[1]["firstVertex"], r[1]["firstNormal"], r[1]["indexOffset"], r[1]["texIndexOffset"]) == (3, 3, 3, 0)
    assert ar["indices"].tolist() == [0, 1, 2, 3, 4, 5, 4, 6, 5]
    assert ar["tbn_indices"].tolist() == [0, 1, 2, 3, 4, 5, 4, 6, 5]
    assert ar["tex_indices"].tolist() == [0, 1, 2, 1, 0xFFFFFFFF, 2]
    p = ar["props"]
    assert p[0]["indicesOffset"] == 3 and p[1]["indicesOffset"] == 0
    assert p[0]["texIndicesOffset"] == 0 and p[1]["texIndicesOffset"] == 0xFFFFFFFF
    assert ar["instances"]["instancePropertiesID"].tolist() == [0, 1]
    assert ar["instances"]["modelID"].tolist() == [1, 0]


def test_bvh_build_host_rejects_null():
    L = reina_rt.lib()
    nb = ctypes.c_uint64(0)
    assert L.reina_rt_bvh_build_host(None, 5, None, ctypes.byref(nb), None, ctypes.byref(nb), None) == abi.ERR_INVALID_ARG


@pytest.mark.parametrize("field,value,word", [("slot_tile", 300, "slot_tile"), ("slot_tile", 8192, "slot_tile"),
                                              ("graph_iters", 7, "graph_iters"), ("graph_iters", 30, "graph_iters"),
                                              ("admit_order", 3, "admit_order")])
def test_server_create_rejects_bad_fields(field, value, word):
    """The server's slot tile, graph length and admission order are checked
    before any device work: INVALID_ARG naming the field, GPU or not."""
    with pytest.raises(reina_rt.ReinaError) as e:
        reina_rt.Server(path_budget_bytes=1 << 26, **{field: value})
    assert e.value.code == abi.ERR_INVALID_ARG and word in str(e.value)


def test_server_budget_counts_slot_tiles():
    """path_budget_bytes must hold 8 slot tiles of the requested size."""
    with pytest.raises(reina_rt.ReinaError) as e:
        reina_rt.Server(path_budget_bytes=8 * 4096 * abi.SERVER_BYTES_PER_PIXEL - 1, slot_tile=4096)
    assert e.value.code == abi.ERR_INVALID_ARG and "8 slot tiles" in str(e.value)
