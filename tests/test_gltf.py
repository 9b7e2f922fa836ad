"""glTF scenes: the library's native loader (reina_rt_gltf_load, replacing
reina::scene::gltf::loadScene, src/scene/gltf/gltfloader.cpp) against the
oracle restatement (oracle/ingest.gltf_load) and closed forms, on assets this
test writes (the reference ships no glTF file, and fastgltf / MikkTSpace are
not in this image).

Pinned here: accessor decoding (float, normalized and plain integers, strided
views, sparse substitution), index widths and generated indices, the tangent
rules (TANGENT w, the MikkTSpace result over zero texture coordinates), the
material mapping of gltfloader.cpp:366-428, node transforms against float64
composition, model / instance order, and images against Pillow.
"""
import base64
import io
import json
import os
import struct

import numpy as np
import pytest

import reina_rt
from reina_rt import gltf
from oracle import ingest

F = np.float32


def _png(rgba):
    from PIL import Image
    b = io.BytesIO()
    Image.fromarray(np.asarray(rgba, np.uint8)).save(b, format="PNG")
    return b.getvalue()


class Builder:
    def __init__(self):
        self.doc = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": [], "meshes": [],
                    "nodes": [], "scenes": [], "materials": [], "textures": [], "images": []}
        self.bin = bytearray()

    def view(self, data, stride=None):
        while len(self.bin) % 4:
            self.bin.append(0)
        v = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.bin += data
        self.doc["bufferViews"].append(v)
        return len(self.doc["bufferViews"]) - 1

    def accessor(self, arr, ctype, kind, normalized=False, view=None, offset=0, count=None, sparse=None):
        dt = {5120: "<i1", 5121: "<u1", 5122: "<i2", 5123: "<u2", 5125: "<u4", 5126: "<f4"}[ctype]
        a = {"componentType": ctype, "type": kind}
        if arr is not None:
            arr = np.asarray(arr).astype(dt)
            a["count"] = len(arr)
            a["bufferView"] = self.view(arr.tobytes()) if view is None else view
        else:
            a["count"] = count
            if view is not None:
                a["bufferView"] = view
        if offset:
            a["byteOffset"] = offset
        if normalized:
            a["normalized"] = True
        if sparse is not None:
            idx, vals = sparse
            a["sparse"] = {"count": len(idx),
                           "indices": {"bufferView": self.view(np.asarray(idx, "<u2").tobytes()), "componentType": 5123},
                           "values": {"bufferView": self.view(np.asarray(vals).astype(dt).tobytes())}}
        self.doc["accessors"].append(a)
        return len(self.doc["accessors"]) - 1

    def glb(self):
        js = json.dumps(self.doc | {"buffers": [{"byteLength": len(self.bin)}]}).encode()
        js += b" " * (-len(js) % 4)
        bn = bytes(self.bin) + b"\0" * (-len(self.bin) % 4)
        out = struct.pack("<III", 0x46546C67, 2, 12 + 8 + len(js) + 8 + len(bn))
        out += struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(bn), 0x004E4942) + bn
        return out

    def gltf_data_uri(self):
        uri = "data:application/octet-stream;base64," + base64.b64encode(bytes(self.bin)).decode()
        return json.dumps(self.doc | {"buffers": [{"uri": uri, "byteLength": len(self.bin)}]}).encode()

    def gltf_files(self, d, name="scene"):
        (d / (name + ".bin")).write_bytes(bytes(self.bin))
        doc = self.doc | {"buffers": [{"uri": name + ".bin", "byteLength": len(self.bin)}]}
        p = d / (name + ".gltf")
        p.write_text(json.dumps(doc))
        return str(p)


def _rot(axis, deg):
    h = np.radians(deg) / 2
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    return [float(a[0] * np.sin(h)), float(a[1] * np.sin(h)), float(a[2] * np.sin(h)), float(np.cos(h))]


def build_asset(external_image_uri=None):
    b = Builder()
    d = b.doc
    tex0 = _png(np.random.default_rng(1).integers(0, 256, (5, 7, 4)))
    tex1 = _png(np.random.default_rng(2).integers(0, 256, (4, 4, 3)))
    d["images"] = [{"bufferView": b.view(tex0), "mimeType": "image/png"},
                   {"uri": external_image_uri or ("data:image/png;base64," + base64.b64encode(tex1).decode())}]
    d["textures"] = [{"source": 0}, {"source": 1}]
    # mesh 0: a quad; u16-normalized UVs, u16 indices, VEC4 tangents with w = -1 on half
    quad_p = [[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]]
    quad_n = [[0, 1, 0]] * 4
    quad_t = [[1, 0, 0, 1], [1, 0, 0, 1], [1, 0, 0, -1], [1, 0, 0, -1]]
    quad_uv = [[0, 0], [65535, 0], [65535, 32768], [0, 65535]]
    prim0 = {"attributes": {"POSITION": b.accessor(quad_p, 5126, "VEC3"), "NORMAL": b.accessor(quad_n, 5126, "VEC3"),
                            "TANGENT": b.accessor(quad_t, 5126, "VEC4"),
                            "TEXCOORD_0": b.accessor(quad_uv, 5123, "VEC2", normalized=True)},
             "indices": b.accessor([0, 1, 2, 0, 2, 3], 5123, "SCALAR"), "material": 0}
    # mesh 1: 7-vertex soup, i16-normalized positions with a sparse override,
    # no TANGENT (the MikkTSpace rule), no TEXCOORD_0, no indices; and a second
    # primitive with no material
    rng = np.random.default_rng(3)
    soup = rng.integers(-32768, 32767, (7, 3))
    nrm = rng.normal(size=(7, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(F)
    prim1 = {"attributes": {"POSITION": b.accessor(soup, 5122, "VEC3", normalized=True,
                                                   sparse=([2, 5], [[100, -200, 300], [-32768, 32767, 0]])),
                            "NORMAL": b.accessor(nrm, 5126, "VEC3")}, "material": 1}
    prim1b = {"attributes": {"POSITION": b.accessor([[0, 0, 0], [1, 0, 0], [0, 1, 0]], 5126, "VEC3"),
                             "NORMAL": b.accessor([[0, 0, 1]] * 3, 5126, "VEC3"),
                             "TANGENT": b.accessor([[1, 0, 0]] * 3, 5126, "VEC3")},
              "indices": b.accessor([0, 1, 2], 5121, "SCALAR")}
    # mesh 2: interleaved position + normal (stride 24), u32 indices, u8 UVs
    inter = np.zeros((4, 6), "<f4")
    inter[:, :3] = [[0, 0, 0], [0, 2, 0], [0, 2, 2], [0, 0, 2]]
    inter[:, 3:] = [1, 0, 0]
    v = b.view(inter.tobytes(), stride=24)
    prim2 = {"attributes": {"POSITION": b.accessor(None, 5126, "VEC3", view=v, count=4),
                            "NORMAL": b.accessor(None, 5126, "VEC3", view=v, offset=12, count=4),
                            "TEXCOORD_0": b.accessor([[0, 0], [255, 0], [255, 255], [0, 128]], 5121, "VEC2",
                                                     normalized=True)},
             "indices": b.accessor([0, 1, 2, 0, 2, 3], 5125, "SCALAR"), "material": 2}
    # mesh 3: used by no node of the default scene
    prim3 = {"attributes": {"POSITION": b.accessor([[0, 0, 0], [1, 0, 0], [0, 1, 0]], 5126, "VEC3"),
                            "NORMAL": b.accessor([[0, 0, 1]] * 3, 5126, "VEC3")}}
    d["meshes"] = [{"primitives": [prim0]}, {"primitives": [prim1, prim1b]}, {"primitives": [prim2]},
                   {"primitives": [prim3]}]
    d["materials"] = [
        {"pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.4, 0.6, 0.5], "metallicFactor": 0.3,
                                  "roughnessFactor": 0.05, "baseColorTexture": {"index": 0}},
         "emissiveFactor": [1.0, 0.5, 0.0], "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 4.0}}},
        {"pbrMetallicRoughness": {"roughnessFactor": 0.9}, "doubleSided": False, "normalTexture": {"index": 1},
         "extensions": {"KHR_materials_transmission": {"transmissionFactor": 0.5}}},
        {"emissiveFactor": [2.0, 2.0, 2.0], "emissiveTexture": {"index": 0}, "doubleSided": True,
         "pbrMetallicRoughness": {"roughnessFactor": 0.4, "metallicFactor": 0.0}},
    ]
    d["nodes"] = [
        {"translation": [1.5, -2.0, 0.25], "rotation": _rot([0, 1, 0], 30), "scale": [2.0, 0.5, 1.0],
         "children": [1, 3]},
        {"matrix": [0, 0, -1, 0, 0, 1, 0, 0, 1, 0, 0, 0, 3, 0, 0, 1], "children": [2]},
        {"mesh": 1, "rotation": _rot([1, 2, 3], 75)},
        {"mesh": 0},
        {"mesh": 0, "translation": [0, 5, 0]},
        {"mesh": 2, "scale": [0.5, 0.5, 0.5]},
        {"mesh": 3},
    ]
    d["scenes"] = [{"nodes": [6]}, {"nodes": [0, 4, 5]}]
    d["scene"] = 1
    return b


def assert_same(lib_result, ref):
    models, images, instances = lib_result
    assert len(models) == len(ref["models"])
    for md, r in zip(models, ref["models"]):
        for k, got in (("vertices", md.vertices), ("tbns", md.tbns), ("texCoords", md.texCoords)):
            assert np.array_equal(np.ascontiguousarray(got, F).view(np.uint32),
                                  np.ascontiguousarray(r[k], F).view(np.uint32)), k
        assert np.array_equal(md.indices, r["indices"])
        assert np.array_equal(md.tbnsIndices, r["indices"]) and np.array_equal(md.texIndices, r["indices"])
    assert len(images) == len(ref["images"])
    for a, b in zip(images, ref["images"]):
        assert np.array_equal(a, b)
    assert len(instances) == len(ref["instances"])
    for (m, t, mat), (rm, rt_, rmat) in zip(instances, ref["instances"]):
        assert m == rm
        assert np.array_equal(t.T.reshape(16).view(np.uint32), rt_.view(np.uint32))
        for k, v in rmat.items():
            got = getattr(mat, k)
            if isinstance(v, tuple):
                assert np.array_equal(np.asarray(got, F), np.asarray(v, F)), k
            else:
                assert np.float32(got) == np.float32(v) if not isinstance(v, bool) else got == v, k


@pytest.mark.parametrize("form", ["glb", "data_uri", "files"])
def test_gltf_matches_oracle(form, tmp_path):
    if form == "files":
        (tmp_path / "tex1.png").write_bytes(_png(np.random.default_rng(2).integers(0, 256, (4, 4, 3))))
        b = build_asset(external_image_uri="tex1.png")
        path = b.gltf_files(tmp_path)
        assert_same(gltf.load(path), ingest.gltf_load(open(path, "rb").read(), str(tmp_path)))
    else:
        b = build_asset()
        data = b.glb() if form == "glb" else b.gltf_data_uri()
        assert_same(gltf.load(data), ingest.gltf_load(data))


@pytest.fixture(scope="module")
def loaded():
    return gltf.load(build_asset().glb())


def test_models_and_instances_follow_the_default_scene(loaded):
    """Scene 1 uses meshes 0, 1, 2 (mesh 3 only in scene 0): models are their
    primitives in mesh order; instances follow the depth-first node walk
    (node 0 -> 1 -> 2 (mesh 1: two primitives), node 3 (mesh 0), node 4
    (mesh 0 again), node 5 (mesh 2))."""
    models, images, instances = loaded
    assert len(models) == 4                                      # 1 + 2 + 1 primitives
    assert [m for m, _, _ in instances] == [1, 2, 0, 0, 3]
    assert len(images) == 2


def test_accessor_decoding(loaded):
    models, _, _ = loaded
    uv = models[0].texCoords
    assert uv[1, 0] == 1.0 and uv[2, 1] == F(32768) / F(65535) and uv[3, 1] == 1.0
    assert models[0].indices.tolist() == [0, 1, 2, 0, 2, 3]
    soup = models[1].vertices[:, :3]
    assert soup[2].tolist() == [F(100) / F(32767), F(-200) / F(32767), F(300) / F(32767)]   # sparse
    assert soup[5].tolist() == [-1.0, 1.0, 0.0]                  # -32768 clamps to -1
    assert models[1].indices.tolist() == list(range(7))          # generated indices
    assert models[2].indices.tolist() == [0, 1, 2]               # u8
    inter = models[3].vertices[:, :3]
    assert inter.tolist() == [[0, 0, 0], [0, 2, 0], [0, 2, 2], [0, 0, 2]]     # stride 24
    assert np.allclose(models[3].tbns[:, 2], [1, 0, 0])
    assert models[3].texCoords[3, 1] == F(128) / F(255)


def test_tangent_rules(loaded):
    """VEC4 tangents: B = cross(N, T) * w; VEC3: cross(N, T); none: the
    MikkTSpace frame over zero texture coordinates -- T = (1,0,0),
    B = (0,1,0) for each whole vertex triple, zeros past the last one."""
    models, _, _ = loaded
    q = models[0].tbns
    assert q[:, 0].tolist() == [[1, 0, 0]] * 4
    assert q[:2, 1].tolist() == [[0, 0, -1]] * 2 and q[2:, 1].tolist() == [[0, 0, 1]] * 2
    s = models[1].tbns
    assert s[:6, 0].tolist() == [[1, 0, 0]] * 6 and s[:6, 1].tolist() == [[0, 1, 0]] * 6
    assert not s[6, :2].any()
    t = models[2].tbns
    assert t[:, 1].tolist() == [[0, 1, 0]] * 3                   # cross((0,0,1), (1,0,0))


def test_material_mapping(loaded):
    """gltfloader.cpp:366-428: Disney, interpolated normals; roughness clamped
    to [0.1, 0.7]; emission = factor * strength unless an emissive texture;
    no culling when doubleSided or transmissive; ior 1.5; texture IDs are image
    indices; a primitive without material gets the default white Disney."""
    _, _, inst = loaded
    m1, m1b, m0, _, m2 = (mat for _, _, mat in inst)
    assert (m0.materialIdx, m0.interpNormals, m0.ior) == (3, True, F(1.5))
    assert m0.roughness == F(0.1) and m0.metallic == F(0.3)
    assert np.array_equal(np.asarray(m0.albedo, F), np.asarray([0.2, 0.4, 0.6], F))
    assert np.array_equal(np.asarray(m0.emission, F), np.asarray([4.0, 2.0, 0.0], F))
    assert m0.cullBackface and m0.textureID == 0 and m0.normalMapID == -1
    assert m1.roughness == F(0.7) and m1.specularTransmission == F(0.5) and not m1.cullBackface
    assert m1.normalMapID == 1 and m1.metallic == F(1.0)
    assert tuple(m2.emission) == (0.0, 0.0, 0.0) and not m2.cullBackface and m2.roughness == F(0.4)
    assert m1b.albedo == (1.0, 1.0, 1.0) and m1b.roughness == 0.0 and m1b.metallic == 0.0
    assert m1b.textureID == -1 and not m1b.cullBackface


def _trs(t, q, s):
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = R * np.asarray(s)[None, :]
    M[:3, 3] = t
    return M


def test_node_transforms_against_float64(loaded):
    """world = parent * local, local = matrix or T * R * S (glTF 2.0 spec),
    composed in float64 from the file's numbers: agrees to float32 rounding."""
    _, _, inst = loaded
    n0 = _trs([1.5, -2.0, 0.25], _rot([0, 1, 0], 30), [2.0, 0.5, 1.0])
    n1 = np.array([0, 0, -1, 0, 0, 1, 0, 0, 1, 0, 0, 0, 3, 0, 0, 1], float).reshape(4, 4).T
    n2 = _trs([0, 0, 0], _rot([1, 2, 3], 75), [1, 1, 1])
    want = [n0 @ n1 @ n2, n0 @ n1 @ n2, n0, _trs([0, 5, 0], [0, 0, 0, 1], [1, 1, 1]),
            _trs([0, 0, 0], [0, 0, 0, 1], [0.5, 0.5, 0.5])]
    for (_, t, _), w in zip(inst, want):
        assert np.allclose(t, w, rtol=0, atol=4e-6 * max(1.0, np.abs(w).max()))
    # axis-aligned parts are exact
    assert inst[3][1].tolist() == _trs([0, 5, 0], [0, 0, 0, 1], [1, 1, 1]).tolist()
    assert inst[4][1].tolist() == _trs([0, 0, 0], [0, 0, 0, 1], [0.5, 0.5, 0.5]).tolist()


def test_images_unflipped_as_pillow(loaded):
    from PIL import Image
    _, images, _ = loaded
    want0 = np.asarray(Image.open(io.BytesIO(_png(np.random.default_rng(1).integers(0, 256, (5, 7, 4))))).convert("RGBA"))
    assert np.array_equal(images[0], want0)
    assert images[1].shape == (4, 4, 4) and (images[1][..., 3] == 255).all()


def test_load_scene_builds_arrays():
    s = gltf.load_scene(build_asset().glb())
    arr = s.build_arrays()
    assert len(arr["models"]) == 4 and len(arr["instances"]) == 5 and len(arr["textures"]) == 2
    # into an existing scene: texture IDs shift past its textures
    base = reina_rt.Scene()
    base.add_texture(np.zeros((2, 2, 4), np.uint8))
    s2 = gltf.load_scene(build_asset().glb(), scene=base)
    assert s2._props[2]["textureID"] == 1 and s2._props[0]["normalMapTexID"] == 2


def test_gltf_errors(tmp_path):
    def code(data, flags_skip=False):
        with pytest.raises(reina_rt.ReinaError) as e:
            gltf.load(data, flags_skip)
        return e.value.code

    b = build_asset()
    b.doc["scenes"] = []
    b.doc.pop("scene")
    assert code(b.glb()) == -1                                     # no scenes
    b = build_asset()
    del b.doc["meshes"][0]["primitives"][0]["attributes"]["NORMAL"]
    assert code(b.glb()) == -1                                     # normals required
    b = build_asset()
    b.doc["extensionsRequired"] = ["KHR_draco_mesh_compression"]
    assert code(b.glb()) == -6
    b = build_asset()
    b.doc["accessors"][0]["count"] = 1000                          # past its bufferView
    assert code(b.glb()) == -1
    b = build_asset()
    from PIL import Image
    jb = io.BytesIO()
    Image.fromarray(np.zeros((8, 8, 3), np.uint8)).save(jb, format="GIF")   # not PNG or JPEG: unsupported
    b.doc["images"][1] = {"uri": "data:image/jpeg;base64," + base64.b64encode(jb.getvalue()).decode()}
    assert code(b.glb()) == -6
    models, images, inst = gltf.load(b.glb(), True)                # skipped: texture ID -1
    assert len(images) == 1 and inst[0][2].normalMapID == -1
    assert code(b'{"asset": {"version": "2.0"}, "scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}], "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1}}]}], "buffers": [{"uri": "x.bin", "byteLength": 4}]}') == -1
    with pytest.raises(reina_rt.ReinaError):
        gltf.load(str(tmp_path / "missing.gltf"))
    with pytest.raises(reina_rt.ReinaError):
        gltf.load(b"{ not json")
