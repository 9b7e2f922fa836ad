"""Scene ingest: the library's native OBJ and PNG loaders (host code, no GPU)
against the oracle restatement (oracle/ingest.py), closed forms and Pillow.

The loaders replace Models::getObjData (src/scene/Models.cpp:117-175, Assimp)
and the stb_image calls of src/graphics/Image.cpp:10-41. Neither Assimp nor
stb is in this image; the OBJ oracle restates Assimp's published algorithm and
is pinned here by closed forms and by the OBJ text itself, the PNG side by
Pillow (an independent decoder) and by images whose samples the test chose.
"""
import glob
import os
import struct
import zlib
from fractions import Fraction

import numpy as np
import pytest

import reina_rt
from reina_rt.scene import Scene, load_image, load_obj
from oracle import ingest

HERE = os.path.dirname(os.path.abspath(__file__))
ASSETS = os.path.join(HERE, "golden", "assets")
REF_MODELS = "/root/reference/models"
REF_TEXTURES = "/root/reference/textures"


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def assert_matches_oracle(md, ref):
    assert same_bits(md.vertices, ref["vertices"])
    assert np.array_equal(md.indices, ref["indices"])
    assert np.array_equal(md.tbnsIndices, ref["indices"]) and np.array_equal(md.texIndices, ref["indices"])
    if ref["tbns"] is None:
        assert len(md.tbns) == 0
    else:
        assert same_bits(md.tbns, ref["tbns"])
    if ref["texCoords"] is None:
        assert md.texCoords is None
    else:
        assert same_bits(md.texCoords, ref["texCoords"])


# ---- library == oracle --------------------------------------------------------------
ASSET_OBJS = sorted(glob.glob(os.path.join(ASSETS, "*.obj")))


@pytest.mark.parametrize("path", ASSET_OBJS, ids=[os.path.basename(p) for p in ASSET_OBJS])
def test_obj_fixture_matches_oracle(path):
    text = open(path, "rb").read()
    assert_matches_oracle(load_obj(path), ingest.obj_load(text))
    assert_matches_oracle(load_obj(text), ingest.obj_load(text))


REF_OBJS = sorted(glob.glob(os.path.join(REF_MODELS, "*.obj")))


@pytest.mark.skipif(not REF_OBJS, reason="the reference's models/ directory is not present")
@pytest.mark.parametrize("name", [os.path.basename(p) for p in REF_OBJS] or ["none"])
def test_reference_models_match_oracle(name):
    path = os.path.join(REF_MODELS, name)
    assert_matches_oracle(load_obj(path), ingest.obj_load(open(path, "rb").read()))


# ---- closed forms ----------------------------------------------------------------
def frames(md):
    """per-vertex (T, B, N) rows of the packed column-major mat3s"""
    return md.tbns.reshape(-1, 3, 3)


@pytest.mark.parametrize("src", ["library", "oracle"])
def test_quad_closed_form(src):
    """models/quad.obj: the convex quad splits at vertex 0; v flips; the frame
    is T = dP/du = +x, B = dP/dv of the file's v = +y, N = +z."""
    path = os.path.join(ASSETS, "quad.obj")
    md = load_obj(path) if src == "library" else _as_md(ingest.obj_load(open(path, "rb").read()))
    assert md.vertices.tolist() == [[-1, -1, 0, 1], [1, -1, 0, 1], [1, 1, 0, 1], [-1, 1, 0, 1]]
    assert md.indices.tolist() == [0, 1, 2, 0, 2, 3]
    assert md.texCoords.tolist() == [[0, 1], [1, 1], [1, 0], [0, 0]]
    for f in frames(md):
        assert f.tolist() == [[1, 0, 0], [0, 1, 0], [0, 0, 1]]


def _as_md(d):
    return reina_rt.ModelData(d["vertices"], d["indices"], np.zeros((0, 3, 3), np.float32) if d["tbns"] is None
                              else d["tbns"], d["indices"], d["texCoords"], d["indices"])


def _obj_faces(text):
    """(v, vt, vn) index triples per face, 0-based, from the OBJ text"""
    v, vt, vn, faces = [], [], [], []
    for line in text.splitlines():
        p = line.split("#")[0].split()
        if not p:
            continue
        if p[0] == "v":
            v.append([float(x) for x in p[1:4]])
        elif p[0] == "vt":
            vt.append([float(x) for x in p[1:3]])
        elif p[0] == "vn":
            vn.append([float(x) for x in p[1:4]])
        elif p[0] == "f":
            face = []
            for t in p[1:]:
                s = (t.split("/") + ["", ""])[:3]
                face.append(tuple(int(x) - 1 if x else None for x in s))
            faces.append(face)
    return np.array(v), np.array(vt) if vt else None, np.array(vn) if vn else None, faces


@pytest.mark.parametrize("name", ["cornell_box.obj", "blender_cube.obj"])
def test_box_frames_closed_form(name):
    """Planar faces with affine UVs: per face, T is the unit dP/du, B the unit
    dP/dv (file v, before the flip), N the file's normal (the projection of
    T and B on the plane is exact). Coincident corners of adjacent faces have
    perpendicular normals, so no smoothing mixes faces."""
    path = os.path.join(ASSETS, name)
    text = open(path).read()
    V, VT, VN, faces = _obj_faces(text)
    md = load_obj(path)
    fr = frames(md)
    tri = md.indices.reshape(-1, 3)
    t = 0
    for face in faces:
        p = V[[c[0] for c in face]]
        uv = VT[[c[1] for c in face]]
        n = VN[face[0][2]]
        # dP = Tu du + Tv dv over the first three corners
        A = np.array([uv[1] - uv[0], uv[2] - uv[0]])
        D = np.array([p[1] - p[0], p[2] - p[0]])
        dPdu, dPdv = np.linalg.solve(A, D)
        T = dPdu / np.linalg.norm(dPdu)
        B = dPdv / np.linalg.norm(dPdv)
        ntri = len(face) - 2
        for k in range(ntri):
            for vi in tri[t + k]:
                assert np.allclose(fr[vi][0], T, atol=1e-6), (name, face)
                assert np.allclose(fr[vi][1], B, atol=1e-6), (name, face)
                assert np.allclose(fr[vi][2], n, atol=1e-6), (name, face)
        t += ntri


def test_concave_quad_splits_at_the_concave_corner():
    """A dart with its reflex corner at vertex 2: the fan starts there, giving
    triangles (2, 3, 0) and (2, 0, 1) -- both inside the quad."""
    text = b"v 0 0 0\nv 2 1 0\nv 0.5 0.5 0\nv 1 2 0\nvn 0 0 1\nf 1//1 2//1 3//1 4//1\n"
    for md in (load_obj(text), _as_md(ingest.obj_load(text))):
        pos = md.vertices[:, :2]
        tri = pos[md.indices.reshape(-1, 3)]
        got = sorted(tuple(map(tuple, t.round(6).tolist())) for t in tri)
        want = sorted([((0.5, 0.5), (1, 2), (0, 0)), ((0.5, 0.5), (0, 0), (2, 1))])
        assert got == want
        # no triangle is inverted (the convex-split mistake gives one)
        e1 = tri[:, 1] - tri[:, 0]
        e2 = tri[:, 2] - tri[:, 0]
        assert (e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0] > 0).all()


CUBE_NO_NORMALS = b"""o cube
v -1 -1 -1
v 1 -1 -1
v 1 1 -1
v -1 1 -1
v -1 -1 1
v 1 -1 1
v 1 1 1
v -1 1 1
f 1 4 3 2
f 5 6 7 8
f 1 2 6 5
f 4 8 7 3
f 1 5 8 4
f 2 3 7 6
"""


@pytest.mark.parametrize("src", ["library", "oracle"])
def test_smooth_normals_of_a_cube(src):
    """GenSmoothNormals without a file normal: every corner gets the
    normalised sum of the face normals at its position, (+-1, +-1, +-1)/sqrt3
    pointing outwards; without UVs there are no tangents (zero T and B) and
    each cube corner is one vertex."""
    md = load_obj(CUBE_NO_NORMALS) if src == "library" else _as_md(ingest.obj_load(CUBE_NO_NORMALS))
    assert len(md.vertices) == 8 and len(md.indices) == 36
    fr = frames(md)
    for v, f in zip(md.vertices, fr):
        assert np.allclose(f[2], v[:3] / np.sqrt(3.0), atol=1e-6)
        assert not f[0].any() and not f[1].any()
    assert md.texCoords is None


def _correctly_rounded(s):
    q = Fraction(s)
    f = np.float32(float(q))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    return min(cands, key=lambda c: (abs(Fraction(float(c)) - q), int(np.float32(c).view(np.uint32)) & 1))


def test_fast_atof_pins():
    """Assimp's fast_atof is float(int part) + float(fraction): within one ulp
    of the decimal, equal to the correctly rounded float for exact binary
    fractions, and not always correctly rounded -- the library follows it
    where strtof would not."""
    rng = np.random.default_rng(5)
    for s in ["0.5", "-0.25", "2.000000", "-1.977276", "0.554358", "1e3", "-2.5E-2", "+7", ".5"]:
        f = ingest.fast_atof(s)
        exact = float(Fraction(s))
        assert abs(float(f) - exact) <= abs(float(np.spacing(np.float32(exact)))), s
    for s in ["0.5", "-0.25", "2.000000", "1024.125", "-3.75"]:
        assert ingest.fast_atof(s) == _correctly_rounded(s)
    # a decimal where the two roundings differ from one correct rounding
    diff = None
    for _ in range(20000):
        s = "%d.%06d" % (rng.integers(1, 9999), rng.integers(0, 10 ** 6))
        if ingest.fast_atof(s) != _correctly_rounded(s):
            diff = s
            break
    assert diff is not None
    text = ("v %s 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3\n" % diff).encode()
    got = load_obj(text).vertices[0, 0]
    assert got == ingest.fast_atof(diff) and got != _correctly_rounded(diff)


@pytest.mark.parametrize("name", ["lowpoly_suzanne.obj", "uv_sphere.obj", "cornell_light.obj"])
def test_vertices_come_from_the_file(name):
    """Every output triangle corner carries the position, normal and (flipped)
    uv of the corresponding face corner of the OBJ text; faces appear in
    order, quads as (0, 1, 2), (0, 2, 3) when convex."""
    path = os.path.join(ASSETS, name)
    V, VT, VN, faces = _obj_faces(open(path).read())
    md = load_obj(path)
    fr = frames(md)
    tri = md.indices.reshape(-1, 3)
    t = 0
    for face in faces:
        split = [(0, 1, 2)] if len(face) == 3 else [(0, 1, 2), (0, 2, 3)]
        for corners in split:
            for c, vi in zip(corners, tri[t]):
                vv, vtt, vnn = face[c]
                assert np.allclose(md.vertices[vi, :3], V[vv], atol=1e-6)
                assert np.allclose(fr[vi][2], VN[vnn], atol=1e-6)
                if VT is not None and vtt is not None:
                    assert np.allclose(md.texCoords[vi], [VT[vtt][0], 1 - VT[vtt][1]], atol=1e-6)
            t += 1
    assert t * 3 == len(md.indices)
    # joined: no two output vertices carry the same attributes
    keys = {tuple(np.concatenate([md.vertices[i], fr[i].ravel(),
                                  md.texCoords[i] if md.texCoords is not None else []]).tolist())
            for i in range(len(md.vertices))}
    assert len(keys) == len(md.vertices)


def test_line_records_follow_the_triangles():
    """The reference copies every face's indices: a polyline of k points adds
    k - 1 two-index faces after the triangles (the index count is then not a
    multiple of 3, and a model has floor(n / 3) triangles)."""
    text = b"v 0 0 0\nv 1 0 0\nv 0 1 0\nv 2 2 0\nvt 0 0\nvt 1 0\nvt 0 1\nvn 0 0 1\n" \
           b"f 1/1/1 2/2/1 3/3/1\nl 2 3 4\n"
    for md in (load_obj(text), _as_md(ingest.obj_load(text))):
        assert len(md.indices) == 3 + 2 * 2
        pos = md.vertices[:, :3]
        seg = md.indices[3:].reshape(2, 2)
        assert np.allclose(pos[seg[0]], [[1, 0, 0], [0, 1, 0]])
        assert np.allclose(pos[seg[1]], [[0, 1, 0], [2, 2, 0]])
        # the line corners have no normal (zero) and NaN tangents
        assert not frames(md)[seg[0][0]][2].any() and np.isnan(frames(md)[seg[0][0]][0]).all()
    s = Scene()
    oid = s.define_object(text)
    assert s._ranges[oid]["indexCount"] == 2


def test_obj_errors():
    for bad in [b"", b"v 0 0 0\n", b"v 0 0 0\nf 1 2 3\n", b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 x\n"]:
        with pytest.raises(reina_rt.ReinaError) as e:
            load_obj(bad)
        assert e.value.code == -1
    with pytest.raises(reina_rt.ReinaError):
        load_obj(os.path.join(ASSETS, "does_not_exist.obj"))


# ---- PNG -----------------------------------------------------------------------------
def _chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)


def _pack_row(samples, depth):
    flat = [int(x) for x in np.asarray(samples).ravel()]
    if depth == 16:
        return b"".join(struct.pack(">H", x) for x in flat)
    if depth == 8:
        return bytes(flat)
    out, acc, nbits = bytearray(), 0, 0
    for x in flat:
        acc = (acc << depth) | x
        nbits += depth
        if nbits == 8:
            out.append(acc)
            acc, nbits = 0, 0
    if nbits:
        out.append(acc << (8 - nbits))
    return bytes(out)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)


def _filter_rows(rows, bpp, rng):
    out, prev = bytearray(), bytes(len(rows[0])) if rows else b""
    for row in rows:
        f = int(rng.integers(0, 5))
        out.append(f)
        for i, x in enumerate(row):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][f]
            out.append((x - pred) & 0xFF)
        prev = row
    return bytes(out)


ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def encode_png(samples, color, depth, interlace=0, palette=None, trns=None, seed=0):
    s = np.asarray(samples)
    h, w, ch = s.shape
    rng = np.random.default_rng(seed)
    bpp = max(1, ch * depth // 8)
    raw = b""
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    for x0, y0, dx, dy in passes:
        sub = s[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        rows = [_pack_row(r, depth) for r in sub]
        raw += _filter_rows(rows, bpp, rng)
    z = zlib.compress(raw, 9)
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, interlace))
    out += _chunk(b"gAMA", struct.pack(">I", 45455))           # an ancillary chunk to skip
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).ravel()))
    if trns is not None:
        if color == 3:
            out += _chunk(b"tRNS", bytes(np.asarray(trns, np.uint8)))
        else:
            out += _chunk(b"tRNS", b"".join(struct.pack(">H", k) for k in trns))
    half = len(z) // 2                                          # two IDAT chunks
    out += _chunk(b"IDAT", z[:half]) + _chunk(b"IDAT", z[half:]) + _chunk(b"IEND", b"")
    return out


PNG_CASES = [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2), (3, 4), (3, 8),
             (4, 8), (4, 16), (6, 8), (6, 16)]


# tRNS is only allowed without an alpha channel
PNG_PARAMS = [(c, d, t) for c, d in PNG_CASES for t in (False, True) if not (t and c in (4, 6))]


@pytest.mark.parametrize("interlace", [0, 1])
@pytest.mark.parametrize("color,depth,with_trns", PNG_PARAMS)
def test_png_decode(color, depth, interlace, with_trns):
    rng = np.random.default_rng(color * 100 + depth + 7 * interlace + 3 * with_trns)
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[color]
    for (h, w) in [(7, 13), (1, 1), (17, 3), (9, 9)]:
        hi = (1 << depth) - 1
        palette = trns = None
        if color == 3:
            n = min(hi + 1, 200)
            palette = rng.integers(0, 256, (n, 3))
            samples = rng.integers(0, n, (h, w, 1))
            if with_trns:
                trns = rng.integers(0, 256, int(rng.integers(1, n + 1)))
        else:
            samples = rng.integers(0, hi + 1, (h, w, ch))
            if with_trns:
                key = samples[0, 0, :ch].tolist()                 # guarantee at least one hit
                trns = [int(k) for k in key]
        data = encode_png(samples, color, depth, interlace, palette, trns, seed=h * w)
        want = ingest.png_expected_rgba(samples, color, depth, palette, trns)
        got = load_image(data)
        assert got.shape == (h, w, 4)
        assert np.array_equal(got, want), (color, depth, interlace, with_trns, h, w)
        # Pillow decodes the same file the same way, except where it differs
        # from stb: 16-bit grey (kept at 16 bits and clipped, not the high
        # byte), tRNS keys of 2/4-bit grey (ignored) and of 16-bit images
        # (compared after the reduction to 8 bits)
        pil_differs = (color in (0, 4) and depth == 16) or \
            (with_trns and color in (0, 2) and (depth in (2, 4) or depth == 16))
        if not pil_differs:
            from PIL import Image
            import io
            pil = np.asarray(Image.open(io.BytesIO(data)).convert("RGBA"))
            assert np.array_equal(got, pil), (color, depth, interlace, with_trns, h, w)


def test_png_path_flips_rows(tmp_path):
    samples = np.random.default_rng(1).integers(0, 256, (5, 4, 4))
    data = encode_png(samples, 6, 8)
    p = tmp_path / "x.png"
    p.write_bytes(data)
    assert np.array_equal(load_image(str(p)), load_image(data)[::-1])


def test_png_asset_matches_pillow():
    from PIL import Image
    path = os.path.join(ASSETS, "cornell_texture.png")
    got = load_image(path)
    pil = np.asarray(Image.open(path).convert("RGBA"))
    assert np.array_equal(got, pil[::-1])


REF_PNGS = sorted(glob.glob(os.path.join(REF_TEXTURES, "**", "*.png"), recursive=True))


@pytest.mark.skipif(not REF_PNGS, reason="the reference's textures/ directory is not present")
def test_reference_textures_match_pillow():
    from PIL import Image
    for path in REF_PNGS:
        im = Image.open(path)
        if im.mode in ("I;16", "I;16B", "I"):
            continue
        assert np.array_equal(load_image(path), np.asarray(im.convert("RGBA"))[::-1]), path


def test_png_errors():
    with pytest.raises(reina_rt.ReinaError) as e:
        load_image(b"GIF89a" + b"\0" * 16)
    assert e.value.code == -6
    good = encode_png(np.zeros((2, 2, 3), np.int64), 2, 8)
    for bad in [good[:40], good.replace(b"IHDR", b"IHDX", 1), b"not a png at all"]:
        with pytest.raises(reina_rt.ReinaError):
            load_image(bad)
    # a palette index beyond PLTE
    data = encode_png(np.full((2, 2, 1), 3), 3, 8, palette=np.zeros((2, 3), np.int64))
    with pytest.raises(reina_rt.ReinaError):
        load_image(data)


def test_define_texture_and_object_from_files():
    """Scene::defineTexture / addObject with paths (src/scene/Scene.cpp:7-25,
    171-177) build the same arrays as the decoded data passed directly."""
    a = Scene()
    tex = a.define_texture(os.path.join(ASSETS, "cornell_texture.png"))
    oid = a.add_object(os.path.join(ASSETS, "cornell_box.obj"), np.eye(4), reina_rt.Material(textureID=tex))
    b = Scene()
    tex_b = b.add_texture(load_image(os.path.join(ASSETS, "cornell_texture.png")))
    oid_b = b.add_object(load_obj(os.path.join(ASSETS, "cornell_box.obj")), np.eye(4), reina_rt.Material(textureID=tex_b))
    assert (tex, oid) == (tex_b, oid_b)
    ra, rb = a.build_arrays(), b.build_arrays()
    for k in ra:
        if isinstance(ra[k], np.ndarray):
            assert np.array_equal(ra[k], rb[k]), k
