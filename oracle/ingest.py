"""ORACLE — CPU restatement of Reina's scene ingest, in Python + numpy.
Test infrastructure only: used by tests/ as the checker of the library's
native loaders (reina_rt_model_load_obj, reina_rt_image_load,
reina_rt_gltf_load), never by the product path.

obj_load(text) restates Models::getObjData (src/scene/Models.cpp:117-175):
Assimp's OBJ importer (a vcpkg dependency of the reference, CMakeLists.txt:8,
version unpinned; its published 5.x algorithm) with aiProcess_FlipUVs,
Triangulate, GenSmoothNormals, CalcTangentSpace and JoinIdenticalVertices, in
the order Assimp's post-process registry runs them, then the ModelData copy of
mMeshes[0]. Float arithmetic is float32, operation by operation, in the order
Assimp's aiVector3D operators evaluate it. Neighbour queries scan a window of a
sort along x (independent of the library's grid); candidates are visited in
ascending vertex order.

png_expected_rgba(...) states stb_image's 4-channel expansion of decoded
PNG samples (src/graphics/Image.cpp:10-41 requests 4 channels).

gltf_load(data) restates reina::scene::gltf::loadScene
(src/scene/gltf/gltfloader.cpp) over fastgltf's accessor conversions and the
glTF 2.0 node transform rules; images are decoded by Pillow.

Pins (tests/test_ingest.py): closed forms (quad and cube tangent frames,
triangulation of convex and concave quads, smooth normals of a cube without
normals, fast_atof against correctly rounded decimals and a case where it
differs), structural invariants against the OBJ text (every output corner is
one of the file's (v, vt, vn) triples, triangles follow the faces), and Pillow
as an independent PNG decoder.
"""
import ctypes
import ctypes.util

import numpy as np

F = np.float32
_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.acosf.argtypes = [ctypes.c_float]
_libm.acosf.restype = ctypes.c_float
_libm.cosf.argtypes = [ctypes.c_float]
_libm.cosf.restype = ctypes.c_float
_libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
_libm.powf.restype = ctypes.c_float

_NEG_POW10 = [0.0, 0.1, 0.01, 0.001, 0.0001, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10, 1e-11, 1e-12, 1e-13, 1e-14,
              1e-15]


class ObjError(ValueError):
    pass


# ---- fast_atof -----------------------------------------------------------------
def _digits(s, i, limit=None):
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    if j == i:
        raise ObjError("cannot parse a number")
    if limit is not None and j - i > limit:
        return int(s[i:i + limit]), limit, j
    return int(s[i:j]), j - i, j


def fast_atof(s):
    """Assimp's fast_atoreal_move<float>: float(integer part) + float(fraction
    digits * 10^-n computed in double), then * powf(10, exponent)."""
    i = 0
    neg = s.startswith("-")
    if neg or s.startswith("+"):
        i = 1
    low = s[i:i + 3].lower()
    if low == "nan":
        return F(np.nan)
    if low == "inf":
        return F(-np.inf) if neg else F(np.inf)
    c0 = s[i] if i < len(s) else ""
    c1 = s[i + 1] if i + 1 < len(s) else ""
    if not (c0.isdigit() or (c0 in ".," and c1.isdigit())):
        raise ObjError("cannot parse a real number")
    f = F(0.0)
    if c0 not in ".,":
        v, _, i = _digits(s, i)
        if v >= 1 << 64:
            raise ObjError("number overflows 64 bits")
        f = F(v)
    if i < len(s) and s[i] in ".," and i + 1 < len(s) and s[i + 1].isdigit():
        v, n, i = _digits(s, i + 1, 15)
        pl = float(v) * _NEG_POW10[n]
        f = F(f + F(pl))
    elif i < len(s) and s[i] == ".":
        i += 1
    if i < len(s) and s[i] in "eE":
        i += 1
        eneg = i < len(s) and s[i] == "-"
        if i < len(s) and s[i] in "+-":
            i += 1
        e, _, i = _digits(s, i)
        e = F(e)
        if eneg:
            e = -e
        f = F(f * F(_libm.powf(10.0, float(e))))
    return F(-f) if neg else f


# ---- parser -------------------------------------------------------------------------
def _tokens(line):
    out = []
    for t in line.split():
        if t.startswith("#"):
            break
        out.append(t)
    return out


def _numeric(t):
    return t[0].isdigit() or t[0] in "+-" or t[:3].lower() in ("nan", "inf")


class _Model:
    def __init__(self):
        self.pos, self.col, self.nrm, self.uv = [], [], [], []
        self.objects = []      # [name, [mesh ids]]
        self.meshes = []       # dict(faces, material, normals, uv_refs)
        self.materials = []
        self.cur_obj = None
        self.cur_mesh = None
        self.cur_mat = None
        self.group = ""

    def mesh(self):
        self.meshes.append(dict(faces=[], material=-1, normals=False, uv_refs=0))
        self.cur_mesh = len(self.meshes) - 1
        if self.cur_obj is not None:
            self.objects[self.cur_obj][1].append(self.cur_mesh)

    def obj(self, name):
        self.objects.append([name, []])
        self.cur_obj = len(self.objects) - 1
        self.mesh()
        if self.cur_mat is not None:
            self.meshes[self.cur_mesh]["material"] = self.materials.index(self.cur_mat)


def _face(m, toks, kind):
    v, vt, vn = [], [], []
    have_vt, have_vn = bool(m.uv), bool(m.nrm)
    has_n = False
    for t in toks:
        for slot, part in enumerate(t.split("/")):
            if part == "":
                continue
            try:
                val = int(part)
            except ValueError:
                raise ObjError("invalid face index")
            s = 2 if (slot == 1 and not have_vt and have_vn) else slot
            if val == 0:
                raise ObjError("invalid face index")
            count = (len(m.pos), len(m.uv), len(m.nrm))[s] if s < 3 else 0
            idx = val - 1 if val > 0 else count + val
            if idx < 0:
                raise ObjError("relative index before the first element")
            if s == 0:
                v.append(idx)
            elif s == 1:
                vt.append(idx)
            elif s == 2:
                vn.append(idx)
                has_n = True
            else:
                raise ObjError("too many indices in a face element")
    if not v:
        return
    if m.cur_obj is None:
        m.obj("defaultobject")
    if m.cur_mesh is None:
        m.mesh()
    me = m.meshes[m.cur_mesh]
    me["uv_refs"] += len(vt)
    me["normals"] = me["normals"] or has_n
    me["faces"].append((kind, v, vt, vn))


def _parse(text):
    m = _Model()
    for raw in text.split("\n"):
        parts = raw.split(None, 1)
        if not parts:
            continue
        key, rest = parts[0], (parts[1] if len(parts) > 1 else "")
        if key == "v":
            n = [fast_atof(t) for t in _tokens(rest) if _numeric(t)]
            if len(n) == 3:
                m.pos.append((n[0], n[1], n[2]))
            elif len(n) == 4:
                if n[3] == 0:
                    raise ObjError("division by zero in a homogeneous vertex")
                m.pos.append((F(n[0] / n[3]), F(n[1] / n[3]), F(n[2] / n[3])))
            elif len(n) == 6:
                m.col.extend([(F(0), F(0), F(0))] * (len(m.pos) - len(m.col)))
                m.pos.append((n[0], n[1], n[2]))
                m.col.append((n[3], n[4], n[5]))
            if m.col and len(m.col) < len(m.pos):
                m.col.extend([(F(0), F(0), F(0))] * (len(m.pos) - len(m.col)))
        elif key == "vt":
            n = [fast_atof(t) for t in _tokens(rest) if _numeric(t)]
            if len(n) not in (2, 3):
                raise ObjError("invalid number of texture coordinate components")
            n = [x if np.isfinite(x) else F(0) for x in n] + [F(0)] * (3 - len(n))
            m.uv.append(tuple(n))
        elif key == "vn":
            t = _tokens(rest)
            if len(t) < 3:
                raise ObjError("normal with fewer than 3 components")
            m.nrm.append(tuple(fast_atof(x) for x in t[:3]))
        elif key in ("f", "l", "p"):
            _face(m, _tokens(rest), {"f": "poly", "l": "line", "p": "point"}[key])
        elif key == "o":
            t = _tokens(rest)
            if t:
                names = [o[0] for o in m.objects]
                if t[0] in names:
                    m.cur_obj = names.index(t[0])
                else:
                    m.obj(t[0])
        elif key == "g":
            name = rest.strip()
            if name and name != m.group:
                m.obj(name)
                m.group = name
        elif key == "usemtl":
            name = rest.strip()
            if m.cur_mat == name:
                continue
            if name not in m.materials:
                m.materials.append(name)
            m.cur_mat = name
            idx = m.materials.index(name)
            need = m.cur_mesh is None
            if not need:
                me = m.meshes[m.cur_mesh]
                need = me["material"] != -1 and me["material"] != idx and bool(me["faces"])
            if need:
                m.mesh()
            m.meshes[m.cur_mesh]["material"] = idx
    return m


# ---- float32 vector helpers on (n, 3) arrays -------------------------------------------
def _dot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def _len(a):
    return np.sqrt(_dot(a, a))


def _divide(a, f):
    """aiVector3D /= f: unchanged for f == 0, else times (1 / f)."""
    f = np.asarray(f, F)
    safe = np.where(f == 0, F(1), f)
    inv = (F(1) / safe).astype(F)
    out = a * inv[..., None]
    return np.where((f == 0)[..., None], a, out).astype(F)


def _normalize(a):
    return _divide(a, _len(a))


def _normalize_safe(a):
    ln = _len(a)
    return np.where((ln > 0)[..., None], _divide(a, ln), a).astype(F)


class _Window:
    """Corners strictly within r of a point: a sweep over positions sorted by x."""

    def __init__(self, pos, r):
        self.pos, self.r = pos, F(r)
        self.r2 = F(self.r * self.r)
        self.order = np.argsort(pos[:, 0], kind="stable")
        self.xs = pos[self.order, 0].astype(np.float64)
        self.margin = float(r) * 1.001 + 1e-30

    def find(self, q):
        lo = np.searchsorted(self.xs, float(q[0]) - self.margin, "left")
        hi = np.searchsorted(self.xs, float(q[0]) + self.margin, "right")
        cand = np.sort(self.order[lo:hi])
        d = (self.pos[cand] - q).astype(F)
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        return cand[d2 < self.r2]


def _eps(pos):
    lo = np.full(3, F(1e10), F)
    hi = np.full(3, F(-1e10), F)
    for c in range(3):
        if len(pos):
            lo[c] = min(F(1e10), pos[:, c].min())
            hi[c] = max(F(-1e10), pos[:, c].max())
    d = (hi - lo).astype(F)
    return F(F(np.sqrt(F(F(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))) * F(1e-4))


def obj_load(text):
    """Returns dict(vertices (n,4), indices, tbns (k,3,3) with rows T, B, N,
    texCoords (j,2) or None) — ModelData of the file's first mesh."""
    if isinstance(text, (bytes, bytearray)):
        text = bytes(text).decode("latin-1")
    m = _parse(text)
    mesh = None
    for _, ids in m.objects:
        for i in ids:
            if m.meshes[i]["faces"]:
                mesh = m.meshes[i]
                break
        if mesh is not None:
            break
    if mesh is None:
        raise ObjError("no faces")
    has_n = bool(m.nrm) and mesh["normals"]
    has_uv = bool(m.uv) and mesh["uv_refs"] > 0
    has_col = bool(m.col)

    # verbose vertex array (createVertexArray)
    P, N, U, C, CA = [], [], [], [], []
    faces = []
    normals_ok = uv_ok = True
    zero = (F(0), F(0), F(0))
    for kind, v, vt, vn in mesh["faces"]:
        k = len(v)
        start = len(P)
        corners = []
        for c in range(k):
            vi = v[c]
            if vi >= len(m.pos):
                raise ObjError("vertex index out of range")
            if not all(np.isfinite(x) for x in m.pos[vi]):
                raise ObjError("non-finite vertex position")
            nn = zero
            if has_n and normals_ok and c < len(vn):
                if vn[c] >= len(m.nrm):
                    normals_ok = False
                else:
                    nn = m.nrm[vn[c]]
            uu = zero
            if has_uv and uv_ok and c < len(vt):
                if vt[c] >= len(m.uv):
                    uv_ok = False
                else:
                    uu = m.uv[vt[c]]
            corners.append((m.pos[vi], nn, uu, m.col[vi] if has_col else zero))
        if kind == "poly":
            for p, nn, uu, cc in corners:
                P.append(p); N.append(nn); U.append(uu); C.append(cc); CA.append(F(1))
            faces.append(list(range(start, start + k)))
        elif kind == "point":
            for p, nn, uu, cc in corners:
                faces.append([len(P)])
                P.append(p); N.append(nn); U.append(uu); C.append(cc); CA.append(F(1))
        else:
            # a polyline of k points fills 2(k-1) slots: point 0 at slot 0,
            # point i > 0 at slot 2i-1, and every inner point's duplicate at
            # slot 2i (position, uv and -- with normal indices -- normal
            # copied; colour black with alpha 0). Segment s runs from point
            # s's own slot to point s+1's duplicate (its own slot for the
            # last point).
            slots = [None] * (2 * (k - 1))
            for i in range(k):
                p, nn, uu, cc = corners[i]
                slots[0 if i == 0 else 2 * i - 1] = (p, nn, uu, cc, F(1))
                if 0 < i < k - 1:
                    slots[2 * i] = (p, nn if vn else zero, uu, zero, F(0))
            for p, nn, uu, cc, ca in slots:
                P.append(p); N.append(nn); U.append(uu); C.append(cc); CA.append(ca)
            for s in range(k - 1):
                a = 0 if s == 0 else 2 * s - 1
                b = 2 * (k - 1) - 1 if s == k - 2 else 2 * s + 2
                faces.append([start + a, start + b])
    pos = np.array(P, F).reshape(-1, 3)
    nrm = np.array(N, F).reshape(-1, 3)
    uv = np.array(U, F).reshape(-1, 3)
    col = np.array(C, F).reshape(-1, 3)
    cola = np.array(CA, F)
    if has_uv:
        uv[:, 1] = (F(1) - uv[:, 1]).astype(F)     # FlipUVs

    # Triangulate
    tri = []
    for f in faces:
        if len(f) <= 3:
            tri.append(f)
        elif len(f) == 4:
            start = 0
            for i in range(4):
                p = pos[f[i]]
                left = _normalize((pos[f[(i + 3) % 4]] - p).astype(F))
                diag = _normalize((pos[f[(i + 2) % 4]] - p).astype(F))
                right = _normalize((pos[f[(i + 1) % 4]] - p).astype(F))
                ang = F(F(_libm.acosf(float(_dot(left, diag)))) + F(_libm.acosf(float(_dot(right, diag)))))
                if ang > F(3.1415926538):
                    start = i
                    break
            tri.append([f[start], f[(start + 1) % 4], f[(start + 2) % 4]])
            tri.append([f[start], f[(start + 2) % 4], f[(start + 3) % 4]])
        else:
            raise ObjError("polygons with more than 4 corners are outside the oracle")
    faces = tri
    has_polys = any(kind == "poly" and len(v) >= 3 for kind, v, _, _ in mesh["faces"])
    nv = len(pos)

    # GenSmoothNormals (no angle limit)
    if not has_n and has_polys:
        fn = np.zeros((nv, 3), F)
        for f in faces:
            if len(f) < 3:
                fn[f] = np.nan
                continue
            p1, p2, p3 = pos[f[0]], pos[f[1]], pos[f[-1]]
            fn[f] = _normalize_safe(_cross((p2 - p1).astype(F), (p3 - p1).astype(F)))
        win = _Window(pos, _eps(pos))
        nrm = np.zeros((nv, 3), F)
        had = np.zeros(nv, bool)
        for i in range(nv):
            if had[i]:
                continue
            found = win.find(pos[i])
            s = np.zeros(3, F)
            for j in found:
                if not np.isnan(fn[j, 0]):
                    s = (s + fn[j]).astype(F)
            s = _normalize_safe(s)
            nrm[found] = s
            had[found] = True
        has_n = True

    # CalcTangentSpace
    has_tan = False
    tan = np.zeros((nv, 3), F)
    bit = np.zeros((nv, 3), F)
    if has_polys and has_n and has_uv:
        done = np.zeros(nv, bool)
        tri_faces = [f for f in faces if len(f) >= 3]
        for f in faces:
            if len(f) < 3:
                done[f] = True
                tan[f] = np.nan
                bit[f] = np.nan
        if tri_faces:
            fa = np.array([f[:3] for f in tri_faces], np.int64)
            v = (pos[fa[:, 1]] - pos[fa[:, 0]]).astype(F)
            w = (pos[fa[:, 2]] - pos[fa[:, 0]]).astype(F)
            sx = (uv[fa[:, 1], 0] - uv[fa[:, 0], 0]).astype(F)
            sy = (uv[fa[:, 1], 1] - uv[fa[:, 0], 1]).astype(F)
            tx = (uv[fa[:, 2], 0] - uv[fa[:, 0], 0]).astype(F)
            ty = (uv[fa[:, 2], 1] - uv[fa[:, 0], 1]).astype(F)
            dirc = np.where((tx * sy - ty * sx) < 0, F(-1), F(1)).astype(F)
            flat = (sx * ty) == (sy * tx)
            sx = np.where(flat, F(0), sx); sy = np.where(flat, F(1), sy)
            tx = np.where(flat, F(1), tx); ty = np.where(flat, F(0), ty)
            T = ((w * sy[:, None] - v * ty[:, None]) * dirc[:, None]).astype(F)
            B = ((w * sx[:, None] - v * tx[:, None]) * dirc[:, None]).astype(F)
            corner, fidx = [], []
            for k, f in enumerate(tri_faces):
                corner.extend(f)
                fidx.extend([k] * len(f))
            corner = np.array(corner, np.int64)
            fidx = np.array(fidx, np.int64)
            t, b, n = T[fidx], B[fidx], nrm[corner]
            lt = (t - n * _dot(t, n)[:, None]).astype(F)
            lb = ((b - n * _dot(b, n)[:, None]) - lt * _dot(b, lt)[:, None]).astype(F)
            lt = _normalize_safe(lt)
            lb = _normalize_safe(lb)
            bad_t = ~np.isfinite(lt).all(axis=1)
            bad_b = ~np.isfinite(lb).all(axis=1)
            fix_t = bad_t & ~bad_b
            fix_b = bad_b & ~bad_t
            if fix_t.any():
                lt[fix_t] = _normalize_safe(_cross(n[fix_t], lb[fix_t]))
            if fix_b.any():
                lb[fix_b] = _normalize_safe(_cross(lt[fix_b], n[fix_b]))
            # a corner belongs to one face: later faces overwrite nothing
            tan[corner] = lt
            bit[corner] = lb
        win = _Window(pos, _eps(pos))
        limit = F(_libm.cosf(float(F(F(45.0) * F(0.0174532925)))))
        for a in range(nv):
            if done[a]:
                continue
            on, ot, ob = nrm[a].copy(), tan[a].copy(), bit[a].copy()
            group = [a]
            for j in win.find(pos[a]):
                if done[j]:
                    continue
                if _dot(nrm[j], on) < F(0.9999):
                    continue
                if _dot(tan[j], ot) < limit:
                    continue
                if _dot(bit[j], ob) < limit:
                    continue
                group.append(j)
                done[j] = True
            st = np.zeros(3, F)
            sb = np.zeros(3, F)
            for j in group:
                st = (st + tan[j]).astype(F)
                sb = (sb + bit[j]).astype(F)
            st = _normalize(st)
            sb = _normalize(sb)
            tan[group] = st
            bit[group] = sb
        has_tan = True

    # JoinIdenticalVertices
    def key(p):
        return tuple(0 if x == 0 else int(np.float32(x).view(np.uint32)) for x in p)

    sq = F(F(1e-5) * F(1e-5))

    def close(a, b):
        d = (a - b).astype(F)
        return not (F(F(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) > sq)

    def same(a, u):
        if not close(pos[a], pos[u]):
            return False
        if has_n and not close(nrm[a], nrm[u]):
            return False
        if has_tan and (not close(tan[a], tan[u]) or not close(bit[a], bit[u])):
            return False
        if has_uv and not close(uv[a], uv[u]):
            return False
        if has_col:
            d = (col[a] - col[u]).astype(F)
            da = F(cola[a] - cola[u])
            if F(F(F(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) + da * da) > sq:
                return False
        return True

    buckets, unique, remap = {}, [], np.zeros(nv, np.int64)
    for a in range(nv):
        cands = buckets.setdefault(key(pos[a]), [])
        match = -1
        for c in reversed(cands):
            if same(a, unique[c]):
                match = c
                break
        if match < 0:
            remap[a] = len(unique)
            cands.append(len(unique))
            unique.append(a)
        else:
            remap[a] = match
    u = np.array(unique, np.int64)
    out_idx = np.array([remap[i] for f in faces for i in f], np.uint32)
    verts = np.concatenate([pos[u], np.ones((len(u), 1), F)], axis=1).astype(F)
    tbns = None
    if has_n:
        tbns = np.stack([tan[u] if has_tan else np.zeros((len(u), 3), F),
                         bit[u] if has_tan else np.zeros((len(u), 3), F), nrm[u]], axis=1).astype(F)
    tcs = uv[u, :2].astype(F) if has_uv else None
    return dict(vertices=verts, indices=out_idx, tbns=tbns, texCoords=tcs)


# ---- PNG: stb_image's 4-channel expansion --------------------------------------
def png_expected_rgba(samples, color, depth, palette=None, trns=None):
    """samples: (h, w, channels) integer samples as stored in the file (before
    any scaling); palette: (n, 3) uint8; trns: palette alphas, or the colour
    key (1 or 3 integers). Returns (h, w, 4) uint8 as stbi_load(..., 4)."""
    s = np.asarray(samples, np.int64)
    h, w = s.shape[:2]
    out = np.zeros((h, w, 4), np.uint8)
    if color == 3:
        pal = np.zeros((len(palette), 4), np.uint8)
        pal[:, :3] = palette
        pal[:, 3] = 255
        if trns is not None:
            pal[:len(trns), 3] = trns
        return pal[s[..., 0]]
    scale = {1: 255, 2: 85, 4: 17, 8: 1, 16: 1}[depth]
    to8 = (s >> 8) if depth == 16 else (s * scale) & 0xFF
    if color in (0, 4):
        out[..., 0] = out[..., 1] = out[..., 2] = to8[..., 0]
        out[..., 3] = to8[..., 1] if color == 4 else 255
    else:
        out[..., :3] = to8[..., :3]
        out[..., 3] = to8[..., 3] if color == 6 else 255
    if trns is not None and color in (0, 2):
        nk = 1 if color == 0 else 3
        if depth == 16:
            hit = np.all(s[..., :nk] == np.asarray(trns[:nk]), axis=-1)
        else:
            keys = np.array([((k & 255) * scale) & 0xFF for k in trns[:nk]])
            hit = np.all(to8[..., :nk] == keys, axis=-1)
        out[..., 3][hit] = 0
    return out


# ---- glTF ----------------------------------------------------------------------------------
_GLTF_SUPPORTED = {"KHR_mesh_quantization", "KHR_texture_transform", "KHR_materials_variants",
                   "KHR_materials_transmission", "KHR_materials_clearcoat", "KHR_materials_emissive_strength"}
_COMP = {5120: ("<i1", 127.0), 5121: ("<u1", 255.0), 5122: ("<i2", 32767.0), 5123: ("<u2", 65535.0),
         5125: ("<u4", 4294967295.0), 5126: ("<f4", None)}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}


class GltfError(ValueError):
    pass


def _gltf_uri(uri, base_dir):
    import base64
    import os
    import urllib.parse
    if uri.startswith("data:"):
        head, _, data = uri.partition(",")
        if ";base64" not in head:
            raise GltfError("only base64 data URIs")
        return base64.b64decode(data)
    if base_dir is None:
        raise GltfError("external file in an in-memory asset")
    return open(os.path.join(base_dir, urllib.parse.unquote(uri)), "rb").read()


def _gltf_doc(data, base_dir):
    import json
    import struct
    bin_chunk = None
    if data[:4] == b"glTF":
        _, version, length = struct.unpack_from("<III", data, 0)
        if version != 2:
            raise GltfError("GLB version")
        off, doc = 12, None
        while off + 8 <= length:
            n, kind = struct.unpack_from("<II", data, off)
            chunk = data[off + 8: off + 8 + n]
            if kind == 0x4E4F534A and doc is None:
                doc = json.loads(chunk.decode("utf-8"))
            elif kind == 0x004E4942 and bin_chunk is None:
                bin_chunk = chunk
            off += 8 + n
    else:
        doc = json.loads(data.decode("utf-8"))
    for e in doc.get("extensionsRequired", []):
        if e not in _GLTF_SUPPORTED:
            raise GltfError("unsupported required extension " + e)
    buffers = []
    for b in doc.get("buffers", []):
        buffers.append(_gltf_uri(b["uri"], base_dir) if "uri" in b else bin_chunk)
    return doc, buffers


def _gltf_view(doc, buffers, view, offset, count, elem):
    v = doc["bufferViews"][view]
    stride = v.get("byteStride", 0) or elem
    base = v.get("byteOffset", 0) + offset
    buf = buffers[v["buffer"]]
    if count and offset + stride * (count - 1) + elem > v["byteLength"]:
        raise GltfError("accessor outside its view")
    return buf, base, stride


def _gltf_accessor(doc, buffers, idx, as_index=False):
    a = doc["accessors"][idx]
    dt, mx = _COMP[a["componentType"]]
    nc = _NCOMP[a["type"]]
    count = a["count"]
    norm = a.get("normalized", False)
    size = np.dtype(dt).itemsize

    def conv(raw):
        if as_index:
            return raw.astype(np.uint32)
        if dt == "<f4":
            return raw.astype(F)
        f = raw.astype(F)
        if norm:
            return np.maximum((f / F(mx)).astype(F), F(-1))
        return f

    out = np.zeros((count, nc), np.uint32 if as_index else F)
    if "bufferView" in a:
        buf, base, stride = _gltf_view(doc, buffers, a["bufferView"], a.get("byteOffset", 0), count, size * nc)
        rows = [np.frombuffer(buf, dt, nc, base + i * stride) for i in range(count)]
        if rows:
            out[:] = conv(np.stack(rows))
    if "sparse" in a:
        sp = a["sparse"]
        n = sp["count"]
        idt, _ = _COMP[sp["indices"]["componentType"]]
        ib, ibase, _ = _gltf_view(doc, buffers, sp["indices"]["bufferView"], sp["indices"].get("byteOffset", 0), n,
                                  np.dtype(idt).itemsize)
        vb, vbase, _ = _gltf_view(doc, buffers, sp["values"]["bufferView"], sp["values"].get("byteOffset", 0), n,
                                  size * nc)
        ids = np.frombuffer(ib, idt, n, ibase).astype(np.int64)
        vals = np.frombuffer(vb, dt, n * nc, vbase).reshape(n, nc)
        out[ids] = conv(vals)
    return out[:, 0] if as_index else out


def _mat_mul(a, b):
    """column-major float32 4x4 product, terms summed k = 0..3 left to right"""
    A = a.reshape(4, 4)   # A[c, r]
    B = b.reshape(4, 4)
    r = np.zeros((4, 4), F)
    for c in range(4):
        for row in range(4):
            r[c, row] = F(F(F(A[0, row] * B[c, 0]) + F(A[1, row] * B[c, 1])) + F(A[2, row] * B[c, 2])) \
                + F(A[3, row] * B[c, 3])
    return r.reshape(16)


def _node_matrix(node, parent):
    if "matrix" in node:
        return _mat_mul(parent, np.array(node["matrix"], np.float64).astype(F))
    t = np.array(node.get("translation", [0, 0, 0]), np.float64).astype(F)
    x, y, z, w = np.array(node.get("rotation", [0, 0, 0, 1]), np.float64).astype(F)
    s = np.array(node.get("scale", [1, 1, 1]), np.float64).astype(F)
    P = parent.reshape(4, 4)
    m = P.copy()
    m[3] = F(F(F(P[0] * t[0]) + F(P[1] * t[1])) + F(P[2] * t[2])) + P[3]
    xx, yy, zz, xy, xz, yz, wx, wy, wz = x * x, y * y, z * z, x * y, x * z, y * z, w * x, w * y, w * z
    one, two = F(1), F(2)
    R = np.eye(4, dtype=F)
    R[0, :3] = [one - two * (yy + zz), two * (xy + wz), two * (xz - wy)]
    R[1, :3] = [two * (xy - wz), one - two * (xx + zz), two * (yz + wx)]
    R[2, :3] = [two * (xz + wy), two * (yz - wx), one - two * (xx + yy)]
    m = _mat_mul(m.reshape(16), R.reshape(16)).reshape(4, 4)
    m[0] = (m[0] * s[0]).astype(F)
    m[1] = (m[1] * s[1]).astype(F)
    m[2] = (m[2] * s[2]).astype(F)
    return m.reshape(16)


def gltf_load(data, base_dir=None):
    """Restates reina::scene::gltf::loadScene (src/scene/gltf/gltfloader.cpp)
    with fastgltf's accessor conversions: returns dict(models=[dict(vertices,
    indices, tbns (rows T, B, N), texCoords)], images=[RGBA8 via Pillow],
    instances=[(model, column-major float32[16], material dict)])."""
    import io

    from PIL import Image

    doc, buffers = _gltf_doc(data, base_dir)
    scenes = doc.get("scenes", [])
    if not scenes:
        raise GltfError("no scenes")
    visits = []

    def walk(ni, parent):
        node = doc["nodes"][ni]
        m = _node_matrix(node, parent)
        visits.append((ni, m))
        for c in node.get("children", []):
            walk(c, m)

    ident = np.eye(4, dtype=F).reshape(16)
    for r in scenes[doc.get("scene", 0)].get("nodes", []):
        walk(r, ident)

    images = []
    for im in doc.get("images", []):
        if "uri" in im:
            raw = _gltf_uri(im["uri"], base_dir)
        else:
            v = doc["bufferViews"][im["bufferView"]]
            buf = buffers[v["buffer"]]
            raw = buf[v.get("byteOffset", 0): v.get("byteOffset", 0) + v["byteLength"]]
        images.append(np.asarray(Image.open(io.BytesIO(raw)).convert("RGBA")))

    used = sorted({doc["nodes"][ni]["mesh"] for ni, _ in visits if "mesh" in doc["nodes"][ni]})
    models, mesh_models = [], {}
    for mi in used:
        mesh_models[mi] = []
        for p in doc["meshes"][mi]["primitives"]:
            at = p["attributes"]
            P = _gltf_accessor(doc, buffers, at["POSITION"])
            if "NORMAL" not in at:
                raise GltfError("no normals")
            N = _gltf_accessor(doc, buffers, at["NORMAL"])
            n = len(P)
            T = np.zeros((n, 3), F)
            B = np.zeros((n, 3), F)
            if "TANGENT" in at:
                Ta = _gltf_accessor(doc, buffers, at["TANGENT"])
                T[:] = Ta[:, :3]
                B[:] = _cross(N, T)
                if Ta.shape[1] == 4:
                    B[:] = (B * Ta[:, 3:4]).astype(F)
            else:
                k = n // 3 * 3
                T[:k, 0] = 1
                B[:k, 1] = 1
            uv = _gltf_accessor(doc, buffers, at["TEXCOORD_0"]) if "TEXCOORD_0" in at else np.zeros((n, 2), F)
            idx = _gltf_accessor(doc, buffers, p["indices"], True) if "indices" in p else np.arange(n, dtype=np.uint32)
            mesh_models[mi].append(len(models))
            models.append(dict(vertices=np.concatenate([P, np.ones((n, 1), F)], 1).astype(F),
                               indices=idx.astype(np.uint32), tbns=np.stack([T, B, N], 1).astype(F),
                               texCoords=uv.astype(F), material=p.get("material", -1)))

    def tex_image(info):
        src = doc["textures"][info["index"]].get("source")
        if src is None:
            raise GltfError("texture without source")
        return src

    def material(mi):
        m = dict(materialIdx=3, textureID=-1, normalMapID=-1, bumpMapID=-1, albedo=(1.0, 1.0, 1.0),
                 emission=(0.0, 0.0, 0.0), roughness=0.0, ior=1.5, interpNormals=True, absorption=0.0,
                 cullBackface=False, anisotropic=0.0, subsurface=0.0, clearcoatGloss=0.0, sheenTint=(1.0, 1.0, 1.0),
                 specularTint=(1.0, 1.0, 1.0), metallic=0.0, clearcoat=0.0, specularTransmission=0.0, sheen=0.0)
        if mi < 0:
            return m
        g = doc["materials"][mi]
        pbr = g.get("pbrMetallicRoughness", {})
        ext = g.get("extensions", {})
        base = [F(c) for c in pbr.get("baseColorFactor", [1, 1, 1, 1])]
        em = [F(c) for c in g.get("emissiveFactor", [0, 0, 0])]
        strength = F(ext.get("KHR_materials_emissive_strength", {}).get("emissiveStrength", 1.0))
        if any(c > 0 for c in em) and "emissiveTexture" not in g:
            m["emission"] = tuple(F(c * strength) for c in em)
        m["metallic"] = F(pbr.get("metallicFactor", 1.0))
        m["roughness"] = F(max(min(F(pbr.get("roughnessFactor", 1.0)), F(0.7)), F(0.1)))
        m["albedo"] = tuple(base[:3])
        m["cullBackface"] = not g.get("doubleSided", False)
        if "KHR_materials_transmission" in ext:
            m["specularTransmission"] = F(ext["KHR_materials_transmission"].get("transmissionFactor", 0.0))
            m["cullBackface"] = False
        if "baseColorTexture" in pbr:
            m["textureID"] = tex_image(pbr["baseColorTexture"])
        if "normalTexture" in g:
            m["normalMapID"] = tex_image(g["normalTexture"])
        return m

    instances = []
    for ni, mat in visits:
        node = doc["nodes"][ni]
        if "mesh" not in node:
            continue
        for model in mesh_models[node["mesh"]]:
            instances.append((model, mat.astype(F), material(models[model]["material"])))
    return dict(models=models, images=images, instances=instances)
