"""Seeded, deterministic procedural scenes for the benchmark configurations.

cornell()      40 triangles: floor and ceiling as two quads each, back, red
               left and green right walls, a ceiling quad light (emission 15,
               at the height of models/cornell_light.obj), a tall box and a
               short box rotated 15 degrees (unit-cube model instanced with
               scale/rotate/translate). Open front; the box spans
               [-1,1]x[0,2]x[-1,1] like models/cornell_box.obj.
icosphere()    icosahedron subdivided 7 times (327,680 triangles), vertices
               displaced radially by 4-octave value noise (amplitude 0.05);
               one model + Disney instance per group of 256 triangles
               (1,280 groups, roughness / metallic / base colour uniform per
               group), a ground quad and two quad lights.
instanced()    copies of a 2,048-triangle displaced UV sphere with random
               scale, rotation and position in a 100^3 box; materials by
               instance index mod 3 (dielectric, metal, Lambertian).
"""
import numpy as np

from .scene import Material, ModelData, Scene, push_consts


# ---- helpers ---------------------------------------------------------------------
def _quad(c, u, v):
    """Quad c +- u +- v with geometric normal along cross(u, v); flat TBN (u, v, n)."""
    c, u, v = (np.asarray(a, np.float64) for a in (c, u, v))
    p = np.array([c - u - v, c + u - v, c + u + v, c - u + v])
    n = np.cross(u, v)
    n /= np.linalg.norm(n)
    t = u / np.linalg.norm(u)
    b = v / np.linalg.norm(v)
    return p, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), (t, b, n)


def _model_from_quads(quads, uv_scale=None):
    """uv_scale: give every quad texture coordinates (0,0), (s,0), (s,s), (0,s)
    at its corners (s > 1 repeats the texture)."""
    verts, idx, tbns, uvs = [], [], [], []
    for p, f, (t, b, n) in quads:
        base = len(verts)
        for k, q in enumerate(p):
            verts.append([q[0], q[1], q[2], 1.0])
            tbns.append(np.stack([t, b, n]))
            if uv_scale is not None:
                uvs.append([uv_scale * (k in (1, 2)), uv_scale * (k in (2, 3))])
        idx.extend((f + base).reshape(-1).tolist())
    verts = np.array(verts, np.float32)
    idx = np.array(idx, np.uint32)
    tbns = np.array(tbns, np.float32)
    if uv_scale is None:
        return ModelData(verts, idx, tbns, idx.copy())
    return ModelData(verts, idx, tbns, idx.copy(), np.array(uvs, np.float32), idx.copy())


def _translate(t):
    m = np.eye(4, dtype=np.float64)
    m[:3, 3] = t
    return m


def _scale(s):
    return np.diag([s[0], s[1], s[2], 1.0])


def _rot_y(deg):
    a = np.radians(deg)
    m = np.eye(4)
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = np.cos(a), np.sin(a), -np.sin(a), np.cos(a)
    return m


def _rot_z(deg):
    a = np.radians(deg)
    m = np.eye(4)
    m[0, 0], m[0, 1], m[1, 0], m[1, 1] = np.cos(a), -np.sin(a), np.sin(a), np.cos(a)
    return m


def _unit_cube():
    h = 0.5
    faces = [
        ((h, 0, 0), (0, h, 0), (0, 0, h)), ((-h, 0, 0), (0, 0, h), (0, h, 0)),
        ((0, h, 0), (0, 0, h), (h, 0, 0)), ((0, -h, 0), (h, 0, 0), (0, 0, h)),
        ((0, 0, h), (h, 0, 0), (0, h, 0)), ((0, 0, -h), (0, h, 0), (h, 0, 0)),
    ]
    return _model_from_quads([_quad(*f) for f in faces])


# ---- C1: Cornell box -------------------------------------------------------------
RED = (0.63, 0.065, 0.05)
GREEN = (0.14, 0.45, 0.091)
GREY = (0.73, 0.73, 0.73)


def cornell(boxes=True):
    """Returns (Scene, camera dict). 40 triangles (16 without the boxes)."""
    s = Scene()
    wall = lambda albedo: Material(materialIdx=0, albedo=albedo, cullBackface=True)
    I = np.eye(4)
    floor = [_quad((0, 0, z), (0, 0, 0.5), (1, 0, 0)) for z in (0.5, -0.5)]
    ceil = [_quad((0, 2, z), (1, 0, 0), (0, 0, 0.5)) for z in (-0.5, 0.5)]
    s.add_object(_model_from_quads(floor), I, wall(GREY))
    s.add_object(_model_from_quads(ceil), I, wall(GREY))
    s.add_object(_model_from_quads([_quad((0, 1, -1), (1, 0, 0), (0, 1, 0))]), I, wall(GREY))
    s.add_object(_model_from_quads([_quad((-1, 1, 0), (0, 1, 0), (0, 0, 1))]), I, wall(RED))
    s.add_object(_model_from_quads([_quad((1, 1, 0), (0, 0, 1), (0, 1, 0))]), I, wall(GREEN))
    light = _model_from_quads([_quad((-0.005, 1.989, -0.03), (0.235, 0, 0), (0, 0, 0.19))])
    s.add_object(light, I, Material(materialIdx=0, albedo=(0.9, 0.9, 0.9), emission=(15.0, 15.0, 15.0),
                                    cullBackface=True))
    cam = dict(eye=(0.0, 1.0, 3.4), center=(0.0, 1.0, 0.0), fov=40.0)
    if not boxes:
        return s, cam
    cube = s.define_object(_unit_cube())
    tall = _translate((-0.35, 0.6, -0.35)) @ _scale((0.6, 1.2, 0.6))
    short = _translate((0.35, 0.3, 0.3)) @ _rot_y(15.0) @ _scale((0.6, 0.6, 0.6))
    s.add_instance(cube, tall, Material(materialIdx=0, albedo=GREY))
    s.add_instance(cube, short, Material(materialIdx=0, albedo=GREY))
    return s, cam


def cornell_push(width=256, height=256, spp=16, max_bounces=8, sample_batch=0):
    _, cam = cornell()
    return push_consts(cam["eye"], cam["center"], cam["fov"], width, height, spp, max_bounces,
                       sample_batch=sample_batch, defocus=0.0, direct_clamp=100.0)


# ---- textured Cornell box (the hit shaders' texture, normal-map and bump-map paths) -
def checker_texture(w=64, h=48, cells=8, seed=1):
    """RGBA8 checkerboard; some cells fully transparent, some half."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    c = ((x * cells // w) + (y * cells // h)) % 2
    img = np.zeros((h, w, 4), np.uint8)
    img[..., 0] = np.where(c, 230, 40)
    img[..., 1] = np.where(c, 200, 90)
    img[..., 2] = np.where(c, 60, 200)
    alpha = np.full((cells, cells), 255, np.uint8)
    alpha[rng.integers(cells, size=4), rng.integers(cells, size=4)] = 0
    alpha[rng.integers(cells, size=4), rng.integers(cells, size=4)] = 128
    img[..., 3] = alpha[y * cells // h, x * cells // w]
    return img


def wave_normal_map(w=32, h=32, amp=0.35):
    """Tangent-space normal map of a wavy surface, encoded rgb = n * 0.5 + 0.5."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    n = np.stack([amp * np.sin(2 * np.pi * x / w), amp * np.cos(2 * np.pi * y / h), np.ones_like(x)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    img = np.zeros((h, w, 4), np.uint8)
    img[..., :3] = np.clip(np.round((n * 0.5 + 0.5) * 255), 0, 255)
    img[..., 3] = 255
    return img


def height_texture(w=32, h=32, seed=5):
    """Height map in the red channel (value noise in [0, 1])."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    p = np.stack([x / w * 4, y / h * 4, np.zeros_like(x)], -1)
    v = value_noise(p.reshape(-1, 3), seed, octaves=3).reshape(h, w)
    v = (v - v.min()) / max(v.max() - v.min(), 1e-9)
    img = np.zeros((h, w, 4), np.uint8)
    img[..., 0] = np.round(v * 255)
    img[..., 3] = 255
    return img


def textured_cornell(white_opaque=False):
    """The Cornell box with texture coordinates on every wall and box.
    Default: the back wall has a checker albedo texture with transparent
    cells, repeated twice (Lambertian); the floor a height map (bump
    mapping) and the checker; the left wall and the ceiling a normal map;
    the right wall is a normal-mapped metal, the tall box a textured,
    bump-mapped dielectric, the short box a textured, normal-mapped Disney
    surface. white_opaque: cornell()'s materials, each with an opaque white
    (255, 255, 255, 255) albedo texture and no normal or bump map: it renders
    exactly like cornell()."""
    s = Scene()
    I = np.eye(4)
    if white_opaque:
        tex = s.add_texture(np.full((4, 4, 4), 255, np.uint8))
        wall = lambda albedo: Material(materialIdx=0, albedo=albedo, cullBackface=True, textureID=tex)
        floor_m = ceil_m = back_m = wall(GREY)
        left_m, right_m = wall(RED), wall(GREEN)
        tall_m = short_m = Material(materialIdx=0, albedo=GREY, textureID=tex)
    else:
        tex = s.add_texture(checker_texture())
        nrm = s.add_texture(wave_normal_map())
        bump = s.add_texture(height_texture())
        floor_m = Material(materialIdx=0, albedo=GREY, cullBackface=True, textureID=tex, bumpMapID=bump)
        ceil_m = Material(materialIdx=0, albedo=GREY, cullBackface=True, normalMapID=nrm)
        back_m = Material(materialIdx=0, albedo=GREY, cullBackface=True, textureID=tex)
        left_m = Material(materialIdx=0, albedo=RED, cullBackface=True, textureID=tex, normalMapID=nrm)
        right_m = Material(materialIdx=1, albedo=GREEN, roughness=0.2, normalMapID=nrm)
        tall_m = Material(materialIdx=2, albedo=(0.9, 0.95, 0.9), ior=1.5, textureID=tex, bumpMapID=bump)
        short_m = Material(materialIdx=3, albedo=GREY, roughness=0.4, metallic=0.3, textureID=tex, normalMapID=nrm,
                           interpNormals=True)
    floor = [_quad((0, 0, z), (0, 0, 0.5), (1, 0, 0)) for z in (0.5, -0.5)]
    ceil = [_quad((0, 2, z), (1, 0, 0), (0, 0, 0.5)) for z in (-0.5, 0.5)]
    s.add_object(_model_from_quads(floor, 1.0), I, floor_m)
    s.add_object(_model_from_quads(ceil, 1.0), I, ceil_m)
    s.add_object(_model_from_quads([_quad((0, 1, -1), (1, 0, 0), (0, 1, 0))], 2.0), I, back_m)
    s.add_object(_model_from_quads([_quad((-1, 1, 0), (0, 1, 0), (0, 0, 1))], 1.0), I, left_m)
    s.add_object(_model_from_quads([_quad((1, 1, 0), (0, 0, 1), (0, 1, 0))], 1.0), I, right_m)
    light = _model_from_quads([_quad((-0.005, 1.989, -0.03), (0.235, 0, 0), (0, 0, 0.19))])
    s.add_object(light, I, Material(materialIdx=0, albedo=(0.9, 0.9, 0.9), emission=(15.0, 15.0, 15.0),
                                    cullBackface=True))
    h = 0.5
    cube_faces = [((h, 0, 0), (0, h, 0), (0, 0, h)), ((-h, 0, 0), (0, 0, h), (0, h, 0)),
                  ((0, h, 0), (0, 0, h), (h, 0, 0)), ((0, -h, 0), (h, 0, 0), (0, 0, h)),
                  ((0, 0, h), (h, 0, 0), (0, h, 0)), ((0, 0, -h), (0, h, 0), (h, 0, 0))]
    cube = s.define_object(_model_from_quads([_quad(*f) for f in cube_faces], 1.0))
    tall = _translate((-0.35, 0.6, -0.35)) @ _scale((0.6, 1.2, 0.6))
    short = _translate((0.35, 0.3, 0.3)) @ _rot_y(15.0) @ _scale((0.6, 0.6, 0.6))
    s.add_instance(cube, tall, tall_m)
    s.add_instance(cube, short, short_m)
    cam = dict(eye=(0.0, 1.0, 3.4), center=(0.0, 1.0, 0.0), fov=40.0)
    return s, cam


def obj_cornell(models_dir, textures_dir=None, subject="uv_sphere_highres.obj"):
    """The OBJ scene of src/Reina.cpp:103-115, loaded through the library's
    OBJ and PNG loaders: models/cornell_box.obj with textures/cornell_texture.png
    (Lambertian, back faces culled), models/cornell_light.obj (emission 16) and
    a rough Disney subject scaled by 0.25 after a (0, 2, 0) translation
    (glm::translate(glm::scale(I, 0.25), (0, 2, 0))). textures_dir defaults to
    models_dir."""
    import os
    textures_dir = textures_dir or models_dir
    s = Scene()
    wall_tex = s.define_texture(os.path.join(textures_dir, "cornell_texture.png"))
    wall = Material(materialIdx=0, textureID=wall_tex, albedo=(0.9, 0.9, 0.9), roughness=0.0, ior=0.0,
                    cullBackface=True, sheenTint=(0.0, 0.0, 0.0))
    subject_m = Material(materialIdx=3, albedo=(1.0, 1.0, 1.0), roughness=0.5, ior=1.5, interpNormals=True,
                         sheenTint=(1.0, 1.0, 1.0))
    light_m = Material(materialIdx=0, albedo=(0.9, 0.9, 0.9), emission=(16.0, 16.0, 16.0), roughness=0.0, ior=0.0,
                       cullBackface=True, sheenTint=(0.0, 0.0, 0.0))
    I = np.eye(4)
    s.add_object(os.path.join(models_dir, "cornell_box.obj"), I, wall)
    s.add_object(os.path.join(models_dir, "cornell_light.obj"), I, light_m)
    s.add_object(os.path.join(models_dir, subject), _scale((0.25, 0.25, 0.25)) @ _translate((0.0, 2.0, 0.0)),
                 subject_m)
    cam = dict(eye=(0.0, 1.0, 3.4), center=(0.0, 1.0, 0.0), fov=40.0)
    return s, cam


# ---- value noise -----------------------------------------------------------------
def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.int64) * 73856093) ^ (iy.astype(np.int64) * 19349663) ^ (iz.astype(np.int64) * 83492791)
    h = (h ^ (seed * 2654435761)) & 0xFFFFFFFF
    h = (h * 747796405 + 2891336453) & 0xFFFFFFFF
    h = ((h >> ((h >> 28) + 4)) ^ h) * 277803737 & 0xFFFFFFFF
    h = (h >> 22) ^ h
    return (h.astype(np.float64) / 4294967295.0) * 2.0 - 1.0


def value_noise(p, seed, octaves=4, freq=2.0):
    """Sum of `octaves` trilinear value-noise octaves (smoothstep weights),
    amplitude halving and frequency doubling, normalised to [-1, 1]."""
    total = np.zeros(len(p))
    amp, norm = 1.0, 0.0
    for o in range(octaves):
        q = p * freq * (2.0 ** o)
        i0 = np.floor(q)
        f = q - i0
        w = f * f * (3.0 - 2.0 * f)
        i0 = i0.astype(np.int64)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    val = _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed + 1013 * o)
                    wx = w[:, 0] if dx else 1.0 - w[:, 0]
                    wy = w[:, 1] if dy else 1.0 - w[:, 1]
                    wz = w[:, 2] if dz else 1.0 - w[:, 2]
                    acc += val * wx * wy * wz
        total += amp * acc
        norm += amp
        amp *= 0.5
    return total / norm


# ---- C2: displaced icosphere -------------------------------------------------------
def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v, f


def subdivided_icosphere(levels):
    """Unit icosphere; children of face i are faces 4i..4i+3 (depth-first order)."""
    v, f = _icosahedron()
    for _ in range(levels):
        nv = len(v)
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        lo, hi = np.minimum(e[:, 0], e[:, 1]), np.maximum(e[:, 0], e[:, 1])
        key = lo * nv + hi
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // nv, uk % nv
        m = v[a] + v[b]
        m /= np.linalg.norm(m, axis=1, keepdims=True)
        mid = nv + inv.reshape(3, -1).T          # (nf, 3): ab, bc, ca
        v = np.concatenate([v, m])
        ab, bc, ca = mid[:, 0], mid[:, 1], mid[:, 2]
        a0, b0, c0 = f[:, 0], f[:, 1], f[:, 2]
        nf = np.stack([np.stack([a0, ab, ca], 1), np.stack([ab, b0, bc], 1), np.stack([ca, bc, c0], 1),
                       np.stack([ab, bc, ca], 1)], 1)
        f = nf.reshape(-1, 3)
    return v, f


def _vertex_frames(v, f):
    """Area-weighted vertex normals and a tangent frame (T, B, N)."""
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    n = np.zeros_like(v)
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    ref = np.where(np.abs(n[:, 1:2]) < 0.99, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(ref, n)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(n, t)
    return np.stack([t, b, n], 1)


def icosphere(levels=7, seed=7, amplitude=0.05, group=256):
    """Returns (Scene, camera dict)."""
    v, f = subdivided_icosphere(levels)
    r = 1.0 + amplitude * value_noise(v, seed)
    vd = v * r[:, None]
    frames = _vertex_frames(vd, f)
    rng = np.random.default_rng(seed)
    ngroups = len(f) // group
    params = rng.random((ngroups, 5))   # roughness, metallic, r, g, b
    s = Scene()
    I = np.eye(4)
    v32 = vd.astype(np.float32)
    fr32 = frames.astype(np.float32)
    for g in range(ngroups):
        ff = f[g * group:(g + 1) * group]
        uv, loc = np.unique(ff.reshape(-1), return_inverse=True)
        verts = np.concatenate([v32[uv], np.ones((len(uv), 1), np.float32)], 1)
        idx = loc.astype(np.uint32)
        md = ModelData(verts, idx, fr32[uv], idx.copy())
        rough, metal, cr, cg, cb = params[g]
        s.add_object(md, I, Material(materialIdx=3, albedo=(cr, cg, cb), roughness=float(rough),
                                     metallic=float(metal), ior=1.5, interpNormals=True,
                                     specularTint=(1.0, 1.0, 1.0)))
    ground = _model_from_quads([_quad((0, -1.1, 0), (0, 0, 20), (20, 0, 0))])
    s.add_object(ground, I, Material(materialIdx=0, albedo=(0.5, 0.5, 0.5)))
    l1 = _model_from_quads([_quad((-2.0, 3.0, 1.0), (0.5, 0, 0), (0, 0, 0.5))])
    l2 = _model_from_quads([_quad((2.5, 2.5, -1.0), (0.5, 0, 0), (0, 0, 0.5))])
    s.add_object(l1, I, Material(materialIdx=0, albedo=(0.9, 0.9, 0.9), emission=(10.0, 10.0, 10.0),
                                 cullBackface=True))
    s.add_object(l2, I, Material(materialIdx=0, albedo=(0.9, 0.9, 0.9), emission=(6.0, 6.0, 8.0),
                                 cullBackface=True))
    cam = dict(eye=(0.0, 0.4, 3.2), center=(0.0, 0.0, 0.0), fov=45.0)
    return s, cam


# ---- C3: instanced displaced spheres ----------------------------------------------
def uv_sphere(segments=32, rings=32, seed=3, amplitude=0.05):
    """segments x rings quads -> 2*segments*rings triangles (pole triangles are
    kept, degenerate), radially displaced by value noise."""
    th = np.linspace(0.0, np.pi, rings + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, segments + 1)[:-1]
    T, P = np.meshgrid(th, ph, indexing="ij")
    v = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    r = 1.0 + amplitude * value_noise(v, seed)
    v = v * r[:, None]
    faces = []
    for i in range(rings):
        for j in range(segments):
            a = i * segments + j
            b = i * segments + (j + 1) % segments
            c = (i + 1) * segments + j
            d = (i + 1) * segments + (j + 1) % segments
            faces.append([a, c, b])
            faces.append([b, c, d])
    f = np.array(faces, np.int64)
    return v, f


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def instanced(copies=4096, seed=11, box=100.0):
    """Returns (Scene, camera dict)."""
    v, f = uv_sphere()
    with np.errstate(invalid="ignore"):
        frames = _vertex_frames(v, f)
    frames = np.nan_to_num(frames)
    md = ModelData(np.concatenate([v, np.ones((len(v), 1))], 1).astype(np.float32), f.reshape(-1).astype(np.uint32),
                   frames.astype(np.float32), f.reshape(-1).astype(np.uint32))
    s = Scene()
    oid = s.define_object(md)
    rng = np.random.default_rng(seed)
    for k in range(copies):
        sc = rng.uniform(0.2, 1.0)
        R = _random_rotation(rng)
        pos = rng.uniform(0.0, box, size=3)
        M = np.eye(4)
        M[:3, :3] = R * sc
        M[:3, 3] = pos
        kind = k % 3
        if kind == 0:
            mat = Material(materialIdx=2, albedo=(0.95, 0.95, 0.95), ior=1.5,
                           absorption=float(rng.uniform(0.0, 2.0)))
        elif kind == 1:
            mat = Material(materialIdx=1, albedo=tuple(rng.uniform(0.3, 1.0, 3)), roughness=float(rng.uniform(0.0, 0.5)))
        else:
            mat = Material(materialIdx=0, albedo=tuple(rng.uniform(0.2, 0.9, 3)))
        s.add_instance(oid, M, mat)
    h = 0.5 * box
    cam = dict(eye=(h, h, -0.4 * box), center=(h, h, h), fov=60.0)
    return s, cam


def disney_lobes():
    """The Cornell box with its cubes replaced by eight spheres (radius 0.18,
    288 triangles each) in two rows, each Disney material (materialIdx 3)
    weighted toward one lobe of brdfDisney.h.glsl: clearcoat, sheen,
    subsurface, anisotropic metal, rough glass (specularTransmission 1),
    a glass / clearcoat mix, tinted specular, and every lobe at once.
    Returns (Scene, camera dict)."""
    s, cam = cornell(boxes=False)
    sphere = s.define_object(_sphere_model(12, 12))
    D = dict(materialIdx=3, ior=1.5, interpNormals=True)
    mats = [Material(albedo=(0.8, 0.3, 0.2), roughness=0.5, clearcoat=1.0, clearcoatGloss=0.9, **D),
            Material(albedo=(0.2, 0.4, 0.8), roughness=0.8, sheen=1.0, sheenTint=(1.0, 0.5, 0.5), **D),
            Material(albedo=(0.9, 0.8, 0.6), roughness=0.4, subsurface=1.0, **D),
            Material(albedo=(0.9, 0.7, 0.3), roughness=0.3, metallic=1.0, anisotropic=0.8, **D),
            Material(albedo=(0.95, 0.95, 0.95), roughness=0.1, specularTransmission=1.0, **D),
            Material(albedo=(0.7, 0.9, 0.7), roughness=0.2, specularTransmission=0.5, metallic=0.2,
                     clearcoat=0.5, clearcoatGloss=0.3, **D),
            Material(albedo=(0.6, 0.6, 0.6), roughness=0.6, metallic=0.5, specularTint=(0.3, 0.6, 1.0), **D),
            Material(albedo=(0.5, 0.7, 0.9), roughness=0.35, metallic=0.3, clearcoat=0.3, clearcoatGloss=0.6,
                     sheen=0.5, sheenTint=(0.5, 1.0, 0.5), subsurface=0.5, anisotropic=0.5,
                     specularTransmission=0.3, **D)]
    for k, m in enumerate(mats):
        s.add_instance(sphere, _translate((-0.6 + 0.4 * (k % 4), 0.45 + 0.8 * (k // 4), -0.1 * (k // 4)))
                       @ _scale((0.18, 0.18, 0.18)), m)
    return s, cam


def _slab(y_top, y_bottom, half=50.0):
    """Two horizontal quads: the top one facing +y, the bottom one -y."""
    return _model_from_quads([_quad((0, y_top, 0), (0, 0, half), (half, 0, 0)),
                              _quad((0, y_bottom, 0), (half, 0, 0), (0, 0, half))])


def _down_camera(fov=0.5):
    return dict(eye=(0.0, 1.0, 0.0), center=(0.0, -10.0, 0.0), fov=fov, up=(0.0, 0.0, 1.0))


def dielectric_stack():
    """Horizontal dielectric slabs (IOR 1, roughness 0: rays at normal
    incidence refract straight through) below a camera looking down -y:
    B [0, -0.3] with absorption 1, C [-0.5, -0.6] emitting 2, D [-1, -3]
    holding E [-1.4, -1.6]. With 6 bounces a path ends just after entering E
    from inside D, 0.4 of distance accumulated and never applied."""
    s = Scene()
    I = np.eye(4)
    glass = dict(materialIdx=2, albedo=(1.0, 1.0, 1.0), ior=1.0, roughness=0.0)
    s.add_object(_slab(0.0, -0.3), I, Material(absorption=1.0, **glass))
    s.add_object(_slab(-0.5, -0.6), I, Material(emission=(2.0, 2.0, 2.0), **glass))
    s.add_object(_slab(-1.0, -3.0), I, Material(**glass))
    s.add_object(_slab(-1.4, -1.6), I, Material(**glass))
    return s, _down_camera()


def nan_floor_slab():
    """A clear glass slab [0, -0.5] (IOR 2.618: normal-incidence reflectance
    0.2) above a Lambertian floor at y = -1 that emits NaN, under a camera
    looking down -y: every sample either leaves upward (the zenith sky) or
    reaches the floor (NaN)."""
    s = Scene()
    I = np.eye(4)
    s.add_object(_slab(0.0, -0.5), I, Material(materialIdx=2, albedo=(1.0, 1.0, 1.0), ior=2.618, roughness=0.0))
    s.add_object(_model_from_quads([_quad((0, -1.0, 0), (0, 0, 50.0), (50.0, 0, 0))]), I,
                 Material(materialIdx=0, albedo=(0.5, 0.5, 0.5), emission=(float("nan"),) * 3))
    return s, _down_camera()


TILTED_MIRROR_XFORM = np.diag([1.5, 1.0, 1.0, 1.0]) @ _rot_z(20.0)


def tilted_mirror():
    """A mirror plane (metal, roughness 0, albedo (0.9, 0.8, 0.7)) through
    the origin, object normal +y, placed by the non-uniform transform
    TILTED_MIRROR_XFORM = diag(1.5, 1, 1) * rot_z(20 deg): its true normal
    normalize(S^-1 R y) is tilted 13.6 deg from +y, the mat3(O2W) one
    normalize(S R y) 28.6 deg. A camera looks down at it from (0, 3, 0.3);
    both reflections leave the plane."""
    s = Scene()
    # small enough that the FP32 intersection error (ulp of the vertex
    # distances) stays well below the 256-ulp origin offset: no self-hits
    plane = _model_from_quads([_quad((0, 0, 0), (0, 0, 2.0), (2.0, 0, 0))])
    s.add_object(plane, TILTED_MIRROR_XFORM, Material(materialIdx=1, albedo=(0.9, 0.8, 0.7), roughness=0.0))
    return s, dict(eye=(0.0, 3.0, 0.3), center=(0.0, 0.0, 0.0), fov=10.0)


# ---- scene pools for batching and serving ------------------------------------------
_SPHERE_CACHE = {}


def _sphere_model(segments, rings):
    key = (segments, rings)
    if key not in _SPHERE_CACHE:
        v, f = uv_sphere(segments, rings)
        with np.errstate(invalid="ignore"):
            frames = np.nan_to_num(_vertex_frames(v, f))
        _SPHERE_CACHE[key] = ModelData(np.concatenate([v, np.ones((len(v), 1))], 1).astype(np.float32),
                                       f.reshape(-1).astype(np.uint32), frames.astype(np.float32),
                                       f.reshape(-1).astype(np.uint32))
    return _SPHERE_CACHE[key]


def pool_scene(ntris, seed):
    """A seeded scene of about `ntris` triangles for the batching / serving
    pools: copies of a displaced sphere (2,048 triangles each, or one finer
    or coarser sphere below 8,192 triangles) at random scale, rotation and
    position in a cube sized for constant density, materials alternating
    Disney (random roughness / metallic / colour) and dielectric (IOR 1.5,
    Beer absorption U[0, 1]), plus a ground quad and a quad light.
    Returns (Scene, camera dict with eye, center, fov, radius)."""
    rng = np.random.default_rng(seed)
    if ntris <= 8192:
        seg = max(4, int(round(np.sqrt(ntris / 2.0))))
        model, copies = _sphere_model(seg, seg), 1
    else:
        model, copies = _sphere_model(32, 32), max(1, int(round(ntris / 2048.0)))
    side = 4.0 * copies ** (1.0 / 3.0)
    s = Scene()
    oid = s.define_object(model)
    for k in range(copies):
        M = np.eye(4)
        M[:3, :3] = _random_rotation(rng) * rng.uniform(0.3, 1.0)
        M[:3, 3] = rng.uniform(0.0, side, size=3) if copies > 1 else 0.5 * side
        if k % 2 == 0:
            mat = Material(materialIdx=3, albedo=tuple(rng.uniform(0.2, 0.95, 3)), roughness=float(rng.uniform()),
                           metallic=float(rng.uniform()), ior=1.5, interpNormals=True,
                           specularTint=(1.0, 1.0, 1.0))
        else:
            mat = Material(materialIdx=2, albedo=(0.95, 0.95, 0.95), ior=1.5,
                           absorption=float(rng.uniform(0.0, 1.0)))
        s.add_instance(oid, M, mat)
    c = 0.5 * side
    ground = _model_from_quads([_quad((c, -1.0, c), (0, 0, 2.0 * side), (2.0 * side, 0, 0))])
    s.add_object(ground, np.eye(4), Material(materialIdx=0, albedo=(0.5, 0.5, 0.5)))
    light = _model_from_quads([_quad((c, side + 2.0, c), (0.25 * side, 0, 0), (0, 0, 0.25 * side))])
    s.add_object(light, np.eye(4), Material(materialIdx=0, albedo=(0.9, 0.9, 0.9), emission=(8.0, 8.0, 8.0)))
    cam = dict(eye=(c, c + 0.3 * side, c - 1.4 * side - 2.0), center=(c, c, c), fov=50.0,
               radius=1.5 * side + 2.0)
    return s, cam


def orbit_camera(cam, rng):
    """A random camera on the sphere of radius cam['radius'] around the
    scene centre (upper hemisphere), looking at the centre."""
    c = np.asarray(cam["center"], np.float64)
    while True:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if d[1] > 0.05 and abs(d[1]) < 0.95:
            break
    eye = c + cam["radius"] * d
    return dict(eye=tuple(eye), center=tuple(c), fov=cam["fov"])


def light_box(height=0.3, albedo=0.7, emission=2.0, half=1.0):
    """A closed box for direct-lighting checks: Lambertian floor y = 0 of
    albedo `albedo`, the whole ceiling y = height a quad light of radiance
    `emission` facing down, black walls; every face culls its back side.
    Camera inside, looking down at the floor. Returns (Scene, camera)."""
    s = Scene()
    I = np.eye(4)
    h, a = float(height), float(half)
    floor = _model_from_quads([_quad((0, 0, 0), (0, 0, a), (a, 0, 0))])            # normal +y
    ceil = _model_from_quads([_quad((0, h, 0), (a, 0, 0), (0, 0, a))])             # normal -y
    walls = _model_from_quads([_quad((0, 0.5 * h, -a), (a, 0, 0), (0, 0.5 * h, 0)),  # normal +z
                               _quad((0, 0.5 * h, a), (0, 0.5 * h, 0), (a, 0, 0)),   # normal -z
                               _quad((-a, 0.5 * h, 0), (0, 0.5 * h, 0), (0, 0, a)),  # normal +x
                               _quad((a, 0.5 * h, 0), (0, 0, a), (0, 0.5 * h, 0))])  # normal -x
    s.add_object(floor, I, Material(materialIdx=0, albedo=(albedo,) * 3, cullBackface=True))
    s.add_object(ceil, I, Material(materialIdx=0, albedo=(0.0,) * 3, emission=(emission,) * 3, cullBackface=True))
    s.add_object(walls, I, Material(materialIdx=0, albedo=(0.0,) * 3, cullBackface=True))
    cam = dict(eye=(0.0, 0.9 * h, 0.9 * a), center=(0.0, 0.0, 0.0), fov=70.0)
    return s, cam
