"""Pins the oracle integrator and shaders with closed forms: white furnace,
sky-only frame, Beer slab, mirror reflection, Disney metallic mirror limit,
the self-intersection offset, cosine-weighted diffuse sampling (chi-square),
and the running-mean blend."""
import numpy as np
import pytest

import oracle
from reina_rt import abi, scenes
from reina_rt.scene import Material, ModelData, Scene, make_desc, push_consts


def sky64(d):
    d = np.asarray(d, np.float64)
    y = d[..., 1] / np.linalg.norm(d, axis=-1)
    t = 0.5 * (y + 1.0)
    lo, hi = np.array([0.1, 0.1, 0.1]), np.array([0.4, 1.7, 2.0])
    return (lo * (1 - t)[..., None] + hi * t[..., None]) * 0.07


def quad_model(c, u, v):
    c, u, v = (np.asarray(a, np.float64) for a in (c, u, v))
    p = np.array([c - u - v, c + u - v, c + u + v, c - u + v])
    n = np.cross(u, v)
    n /= np.linalg.norm(n)
    verts = np.concatenate([p, np.ones((4, 1))], 1).astype(np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    tbn = np.tile(np.stack([u / np.linalg.norm(u), v / np.linalg.norm(v), n])[None], (4, 1, 1)).astype(np.float32)
    return ModelData(verts, idx, tbn, idx.copy())


def box_model(lo, hi, inward):
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    c, h = 0.5 * (lo + hi), 0.5 * (hi - lo)
    ex, ey, ez = np.eye(3)
    faces = [(c + h[0] * ex, h[1] * ey, h[2] * ez), (c - h[0] * ex, h[2] * ez, h[1] * ey),
             (c + h[1] * ey, h[2] * ez, h[0] * ex), (c - h[1] * ey, h[0] * ex, h[2] * ez),
             (c + h[2] * ez, h[0] * ex, h[1] * ey), (c - h[2] * ez, h[1] * ey, h[0] * ex)]
    quads = []
    for fc, u, v in faces:
        if inward:
            u, v = v, u
        quads.append(quad_model(fc, u, v))
    verts = np.concatenate([q.vertices for q in quads])
    idx = np.concatenate([q.indices + 4 * k for k, q in enumerate(quads)]).astype(np.uint32)
    tbns = np.concatenate([q.tbns for q in quads])
    return ModelData(verts, idx, tbns, idx.copy())


def render(scene, pc, W, H, use_bvh=False):
    arrays = scene.build_arrays()
    img, rays = oracle.OracleScene(arrays, make_desc).render(pc, W, H, use_bvh=use_bvh, nthreads=4)
    return img


def test_white_furnace():
    """Closed Lambertian box, albedo a, emission E everywhere: every segment
    hits a wall, so each sample is E * sum_{k<D} a^k (raytrace.rgen.glsl:143-178)."""
    a, E, D = 0.7, 0.5, 6
    s = Scene()
    s.add_object(box_model((-1, -1, -1), (1, 1, 1), inward=True), np.eye(4),
                 Material(materialIdx=0, albedo=(a, a, a), emission=(E, E, E)))
    W, H = 16, 12
    pc = push_consts((0.1, 0.2, 0.3), (0.0, 0.0, -1.0), 60.0, W, H, spp=3, max_bounces=D, direct_clamp=1e6)
    img = render(s, pc, W, H)
    expect = E * (1 - a ** D) / (1 - a)
    assert np.allclose(img[..., :3], expect, rtol=2e-6)


def test_sky_only_frame():
    """No geometry: every sample is sky(dir) (raytrace.rmiss.glsl:10-19)."""
    s = Scene()
    W, H = 8, 6
    pc = push_consts((0, 0, 0), (0.3, 0.5, -1.0), 50.0, W, H, spp=1, max_bounces=4)
    img = render(s, pc, W, H)
    for y in range(H):
        for x in range(W):
            _, d, _ = oracle.camera_ray(pc, x, y, W, H, (0 * H + y) * W + x)
            assert np.allclose(img[y, x, :3], sky64(d), rtol=1e-6, atol=1e-8)


def test_beer_slab():
    """IOR-1 slab of thickness d, roughness 0: transmittance exp(-sigma d / cos) * albedo
    times the sky behind it (dielectric.rchit.glsl:75-100)."""
    sigma, d, alb = 1.3, 0.4, 0.9
    s = Scene()
    s.add_object(box_model((-50, -50, -d), (50, 50, 0.0), inward=False), np.eye(4),
                 Material(materialIdx=2, albedo=(alb, alb, alb), ior=1.0, roughness=0.0, absorption=sigma))
    W, H = 6, 5
    pc = push_consts((0, 0, 3.0), (0.0, 0.0, 0.0), 20.0, W, H, spp=1, max_bounces=6)
    img = render(s, pc, W, H)
    for y in range(H):
        for x in range(W):
            _, dd, _ = oracle.camera_ray(pc, x, y, W, H, y * W + x)
            cos = abs(dd[2]) / np.linalg.norm(dd)
            expect = sky64(dd) * np.exp(-sigma * d / cos) * alb
            # the two self-intersection offsets move the path ends by <= 1/65536 each
            assert np.allclose(img[y, x, :3], expect, rtol=1e-4)


def test_metal_mirror():
    """Metal at roughness 0 is a mirror: albedo * sky(reflect(d)) (metal.rchit.glsl:53)."""
    c = np.array([0.9, 0.6, 0.3])
    s = Scene()
    s.add_object(quad_model((0, 0, 0), (0, 0, 100), (100, 0, 0)), np.eye(4),
                 Material(materialIdx=1, albedo=tuple(c), roughness=0.0))
    W, H = 6, 4
    pc = push_consts((0, 2.0, 2.0), (0.0, 0.0, 0.0), 30.0, W, H, spp=1, max_bounces=4)
    img = render(s, pc, W, H)
    for y in range(H):
        for x in range(W):
            _, dd, _ = oracle.camera_ray(pc, x, y, W, H, y * W + x)
            r = dd.astype(np.float64) * np.array([1, -1, 1])
            assert np.allclose(img[y, x, :3], c * sky64(r), rtol=1e-5)


def test_disney_metallic_mirror_limit():
    """Disney with metallic = 1 and roughness -> 0 (alpha = 1e-4) reflects like a
    mirror with Fresnel ~ baseColor at near-normal incidence."""
    c = np.array([0.8, 0.5, 0.2])
    s = Scene()
    s.add_object(quad_model((0, 0, 0), (0, 0, 100), (100, 0, 0)), np.eye(4),
                 Material(materialIdx=3, albedo=tuple(c), roughness=0.0, metallic=1.0))
    W, H = 4, 4
    pc = push_consts((0, 3.0, 0.3), (0.0, 0.0, 0.0), 10.0, W, H, spp=1, max_bounces=4)
    img = render(s, pc, W, H)
    for y in range(H):
        for x in range(W):
            _, dd, _ = oracle.camera_ray(pc, x, y, W, H, y * W + x)
            r = dd.astype(np.float64) * np.array([1, -1, 1])
            assert np.allclose(img[y, x, :3], c * sky64(r), rtol=2e-3)


def test_offset_position_closed_form():
    """closestHitCommon.h.glsl:156-177: p + sign(p) * int(256 n) ulps, or
    p + n / 65536 when |p| < 1/32."""
    rng = np.random.default_rng(6)
    for _ in range(2000):
        p = rng.normal(size=3).astype(np.float32) * np.float32(rng.choice([0.01, 1.0, 100.0]))
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        got = oracle.offset_position(p, n)
        for k in range(3):
            if abs(p[k]) < 1.0 / 32.0:
                exp = np.float32(p[k] + np.float32(1.0 / 65536.0) * n[k])
            else:
                off = int(np.trunc(np.float32(256.0) * n[k]))
                bits = int(np.array([p[k]], np.float32).view(np.int32)[0])
                exp = np.array([bits + (off if p[k] >= 0 else -off)], np.int32).view(np.float32)[0]
            assert got[k] == exp
    # dielectric variant flips n when dot(n, dir) < 0
    p, n = np.array([1.0, 2.0, 3.0], np.float32), np.array([0.0, 1.0, 0.0], np.float32)
    a = oracle.offset_position(p, n, dielectric=True, d=np.array([0, -1, 0], np.float32))
    b = oracle.offset_position(p, -n)
    assert np.array_equal(a, b)


def test_diffuse_sampling_chi2():
    """diffuseReflection (closestHitCommon.h.glsl:215-222) is cosine-weighted:
    P(cos theta in [a, b]) = b^2 - a^2."""
    n = np.array([0.0, 0.0, 1.0], np.float32)
    rng = np.random.default_rng(7)
    cos = []
    for s in rng.integers(0, 2 ** 32, 40000):
        d, _ = oracle.diffuse_dir(n, int(s))
        cos.append(d[2])
    cos = np.clip(np.array(cos), 0, 1)
    edges = np.linspace(0, 1, 11)
    obs, _ = np.histogram(cos, edges)
    exp = (edges[1:] ** 2 - edges[:-1] ** 2) * len(cos)
    chi2 = ((obs - exp) ** 2 / exp).sum()
    assert chi2 < 30.0   # 9 dof: p ~ 4e-4


def test_running_mean_blend(cornell_arrays):
    """sampleBatch b > 0: (prev * b + avg) / (b + 1) per channel."""
    arrays, _ = cornell_arrays
    W, H = 8, 8
    orc = oracle.OracleScene(arrays, make_desc)
    prev = np.full((H, W, 4), 0.25, np.float32)
    pc1 = scenes.cornell_push(W, H, spp=2, max_bounces=3, sample_batch=3)
    fresh, _ = orc.render(pc1, W, H, img=np.zeros((H, W, 4), np.float32))   # blends with 0
    blended, _ = orc.render(pc1, W, H, img=prev.copy())
    # fresh = (0 * 3 + avg) / 4 -> avg = 4 * fresh; blended = (0.25 * 3 + avg) / 4
    avg = fresh[..., :3].astype(np.float64) * 4
    assert np.allclose(blended[..., :3], (0.25 * 3 + avg) / 4, rtol=1e-6, atol=1e-7)
