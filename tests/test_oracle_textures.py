"""Textured hits in the oracle: the hit shaders' texture(textures[id], uv)
(RGBA8 UNORM, linear, repeat, base level: a ray-tracing stage has no
derivatives), the parallax search of texutils.h.glsl:4-41, and the texture /
normal-map / bump-map branches of the four hit shaders. Pinned by exact
texel-centre values, repeat periodicity, an independent float64 evaluation
of the Vulkan bilinear formula, the closed form of the parallax search over
a constant height field, and a white opaque texture rendering exactly like
no texture."""
import numpy as np
import pytest

import oracle
from reina_rt import scenes
from reina_rt.scene import make_desc


def _oracle_with(textures):
    s, _ = scenes.cornell()
    for t in textures:
        s.add_texture(t)
    return oracle.OracleScene(s.build_arrays(), make_desc)


def vk_bilinear64(img, u, v):
    """The Vulkan linear filter with repeat addressing, in float64."""
    h, w = img.shape[:2]
    x, y = u * w - 0.5, v * h - 0.5
    fx, fy = np.floor(x), np.floor(y)
    a, b = x - fx, y - fy
    i0, j0 = int(fx) % w, int(fy) % h
    i1, j1 = (i0 + 1) % w, (j0 + 1) % h
    t = img.astype(np.float64) / 255.0
    return ((1 - a) * (1 - b) * t[j0, i0] + a * (1 - b) * t[j0, i1] + (1 - a) * b * t[j1, i0] + a * b * t[j1, i1])


def test_texel_centres_are_exact():
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (4, 8, 4)).astype(np.uint8)
    orc = _oracle_with([img])
    for j in range(4):
        for i in range(8):
            got = orc.tex_sample(0, (i + 0.5) / 8, (j + 0.5) / 4)
            assert np.array_equal(got, (img[j, i].astype(np.float32) / np.float32(255)))


def test_constant_texture_and_repeat():
    img = np.zeros((5, 7, 4), np.uint8)
    img[...] = (10, 200, 77, 255)
    orc = _oracle_with([img, np.random.default_rng(2).integers(0, 256, (8, 16, 4)).astype(np.uint8)])
    ref = np.array([10, 200, 77, 255], np.float32) / np.float32(255)
    rng = np.random.default_rng(3)
    for u, v in rng.uniform(-5, 5, (200, 2)):
        assert np.array_equal(orc.tex_sample(0, u, v), ref)   # interpolating equal texels is exact
    # dyadic coordinates: u, u + 1 and u - 2 sample the same texels with the same weights
    for k in range(-64, 64, 3):
        u, v = k / 64.0, (k * 5 % 64) / 64.0
        s0 = orc.tex_sample(1, u, v)
        assert np.array_equal(s0, orc.tex_sample(1, u + 1.0, v))
        assert np.array_equal(s0, orc.tex_sample(1, u - 2.0, v + 3.0))


def test_bilinear_matches_float64_formula():
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (23, 37, 4)).astype(np.uint8)
    orc = _oracle_with([img])
    for u, v in rng.uniform(-3, 3, (2000, 2)).astype(np.float32):
        # float32 u * w - 0.5 carries up to 2^-23 |x| of rounding into the weights
        tol = 2.0 ** -22 * (abs(float(u)) * 37 + abs(float(v)) * 23 + 2) + 1e-6
        assert np.allclose(orc.tex_sample(0, u, v), vk_bilinear64(img, float(u), float(v)), rtol=0, atol=tol)


def test_bump_mapping_constant_height_closed_form():
    """Over a constant height field c, the layer search stops after the
    first n layers with n * layerDepth >= h (h = 0.2 c), and the reference
    then returns mix(prevUV, currUV, weight) with weight = n - h/layerDepth
    (texutils.h.glsl:36-40: its blend is the reverse of the textbook
    parallax-occlusion one, kept as it is): uv - (2n - 1 - h/layerDepth) *
    deltaUV. A view from below the surface (V.z <= 0) leaves uv unchanged."""
    c = 150
    img = np.zeros((4, 4, 4), np.uint8)
    img[..., 0] = c
    img[..., 3] = 255
    orc = _oracle_with([img])
    tbn = np.eye(3)
    rng = np.random.default_rng(5)
    for _ in range(50):
        d = rng.normal(size=3)
        d[2] = -abs(d[2]) - 0.2          # looking down onto the surface (V.z > 0)
        d /= np.linalg.norm(d)
        u, v = rng.uniform(0, 1, 2)
        V = -d / np.linalg.norm(d)
        nl = 512.0 * (1 - np.clip(V[2], 0, 1)) + 64.0 * np.clip(V[2], 0, 1)
        ld = 1.0 / nl
        delta = V[:2] / V[2] * 0.2 / nl
        h = (np.float32(c) / np.float32(255)) * np.float32(0.2)
        n, depth = 0, np.float32(0)
        while depth < h:                  # the search's own float32 accumulation
            depth = np.float32(depth + np.float32(ld))
            n += 1
        want = np.array([u, v]) - (2 * n - 1 - float(h) / ld) * delta
        got = orc.bump(0, u, v, d.astype(np.float32), tbn)
        assert np.allclose(got, want, rtol=0, atol=2e-5), (got, want)
        up = orc.bump(0, u, v, -d.astype(np.float32), tbn)   # from below
        assert np.array_equal(up, np.array([u, v], np.float32))


def test_white_opaque_texture_renders_like_no_texture():
    """An opaque white albedo texture multiplies by exactly 1 and draws no
    random number for the alpha test: the textured box renders bit for bit
    like the untextured one."""
    a_s, _ = scenes.cornell()
    b_s, _ = scenes.textured_cornell(white_opaque=True)
    pc = scenes.cornell_push(24, 24, spp=3, max_bounces=6)
    ia, na = oracle.OracleScene(a_s.build_arrays(), make_desc).render(pc, 24, 24, use_bvh=False)
    ib, nb = oracle.OracleScene(b_s.build_arrays(), make_desc).render(pc, 24, 24, use_bvh=False)
    assert na == nb
    assert np.array_equal(ia.view(np.uint32), ib.view(np.uint32))


def test_textured_box_differs_and_is_finite():
    """The textured box (checker with transparent cells, normal and bump
    maps on every material) renders finite values and differs from the
    untextured box."""
    a_s, _ = scenes.cornell()
    b_s, _ = scenes.textured_cornell()
    pc = scenes.cornell_push(24, 24, spp=2, max_bounces=6)
    ia, _ = oracle.OracleScene(a_s.build_arrays(), make_desc).render(pc, 24, 24, use_bvh=False)
    ib, _ = oracle.OracleScene(b_s.build_arrays(), make_desc).render(pc, 24, 24, use_bvh=False)
    assert np.isfinite(ib).all()
    assert not np.array_equal(ia, ib)
