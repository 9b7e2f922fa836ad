"""Pins the oracle's camera rays (raytrace.rgen.glsl:194-247): the Gaussian
pixel jitter takes exactly two PCG draws, the hexagon lens sample reads a copy
of the state left after them and does not advance it, every defocused ray
passes through the pinhole ray's focal point at focusDist, and lens offsets
are uniform over the regular hexagon with vertices (+-1, 0) scaled by
defocusMultiplier (area 3*sqrt(3)/2; the band |x| <= 1/2 holds 2/3 of it)."""
import numpy as np
import pytest

import oracle
from reina_rt.scene import push_consts

A, C, M = 747796405, 1, 2 ** 32
SQ3 = np.sqrt(3.0)


def _setup(defocus, focus=2.2):
    W, H = 64, 48                       # the Cornell scenes' camera
    pc = push_consts((0.0, 1.0, 3.4), (0.0, 1.0, 0.0), 40.0, W, H, spp=1, max_bounces=1, defocus=defocus, focus_dist=focus)
    return pc, W, H


def _basis(pc):
    iv = np.asarray(pc["invView"], np.float64).reshape(4, 4)   # column-major: row k = column k
    right = iv[0, :3] / np.linalg.norm(iv[0, :3])
    up = iv[1, :3] / np.linalg.norm(iv[1, :3])
    return iv[3, :3], right, up


def _rays(pc, W, H, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x, y = int(rng.integers(0, W)), int(rng.integers(0, H))
        s = int(rng.integers(0, 2 ** 32))
        o, d, st = oracle.camera_ray(pc, x, y, W, H, s)
        out.append((x, y, s, o.astype(np.float64), d.astype(np.float64), st))
    return out


@pytest.mark.parametrize("defocus", [0.0, 0.015, 0.3])
def test_camera_ray_takes_two_draws(defocus):
    pc, W, H = _setup(defocus)
    for _, _, s, _, _, st in _rays(pc, W, H, 200, 1):
        assert st == (A * A * s + C * (A + 1)) % M


def test_lens_sample_uses_post_jitter_state_by_value():
    """The offset equals the first accepted hexagon pair of the stream that
    starts at the state after the jitter, and that state comes back
    unchanged; a lens draw from the pre-jitter state or one that advanced the
    state fails."""
    defocus = 0.3
    pc, W, H = _setup(defocus)
    eye, right, up = _basis(pc)
    for _, _, s, o, _, st in _rays(pc, W, H, 300, 2):
        _, vals = oracle.pcg_stream(st, 64)
        vals = vals.astype(np.float64)
        for k in range(0, 64, 2):
            hx, hy = 2.0 * vals[k] - 1.0, (vals[k + 1] - 0.5) * SQ3
            if abs(hy) <= SQ3 / 2 and SQ3 * abs(hx) + abs(hy) <= SQ3:
                break
        off = o - eye
        np.testing.assert_allclose([off @ right, off @ up], [hx * defocus, hy * defocus], atol=2e-6)


@pytest.mark.parametrize("defocus,focus", [(0.015, 2.2), (0.3, 2.2), (0.3, 5.0)])
def test_defocused_rays_meet_the_pinhole_focal_point(defocus, focus):
    pc, W, H = _setup(defocus, focus)
    pin, _, _ = _setup(0.0, focus)
    eye, right, up = _basis(pc)
    fwd = np.cross(up, right)
    for x, y, s, o, d, st in _rays(pc, W, H, 300, 3):
        o0, d0, st0 = oracle.camera_ray(pin, x, y, W, H, s)
        assert st == st0
        assert np.array_equal(o0.astype(np.float64), eye)        # pinhole: no offset
        focal = eye + d0.astype(np.float64) * focus
        t = np.linalg.norm(focal - o)
        np.testing.assert_allclose(o + d * t, focal, atol=5e-6 * max(1.0, focus))
        assert np.isclose(np.linalg.norm(d), 1.0, atol=1e-6)
        off = o - eye
        assert abs(off @ fwd) < 1e-6                               # the lens is the image plane
        lx, ly = off @ right / defocus, off @ up / defocus
        assert abs(ly) <= SQ3 / 2 + 1e-5 and SQ3 * abs(lx) + abs(ly) <= SQ3 + 1e-5


def test_lens_offsets_uniform_over_hexagon():
    pc, W, H = _setup(1.0)
    eye, right, up = _basis(pc)
    off = np.array([[(r[3] - eye) @ right, (r[3] - eye) @ up] for r in _rays(pc, W, H, 6000, 4)])
    # band |x| <= 1/2 holds 2/3 of the area; binomial sd = sqrt(2/9/6000) = 0.006
    assert abs(np.mean(np.abs(off[:, 0]) <= 0.5) - 2 / 3) < 0.025
    # the corner triangles x > 1/2 each hold 1/6
    assert abs(np.mean(off[:, 0] > 0.5) - 1 / 6) < 0.02
    assert abs(off[:, 0].mean()) < 0.02 and abs(off[:, 1].mean()) < 0.02
    assert off[:, 0].max() > 0.9 and off[:, 1].max() > 0.8
