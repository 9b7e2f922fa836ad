"""Pins the oracle's post-processing chain (or_post.c: blurCommon.h.glsl:15-56,
combine.comp.glsl:15-27, tonemapping.comp.glsl:17-84) by closed forms, and
shows that each check fails on a copy of the oracle built with the
corresponding mistake planted (oracle.planted, OR_PLANT in or_post.c):

5 ACES matrices applied as column vectors  6 weight sum over in-bounds taps
only  7 threshold in both blur passes  8 UNORM8 by truncation  9 sigma = the
radius in pixels instead of the config's percent value (Q11)

- blur: the taps span k = int(3 * size * radius / 100 + 0.5) pixels per axis
  (size = W for the x pass, H for the y pass) with Gaussian weights
  exp(-i^2 / (2 radius^2)), and every tap's weight enters the sum, also
  out-of-bounds and below-threshold ones. A constant image above the
  threshold therefore blurs to c * S_x(x) S_y(y) / (W_x W_y), with S the
  in-bounds and W the total weight per axis; a single bright pixel spreads as
  w(dx) w(dy) / (W_x W_y); an image below the threshold has no bloom.
- tonemap: exposure 2^e, then ACES fitted with the two published matrices as
  row-vector products (color * M of the GLSL mat3 built from columns, i.e.
  the matrices' rows dotted with the colour), the RRT + ODT rational fit,
  clamp, UNORM8 = floor(c * 255 + 0.5); alpha 255.
"""
import numpy as np
import pytest

import oracle

ACES_IN = np.array([[0.59719, 0.35458, 0.04823], [0.07600, 0.90834, 0.01566], [0.02840, 0.13383, 0.83777]])
ACES_OUT = np.array([[1.60475, -0.53108, -0.07367], [-0.10208, 1.10813, -0.00605], [-0.00327, -0.07276, 1.07602]])


def _taps(size, radius):
    rpx = np.float32(np.float32(size) * np.float32(radius)) / np.float32(100.0)
    return int(np.float32(rpx * np.float32(3.0)) + np.float32(0.5))


def _axis_weights(size, radius, sigma=None):
    """In-bounds weight sum S(p) for p in [0, size) and the total weight."""
    k = _taps(size, radius)
    sigma = radius if sigma is None else sigma
    i = np.arange(-k, k + 1)
    w = np.exp(-i * i / (2.0 * sigma * sigma))
    p = np.arange(size)
    inb = (p[:, None] + i[None, :] >= 0) & (p[:, None] + i[None, :] < size)
    return (w[None, :] * inb).sum(1), w.sum(), w, k


W, H, RADIUS, THRESH = 37, 23, 5.0, 1.0


def blur_mismatches():
    bad = []
    sx, wx, gx, kx = _axis_weights(W, RADIUS)
    sy, wy, gy, ky = _axis_weights(H, RADIUS)
    assert (kx, ky) == (6, 3)
    # a constant image above the threshold
    c = np.array([2.0, 1.5, 1.0], np.float32)
    img = np.zeros((H, W, 4), np.float32)
    img[..., :3] = c
    img[..., 3] = 1.0
    _, comb, bloom = oracle.postprocess(img, radius=RADIUS, threshold=THRESH, intensity=0.25, exposure=0.0,
                                        return_bloom=True)
    expect = c[None, None, :] * (sy[:, None, None] / wy) * (sx[None, :, None] / wx)
    if not np.allclose(bloom[..., :3], expect, rtol=1e-5):
        bad.append(("constant", float(np.abs(bloom[..., :3] / expect - 1).max())))
    if not np.array_equal(comb[..., :3], img[..., :3] + bloom[..., :3] * np.float32(0.25)):
        bad.append("combine")
    # one bright pixel
    x0, y0 = 18, 11
    cd = np.array([3.0, 2.0, 1.0], np.float32)
    img = np.zeros((H, W, 4), np.float32)
    img[y0, x0, :3] = cd
    _, _, bloom = oracle.postprocess(img, radius=RADIUS, threshold=THRESH, intensity=0.25, exposure=0.0,
                                     return_bloom=True)
    ex = np.zeros(W)
    ey = np.zeros(H)
    ex[x0 - kx:x0 + kx + 1] = gx / wx
    ey[y0 - ky:y0 + ky + 1] = gy / wy
    expect = cd[None, None, :] * ey[:, None, None] * ex[None, :, None]
    if not np.allclose(bloom[..., :3], expect, rtol=1e-5, atol=1e-12):
        bad.append(("delta", float(np.abs(bloom[..., :3] - expect).max())))
    # everything below the threshold: no bloom, the output is the input
    rng = np.random.default_rng(3)
    img = np.zeros((H, W, 4), np.float32)
    img[..., :3] = rng.uniform(0.0, 0.95, (H, W, 3))
    _, comb, bloom = oracle.postprocess(img, radius=RADIUS, threshold=THRESH, intensity=0.25, exposure=0.0,
                                        return_bloom=True)
    if np.any(bloom[..., :3] != 0.0) or not np.array_equal(comb[..., :3], img[..., :3]):
        bad.append("below threshold")
    return bad


def test_blur_closed_forms():
    assert blur_mismatches() == []


@pytest.mark.parametrize("plant", [6, 7, 9])
def test_blur_check_catches_planted_mistake(plant):
    with oracle.planted(plant):
        assert blur_mismatches() != []


def _aces(c, exposure):
    v = c * 2.0 ** exposure
    v = v @ ACES_IN.T
    v = (v * (v + 0.0245786) - 0.000090537) / (v * (0.983729 * v + 0.4329510) + 0.238081)
    v = v @ ACES_OUT.T
    return np.clip(v, 0.0, 1.0)


def tonemap_mismatches():
    bad = []
    rng = np.random.default_rng(11)
    n = 4096
    for exposure in (0.0, 1.0, -0.5):
        c = rng.uniform(0.0, 4.0, (n, 3)) * rng.uniform(0.0, 1.0, (n, 1))
        c[0] = 0.0
        c[1] = 100.0
        img = np.zeros((1, n, 4), np.float32)
        img[0, :, :3] = c
        out, _ = oracle.postprocess(img, radius=RADIUS, threshold=THRESH, intensity=0.0, exposure=exposure)
        x = _aces(img[0, :, :3].astype(np.float64), exposure) * 255.0 + 0.5
        expect = np.floor(x)
        near = np.abs(x - np.round(x)) < 2e-3        # float32 rounding can go either way there
        got = out[0, :, :3].astype(np.float64)
        if np.any((got != expect) & ~near) or np.any(np.abs(got - expect) > 1):
            bad.append((exposure, int(np.sum((got != expect) & ~near))))
        if not (np.all(out[0, 0, :3] == 0) and np.all(out[0, 1, :3] == 255) and np.all(out[0, :, 3] == 255)):
            bad.append((exposure, "limits", out[0, :2].tolist()))
        if near.sum() > 0.01 * near.size:
            bad.append((exposure, "too many ties", int(near.sum())))
    return bad


def test_tonemap_closed_form():
    assert tonemap_mismatches() == []


@pytest.mark.parametrize("plant", [5, 8])
def test_tonemap_check_catches_planted_mistake(plant):
    with oracle.planted(plant):
        assert tonemap_mismatches() != []
