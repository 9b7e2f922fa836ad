"""Pins three behaviours of the reference's raygen loop in the oracle with
closed forms on constructed scenes:

- Q8: accumulatedDistance is reset only by the Lambertian / metal / Disney
  hit shaders and by a dielectric exit, never at the start of a path
  (raytrace.rgen.glsl:98 resets only insideDielectric), so distance travelled
  inside glass by a path that ends there is charged to the next sample's
  first exit.
- Q15: a NaN sample is dropped from the pixel's average
  (raytrace.rgen.glsl:268-277), not counted as zero.
- Q6: directClamp clamps the whole path sample (raytrace.rgen.glsl:266);
  indirectClamp and totalEmissiveWeight do nothing on this path.
"""
import numpy as np
import pytest

import oracle
from reina_rt import scenes
from reina_rt.scene import make_desc, push_consts


def _render(scene, cam, W, H, spp, bounces, **kw):
    arrays = scene.build_arrays()
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, spp, bounces, up=cam.get("up", (0.0, 1.0, 0.0)),
                     **kw)
    img, _ = oracle.OracleScene(arrays, make_desc).render(pc, W, H, use_bvh=False, nthreads=4)
    return img


@pytest.mark.parametrize("spp", [1, 2, 4, 7])
def test_dielectric_distance_leaks_into_next_sample(spp):
    s, cam = scenes.dielectric_stack()
    img = _render(s, cam, 6, 5, spp, 6)
    # sample 1 exits B having travelled 0.3 inside it; each later sample also
    # carries the 0.4 its predecessor travelled inside D before the path ended
    expect = 2.0 * (np.exp(-0.3) + (spp - 1) * np.exp(-0.7)) / spp
    assert np.allclose(img[..., :3], expect, rtol=1e-3), (img[..., 0].min(), img[..., 0].max(), expect)
    # a seventh segment leaves E through its bottom face: an exit (the flag
    # insideDielectric is one bit, so leaving E counts as leaving glass while
    # the ray is still in D), which applies and clears the distance: no leak
    img7 = _render(s, cam, 6, 5, spp, 7)
    assert np.allclose(img7[..., :3], 2.0 * np.exp(-0.3), rtol=1e-3)


def test_nan_samples_are_dropped():
    s, cam = scenes.nan_floor_slab()
    W, H, spp = 8, 6, 16
    img = _render(s, cam, W, H, spp, 12)
    zenith = np.array([0.4, 1.7, 2.0]) * 0.07            # the miss shader's sky at d = +y
    finite = ~np.isnan(img[..., 0])
    # 20% of samples reflect at the first surface: a pixel is NaN only if all
    # 16 of its samples reached the floor (probability 0.8^16 = 2.8%)
    assert finite.mean() > 0.85
    assert np.allclose(img[finite][:, :3], zenith, rtol=1e-4), img[finite][:, :3]
    assert np.isnan(img[~finite][:, :3]).all()


def test_direct_clamp_whole_sample_and_unused_constants():
    s, cam = scenes.cornell()
    W, H = 64, 64                                          # the light spans about 3 rows
    base = dict(direct_clamp=0.5)
    img = _render(s, cam, W, H, 4, 6, **base)
    assert img[..., :3].max() == np.float32(0.5)           # the light's pixels: every sample clamps
    assert (img[..., :3] <= 0.5).all()
    free = _render(s, cam, W, H, 4, 6, direct_clamp=1e9)
    assert free[..., :3].max() > 5.0
    bits = img.view(np.uint32)
    for kw in (dict(indirect_clamp=1e-3), dict(indirect_clamp=1e9)):
        assert np.array_equal(_render(s, cam, W, H, 4, 6, **base, **kw).view(np.uint32), bits)
    arrays = s.build_arrays()
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 4, 6, **base)
    pc["totalEmissiveWeight"] = 123.0
    other, _ = oracle.OracleScene(arrays, make_desc).render(pc, W, H, use_bvh=False, nthreads=4)
    assert np.array_equal(other.view(np.uint32), bits)
