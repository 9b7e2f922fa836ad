"""Pins three behaviours of the reference's raygen loop in the oracle with
closed forms on constructed scenes:

- accumulatedDistance is reset only by the Lambertian / metal / Disney
  hit shaders and by a dielectric exit, never at the start of a path
  (raytrace.rgen.glsl:98 resets only insideDielectric), so distance travelled
  inside glass by a path that ends there is charged to the next sample's
  first exit.
- a NaN sample is dropped from the pixel's average
  (raytrace.rgen.glsl:268-277), not counted as zero.
- directClamp clamps the whole path sample (raytrace.rgen.glsl:266);
  indirectClamp and totalEmissiveWeight do nothing on this path.
- the world normal is mat3(O2W) applied to the object normal
  (closestHitCommon.h.glsl:113-114), by a mirror under non-uniform scale.

And the library's REINA_RT_TRACE_FIX_* corrections of these behaviours, each
against its own closed form.
"""
import numpy as np
import pytest

import oracle
from reina_rt import abi, scenes
from reina_rt.scene import Material, Scene, make_desc, push_consts


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


# ---- the REINA_RT_TRACE_FIX_* corrections ---------------------------------------------
def _render_fixed(scene, cam, W, H, spp, bounces, fixes, **kw):
    arrays = scene.build_arrays()
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, spp, bounces, up=cam.get("up", (0.0, 1.0, 0.0)),
                     **kw)
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_fixes(fixes)
    img, _ = orc.render(pc, W, H, use_bvh=False, nthreads=4)
    return img, pc


def _sky64(d):
    d = np.asarray(d, np.float64)
    t = 0.5 * (d[..., 1] / np.linalg.norm(d, axis=-1) + 1.0)
    lo, hi = np.array([0.1, 0.1, 0.1]), np.array([0.4, 1.7, 2.0])
    return (lo * (1 - t)[..., None] + hi * t[..., None]) * 0.07


@pytest.mark.parametrize("spp", [1, 4])
def test_fix_medium_reset_clears_the_leak(spp):
    s, cam = scenes.dielectric_stack()
    img, _ = _render_fixed(s, cam, 6, 5, spp, 6, abi.TRACE_FIX_MEDIUM_RESET)
    assert np.allclose(img[..., :3], 2.0 * np.exp(-0.3), rtol=1e-3)


def test_fix_lens_rng_advances_past_the_hexagon_draws():
    _, pc = _render_fixed(scenes.cornell()[0], scenes.cornell()[1], 4, 4, 1, 1, 0, defocus=0.3)
    rng = np.random.default_rng(8)
    for _ in range(200):
        s = int(rng.integers(0, 2 ** 32))
        o0, d0, st0 = oracle.camera_ray(pc, 2, 1, 4, 4, s)
        o1, d1, st1 = oracle.camera_ray(pc, 2, 1, 4, 4, s, lens_advances=True)
        assert np.array_equal(o0, o1) and np.array_equal(d0, d1)      # the same ray
        states, vals = oracle.pcg_stream(st0, 64)                       # draws after the jitter
        vals = vals.astype(np.float64)
        for k in range(0, 64, 2):
            hx, hy = 2.0 * vals[k] - 1.0, (vals[k + 1] - 0.5) * np.sqrt(3.0)
            if abs(hy) <= np.sqrt(3.0) / 2 and np.sqrt(3.0) * abs(hx) + abs(hy) <= np.sqrt(3.0):
                break
        assert st1 == int(states[k + 1])                                # just past the accepted pair


@pytest.mark.parametrize("fix", [False, True], ids=["as_reference", "inverse_transpose"])
def test_tilted_mirror_normal_closed_form(fix):
    """Reflection off a mirror placed by diag(1.5, 1, 1) * rot_z(20 deg): the
    reference's mat3(O2W) normal normalize(S R y), or with FIX_NORMALS the
    true normal normalize(S^-1 R y)."""
    s, cam = scenes.tilted_mirror()
    W, H = 6, 5
    img, pc = _render_fixed(s, cam, W, H, 1, 2, abi.TRACE_FIX_NORMALS if fix else 0)
    M = scenes.TILTED_MIRROR_XFORM[:3, :3]
    true_n = np.linalg.inv(M).T @ np.array([0.0, 1.0, 0.0])
    true_n /= np.linalg.norm(true_n)
    n = true_n if fix else M @ np.array([0.0, 1.0, 0.0])
    n /= np.linalg.norm(n)
    albedo = np.array([0.9, 0.8, 0.7])
    leaves = 0
    for y in range(H):
        for x in range(W):
            _, d, _ = oracle.camera_ray(pc, x, y, W, H, y * W + x)
            d = d.astype(np.float64) / np.linalg.norm(d)
            r = d - 2.0 * (d @ n) * n
            sky = albedo * _sky64(r)
            if r @ true_n > 0:
                leaves += 1
                assert np.allclose(img[y, x, :3], sky, rtol=2e-4), (x, y)
            else:
                # into the mirror's half-space from an origin offset by ~1e-6:
                # either it hits the plane again (two segments: dark) or the
                # FP32 test misses that hit at t ~ 1e-6 and it reaches the sky
                assert (np.allclose(img[y, x, :3], sky, rtol=2e-4) or (img[y, x, :3] == 0).all()), (x, y)
    assert leaves == W * H


def test_fix_metal_pdf():
    """With k = 1 the GGX reflection pdf integrates to 1 over reflected
    directions (it is the VNDF sampler's density); a fully metallic Disney
    surface then weights every sample by at least as much as with the
    bounded pdf (same directions, smaller pdf)."""
    from test_oracle_disney import _dirs, _slope_grid
    for ax, ay in [(0.5, 0.5), (0.3, 0.1)]:
        for th in (0, 40, 70):
            i = _dirs(th, 60)
            m, _, _, w = _slope_grid(ax, ay, n_psi=2000, n_phi=128)
            im = m @ i
            o = 2.0 * im[:, None] * m - i
            pdf, _ = oracle.disney_probe("pdf_refl", len(m), a=i, b=o, prm=(ax, ay, 1.0))
            integral = np.sum(pdf[:, 0].astype(np.float64) * 4.0 * np.maximum(im, 0.0) * m[:, 2] ** 3 * w)
            assert abs(integral - 1.0) < 5e-4, (ax, ay, th, integral)
    s = Scene()
    s.add_object(scenes._sphere_model(16, 16), np.eye(4),
                 Material(materialIdx=3, albedo=(1.0, 1.0, 1.0), roughness=0.6, metallic=1.0, ior=1.5))
    cam = dict(eye=(0.0, 0.0, 3.0), center=(0.0, 0.0, 0.0), fov=40.0)
    ref, _ = _render_fixed(s, cam, 16, 12, 2, 2, 0)
    fixed, _ = _render_fixed(s, cam, 16, 12, 2, 2, abi.TRACE_FIX_METAL_PDF)
    assert (fixed[..., :3] >= ref[..., :3]).all()
    assert (fixed[..., :3] > ref[..., :3] * 1.01).any()


# ---- independent sample streams ----------------------------------------------------
@pytest.mark.parametrize("log2g,batch", [(0, 0), (2, 0), (3, 5)])
def test_sample_groups_seed_each_group_as_a_batch(log2g, batch):
    """With G = 2^n sample groups and spp = G (one sample per group), a pixel
    of a scene every ray misses is the mean of the sky over the G camera rays
    drawn from the reference seeds of batches batch * G + g."""
    G = 1 << log2g
    s = Scene()
    s.add_object(scenes._model_from_quads([scenes._quad((0, 0, 50.0), (0.1, 0, 0), (0, 0.1, 0))]), np.eye(4),
                 Material(materialIdx=0))                       # behind the camera: never hit
    W, H = 5, 4
    pc = push_consts((0.0, 0.0, 0.0), (0.0, 0.3, -1.0), 50.0, W, H, G, 3, sample_batch=0)
    orc = oracle.OracleScene(s.build_arrays(), make_desc)
    orc.set_fixes(abi.TRACE_SAMPLE_GROUPS(log2g))
    pc["sampleBatch"] = batch
    img, _ = orc.render(pc, W, H, use_bvh=False)                 # blends into a zero image when batch > 0
    for y in range(H):
        for x in range(W):
            sky = np.mean([_sky64(oracle.camera_ray(pc, x, y, W, H, ((batch * G + g) * H + y) * W + x)[1])
                           for g in range(G)], axis=0)
            expect = sky / (batch + 1) if batch else sky          # running mean over a zero image
            assert np.allclose(img[y, x, :3], expect, rtol=2e-6), (x, y)


def test_sample_groups_one_group_is_the_reference(cornell_arrays):
    arrays, cam = cornell_arrays
    W, H = 12, 9
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 8, 6)
    a = oracle.OracleScene(arrays, make_desc).render(pc, W, H, use_bvh=False)[0]
    o = oracle.OracleScene(arrays, make_desc)
    o.set_fixes(abi.TRACE_SAMPLE_GROUPS(0))
    assert np.array_equal(o.render(pc, W, H, use_bvh=False)[0].view(np.uint32), a.view(np.uint32))
    o.set_fixes(abi.TRACE_SAMPLE_GROUPS(3))
    b, nrays = o.render(pc, W, H, use_bvh=False)
    assert not np.array_equal(b.view(np.uint32), a.view(np.uint32)) and np.isfinite(b).all()
