"""Pins the oracle's next-event estimation (the library's REINA_RT_TRACE_NEE):
light tables against closed forms, shadow-ray visibility on constructed
cases, and the NEE + MIS estimator against the BSDF-only estimator (equal
expectation; the BSDF-only path is pinned by the furnace and closed-form
tests)."""
import numpy as np
import pytest

import oracle
from reina_rt import abi, scenes
from reina_rt.scene import make_desc, push_consts

from conftest import world_triangles


def parse_tables(buf):
    n_emi, n_cdf, n_inst, n_tri = buf[:16].view(np.uint32)
    p = 16
    rec = buf[p:p + 32 * n_emi].reshape(-1, 32)
    p += 32 * n_emi
    emi = {"inst": rec[:, 0:4].copy().view(np.uint32)[:, 0], "ntri": rec[:, 4:8].copy().view(np.uint32)[:, 0],
           "cdf_base": rec[:, 8:12].copy().view(np.uint32)[:, 0], "cdf": rec[:, 12:16].copy().view(np.float32)[:, 0],
           "pdf_area": rec[:, 16:20].copy().view(np.float32)[:, 0]}
    tri_cdf = buf[p:p + 4 * n_cdf].copy().view(np.float32)
    p += 4 * n_cdf
    inst_emi = buf[p:p + 4 * n_inst].copy().view(np.int32)
    p += 4 * n_inst
    flags = buf[p:p + n_tri].copy()
    return emi, tri_cdf, inst_emi, flags


def tri_areas(arrays):
    w = world_triangles(arrays).reshape(-1, 3, 3).astype(np.float64)
    return 0.5 * np.linalg.norm(np.cross(w[:, 1] - w[:, 0], w[:, 2] - w[:, 0]), axis=1)


def instance_tri_ranges(arrays):
    out, base = [], 0
    for inst in arrays["instances"]:
        n = int(arrays["models"][inst["modelID"]]["indexCount"])
        out.append((base, n))
        base += n
    return out


@pytest.mark.parametrize("which", ["cornell", "icosphere"])
def test_light_tables_closed_form(which, cornell_arrays, icosphere_arrays):
    arrays, _ = cornell_arrays if which == "cornell" else icosphere_arrays
    orc = oracle.OracleScene(arrays, make_desc)
    emi, tri_cdf, inst_emi, flags = parse_tables(orc.export_emissive())
    areas = tri_areas(arrays)
    ranges = instance_tri_ranges(arrays)
    props = arrays["props"]
    lights = [i for i, inst in enumerate(arrays["instances"])
              if (props[inst["instancePropertiesID"]]["emission"].astype(np.float64) ** 2).sum() > 1e-10]
    assert list(emi["inst"]) == lights and len(lights) == (1 if which == "cornell" else 2)
    w = []
    for k, i in enumerate(lights):
        b, n = ranges[i]
        a = areas[b:b + n]
        e = props[arrays["instances"][i]["instancePropertiesID"]]["emission"].astype(np.float64)
        w.append(a.sum() * (0.2126 * e[0] + 0.7152 * e[1] + 0.0722 * e[2]))
        cdf = tri_cdf[emi["cdf_base"][k]:emi["cdf_base"][k] + n]
        np.testing.assert_allclose(cdf, np.cumsum(a) / a.sum(), rtol=1e-6)
        assert cdf[-1] == 1.0 and inst_emi[i] == k
    w = np.array(w)
    np.testing.assert_allclose(emi["cdf"], np.cumsum(w) / w.sum(), rtol=1e-6)
    for k, i in enumerate(lights):
        b, n = ranges[i]
        np.testing.assert_allclose(emi["pdf_area"][k], (w[k] / w.sum()) / areas[b:b + n].sum(), rtol=1e-6)
    # culling flags follow cullBackface of each instance's properties
    for i, (b, n) in enumerate(ranges):
        cull = props[arrays["instances"][i]["instancePropertiesID"]]["cullBackface"] != 0
        assert ((flags[b:b + n] & 1) == cull).all()


def test_shadow_visibility_cases(cornell_arrays):
    """Through the Cornell box: a ray crossing a culling wall from behind
    is unoccluded; the same ray the other way is occluded; tmax stops short
    of a blocker; the tall box blocks a ray through it."""
    arrays, _ = cornell_arrays
    orc = oracle.OracleScene(arrays, make_desc)

    def ray(o, d, tmax):
        r = np.zeros((), abi.RAY)
        r["ox"], r["oy"], r["oz"] = o
        r["dx"], r["dy"], r["dz"] = np.asarray(d, np.float32) / np.linalg.norm(d)
        r["tmin"], r["tmax"] = 0.0, tmax
        return r

    # back wall z = -1 faces +z (into the box); from behind (z = -2) going +z it is a back face
    assert not orc.occluded(ray((0.0, 1.0, -2.0), (0, 0, 1), 0.9))
    assert not orc.occluded(ray((0.0, 1.5, -2.0), (0, 0, 1), 1.5))      # passes the back wall, stops before anything
    assert orc.occluded(ray((0.0, 1.0, 0.5), (0, 0, -1), 3.0))           # front face of the back wall
    assert not orc.occluded(ray((0.0, 1.0, 0.5), (0, 0, -1), 1.4))       # tmax short of it
    assert orc.occluded(ray((-0.35, 0.6, 0.9), (0, 0, -1), 0.95))        # into the tall box


def test_nee_equals_bsdf_only_in_expectation(cornell_arrays):
    """Per pixel, NEE + MIS and BSDF-only sampling estimate the same value
    (Cornell box: Lambertian walls and boxes, culled back faces, one quad
    light; depth 4, no clamping). z-scores over batch means."""
    arrays, cam = cornell_arrays
    W = H = 10
    orc = oracle.OracleScene(arrays, make_desc)
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 16, 4, direct_clamp=1e9)
    orc.set_nee(False)
    sa, qa, ca = orc.render_sum64(pc, W, H, 1, 256)
    orc.set_nee(True)
    sb, qb, cb = orc.render_sum64(pc, W, H, 5000, 128)
    orc.set_nee(False)
    na, nb = 256.0, 128.0
    ma, mb = sa / na, sb / nb
    va = np.maximum(qa / na - ma * ma, 1e-12) / na
    vb = np.maximum(qb / nb - mb * mb, 1e-12) / nb
    z = (mb - ma) / np.sqrt(va + vb)
    assert np.abs(z).max() < 5.0, np.abs(z).max()
    assert (z * z).mean() < 1.6, (z * z).mean()
    # NEE reduces the variance of the estimate substantially
    assert np.median(vb * nb) < 0.5 * np.median(va * na)
    # image mean agrees within 3 sigma
    assert abs(mb.mean() - ma.mean()) < 3.0 * np.sqrt(va.sum() + vb.sum()) / ma.size


def rect_irradiance(px, pz, h, x1, x2, z1, z2, le):
    """Irradiance at (px, 0, pz) (normal +y) from a Lambertian rectangle of
    radiance le on y = h facing down, [x1, x2] x [z1, z2]: the
    differential-area-to-parallel-rectangle form factor summed over the four
    corner rectangles, E = pi * le * F."""
    def f(X, Y):
        a, b = np.sqrt(1 + X * X), np.sqrt(1 + Y * Y)
        return X / a * np.arctan(Y / a) + Y / b * np.arctan(X / b)
    X1, X2 = (x1 - px) / h, (x2 - px) / h
    Z1, Z2 = (z1 - pz) / h, (z2 - pz) / h
    return 0.5 * le * (f(X2, Z2) - f(X1, Z2) - f(X2, Z1) + f(X1, Z1))


def test_rect_irradiance_formula():
    """The corner form-factor formula against a direct quadrature of
    le * cos * cos' / d^2 over the rectangle."""
    h, le = 0.3, 2.0
    u = (np.arange(2000) + 0.5) / 2000 * 2 - 1
    U, V = np.meshgrid(u, u, indexing="ij")
    for px, pz in ((0.0, 0.0), (0.7, -0.2), (-0.95, 0.9)):
        d2 = (U - px) ** 2 + (V - pz) ** 2 + h * h
        q = (le * h * h / d2 ** 2).sum() * (2.0 / 2000) ** 2
        assert abs(q - rect_irradiance(px, pz, h, -1, 1, -1, 1, le)) < 2e-4 * q


@pytest.mark.parametrize("nee", [True, False])
def test_direct_light_closed_form(nee):
    """Closed box, floor of albedo a under a ceiling-wide light, black walls,
    depth 2 (camera -> floor -> light): given each sample's camera ray, its
    expectation is a/pi * E(x) on the floor, le on the light, 0 on a wall.
    The sum over 1-spp batches of the rendered samples matches the sum of
    these expectations, with NEE and with BSDF sampling alone."""
    h, alb, le = 0.3, 0.7, 2.0
    sc, cam = scenes.light_box(h, alb, le)
    arrays = sc.build_arrays()
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_nee(nee)
    W = H = 12
    N = 192
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 1, 2, direct_clamp=1e9)
    s, sq, cnt = orc.render_sum64(pc, W, H, 1, N)
    orc.set_nee(False)
    rays = np.zeros((N, H, W), abi.RAY)
    for b in range(N):
        pcb = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 1, 2, sample_batch=b + 1, direct_clamp=1e9)
        for y in range(H):
            for x in range(W):
                o, d, _ = oracle.camera_ray(pcb, x, y, W, H, ((b + 1) * H + y) * W + x)
                r = rays[b, y, x]
                r["ox"], r["oy"], r["oz"] = o
                r["dx"], r["dy"], r["dz"] = d
                r["tmax"] = 10000.0
    flat = rays.reshape(-1)
    hit = orc.brute32(flat)
    p = np.stack([flat["ox"], flat["oy"], flat["oz"]], 1).astype(np.float64) + \
        hit["t"][:, None].astype(np.float64) * np.stack([flat["dx"], flat["dy"], flat["dz"]], 1)
    exp = np.zeros(len(flat))
    floor = hit["prim"] < 2
    light = (hit["prim"] >= 2) & (hit["prim"] < 4)
    exp[floor] = alb / np.pi * rect_irradiance(p[floor, 0], p[floor, 2], h, -1, 1, -1, 1, le)
    exp[light] = le
    exp = exp.reshape(N, H, W).sum(0)
    got = s[..., 0]
    var = np.maximum(sq[..., 0] - s[..., 0] ** 2 / N, 0.0)          # sum of squared deviations
    err_tot = np.sqrt(var.sum())
    assert abs(got.sum() - exp.sum()) < 4.0 * err_tot + 1e-6 * exp.sum(), (got.sum(), exp.sum(), err_tot)
    z = (got - exp) / np.sqrt(np.maximum(var, 1e-12))
    assert (z * z).mean() < 1.8


def disney_nee_zscores(fixes, nb_batches=192):
    sc, cam = scenes.disney_lobes()
    arrays = sc.build_arrays()
    W = H = 10
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_fixes(fixes)
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 16, 4, direct_clamp=1e9)
    orc.set_nee(False)
    sa, qa, ca = orc.render_sum64(pc, W, H, 1, 256)
    orc.set_nee(True)
    sb, qb, cb = orc.render_sum64(pc, W, H, 5000, nb_batches)
    na, nb = 256.0, float(nb_batches)
    ma, mb = sa / na, sb / nb
    va = np.maximum(qa / na - ma * ma, 1e-12) / na
    vb = np.maximum(qb / nb - mb * mb, 1e-12) / nb
    return (mb - ma) / np.sqrt(va + vb), ma, mb, va, vb, na, nb


@pytest.mark.parametrize("fixes", [0, abi.TRACE_FIX_DISNEY_MIX | abi.TRACE_FIX_METAL_PDF], ids=["reference", "fixed"])
def test_nee_disney_equals_bsdf_only_in_expectation(fixes):
    """Disney hits (eight spheres, one per lobe and mixes, in the Cornell box
    under its ceiling light): NEE + MIS and BSDF-only sampling estimate the
    same value per pixel, with the reference's evalDisney mixture and with
    the fixed one (z-scores over batch means; depth 4, no clamping). The
    light sample's weight is the hit shader's own expected continuation
    weight, so NEE changes only the variance, including on glass, whose
    continuation the raygen carries with weight 1 (Q4)."""
    z, ma, mb, va, vb, na, nb = disney_nee_zscores(fixes)
    assert np.abs(z).max() < 5.0, np.abs(z).max()
    assert (z * z).mean() < 1.6, (z * z).mean()
    assert np.median(vb * nb) < 0.8 * np.median(va * na)
    assert abs(mb.mean() - ma.mean()) < 3.0 * np.sqrt(va.sum() + vb.sum()) / ma.size

