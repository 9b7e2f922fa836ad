"""Pins the oracle's Disney BSDF pieces (brdfDisney.h.glsl) against the
mathematics they come from, each through the oracle's own float32 functions
(or_disney_probe, tangent frame, normal +z):

- GGX D (both restatements, D_GGX_Aniso :42-48 and evalDm :224-229) is
  normalised: integral of D(m) m.z over the hemisphere = 1, by quadrature in
  scaled slope space, where the anisotropic GGX is 1 / (pi (1 + |u|^2)^2).
- Smith G1 (smithG :231-239 and SmithGAniso :433-439) satisfies the masking
  identity integral of D(m) max(0, v.m) dm = v.z / G1(v), with
  Lambda = (sqrt(1 + (ax^2 vx^2 + ay^2 vy^2) / vz^2) - 1) / 2 (Heitz 2014).
- sampleGGXVNDF (:9-40, Heitz 2018 / GLSL-PathTracer) draws normals with the
  visible-normal density G1(v) max(0, v.m) D(m) / v.z (binned chi-square).
- pdfGGXReflection (:55-73) is the bounded-VNDF pdf of Eto and Tokuyoshi
  (SIGGRAPH Asia 2023, doi 10.1145/3610543.3626163): over the sphere of
  reflected directions it integrates to (i.z + t) / (k i.z + t), t =
  |(ax ix, ay iy, iz)|, k from the paper's Eq. 5-6; the renderer samples the
  unbounded VNDF, whose density is this pdf times (k i.z + t) / (i.z + t).
- GTR1 (evalDc :328-333) is normalised and sampleGTR1_attempt2 (:96-111)
  draws cos(theta_h) with the CDF ln(1 - (1 - a^2) mu^2) / ln(a^2) obtained by
  integrating it; pdfClearcoat (:368-383) has as much mass over the upper
  hemisphere as sampleClearcoat (:356-366) puts there.
- Burley's diffuse and subsurface terms (evalDiffuse :124-168) meet their
  published closed forms (Burley 2012); the sheen term (:587-594) vanishes
  at h.wo = 1 and follows its tint limits.
"""
import numpy as np
import pytest

import oracle

ALPHAS = [(0.5, 0.5), (0.3, 0.1), (0.05, 0.2), (1.0, 0.6), (0.4, 0.002)]


def _dirs(theta_deg, phi_deg):
    t, p = np.radians(theta_deg), np.radians(phi_deg)
    return np.array([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)])


def _slope_grid(ax, ay, n_psi=3000, n_phi=128):
    """Midpoint grid over scaled slopes u = (sx / ax, sy / ay) in polar form
    (|u| = tan psi). Returns unit normals m (float64), u and the d^2 s measure
    of each cell; d omega_m = m.z^3 d^2 s."""
    psi = (np.arange(n_psi) + 0.5) * (0.5 * np.pi / n_psi)
    phi = (np.arange(n_phi) + 0.5) * (2 * np.pi / n_phi)
    P, F = np.meshgrid(psi, phi, indexing="ij")
    rho = np.tan(P)
    u1, u2 = rho * np.cos(F), rho * np.sin(F)
    m = np.stack([-ax * u1, -ay * u2, np.ones_like(u1)], -1)
    m /= np.linalg.norm(m, axis=-1, keepdims=True)
    w = ax * ay * rho / np.cos(P) ** 2 * (0.5 * np.pi / n_psi) * (2 * np.pi / n_phi)
    return m.reshape(-1, 3), u1.reshape(-1), u2.reshape(-1), w.reshape(-1)


def _lambda(v, ax, ay):
    return (np.sqrt(1.0 + (ax * ax * v[0] ** 2 + ay * ay * v[1] ** 2) / v[2] ** 2) - 1.0) / 2.0


@pytest.mark.parametrize("ax,ay", ALPHAS)
def test_ggx_d_normalised(ax, ay):
    m, _, _, w = _slope_grid(ax, ay)
    d, _ = oracle.disney_probe("ggx_d", len(m), a=m, prm=(ax, ay))
    for col in (0, 1):
        integral = np.sum(d[:, col].astype(np.float64) * m[:, 2] ** 4 * w)
        assert abs(integral - 1.0) < 2e-4, (col, integral)


@pytest.mark.parametrize("ax,ay", ALPHAS[:4])
def test_smith_g1_masking_identity(ax, ay):
    m, _, _, w = _slope_grid(ax, ay)
    d, _ = oracle.disney_probe("ggx_d", len(m), a=m, prm=(ax, ay))
    for th, ph in [(0, 0), (30, 20), (60, 100), (80, 230), (88, 45)]:
        v = _dirs(th, ph)
        visible = np.sum(d[:, 0].astype(np.float64) * np.maximum(0.0, m @ v) * m[:, 2] ** 3 * w)
        g1 = 1.0 / (1.0 + _lambda(v, ax, ay))
        assert abs(visible - v[2] / g1) < 3e-4 * max(1.0, v[2] / g1), (th, visible, v[2] / g1)
        g, _ = oracle.disney_probe("smith_g1", 1, a=v, prm=(ax, ay))
        assert np.allclose(g[0], g1, rtol=2e-5, atol=1e-6), (th, g[0], g1)


@pytest.mark.parametrize("ax,ay", [(0.5, 0.5), (0.3, 0.1), (0.05, 0.2)])
@pytest.mark.parametrize("th,ph", [(0, 0), (45, 30), (75, 200)])
def test_vndf_sampling_density(ax, ay, th, ph):
    v = _dirs(th, ph)
    n = 200000
    h, _ = oracle.disney_probe("vndf", n, a=v, prm=(ax, ay), seed=7 + th)
    h = h.astype(np.float64)
    assert np.all(h[:, 2] >= 0.0) and np.allclose(np.linalg.norm(h, axis=1), 1.0, atol=1e-5)
    assert np.mean(h @ v >= 0.0) > 0.999                      # visible normals only
    with np.errstate(divide="ignore", invalid="ignore"):
        u1 = -h[:, 0] / (h[:, 2] * ax)
        u2 = -h[:, 1] / (h[:, 2] * ay)
    psi = np.where(h[:, 2] > 0, np.arctan(np.hypot(u1, u2)), 0.5 * np.pi)
    phi = np.mod(np.arctan2(u2, u1), 2 * np.pi)
    K, J = 16, 16
    ib = np.minimum((psi / (0.5 * np.pi) * K).astype(int), K - 1)
    jb = np.minimum((phi / (2 * np.pi) * J).astype(int), J - 1)
    counts = np.bincount(ib * J + jb, minlength=K * J)
    # expected: visible-normal density over scaled slopes, from the closed forms
    m, g1u, g2u, w = _slope_grid(1.0, 1.0, n_psi=K * 100, n_phi=J * 16)
    g1 = 1.0 / (1.0 + _lambda(v, ax, ay))
    dens = g1 / v[2] * np.maximum(0.0, v[2] - ax * v[0] * g1u - ay * v[1] * g2u) / (np.pi * (1 + g1u ** 2 + g2u ** 2) ** 2)
    p = (dens * w).reshape(K, 100, J, 16).sum(axis=(1, 3)).reshape(-1)
    assert abs(p.sum() - 1.0) < 1e-3, p.sum()
    e = n * p / p.sum()
    ok = e >= 5
    chi2 = np.sum((counts[ok] - e[ok]) ** 2 / e[ok])
    dof = ok.sum() - 1
    assert chi2 < dof + 6 * np.sqrt(2 * dof), (chi2, dof)
    assert counts[~ok].sum() < 5 * max(1, (~ok).sum()) + e[~ok].sum() * 3


@pytest.mark.parametrize("ax,ay", ALPHAS)
@pytest.mark.parametrize("th,ph", [(0, 0), (40, 60), (70, 300)])
def test_bounded_vndf_pdf_integral(ax, ay, th, ph):
    i = _dirs(th, ph)
    m, _, _, w = _slope_grid(ax, ay, n_psi=2000, n_phi=128)
    im = m @ i
    o = 2.0 * im[:, None] * m - i
    pdf, _ = oracle.disney_probe("pdf_refl", len(m), a=i, b=o, prm=(ax, ay))
    integral = np.sum(pdf[:, 0].astype(np.float64) * 4.0 * np.maximum(im, 0.0) * m[:, 2] ** 3 * w)
    t = np.sqrt((ax * i[0]) ** 2 + (ay * i[1]) ** 2 + i[2] ** 2)
    a = min(ax, ay, 1.0)
    s = 1.0 + np.hypot(i[0], i[1])
    k = (1 - a * a) * s * s / (s * s + a * a * i[2] ** 2)
    expect = (i[2] + t) / (k * i[2] + t)
    assert abs(integral - expect) < 5e-4 * expect, (integral, expect)
    # away from the horizon cutoff the pdf is the unbounded VNDF reflection pdf
    # G1(i) D(m) / (4 i.z) times that same constant
    sel = (im > 0.05) & (m[:, 2] > 0.05)
    g1 = 1.0 / (1.0 + _lambda(i, ax, ay))
    d = 1.0 / (np.pi * ax * ay * ((m[sel, 0] / ax) ** 2 + (m[sel, 1] / ay) ** 2 + m[sel, 2] ** 2) ** 2)
    ratio = pdf[sel, 0] / (g1 * d / (4.0 * i[2]))
    assert np.allclose(ratio, expect, rtol=2e-3), (ratio.min(), ratio.max(), expect)


def _x_grid(n=40000):
    """x = 1 - cos^2(theta_h) on a log grid (GTR1 peaks at x = 0 with width a^2)."""
    lx = np.linspace(np.log(1e-9), 0.0, n)
    x = np.exp(lx)
    return x, np.gradient(lx) * x


@pytest.mark.parametrize("cg", [0.0, 0.5, 0.9])
def test_gtr1_normalised_and_sampled(cg):
    ag = (1 - cg) * 0.1 + cg * 0.001
    x, dx = _x_grid()
    hh = np.stack([np.sqrt(x), np.zeros_like(x), np.sqrt(1 - x)], -1)
    dc, _ = oracle.disney_probe("dc", len(x), a=hh, prm=(ag,))
    # integral of Dc(h) h.z d omega = pi * integral of Dc dx (x = 1 - mu^2)
    integral = np.pi * np.sum(dc[:, 0].astype(np.float64) * dx) + np.pi * dc[0, 0] * 1e-9
    assert abs(integral - 1.0) < 2e-3, integral
    n = 200000
    h, _ = oracle.disney_probe("gtr1", n, prm=(ag,), seed=3)
    h = h.astype(np.float64)
    mu = np.sort(h[:, 2])
    cdf = np.log1p(-(1 - ag * ag) * mu ** 2) / np.log(ag * ag)
    ks = np.max(np.abs(cdf - (np.arange(1, n + 1) - 0.5) / n))
    assert ks < 1.95 / np.sqrt(n) + 1.0 / n, ks
    az = np.mod(np.arctan2(h[:, 1], h[:, 0]), 2 * np.pi)
    hist = np.bincount((az / (2 * np.pi) * 16).astype(int) % 16, minlength=16)
    assert np.all(np.abs(hist - n / 16) < 6 * np.sqrt(n / 16))


@pytest.mark.parametrize("cg", [0.0, 0.7])
@pytest.mark.parametrize("th", [0, 50, 80])
def test_clearcoat_pdf_matches_sampling(cg, th):
    wi = _dirs(th, 35)
    ag = (1 - cg) * 0.1 + cg * 0.001
    n = 200000
    wo, _ = oracle.disney_probe("clearcoat", n, a=wi, prm=(cg,), seed=11 + th)
    wo = wo.astype(np.float64)
    up = wo[:, 2] > 0
    # quadrature over half vectors: d omega_o = 4 (wi.h) d omega_h
    m, _, _, w = _slope_grid(ag, ag, n_psi=3000, n_phi=128)
    im = m @ wi
    o = 2.0 * im[:, None] * m - wi
    pdf, _ = oracle.disney_probe("pdf_clearcoat", len(m), a=wi, b=o, prm=(cg,))
    mass = pdf[:, 0].astype(np.float64) * 4.0 * np.maximum(im, 0.0) * m[:, 2] ** 3 * w
    se = np.sqrt(up.mean() * (1 - up.mean()) / n)
    assert abs(mass.sum() - up.mean()) < 5 * se + 1e-3, (mass.sum(), up.mean())
    for comp in (2, 0):                       # E[wo.z], E[wo.x] over the upper hemisphere
        gs = np.where(up, wo[:, comp], 0.0)
        assert abs(np.sum(mass * o[:, comp]) - gs.mean()) < 5 * gs.std() / np.sqrt(n) + 1e-3, comp


def test_burley_diffuse_closed_forms():
    bc = (0.8, 0.4, 0.2)
    for r in (0.0, 0.3, 1.0):
        for th in (0.0, 30.0, 60.0, 85.0):
            wi, wo = _dirs(th, 40), _dirs(th, 220)
            c = wi[2]
            for a_, b_, cd in [(wi, wo, c), (wi, wi, 1.0)]:
                # mirrored pair: h = n, cos(theta_d) = c; equal pair: h = wi, cos(theta_d) = 1
                f, _ = oracle.disney_probe("diffuse", 1, a=a_, b=b_, prm=(r, 0.0) + bc)
                fd = 1.0 + (0.5 + 2 * r * cd * cd - 1.0) * (1 - c) ** 5
                assert np.allclose(f[0], np.array(bc) / np.pi * fd * fd, rtol=1e-5), (r, th, cd)
                fs, _ = oracle.disney_probe("diffuse", 1, a=a_, b=b_, prm=(r, 1.0) + bc)
                fss = 1.0 + (r * cd * cd - 1.0) * (1 - c) ** 5
                expect = 1.25 * np.array(bc) / np.pi * (fss * fss * (1.0 / (2 * c) - 0.5) + 0.5)
                assert np.allclose(fs[0], expect, rtol=1e-5), (r, th, cd)
        # reciprocity over random pairs
        rng = np.random.default_rng(5)
        wi = rng.normal(size=(64, 3)); wi[:, 2] = np.abs(wi[:, 2]); wi /= np.linalg.norm(wi, axis=1, keepdims=True)
        wo = rng.normal(size=(64, 3)); wo[:, 2] = np.abs(wo[:, 2]); wo /= np.linalg.norm(wo, axis=1, keepdims=True)
        for ss in (0.0, 0.5):
            a, _ = oracle.disney_probe("diffuse", 64, a=wi, b=wo, prm=(r, ss) + bc)
            b, _ = oracle.disney_probe("diffuse", 64, a=wo, b=wi, prm=(r, ss) + bc)
            assert np.allclose(a, b, rtol=1e-5)


def test_sheen_limits():
    bc = (0.8, 0.4, 0.2)
    lum = 0.2126 * bc[0] + 0.7152 * bc[1] + 0.0722 * bc[2]
    w = _dirs(40, 10)
    s0, _ = oracle.disney_probe("sheen", 1, a=w, b=w, prm=(0, 0) + bc + (1.0, 1.0, 1.0))
    assert np.all(s0 == 0.0)                                   # h.wo = 1
    wi, wo = _dirs(70, 0), _dirs(70, 180)                      # h = n, h.wo = cos 70
    grey, _ = oracle.disney_probe("sheen", 1, a=wi, b=wo, prm=(0, 0) + bc + (0.0, 0.0, 0.0))
    tint, _ = oracle.disney_probe("sheen", 1, a=wi, b=wo, prm=(0, 0) + bc + (1.0, 1.0, 1.0))
    c = np.cos(np.radians(70))
    assert np.allclose(grey[0], (1 - c) ** 5 * c, rtol=1e-5)
    assert np.allclose(tint[0], np.array(bc) / lum * (1 - c) ** 5 * c, rtol=1e-5)


# ---- the whole mixture: sampleDisney against evalDisney's pdf ----------------------
MATERIALS = {   # bc, specularTint, sheenTint, anisotropic, roughness, subsurface, cg, metallic, clearcoat, st, sheen
    "clearcoat_plastic": ((0.8, 0.3, 0.2), (1, 1, 1), (0, 0, 0), 0.0, 0.5, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0),
    "anisotropic_metal": ((0.9, 0.7, 0.3), (1, 1, 1), (0, 0, 0), 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0),
    "four_lobes": ((0.5, 0.7, 0.9), (1, 1, 1), (0.5, 1, 0.5), 0.3, 0.5, 0.5, 0.0, 0.3, 0.5, 0.4, 0.5),
    "glass": ((0.95, 0.95, 0.95), (1, 1, 1), (0, 0, 0), 0.0, 0.4, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0),
}


def _material_prm(name, fixes, eta=1.0 / 1.5):
    bc, stint, shtint, an, rough, ss, cg, metal, cc, st, sheen = MATERIALS[name]
    return tuple(bc) + tuple(stint) + tuple(shtint) + (an, rough, ss, cg, eta, metal, cc, st, sheen, float(fixes))


def _mixture_chi2(name, fixes, th, n=300000, K=12, J=12, S=24):
    wi = _dirs(th, 25)
    prm = _material_prm(name, fixes)
    smp, _ = oracle.disney_probe("disney_sample", n, a=wi, prm=prm, seed=5 + th)
    wo = smp[:, :3].astype(np.float64)
    up = wo[:, 2] > 0
    mu = np.clip(wo[up, 2] / np.linalg.norm(wo[up], axis=1), 0, 1 - 1e-12)
    phi = np.mod(np.arctan2(wo[up, 1], wo[up, 0]), 2 * np.pi)
    counts = np.bincount(np.minimum((mu * K).astype(int), K - 1) * J + np.minimum((phi / (2 * np.pi) * J).astype(int), J - 1),
                         minlength=K * J)
    # expected: the pdf over each equal-area bin (d omega = d mu d phi)
    mg = (np.arange(K * S) + 0.5) / (K * S)
    pg = (np.arange(J * S) + 0.5) / (J * S) * 2 * np.pi
    M, P = np.meshgrid(mg, pg, indexing="ij")
    sn = np.sqrt(1 - M ** 2)
    dirs = np.stack([sn * np.cos(P), sn * np.sin(P), M], -1).reshape(-1, 3)
    pdf, _ = oracle.disney_probe("disney_pdf", len(dirs), a=wi, b=dirs, prm=prm)
    cell = (1.0 / (K * S)) * (2 * np.pi / (J * S))
    p = (pdf[:, 0].astype(np.float64) * cell).reshape(K, S, J, S).sum(axis=(1, 3)).reshape(-1)
    e = n * p
    ok = e >= 5
    chi2 = np.sum((counts[ok] - e[ok]) ** 2 / e[ok])
    return chi2, ok.sum() - 1, up.mean(), p.sum()


@pytest.mark.parametrize("name", list(MATERIALS))
@pytest.mark.parametrize("th", [20, 60])
def test_fixed_mixture_pdf_is_the_sampler_density(name, th):
    """With REINA_RT_TRACE_FIX_METAL_PDF | FIX_DISNEY_MIX, evalDisney's pdf
    is the density of sampleDisney's directions over the upper hemisphere
    (where the estimator's cosine is nonzero): binned chi-square, and the
    sampled upper-hemisphere fraction equals the pdf's mass there."""
    import reina_rt.abi as abi
    chi2, dof, frac_up, mass_up = _mixture_chi2(name, abi.TRACE_FIX_METAL_PDF | abi.TRACE_FIX_DISNEY_MIX, th)
    assert chi2 < dof + 6 * np.sqrt(2 * dof), (chi2, dof)
    assert abs(frac_up - mass_up) < 0.01, (frac_up, mass_up)


def test_reference_mixture_is_not_the_sampler_density():
    """The reference's weights (clearcoat outside the normalised sum) do not
    give the sampler's density: the same test fails by far."""
    chi2, dof, frac_up, mass_up = _mixture_chi2("clearcoat_plastic", 0x40, 20)
    assert chi2 > 10 * dof, (chi2, dof)
