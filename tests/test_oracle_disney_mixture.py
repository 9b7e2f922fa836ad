"""Pins the reference's evalDisney mixture (brdfDisney.h.glsl:656-711) as the
oracle restates it, by closed forms built from lobes pinned in
test_oracle_disney.py. The mixture has three terms that no textbook mixture
has: clearcoat outside the normalising sum (:682, :704, :709),
specularTransmission passed as evalMetal's `specular` (:692), and a glass pdf
without the probability of the sampled branch (:571-580). Each check below
fails on a copy of the oracle built with the corresponding mistake planted
(oracle.planted, OR_PLANT in or_shade.c):

1 clearcoat inside the sum  2 specular = 1  3 glass pdf times F or 1 - F
4 the refraction Jacobian without the eta^2 the reference writes (:459-461)

- normal incidence (wi = wo = n): h = n, so every Fresnel term is its F0,
  G = 1, D = 1/(pi ax ay), the bounded-VNDF pdf is D (1 + a^2) / 4 with
  a = min(ax, ay) (Eto and Tokuyoshi 2023, Eq. 5-8 at i = n), GTR1 is
  (ag^2 - 1) / (pi ln(ag^2) ag^2), Burley's diffuse is bc/pi and its
  subsurface term 0.625 bc/pi. f and pdf follow in float64 from these.
- pure glass: over the sphere the pdf integrates to the bounded-VNDF constant
  times the VNDF mass reflected above the horizon plus eta^2 times the VNDF
  mass refracted below it (Walter et al. 2007's half-vector Jacobian), not 1.
- metallic = 0 and no transmission: the pdf is (1 + clearcoat / 4) times the
  density of sampleDisney's directions (binned chi-square).
"""
import numpy as np
import pytest

import oracle
from test_oracle_disney import _dirs, _lambda, _slope_grid

BC = (0.8, 0.4, 0.2)
STINT = (0.3, 0.6, 0.9)
ETA = 1.0 / 1.5


def _prm(metallic, clearcoat, st, roughness=0.5, anisotropic=0.0, subsurface=0.3, cg=0.4, sheen=0.7, fixes=0):
    return BC + STINT + (1.0, 1.0, 1.0) + (anisotropic, roughness, subsurface, cg, ETA, metallic, clearcoat, st, sheen,
                                           float(fixes))


def _normal_incidence_expect(metallic, clearcoat, st, roughness, anisotropic, subsurface, cg):
    bc = np.array(BC)
    stint = np.array(STINT)
    aspect = np.sqrt(1.0 - 0.9 * anisotropic)
    ax = max(1e-4, roughness * roughness / aspect)
    ay = max(1e-4, roughness * roughness * aspect)
    d0 = 1.0 / (np.pi * ax * ay)
    a = min(ax, ay, 1.0)
    pdf_refl = d0 * (1.0 + a * a) / 4.0                 # bounded VNDF at i = o = n
    r0_eta = ((ETA - 1.0) / (ETA + 1.0)) ** 2
    lum = 0.2126 * bc[0] + 0.7152 * bc[1] + 0.0722 * bc[2]
    ctint = bc / lum
    ks = (1.0 - ctint) + stint * ctint
    f0_metal = ks * (st * r0_eta) * (1.0 - metallic) + bc * metallic     # specular = specularTransmission
    ag = (1.0 - cg) * 0.1 + cg * 0.001
    dc = (ag * ag - 1.0) / (np.pi * np.log(ag * ag) * ag * ag)
    f_clearcoat = 0.25 * 0.04 * dc                      # F0 of IOR 1.5, G = 1
    f_diffuse = bc / np.pi * ((1.0 - subsurface) + 0.625 * subsurface)
    wd, wm, wc, wg = (1.0 - st) * (1.0 - metallic), metallic, 0.25 * clearcoat, (1.0 - metallic) * st
    ws = wd + wm + wg                                   # clearcoat outside the sum
    f = wd / ws * f_diffuse + wm / ws * f0_metal * d0 / 4.0 + wc * f_clearcoat + wg / ws * bc * d0 / 4.0
    pdf = wd / ws / np.pi + (wm + wg) / ws * pdf_refl + wc * dc / 4.0
    return f, pdf


CASES = [  # metallic, clearcoat, specularTransmission, roughness, anisotropic
    (0.0, 0.0, 0.0, 0.5, 0.0), (0.3, 0.0, 0.0, 0.5, 0.0), (0.3, 0.8, 0.5, 0.5, 0.0), (0.7, 0.0, 0.5, 0.3, 0.4),
    (0.0, 1.0, 1.0, 0.6, 0.0), (1.0, 0.5, 0.0, 0.4, 0.0), (0.5, 0.5, 1.0, 0.5, 0.7), (0.2, 0.3, 0.25, 0.2, 0.0),
]


def normal_incidence_mismatches():
    n = np.array([0.0, 0.0, 1.0])
    bad = []
    for (m, cc, st, r, an) in CASES:
        out, _ = oracle.disney_probe("disney_eval", 1, a=n, b=n, prm=_prm(m, cc, st, roughness=r, anisotropic=an))
        f, pdf = _normal_incidence_expect(m, cc, st, r, an, 0.3, 0.4)
        if not (np.allclose(out[0, :3], f, rtol=2e-5) and np.isclose(out[0, 3], pdf, rtol=2e-5)):
            bad.append(((m, cc, st, r, an), out[0].tolist(), f.tolist(), pdf))
    return bad


def test_reference_mixture_at_normal_incidence():
    """With specularTransmission = 0 the metal lobe's Fresnel term at
    h.wo = 1 is exactly bc * metallic; otherwise its F0 takes
    specularTransmission * R0(eta) for the dielectric part."""
    assert normal_incidence_mismatches() == []


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_normal_incidence_check_catches_planted_mistake(plant):
    with oracle.planted(plant):
        assert normal_incidence_mismatches() != []


def _refract(i, m, eta):
    """Snell refraction of the incoming direction -i about the normals m
    (float64; i one unit vector, m rows of unit vectors with i.m > 0): the
    directions and a TIR flag per row."""
    c = m @ i
    k = 1.0 - eta * eta * (1.0 - c * c)
    o = -eta * i + (eta * c - np.sqrt(np.maximum(k, 0.0)))[..., None] * m
    return o, k < 0.0


def glass_pdf_integrals(th, roughness, ph=25.0):
    """(oracle upper, expected upper, oracle lower, expected lower) for a pure
    glass material (metallic 0, specularTransmission 1) seen from wi at
    polar angle th, integrated over the half-vector grid of the lobe."""
    i = _dirs(th, ph)
    ax = ay = max(1e-4, roughness * roughness)
    m, _, _, w = _slope_grid(ax, ay, n_psi=2000, n_phi=128)
    dwm = m[:, 2] ** 3 * w                                       # d omega_m
    im = m @ i
    vis = im > 0.0
    d = 1.0 / (np.pi * ax * ay * ((m[:, 0] / ax) ** 2 + (m[:, 1] / ay) ** 2 + m[:, 2] ** 2) ** 2)
    g1 = 1.0 / (1.0 + _lambda(i, ax, ay))
    dv = np.where(vis, g1 * np.maximum(im, 0.0) * d / i[2], 0.0)  # Heitz's visible-normal density
    prm = _prm(0.0, 0.0, 1.0, roughness=roughness, subsurface=0.0, sheen=0.0)
    # reflection above the horizon: d omega_o = 4 (i.m) d omega_m
    o_r = 2.0 * im[:, None] * m - i
    up = vis & (o_r[:, 2] > 1e-6)
    pr, _ = oracle.disney_probe("disney_pdf", int(up.sum()), a=i, b=o_r[up], prm=prm)
    got_up = np.sum(pr[:, 0].astype(np.float64) * 4.0 * im[up] * dwm[up])
    t = np.sqrt((ax * i[0]) ** 2 + (ay * i[1]) ** 2 + i[2] ** 2)
    a = min(ax, ay, 1.0)
    s = 1.0 + np.hypot(i[0], i[1])
    k = (1 - a * a) * s * s / (s * s + a * a * i[2] ** 2)
    expect_up = (i[2] + t) / (k * i[2] + t) * np.sum(dv[up] * dwm[up])
    # refraction below the horizon: d omega_o = (eta i.m + o.m)^2 / |o.m| d omega_m
    o_t, tir = _refract(i, m, ETA)
    down = vis & ~tir & (o_t[:, 2] < -1e-6)
    pt, _ = oracle.disney_probe("disney_pdf", int(down.sum()), a=i, b=o_t[down], prm=prm)
    om = np.sum(o_t[down] * m[down], -1)
    jac = (ETA * im[down] + om) ** 2 / np.abs(om)
    got_down = np.sum(pt[:, 0].astype(np.float64) * jac * dwm[down])
    expect_down = ETA * ETA * np.sum(dv[down] * dwm[down])
    return got_up, expect_up, got_down, expect_down


GLASS = [(15.0, 0.5), (50.0, 0.5), (35.0, 0.3)]


def glass_mismatches():
    bad = []
    for th, r in GLASS:
        gu, eu, gd, ed = glass_pdf_integrals(th, r)
        if abs(gu - eu) > 2e-5 * eu or abs(gd - ed) > 2e-5 * ed:
            bad.append((th, r, gu, eu, gd, ed))
    return bad


def test_reference_glass_pdf_mass():
    """The reference's glass pdf (no branch probability, eta^2 in the
    refraction Jacobian) integrates to the reflected plus eta^2 times the
    refracted visible-normal mass: about 1.4-1.5 here, not 1."""
    assert glass_mismatches() == []
    gu, eu, gd, ed = glass_pdf_integrals(15.0, 0.5)
    assert 1.3 < gu + gd < 1.6, (gu, gd)


@pytest.mark.parametrize("plant", [3, 4])
def test_glass_check_catches_planted_mistake(plant):
    with oracle.planted(plant):
        assert glass_mismatches() != []


def sampler_ratio_chi2(clearcoat, cg, th, scale, n=300000, K=12, J=12, S=24):
    """Binned chi-square of sampleDisney's directions against evalDisney's pdf
    divided by `scale` (metallic 0, no transmission, upper hemisphere)."""
    wi = _dirs(th, 25)
    prm = _prm(0.0, clearcoat, 0.0, roughness=0.5, cg=cg)
    smp, _ = oracle.disney_probe("disney_sample", n, a=wi, prm=prm, seed=17 + th)
    wo = smp[:, :3].astype(np.float64)
    up = wo[:, 2] > 0
    mu = np.clip(wo[up, 2] / np.linalg.norm(wo[up], axis=1), 0, 1 - 1e-12)
    phi = np.mod(np.arctan2(wo[up, 1], wo[up, 0]), 2 * np.pi)
    counts = np.bincount(np.minimum((mu * K).astype(int), K - 1) * J + np.minimum((phi / (2 * np.pi) * J).astype(int), J - 1),
                         minlength=K * J)
    mg = (np.arange(K * S) + 0.5) / (K * S)
    pg = (np.arange(J * S) + 0.5) / (J * S) * 2 * np.pi
    M, P = np.meshgrid(mg, pg, indexing="ij")
    sn = np.sqrt(1 - M ** 2)
    dirs = np.stack([sn * np.cos(P), sn * np.sin(P), M], -1).reshape(-1, 3)
    pdf, _ = oracle.disney_probe("disney_pdf", len(dirs), a=wi, b=dirs, prm=prm)
    cell = (1.0 / (K * S)) * (2 * np.pi / (J * S))
    p = (pdf[:, 0].astype(np.float64) * cell).reshape(K, S, J, S).sum(axis=(1, 3)).reshape(-1) / scale
    e = n * p
    ok = e >= 5
    chi2 = np.sum((counts[ok] - e[ok]) ** 2 / e[ok])
    dof = ok.sum() - 1
    return chi2, dof, up.mean(), p.sum()


SAMPLER = [(1.0, 0.0, 20), (0.6, 0.3, 60)]


def sampler_ratio_mismatches():
    bad = []
    for cc, cg, th in SAMPLER:
        chi2, dof, frac_up, mass_up = sampler_ratio_chi2(cc, cg, th, 1.0 + 0.25 * cc)
        if not (chi2 < dof + 6 * np.sqrt(2 * dof) and abs(frac_up - mass_up) < 0.01):
            bad.append((cc, cg, th, chi2, dof, frac_up, mass_up))
    return bad


def test_reference_pdf_is_scaled_sampler_density():
    """metallic = 0, specularTransmission = 0: sampleDisney picks diffuse or
    clearcoat in proportion 1 : clearcoat/4, while evalDisney divides only
    the diffuse weight by the sum (which is 1), so its pdf is exactly
    (1 + clearcoat/4) times the sampler's density."""
    assert sampler_ratio_mismatches() == []


def test_sampler_ratio_check_catches_planted_mistake():
    with oracle.planted(1):
        assert sampler_ratio_mismatches() != []
