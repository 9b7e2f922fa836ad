"""Pins the oracle's next-event estimation at Disney hits (disney_density,
disney_nee_weight in or_shade.c; the library's shade.h restates them).

The Disney hit shader (disney.rchit.glsl:36-198) draws its continuation with
sampleDisney and weights it f cos / pdf with evalDisney's pdf, except after
the glass lobe, whose bounce the raygen carries with weight 1 (Q4). For NEE +
MIS to estimate the same integral, two functions must be right:

- disney_density(w) is the density of sampleDisney's directions: binned
  chi-square of sampled directions over the whole sphere (refracted
  directions below the surface included), for opaque, metal, clearcoat,
  four-lobe and glass materials, entering and leaving the glass.
- disney_nee_weight g(w) is the expected continuation weight per unit solid
  angle: over cones of directions, the quadrature of g equals the mean of
  (continuation weight x [direction in the cone]) over hit-shader samples,
  within 5 standard errors. Two planted mistakes each fail it: g = f cos (the
  textbook weight, wrong because evalDisney's pdf is not the sampler's
  density and glass continuations carry weight 1), and g without its glass
  term.
"""
import numpy as np
import pytest

import oracle
from test_oracle_disney import _dirs

MATS = {   # bc, specularTint, sheenTint, anisotropic, roughness, subsurface, cg, eta, metallic, clearcoat, st, sheen
    "plastic": ((0.8, 0.3, 0.2), (1, 1, 1), (0, 0, 0), 0.0, 0.5, 0.2, 0.5, 1 / 1.5, 0.0, 1.0, 0.0, 0.3),
    "metal": ((0.9, 0.7, 0.3), (1, 1, 1), (0, 0, 0), 0.5, 0.5, 0.0, 0.0, 1 / 1.5, 1.0, 0.0, 0.0, 0.0),
    "four_lobes": ((0.5, 0.7, 0.9), (0.3, 0.6, 1.0), (0.5, 1, 0.5), 0.3, 0.5, 0.5, 0.3, 1 / 1.5, 0.3, 0.5, 0.4, 0.5),
    "glass_in": ((0.95, 0.9, 0.8), (1, 1, 1), (0, 0, 0), 0.0, 0.45, 0.0, 0.0, 1 / 1.5, 0.0, 0.0, 1.0, 0.0),
    "glass_out": ((0.95, 0.9, 0.8), (1, 1, 1), (0, 0, 0), 0.0, 0.45, 0.0, 0.0, 1.5, 0.0, 0.0, 1.0, 0.0),
    "glass_mix": ((0.7, 0.9, 0.7), (1, 1, 1), (0, 0, 0), 0.0, 0.4, 0.0, 0.3, 1 / 1.5, 0.2, 0.5, 0.5, 0.0),
}


def _prm(name, fixes=0):
    bc, stint, shtint, an, rough, ss, cg, eta, metal, cc, st, sheen = MATS[name]
    return tuple(bc) + tuple(stint) + tuple(shtint) + (an, rough, ss, cg, eta, metal, cc, st, sheen, float(fixes))


def density_chi2(name, th, n=400000, K=24, J=12, S=16):
    wi = _dirs(th, 25)
    prm = _prm(name)
    smp, _ = oracle.disney_probe("disney_sample", n, a=wi, prm=prm, seed=31 + th)
    wo = smp[:, :3].astype(np.float64)
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    ok = np.all(np.isfinite(wo), axis=1)
    mu = np.clip((wo[ok, 2] + 1.0) * 0.5, 0, 1 - 1e-12)
    phi = np.mod(np.arctan2(wo[ok, 1], wo[ok, 0]), 2 * np.pi)
    counts = np.bincount(np.minimum((mu * K).astype(int), K - 1) * J + np.minimum((phi / (2 * np.pi) * J).astype(int), J - 1),
                         minlength=K * J)
    mg = (np.arange(K * S) + 0.5) / (K * S) * 2.0 - 1.0
    pg = (np.arange(J * S) + 0.5) / (J * S) * 2 * np.pi
    M, P = np.meshgrid(mg, pg, indexing="ij")
    sn = np.sqrt(1 - M ** 2)
    dirs = np.stack([sn * np.cos(P), sn * np.sin(P), M], -1).reshape(-1, 3)
    dens, _ = oracle.disney_probe("disney_nee", len(dirs), a=wi, b=dirs, prm=prm)
    cell = (2.0 / (K * S)) * (2 * np.pi / (J * S))
    p = (dens[:, 3].astype(np.float64) * cell).reshape(K, S, J, S).sum(axis=(1, 3)).reshape(-1)
    e = n * p
    sel = e >= 5
    chi2 = np.sum((counts[sel] - e[sel]) ** 2 / e[sel])
    dof = sel.sum() - 1
    return chi2, dof, p.sum(), ok.mean(), counts[~sel].sum(), e[~sel].sum()


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("th", [15, 60])
def test_disney_density_is_sampler_density(name, th):
    chi2, dof, mass, finite, c_low, e_low = density_chi2(name, th)
    assert finite > 0.999
    assert chi2 < dof + 6 * np.sqrt(2 * dof), (chi2, dof)
    assert abs(mass - 1.0) < 0.01, mass
    assert c_low < 3 * e_low + 60, (c_low, e_low)


def _cone(axis, half_deg, n_t=120, n_p=240):
    """Directions within half_deg of axis and their solid-angle weights."""
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    t = (np.arange(n_t) + 0.5) / n_t * np.radians(half_deg)
    p = (np.arange(n_p) + 0.5) / n_p * 2 * np.pi
    T, P = np.meshgrid(t, p, indexing="ij")
    u = np.cross(axis, [1.0, 0.0, 0.0] if abs(axis[0]) < 0.9 else [0.0, 1.0, 0.0])
    u /= np.linalg.norm(u)
    v = np.cross(axis, u)
    d = (np.cos(T)[..., None] * axis + (np.sin(T) * np.cos(P))[..., None] * u + (np.sin(T) * np.sin(P))[..., None] * v)
    w = np.sin(T) * (np.radians(half_deg) / n_t) * (2 * np.pi / n_p)
    return d.reshape(-1, 3), w.reshape(-1), np.cos(np.radians(half_deg))


def weight_mismatches(n=600000):
    bad = []
    for name, th, cones in [
            ("plastic", 30, [((0, 0, 1), 30), ("mirror", 15), ((0.8, 0.0, 0.3), 25)]),
            ("four_lobes", 40, [((0, 0, 1), 35), ("mirror", 20), ("refract", 30)]),
            ("glass_in", 20, [("mirror", 25), ("refract", 25)]),
            ("glass_mix", 50, [("mirror", 20), ("refract", 30), ((0, 0, 1), 30)])]:
        wi = _dirs(th, 25)
        prm = _prm(name)
        smp, _ = oracle.disney_probe("disney_continue", n, a=wi, prm=prm, seed=77 + th)
        wo = smp[:, :3].astype(np.float64)
        wgt = np.nan_to_num(smp[:, 3:6].astype(np.float64), nan=0.0, posinf=0.0, neginf=0.0)
        wo /= np.linalg.norm(wo, axis=1, keepdims=True)
        for axis, half in cones:
            if axis == "mirror":
                axis = (-wi[0], -wi[1], wi[2])
            elif axis == "refract":
                axis = (-0.5 * wi[0], -0.5 * wi[1], -1.0)
            dirs, dw, cmin = _cone(axis, half)
            ax = np.asarray(axis, np.float64) / np.linalg.norm(axis)
            inside = (wo @ ax) >= cmin
            contrib = wgt * inside[:, None]
            a_mean = contrib.mean(0)
            a_se = contrib.std(0) / np.sqrt(n)
            g, _ = oracle.disney_probe("disney_nee", len(dirs), a=wi, b=dirs, prm=prm)
            b_int = (g[:, :3].astype(np.float64) * dw[:, None]).sum(0)
            if np.any(np.abs(a_mean - b_int) > 5 * a_se + 2e-3 * np.abs(b_int) + 1e-5):
                bad.append((name, th, tuple(np.round(ax, 3)), half, a_mean.tolist(), b_int.tolist(), a_se.tolist()))
    return bad


def test_disney_nee_weight_is_expected_continuation_weight():
    assert weight_mismatches() == []


@pytest.mark.parametrize("plant", [10, 11])
def test_nee_weight_check_catches_planted_mistake(plant):
    with oracle.planted(plant):
        assert weight_mismatches() != []
