"""Pins the oracle's restatement of the table consumer (skyRadiance,
inScattering, sunRadiance, skyIrradiance of
src/atmo/examples/atmo/atmosphereShader.glsl) to facts that do not come from
its own code: closed forms on constant tables, texel-centre round trips
through the full camera/sun geometry, the multiplicative and additive
structure of the aerial-perspective formula, translation along the view ray,
and continuity at the HORIZON_HACK band edges.  Also run on a non-Earth
atmosphere, since the consumer takes every constant from the parameters."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from prolandatmo.params import SEED_BATCH, draw_sets, earth, with_res

PI_REF = 3.141592657


def tiny(p=None, R=8, MU=32, MUS=8, NU=4):
    q = with_res(p or earth(), R, MU, MUS, NU)
    q.TRANSMITTANCE_W, q.TRANSMITTANCE_H, q.SKY_W, q.SKY_H = 64, 16, 16, 8
    return q


ATMOS = [tiny(), tiny(draw_sets(SEED_BATCH, 1, res=(8, 32, 8, 4))[0][0])]


def combine(p, s, nu):
    """rgb * phaseR + getMie * phaseM, times SUN_INTENSITY, written from the
    shader's definitions (atmosphereShader.glsl:290-302, :414-416)."""
    g = p.mieG
    pr = 3.0 / (16.0 * PI_REF) * (1 + nu * nu)
    pm = 1.5 / (4 * PI_REF) * (1 - g * g) * (1 + g * g - 2 * g * nu) ** -1.5 * (1 + nu * nu) / (2 + g * g)
    bR = np.array([p.betaR.x, p.betaR.y, p.betaR.z])
    mie = s[:3] * s[3] / max(s[0], 1e-4) * (bR[0] / bR)
    return (s[:3] * pr + mie * pm) * 100.0


def sun_for(mu, muS, nu):
    """unit sun vector with s.z = muS and v.s = nu for v = (sqrt(1-mu^2), 0, mu)."""
    if abs(mu) > 1 - 1e-9:
        return None
    vx = math.sqrt(1 - mu * mu)
    sx = (nu - mu * muS) / vx
    sy2 = 1 - sx * sx - muS * muS
    if sy2 < 0:
        return None
    return np.array([sx, math.sqrt(sy2), muS])


@pytest.mark.parametrize("pi", range(len(ATMOS)))
def test_constant_tables_give_the_closed_form(pi):
    p = ATMOS[pi]
    rng = np.random.default_rng(1)
    sval = np.array([0.3, 0.5, 0.9, 0.2])
    S = np.broadcast_to(sval, (p.RES_R, p.RES_MU, p.RES_MU_S * p.RES_NU, 4)).copy()
    T = np.full((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3), 0.7)
    cam = np.array([0.0, 0.0, p.Rg + 0.3 * (p.Rt - p.Rg)])
    dirs = rng.normal(size=(20, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    sun = np.array([0.6, 0.0, 0.8])
    rad, ext = O.sky_radiance(p, T, S, cam, dirs, sun)
    for d, r in zip(dirs, rad):
        assert np.allclose(r, combine(p, sval, float(d @ sun)), rtol=1e-12, atol=0)
    assert np.allclose(ext, 0.7, rtol=1e-13)
    # shafts multiply by a T ratio of constant tables: 1
    rad2, _ = O.sky_radiance(p, T, S, cam, dirs, sun, shaft_width=0.5)
    assert np.allclose(rad2, rad, rtol=1e-12)


@pytest.mark.parametrize("pi", range(len(ATMOS)))
def test_sky_radiance_at_texel_centres_reads_the_texel(pi):
    """A camera at a layer radius looking along a texel's (mu, muS, nu) reads
    exactly that texel: the geometry chain camera/direction/sun -> texture4D
    coordinates inverts getMuMuSNu."""
    p = ATMOS[pi]
    R, MU, MUS, NU = p.RES_R, p.RES_MU, p.RES_MU_S, p.RES_NU
    rng = np.random.default_rng(2)
    S = rng.random((R, MU, MUS * NU, 4)) + 0.1
    T = np.ones((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3))
    Rg, Rt = float(p.Rg), float(p.Rt)
    n = 0
    for l in range(1, R - 1):
        r = math.sqrt(Rg * Rg + (l / (R - 1)) ** 2 * (Rt * Rt - Rg * Rg))
        dhdH = np.array([Rt - r, math.sqrt(r * r - Rg * Rg) + math.sqrt(Rt * Rt - Rg * Rg), r - Rg,
                         math.sqrt(r * r - Rg * Rg)])
        for y in list(range(2, MU // 2 - 2)) + list(range(MU // 2 + 1, MU - 2)):
            for x in range(MUS * NU):
                if x % MUS == 0:
                    continue  # muS of column 0 is clamped by texture4D
                mu, muS, nu = O.mumusnu(p, r, dhdH, x, y)
                s = sun_for(mu, muS, nu)
                if s is None:
                    continue
                if y < MU // 2:
                    # getMuMuSNu nudges near-horizon ground rows to mu_h - 0.001
                    # (common.glsl:184): those do not map back to their centre
                    d = (1.0 - y / (MU / 2 - 1.0)) * dhdH[3]
                    if abs(mu - (Rg * Rg - r * r - d * d) / (2 * r * d)) > 1e-12:
                        continue
                v = np.array([math.sqrt(1 - mu * mu), 0.0, mu])
                rad, _ = O.sky_radiance(p, T, S, [0, 0, r], v[None], s)
                assert np.allclose(rad[0], combine(p, S[l, y, x], nu), rtol=1e-7), (l, y, x)
                n += 1
    assert n > 100


@pytest.mark.parametrize("pi", range(len(ATMOS)))
def test_camera_outside_equals_camera_at_the_top_and_misses_are_black(pi):
    p = ATMOS[pi]
    T, E, S = O.precompute(p)
    Rt = float(p.Rt)
    rng = np.random.default_rng(3)
    sun = np.array([0.0, 0.6, 0.8])
    cam = np.array([0.0, 0.0, Rt + 50.0])
    for _ in range(10):
        v = rng.normal(size=3)
        v[2] = -abs(v[2]) - 0.5
        v /= np.linalg.norm(v)
        b = cam @ v
        disc = b * b - (cam @ cam - Rt * Rt)
        if disc <= 0:
            continue
        entry = cam + (-b - math.sqrt(disc)) * v
        r1, e1 = O.sky_radiance(p, T, S, cam, v[None], sun)
        r2, e2 = O.sky_radiance(p, T, S, entry, v[None], sun)
        assert np.allclose(r1, r2, rtol=1e-9) and np.allclose(e1, e2, rtol=1e-9)
    up = np.array([[0.0, 1.0, 0.0]])  # tangent from above the atmosphere: misses
    rad, ext = O.sky_radiance(p, T, S, cam, up, sun)
    assert np.all(rad == 0) and np.all(ext == 1)


def _segment_case(p):
    T, E, S = O.precompute(p)
    S = S.copy()
    S[..., 3] = 0.0  # no Mie channel: the radiance is linear in inScatter.rgb
    Rg = float(p.Rg)
    cam = np.array([0.0, 0.0, Rg + 2.0])
    v = np.array([0.6, 0.0, 0.8])  # upward: one transmittance branch throughout
    sun = np.array([0.0, 0.6, 0.8])
    return T, S, cam, v, sun


@pytest.mark.parametrize("pi", range(len(ATMOS)))
def test_aerial_perspective_is_multiplicative_and_additive(pi):
    """ext(c->p2) = ext(c->p1) ext(p1->p2) and, without clamping,
    L(c->p2) = L(c->p1) + ext(c->p1) L(p1->p2) (atmosphereShader.glsl:478-517)."""
    p = ATMOS[pi]
    T, S, cam, v, sun = _segment_case(p)
    p1, p2 = cam + 5.0 * v, cam + 20.0 * v
    L02, e02 = O.in_scattering(p, T, S, cam, p2[None], sun)
    L01, e01 = O.in_scattering(p, T, S, cam, p1[None], sun)
    L12, e12 = O.in_scattering(p, T, S, p1, p2[None], sun)
    assert np.all(e02 < 1) and np.all(e01 < 1)
    assert np.allclose(e02, e01 * e12, rtol=1e-12)
    assert np.allclose(L02, L01 + e01 * L12, rtol=1e-9)
    assert np.all(L02 > L01) and np.all(L01 > 0)
    # a vanishing segment scatters nothing
    L00, e00 = O.in_scattering(p, T, S, cam, (cam + 1e-7 * v)[None], sun)
    assert np.all(np.abs(L00) <= 1e-6 * np.abs(L02)) and np.allclose(e00, 1, atol=1e-8)


@pytest.mark.parametrize("pi", range(len(ATMOS)))
def test_horizon_hack_is_continuous_at_its_band_edges(pi):
    p = ATMOS[pi]
    T, E, S = O.precompute(p)
    Rg = float(p.Rg)
    r = Rg + 3.0
    lim = -math.sqrt(1 - (Rg / r) ** 2)
    sun = np.array([0.0, 0.6, 0.8])
    cam = np.array([0.0, 0.0, r])
    for edge in (lim - 0.004, lim + 0.004):
        out = []
        for mu in (edge - 1e-10, edge + 1e-10):
            v = np.array([math.sqrt(1 - mu * mu), 0.0, mu])
            L, _ = O.in_scattering(p, T, S, cam, (cam + 8.0 * v)[None], sun)
            out.append(L[0])
        assert np.allclose(out[0], out[1], rtol=1e-5, atol=1e-9 * np.abs(out[0]).max())
    # inside the band the result is the blend of the two edge evaluations
    v = np.array([math.sqrt(1 - lim * lim), 0.0, lim])
    L, _ = O.in_scattering(p, T, S, cam, (cam + 8.0 * v)[None], sun)
    assert np.all(np.isfinite(L)) and np.all(L >= 0)


@pytest.mark.parametrize("pi", range(len(ATMOS)))
def test_sun_radiance_and_sky_irradiance(pi):
    p = ATMOS[pi]
    Rg, Rt = float(p.Rg), float(p.Rt)
    T = np.full((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3), 0.25)
    # E affine in its texel coordinates: bilinear filtering reproduces it inside the table
    j, i = np.meshgrid(np.arange(p.SKY_H), np.arange(p.SKY_W), indexing="ij")
    E = np.stack([1 + 0.5 * i + 0.25 * j, 2 + 0.1 * i, 3 + 0.3 * j], axis=-1).astype(np.float64)
    r = Rg + 0.4 * (Rt - Rg)
    lim = -math.sqrt(1 - (Rg / r) ** 2)
    rmus = np.array([[r, 0.5], [r, lim + 1e-3], [r, lim - 1e-3], [r, -0.05]])
    sun, sky = O.sun_sky_irradiance(p, T, E, rmus)
    assert np.allclose(sun[0], 25.0) and np.allclose(sun[1], 25.0)
    assert np.all(sun[2] == 0)
    for (rr, m), e in zip(rmus, sky):
        x = (m + 0.2) / 1.2 * p.SKY_W - 0.5      # texel coordinates of getIrradianceUV
        y = (rr - Rg) / (Rt - Rg) * p.SKY_H - 0.5
        want = np.array([1 + 0.5 * x + 0.25 * y, 2 + 0.1 * x, 3 + 0.3 * y]) * 100.0
        assert np.allclose(e, want, rtol=1e-12)


@pytest.mark.parametrize("pi", range(len(ATMOS)))
def test_aerial_mie_channel_uses_red_extinction_and_sun_fade(pi):
    """With S.a = S.r the alpha difference equals the red one (the shader
    extinguishes alpha with ext.r, :478-517), so getMie = rgb * betaR.r/betaR and
    radiance / radiance(alpha = 0) = 1 + pM betaR.r / (pR betaR); the Mie part
    fades with smoothstep(0, 0.02, muS) (:519) and vanishes below the horizon."""
    p = ATMOS[pi]
    T, S0, cam, v, _ = _segment_case(p)
    Sw = S0.copy()
    Sw[..., 3] = Sw[..., 0]
    bR = np.array([p.betaR.x, p.betaR.y, p.betaR.z])
    g = p.mieG
    for muS, fade in ((0.5, 1.0), (0.01, 0.5), (-0.01, 0.0)):
        sun = np.array([math.sqrt(1 - muS * muS), 0.0, muS])
        pt = (cam + 12.0 * v)[None]
        L0, _ = O.in_scattering(p, T, S0, cam, pt, sun)
        Lw, _ = O.in_scattering(p, T, Sw, cam, pt, sun)
        nu = float(v @ sun)
        pr = 3.0 / (16.0 * PI_REF) * (1 + nu * nu)
        pm = 1.5 / (4 * PI_REF) * (1 - g * g) * (1 + g * g - 2 * g * nu) ** -1.5 * (1 + nu * nu) / (2 + g * g)
        assert np.all(L0[0] > 0)
        assert np.allclose(Lw[0] / L0[0], 1 + fade * pm * bR[0] / (pr * bR), rtol=1e-9)
