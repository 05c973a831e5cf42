"""Pins the FP64 CPU oracle (oracle/) to facts that do not come from its own code.

The reference ships no tests, fixtures or golden tables for this path, and it
cannot run here (Ork + OpenGL).  Each oracle function is therefore checked
against closed forms, independent quadrature (scipy), geometric invariants,
inverse-function round trips and brute force, chosen so that a dropped term,
a wrong sign/index or a swapped operand fails at least one of them.
"""
import math
import os
import subprocess

import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as O
from prolandatmo.params import earth, options, with_res, Vec3f

PI_REF = 3.141592657  # common.glsl:99


@pytest.fixture(scope="module")
def P():
    return earth()


def tiny(p=None, R=8, MU=32, MUS=8, NU=4, TW=64, TH=16, SW=16, SH=8):
    q = with_res(p or earth(), R, MU, MUS, NU)
    q.TRANSMITTANCE_W, q.TRANSMITTANCE_H, q.SKY_W, q.SKY_H = TW, TH, SW, SH
    return q


# ---------------------------------------------------------------- phase functions
@pytest.mark.parametrize("g", [0.0, 0.6, 0.76, 0.8, 0.9])
def test_phase_functions_are_normalised(P, g):
    q = P.copy()
    q.mieG = g
    # 2*pi * int_{-1}^{1} p(mu) dmu = pi / PI_REF (the reference's pi constant)
    iR = 2 * math.pi * integrate.quad(O.phaseR, -1, 1, epsabs=0, epsrel=1e-13)[0]
    iM = 2 * math.pi * integrate.quad(lambda m: O.phaseM(q, m), -1, 1, points=[0.9, 0.99],
                                      epsabs=0, epsrel=1e-13, limit=200)[0]
    assert abs(iR - math.pi / PI_REF) < 1e-11
    assert abs(iM - math.pi / PI_REF) < 1e-9
    if g == 0.0:
        for m in np.linspace(-1, 1, 11):
            assert abs(O.phaseM(q, m) - O.phaseR(m)) < 1e-15


def test_phase_R_closed_form():
    for m in (-1.0, -0.3, 0.0, 0.5, 1.0):
        assert O.phaseR(m) == pytest.approx(3.0 / (16.0 * PI_REF) * (1 + m * m), rel=1e-15)


# ---------------------------------------------------------------- geometry
def test_limit_special_rays(P):
    Rg, RL = P.Rg, P.RL
    for r in (Rg + 0.01, Rg + 5, P.Rt - 0.001):
        assert O.limit(P, r, 1.0) == pytest.approx(RL - r, rel=1e-12)
        assert O.limit(P, r, -1.0) == pytest.approx(r - Rg, rel=1e-9)
        assert O.limit(P, r, 0.0) == pytest.approx(math.sqrt(RL * RL - r * r), rel=1e-12)


def test_limit_brute_force_march(P):
    """The ray length ends on the ground sphere (if hit) or on the RL sphere,
    and no earlier crossing exists (checked by marching the ray)."""
    rng = np.random.default_rng(1)
    for _ in range(200):
        r = P.Rg + rng.uniform(0.001, P.Rt - P.Rg)
        mu = rng.uniform(-1, 1)
        d = O.limit(P, r, mu)
        rend = math.sqrt(r * r + d * d + 2 * r * mu * d)
        assert min(abs(rend - P.Rg), abs(rend - P.RL)) < 1e-6
        ts = np.linspace(0, d, 4001)[:-1]
        rs = np.sqrt(r * r + ts * ts + 2 * r * mu * ts)
        assert np.all(rs >= P.Rg - 1e-6) and np.all(rs <= P.RL + 1e-6)


def test_layer_radii(P):
    R = P.RES_R
    H2 = P.Rt ** 2 - P.Rg ** 2
    for l in range(R):
        r, d = O.layer(P, l)
        exact = math.sqrt(P.Rg ** 2 + (l / (R - 1.0)) ** 2 * H2) + (0.01 if l == 0 else (-0.001 if l == R - 1 else 0))
        assert r == np.float32(exact)  # FP32 uniform
        assert d[0] == np.float32(P.Rt - exact)
        assert d[2] == np.float32(exact - P.Rg)
        assert d[3] == np.float32(math.sqrt(exact ** 2 - P.Rg ** 2))
        assert d[1] == np.float32(math.sqrt(exact ** 2 - P.Rg ** 2) + math.sqrt(H2))


# ---------------------------------------------------------------- transmittance integral
def _geom_trapezoid_vertical(P, H, r, n):
    """Closed form of the n-step trapezoid of exp(-(r+x-Rg)/H) on [0, RL-r]."""
    L = P.RL - r
    h = L / n
    q = math.exp(-h / H)
    y0 = math.exp(-(r - P.Rg) / H)
    s = y0 * (1 - q ** (n + 1)) / (1 - q)  # sum_{i=0..n} y_i
    return h * (s - 0.5 * y0 - 0.5 * y0 * q ** n)


@pytest.mark.parametrize("H", [8.0, 1.2])
@pytest.mark.parametrize("alt", [0.01, 3.0, 59.0])
def test_optical_depth_vertical_closed_form(P, H, alt):
    r = P.Rg + alt
    od = O.optical_depth(P, H, r, 1.0)
    assert od == pytest.approx(_geom_trapezoid_vertical(P, H, r, 500), rel=1e-12)
    exact = H * math.exp(-alt / H) * (1 - math.exp(-(P.RL - r) / H))
    h = (P.RL - r) / 500
    bound = (P.RL - r) * h * h / (12 * H * H) * math.exp(-alt / H)  # trapezoid error bound
    assert 0 <= od - exact <= bound * 1.0001


def test_optical_depth_slant_vs_quadrature(P):
    rng = np.random.default_rng(2)
    for _ in range(12):
        r = P.Rg + rng.uniform(0.01, 59.0)
        muh = -math.sqrt(1 - (P.Rg / r) ** 2)
        mu = rng.uniform(muh + 0.01, 1.0)
        for H in (8.0, 1.2):
            L = O.limit(P, r, mu)
            f = lambda x: math.exp(-(math.sqrt(r * r + x * x + 2 * x * r * mu) - P.Rg) / H)
            ex = integrate.quad(f, 0, L, epsabs=0, epsrel=1e-12, limit=500)[0]
            fine = O.optical_depth(P, H, r, mu, options(transmittance_samples=40000))
            assert fine == pytest.approx(ex, rel=2e-8)
            coarse = O.optical_depth(P, H, r, mu)
            assert abs(coarse - ex) / ex < 2e-2


def test_optical_depth_below_horizon_is_opaque(P):
    r = P.Rg + 2.0
    muh = -math.sqrt(1 - (P.Rg / r) ** 2)
    assert O.optical_depth(P, 8.0, r, muh - 1e-3) == 1e9


def _chapman(P, H, r, mu, d):
    """The reference's own analytic optical depth (common.glsl:252-260), an
    independent approximation of the same integral."""
    a = math.sqrt((0.5 / H) * r)
    a01 = np.array([a * mu, a * (mu + d / r)])
    s = np.sign(a01)
    sq = a01 * a01
    x = math.exp(sq[0]) if s[1] > s[0] else 0.0
    y = s / (2.3193 * np.abs(a01) + np.sqrt(1.52 * sq + 4.0)) * np.array([1.0, math.exp(-d / H * (d / (2.0 * r) + mu))])
    return math.sqrt((6.2831 * H) * r) * math.exp((P.Rg - r) / H) * (x + y[0] - y[1])


def test_transmittance_table_vs_chapman(P):
    T = O.transmittance(P)
    assert T.shape == (64, 256, 3)
    assert np.all(T >= 0) and np.all(T <= 1)
    worst = 0.0
    for j in range(0, 64, 7):
        for i in range(0, 256, 9):
            r, mu = O.transmittance_rmu(P, i, j)
            if mu < -math.sqrt(1 - (P.Rg / r) ** 2) + 0.02:
                continue
            d = O.limit(P, r, mu)
            tau = np.array([P.betaR.x, P.betaR.y, P.betaR.z]) * _chapman(P, P.HR, r, mu, d) + \
                np.array([P.betaMEx.x] * 3) * _chapman(P, P.HM, r, mu, d)
            worst = max(worst, float(np.max(np.abs(np.log(T[j, i]) + tau) / np.maximum(tau, 1e-3))))
    assert worst < 0.05


def test_transmittance_beta_zero_is_one(P):
    q = P.copy()
    q.betaR = Vec3f(0, 0, 0)
    q.betaMEx = Vec3f(0, 0, 0)
    T = O.transmittance(q)
    # below-horizon texels have depth 1e9 * 0 = 0 -> T = 1 as well
    assert np.all(T == 1.0)


# ---------------------------------------------------------------- parameterisation round trips
def test_transmittance_parameterisation_round_trip(P):
    for j in range(P.TRANSMITTANCE_H):
        for i in range(0, P.TRANSMITTANCE_W, 5):
            r, mu = O.transmittance_rmu(P, i, j)
            u, v = O.transmittance_uv(P, r, mu)
            assert u == pytest.approx((i + 0.5) / P.TRANSMITTANCE_W, abs=1e-12)
            assert v == pytest.approx((j + 0.5) / P.TRANSMITTANCE_H, abs=1e-12)


def test_irradiance_parameterisation(P):
    for j in range(P.SKY_H):
        for i in range(P.SKY_W):
            r, muS = O.irradiance_rmus(P, i, j)
            assert r == pytest.approx(P.Rg + j / (P.SKY_H - 1) * (P.Rt - P.Rg), rel=1e-15)
            assert muS == pytest.approx(-0.2 + 1.2 * i / (P.SKY_W - 1), abs=1e-15)


@pytest.mark.parametrize("res", [(32, 128, 32, 8), (8, 32, 8, 4)])
def test_texel_coordinates_round_trip(P, res):
    """getMuMuSNu (texel -> r, mu, muS, nu) composed with texture4D's
    coordinates (-> 3D texture coords) returns the texel's own centre, except
    where getMuMuSNu clamps on purpose."""
    p = with_res(P, *res)
    R, MU, MUS, NU = res
    for l in range(R):
        r, dhdH = O.layer(p, l)
        for y in range(MU):
            for x in range(0, MUS * NU, 3 if MUS > 8 else 1):
                mu, muS, nu = O.mumusnu(p, r, dhdH, x, y)
                uR, uMu, uMuS, lerp = O.texture4D_coords(p, r, mu, muS, nu)
                tr = uR * R - 0.5
                if l == 0:
                    # +0.01 km offset of layer 0 (FP32 r: altitude 0.009766 km): texel coordinate 0.395
                    assert 0.0 < tr < 0.45
                    if R == 32:
                        assert tr == pytest.approx(0.395, abs=2e-3)
                elif l == R - 1:
                    assert R - 1.001 < tr < R - 1
                else:
                    assert tr == pytest.approx(l, abs=5e-3)  # FP32 uniform r
                if x % MUS == 0:
                    # muS(column 0) = tan(-0.814)/tan(1.386) = -0.1978 < -0.1975: clamped
                    assert muS < -0.1975 and 0.0 < uMuS * MUS - 0.5 < 0.05
                else:
                    assert uMuS * MUS - 0.5 == pytest.approx(x % MUS, abs=1e-9)
                assert lerp == pytest.approx(x // MUS, abs=1e-12)
                ty = uMu * MU - 0.5
                # rows whose distance d getMuMuSNu clamps (to [dmin, 0.999 dmax]) or
                # whose mu it nudges below the horizon do not map back to their centre
                half = MU / 2.0
                if y < half:
                    draw = (1.0 - y / (half - 1.0)) * dhdH[3]
                    lo, hi = dhdH[2], dhdH[3] * 0.999
                    muh = -math.sqrt(1.0 - (P.Rg / r) ** 2)
                    free = lo < draw < hi and mu < muh - 0.001 - 1e-12
                    free = free and abs(mu - (P.Rg ** 2 - r * r - draw * draw) / (2 * r * draw)) < 1e-9
                else:
                    draw = (y - half) / (half - 1.0) * dhdH[1]
                    free = dhdH[0] < draw < dhdH[1] * 0.999
                if free:
                    # layers 0-1: r is an FP32 uniform while dhdH comes from the double
                    # radius (PreprocessAtmo.cpp:257-268); near the ground that shifts
                    # the ground-half rows by up to 0.26 texel (layer 0), 0.06 (layer 1)
                    tol = 0.3 if l == 0 else (0.07 if l == 1 else 6e-3)
                    assert ty == pytest.approx(y, abs=tol)
                else:
                    assert -0.5 <= ty <= MU - 0.5


def test_texture4D_stable_form_equals_literal(P):
    rng = np.random.default_rng(3)
    n = 0
    for _ in range(3000):
        r = P.Rg + rng.uniform(0.05, P.RL - P.Rg)
        mu, muS, nu = rng.uniform(-1, 1, 3)
        a = O.texture4D_coords(P, r, mu, muS, nu)
        b = O.texture4D_coords(P, r, mu, muS, nu, literal=True)
        if r > P.Rt and (r * mu) ** 2 - r * r + P.Rt ** 2 < 0:
            # above Rt with an impact parameter > Rt (no ray of the passes does
            # this): the literal form takes sqrt of a negative number, D1 clamps
            assert np.isnan(b[1]) and np.isfinite(a[1])
            continue
        assert np.allclose(a, b, rtol=0, atol=1e-9)
        n += r * mu < 0 and (r * mu) ** 2 - r * r + P.Rg ** 2 > 0
    assert n > 100  # ground branch exercised


# ---------------------------------------------------------------- GL filtering emulation
def test_bilinear_trilinear_reproduce_affine_functions():
    rng = np.random.default_rng(4)
    W, H, D = 7, 5, 4
    a, b, c, d = rng.normal(size=4)
    k, j, i = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    tex3 = (a + b * i + c * j + d * k)[..., None].repeat(2, -1)
    for _ in range(200):
        # interior texel coordinates in [0, N-1]
        x, y, z = rng.uniform(0, W - 1), rng.uniform(0, H - 1), rng.uniform(0, D - 1)
        v = O.sample3D(tex3, (x + 0.5) / W, (y + 0.5) / H, (z + 0.5) / D)
        assert v[0] == pytest.approx(a + b * x + c * y + d * z, abs=1e-12)
    tex2 = tex3[0]
    for _ in range(50):
        x, y = rng.uniform(0, W - 1), rng.uniform(0, H - 1)
        assert O.sample2D(tex2, (x + 0.5) / W, (y + 0.5) / H)[1] == pytest.approx(a + b * x + c * y, abs=1e-12)
    # CLAMP_TO_EDGE: outside the texel-centre range the edge texel is returned
    assert O.sample2D(tex2, -3.0, 0.5 / H)[0] == pytest.approx(a, abs=1e-12)
    assert O.sample2D(tex2, 5.0, 0.5 / H)[0] == pytest.approx(a + b * (W - 1), abs=1e-12)
    assert O.sample2D(tex2, 0.25 / W, 0.5 / H)[0] == pytest.approx(a, abs=1e-12)


def test_texture4D_constant_and_nu_slices(P):
    p = with_res(P, 4, 8, 4, 3)
    tab = np.full((4, 8, 12, 3), 2.5)
    assert np.allclose(O.texture4D(p, tab, P.Rg + 7, 0.3, 0.2, 0.1), 2.5, atol=1e-14)
    # a table that is constant per nu slice interpolates linearly in nu
    tab = np.zeros((4, 8, 12, 3))
    for k in range(3):
        tab[:, :, k * 4:(k + 1) * 4, :] = k
    for nu in (-1.0, -0.5, 0.2, 0.9):
        assert O.texture4D(p, tab, P.Rg + 7, 0.3, 0.2, nu)[0] == pytest.approx((nu + 1) / 2 * 2, abs=1e-12)


# ---------------------------------------------------------------- pass-level closed forms
def test_irradiance1_is_T_times_cos(P):
    p = tiny(P)
    T = np.ones((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3))
    dE = O.irradiance1(p, T)
    for i in range(p.SKY_W):
        muS = -0.2 + 1.2 * i / (p.SKY_W - 1)
        assert np.allclose(dE[:, i, :], max(muS, 0.0), atol=1e-15)


def test_inscatterN_constant_J_gives_ray_length(P):
    p = tiny(P)
    T = np.ones((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3))
    dJ = np.full((p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 3), 0.7)
    dS = O.inscatterN(p, T, dJ)
    for l in range(p.RES_R):
        r, dhdH = O.layer(p, l)
        for y in range(p.RES_MU):
            mu, _, _ = O.mumusnu(p, r, dhdH, 0, y)
            assert np.allclose(dS[l, y], 0.7 * O.limit(p, r, mu), rtol=1e-12)


def test_inscatter1_with_unit_transmittance_matches_quadrature(P):
    """T = 1: dSR = betaR * int_0^limit exp(-h/HR) [sun visible] dt."""
    p = tiny(P)
    o = options(inscatter_samples=6000)
    T = np.ones((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3))
    dSR, dSM = O.inscatter1(p, T, o)
    rng = np.random.default_rng(5)
    for _ in range(25):
        l, y, x = rng.integers(p.RES_R), rng.integers(p.RES_MU), rng.integers(p.RES_NU * p.RES_MU_S)
        r, dhdH = O.layer(p, l)
        mu, muS, nu = O.mumusnu(p, r, dhdH, x, y)
        L = O.limit(p, r, mu)

        def f(t, H):
            ri = math.sqrt(r * r + t * t + 2 * r * mu * t)
            muSi = (nu * t + muS * r) / ri
            ri = max(P.Rg, ri)
            vis = muSi >= -math.sqrt(1 - P.Rg ** 2 / ri ** 2)
            return math.exp(-(ri - P.Rg) / H) if vis else 0.0

        for H, tab, beta in ((P.HR, dSR, P.betaR.x), (P.HM, dSM, P.betaMSca.x)):
            ex = beta * integrate.quad(lambda t: f(t, H), 0, L, limit=400, epsrel=1e-10)[0]
            # trapezoid with a possible jump (sun setting along the ray): error <= one step
            step = L / o.inscatter_samples * beta
            assert abs(tab[l, y, x, 0] - ex) <= 1e-6 * ex + step


def _sphere_sum(NS, f):
    """sum over the reference's midpoint grid (inscatterS.glsl:109-110,127-145)"""
    dth = dph = PI_REF / NS
    s = 0.0
    for it in range(NS):
        th = (it + 0.5) * dth
        s += f(math.cos(th)) * dth * math.sin(th)
    return s * 2 * NS * dph


def test_inscatterS_constant_dS_zenith_texels(P):
    """dS = C, first = 0, dE = 0: for texels looking at the zenith the phi sum
    collapses to 2*PI_REF and J = C * 2pi sum_theta (bR e^-h/HR pR + bM e^-h/HM pM) sin dtheta."""
    p = tiny(P)
    T = np.ones((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3))
    dE = np.zeros((p.SKY_H, p.SKY_W, 3))
    C = 0.37
    dS = np.full((p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 3), C)
    dJ = O.inscatterS(p, T, dE, dS, dS * 0, False)
    y = p.RES_MU // 2  # d = Rt - r -> mu = 1
    for l in range(p.RES_R):
        r, dhdH = O.layer(p, l)
        mu, _, _ = O.mumusnu(p, r, dhdH, 0, y)
        if l == p.RES_R - 1:
            # r - Rt differs from dhdH.x after FP32 rounding: mu = 0.977, not 1
            assert mu < 0.99
            continue
        assert mu == pytest.approx(1.0, abs=1e-4)  # clamped to 1 by the pass
        eR, eM = math.exp(-(r - P.Rg) / P.HR), math.exp(-(r - P.Rg) / P.HM)
        for c, bR in enumerate((P.betaR.x, P.betaR.y, P.betaR.z)):
            bM = P.betaMSca.x
            ex = C * _sphere_sum(16, lambda ct: bR * eR * O.phaseR(ct) + bM * eM * O.phaseM(P, ct))
            assert np.allclose(dJ[l, y, :, c], ex, rtol=2e-5)


def test_inscatterS_quadrature_weight_closed_form(P):
    """mieG = 0 makes both phase functions Rayleigh; with dS = C the result is
    C (bR e^-h/HR + bM e^-h/HM) * sum(pR dw) for every texel, and the grid sums
    of (1 + cos^2) are known in closed form up to the quadrature error."""
    p = tiny(P)
    p.mieG = 0.0
    T = np.ones((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3))
    dE = np.zeros((p.SKY_H, p.SKY_W, 3))
    dS = np.full((p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 3), 1.0)
    dJ = O.inscatterS(p, T, dE, dS, dS * 0, False)
    sum_dw = 2 * PI_REF * (PI_REF / 16) / math.sin(PI_REF / 32) * math.sin(PI_REF / 2) ** 2
    assert sum_dw / (4 * math.pi) - 1 == pytest.approx(1.608e-3, abs=5e-6)
    for l in range(p.RES_R):
        r, _ = O.layer(p, l)
        k = P.betaR.x * math.exp(-(r - P.Rg) / P.HR) + P.betaMSca.x * math.exp(-(r - P.Rg) / P.HM)
        ratio = dJ[l, :, :, 0] / k
        # integral of pR over the sphere is pi/PI_REF; midpoint error < 1%
        assert np.all(np.abs(ratio - 1.0) < 1e-2)


def test_inscatterS_first_order_rayleigh_double_phase(P):
    """first = 1, dSR = C, dSM = 0, betaM = 0: J = C bR e^-h/HR int pR(s.w) pR(v.w) dw
    = C bR e^-h/HR (3/16pi)^2 4pi (5/3 + (1 + 2 nu^2)/15)  (nu = s.v)."""
    p = tiny(P)
    p.betaMSca = Vec3f(0, 0, 0)
    T = np.ones((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3))
    dE = np.zeros((p.SKY_H, p.SKY_W, 3))
    dSR = np.full((p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 3), 1.0)
    dJ = O.inscatterS(p, T, dE, dSR, dSR * 0, True)
    k = (3 / (16 * PI_REF)) ** 2 * 4 * math.pi
    for l in range(1, p.RES_R - 1):
        r, dhdH = O.layer(p, l)
        e = P.betaR.x * math.exp(-(r - P.Rg) / P.HR)
        for y in range(1, p.RES_MU - 1):
            for x in range(p.RES_NU * p.RES_MU_S):
                mu, muS, nu = O.mumusnu(p, r, dhdH, x, y)
                mu, muS = max(-1, min(1, mu)), max(-1, min(1, muS))
                var = math.sqrt(1 - mu * mu) * math.sqrt(1 - muS * muS)
                nu = max(muS * mu - var, min(muS * mu + var, nu))
                ex = e * k * (5.0 / 3.0 + (1 + 2 * nu * nu) / 15.0)
                assert dJ[l, y, x, 0] == pytest.approx(ex, rel=6e-3)


def test_inscatterS_ground_term_zenith(P):
    """dS = 0, dE = E0 constant, T = 1: only the ground-reflection term remains.
    For zenith texels J = (alpha/PI_REF) E0 * 2pi sum_{ground rows} coef(cos) sin dtheta."""
    p = tiny(P)
    T = np.ones((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3))
    E0 = 0.25
    dE = np.full((p.SKY_H, p.SKY_W, 3), E0)
    z = np.zeros((p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 3))
    dJ = O.inscatterS(p, T, dE, z, z, False)
    y = p.RES_MU // 2
    for l in range(p.RES_R - 1):
        r, _ = O.layer(p, l)
        cmin = -math.sqrt(1 - (P.Rg / r) ** 2)
        eR, eM = math.exp(-(r - P.Rg) / P.HR), math.exp(-(r - P.Rg) / P.HM)
        f = lambda ct: (P.betaR.x * eR * O.phaseR(ct) + P.betaMSca.x * eM * O.phaseM(P, ct)) * (ct < cmin)
        ex = P.AVERAGE_GROUND_REFLECTANCE / PI_REF * E0 * _sphere_sum(16, f)
        assert np.allclose(dJ[l, y, :, 0], ex, rtol=2e-5)


def test_irradianceN_closed_forms(P):
    p = tiny(P)
    C = 0.9
    dS = np.full((p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 3), C)
    dE = O.irradianceN(p, dS, dS * 0, False)
    x = PI_REF / 32
    ex = C * 2 * PI_REF * (x / 2) * math.sin(16 * x) ** 2 / math.sin(x)
    assert ex / (C * math.pi) - 1 == pytest.approx(1.608e-3, abs=5e-6)
    assert np.allclose(dE, ex, rtol=1e-12)
    # first order: int_{upper} pR(s.w) w_z dw = (3/16pi)(pi + pi/4 (1 + muS^2))
    dE1 = O.irradianceN(p, dS, dS * 0, True)
    for i in range(p.SKY_W):
        muS = -0.2 + 1.2 * i / (p.SKY_W - 1)
        ex = C * 3 / (16 * PI_REF) * (math.pi + math.pi / 4 * (1 + muS * muS))
        assert np.allclose(dE1[:, i, :], ex, rtol=4e-3)


# ---------------------------------------------------------------- whole chain
def test_zero_scattering_gives_empty_tables(P):
    p = tiny(P)
    p.betaR = Vec3f(0, 0, 0)
    p.betaMSca = Vec3f(0, 0, 0)
    p.betaMEx = Vec3f(0, 0, 0)
    T, E, S = O.precompute(p)
    assert np.all(T == 1.0)
    assert np.all(E == 0.0)
    assert np.all(S == 0.0)


def test_orders_decrease_and_tables_are_physical(P):
    p = tiny(P)
    st = O.precompute_stages(p, options(orders=5))
    m = [np.abs(st["dSR1"]).max()] + [np.abs(st["dS%d" % n]).max() for n in range(2, 6)]
    assert all(m[i + 1] < m[i] for i in range(len(m) - 1))
    e = [np.abs(st["dE%d" % n]).max() for n in range(2, 6)]
    assert all(e[i + 1] < e[i] for i in range(len(e) - 1))
    assert np.all(st["S"] >= 0) and np.all(st["E"] >= 0)
    assert np.all((st["T"] >= 0) & (st["T"] <= 1))
    T, E, S = O.precompute(p, options(orders=5))
    # the chain evaluates its segment ratios from the optical depths (exp/log
    # of the same values as the pass-by-pass quotients): equal to rounding
    assert np.array_equal(T, st["T"])
    assert np.allclose(S, st["S"], rtol=1e-13, atol=0)
    assert np.allclose(E, st["E"], rtol=1e-13, atol=0)


def test_rayleigh_only_atmosphere_has_no_mie_alpha(P):
    p = tiny(P)
    p.betaMSca = Vec3f(0, 0, 0)
    p.betaMEx = Vec3f(0, 0, 0)
    T, E, S = O.precompute(p)
    assert np.all(S[..., 3] == 0.0)
    assert S[..., :3].max() > 0


def test_sample_count_convergence(P):
    """Brute force: raising every sample count moves the tables by less than
    the coarse grid's error estimate, and the change shrinks as the grid refines."""
    p = tiny(P, R=6, MU=16, MUS=6, NU=3, TW=32, TH=8, SW=8, SH=4)
    base = options(orders=2)
    T0, E0, S0 = O.precompute(p, base)
    T1, E1, S1 = O.precompute(p, options(orders=2, transmittance_samples=2000, inscatter_samples=200))
    T2, E2, S2 = O.precompute(p, options(orders=2, transmittance_samples=8000, inscatter_samples=800))
    d01 = np.abs(S1 - S0).max() / S1.max()
    d12 = np.abs(S2 - S1).max() / S2.max()
    assert d12 < d01 / 3
    assert np.abs(T1 - T0).max() > np.abs(T2 - T1).max()


def test_segment_ratio_from_optical_depths_beyond_fp64_range(P):
    """Deviation D6: the chain's segment ratio T(x)/T(y) from the optical-depth
    table.  The quotient of two bilinear lookups is unchanged when the table is
    scaled per channel, so the plain quotient over exp(c - tau), c the depth at
    the start texel (the taps then lie in FP64's range), is the exact value.
    In a dense atmosphere whose grazing depths exceed 745 the plain quotient
    over T itself is 0, 0/0 (1 under D5) or loses digits in subnormals, while
    the depth form stays exact."""
    q = tiny(P, TW=256, TH=64)
    Rg = q.Rg = 46427.34375
    q.Rt, q.RL = Rg + 64.5234375, Rg + 64.5234375 + 0.01171875
    q.HR, q.HM = 42.71, 13.96
    q.betaR = Vec3f(5.8e-3 * 14.77, 1.35e-2 * 14.77, 3.31e-2 * 14.77)
    q.betaMSca = Vec3f(3.165e-4, 3.165e-4, 3.165e-4)
    q.betaMEx = Vec3f(3.165e-4 / 0.9, 3.165e-4 / 0.9, 3.165e-4 / 0.9)
    tau = O.optical_depth_table(q)
    T = O.transmittance(q)
    assert np.allclose(T, np.exp(-tau), rtol=1e-13, atol=0)
    beyond = 0
    for (h, mu, d) in ((0.05, 0.0002, 5.0), (0.2, 0.0005, 20.0), (0.5, 0.0002, 20.0),
                       (1.0, 0.0005, 5.0), (0.5, 0.001, 20.0), (0.2, -0.004, 20.0),
                       (5.0, -0.05, 100.0), (30.0, 0.3, 20.0), (0.05, -0.3, 0.1)):
        r = Rg + h
        u, v = O.transmittance_uv(q, r, abs(mu))
        i = min(int(u * q.TRANSMITTANCE_W), q.TRANSMITTANCE_W - 1)
        j = min(int(v * q.TRANSMITTANCE_H), q.TRANSMITTANCE_H - 1)
        scaled = np.exp(np.minimum(tau[j, i][None, None, :] - tau, 700.0))
        exact = O.transmittance_segment(q, scaled, r, mu, d)
        got = O.transmittance_segment_tau(q, tau, r, mu, d)
        assert np.allclose(got, exact, rtol=1e-13, atol=0), (h, mu, d)
        plain = O.transmittance_segment(q, T, r, mu, d)
        for c in range(3):
            if plain[c] != pytest.approx(exact[c], rel=1e-6, abs=0):
                # lookups at or below FP64's subnormal range: the plain
                # quotient is 0, 1 (0/0 under D5) or off by the lost digits
                beyond += 1
                assert 0.0 < got[c] < 1.0
    assert beyond >= 4


# ---------------------------------------------------------------- segment transmittance
def _segment_tables(P):
    """A fine T table (1024x256, 2000 samples per optical depth) and its depth
    table, so that the lookups' own error (bilinear between texels, trapezoid)
    stays below 2.2e-3 of the optical depth on every segment below."""
    q = P.copy()
    q.TRANSMITTANCE_W, q.TRANSMITTANCE_H = 1024, 256
    o = options(transmittance_samples=2000)
    return q, o, O.transmittance(q, o), O.optical_depth_table(q, o)


def _segment_exact_depth(q, r, mu, d):
    """Optical depth of the segment from radius r along mu over length d:
    sum_i beta_i * int_0^d exp(-(|x + t v| - Rg) / H_i) dt by adaptive quadrature."""
    bR = np.array([q.betaR.x, q.betaR.y, q.betaR.z])
    bM = np.array([q.betaMEx.x, q.betaMEx.y, q.betaMEx.z])

    def rho(t, H):
        return math.exp(-(math.sqrt(r * r + t * t + 2.0 * r * mu * t) - q.Rg) / H)

    iR = integrate.quad(lambda t: rho(t, q.HR), 0, d, epsabs=0, epsrel=1e-12, limit=200)[0]
    iM = integrate.quad(lambda t: rho(t, q.HM), 0, d, epsabs=0, epsrel=1e-12, limit=200)[0]
    return bR * iR + bM * iM


def _segment_cases(q):
    """Upward and downward segments at three altitudes, none reaching the
    ground: 10, 40 and 90 % of the ray length limit(r, mu), including grazing
    rays just above the horizon that pass their lowest point and rise again."""
    out = []
    for alt in (0.5, 5.0, 30.0):
        r = q.Rg + alt
        muh = -math.sqrt(1.0 - (q.Rg / r) ** 2)
        for mu in (0.9, 0.3, 0.05, -0.05, muh + 0.02, -0.5):
            L = O.limit(q, r, mu)
            for f in (0.1, 0.4, 0.9):
                out.append((r, mu, f * L))
    return out


_SEGMENT_CASES = None


def _segment_worst(P):
    """Worst |ln T_segment + tau| / (3e-3 tau + 2e-5) over the cases, for both
    the plain quotient of T lookups and the chain's form from optical depths
    (D6).  <= 1 passes; the lookups' own error reaches 2.2e-3 tau."""
    global _SEGMENT_CASES
    if _SEGMENT_CASES is None:
        q, o, T, tau = _segment_tables(P)
        _SEGMENT_CASES = (q, o, T, tau, [(r, mu, d, _segment_exact_depth(q, r, mu, d))
                                         for (r, mu, d) in _segment_cases(q)])
    q, o, T, tau, cases = _SEGMENT_CASES
    worst = 0.0
    for (r, mu, d, ex) in cases:
        for got in (O.transmittance_segment(q, T, r, mu, d),
                    O.transmittance_segment_tau(q, tau, r, mu, d, o)):
            with np.errstate(divide="ignore"):
                depth = -np.log(got)
            err = np.max(np.abs(depth - ex) / (3e-3 * ex + 2e-5))
            worst = max(worst, float(err) if np.isfinite(err) else math.inf)
    return worst


def test_segment_transmittance_matches_quadrature(P):
    """common.glsl:272-282: T(x, x + d v) = min(T(r, mu) / T(r1, mu1), 1) for
    mu > 0 and T(r1, -mu1) / T(r, -mu) otherwise, against exp(-optical depth)
    of the segment by quadrature, with T tables that are not identically 1."""
    assert _segment_worst(P) <= 1.0


# ---------------------------------------------------------------- copyInscatterN
def _copy_inscatterN_worst(P):
    """copyInscatterN.glsl:105-110 with the ADD blend (PreprocessAtmo.cpp:336,350):
    with dS = c * phaseR(nu_k), nu_k = -1 + 2k / (RES_NU - 1) the nu of column
    x = k RES_MU_S + j (common.glsl:199), S.rgb gains exactly c and S.a gains 0.
    Returns the worst relative deviation."""
    p = tiny(P, R=4, MU=16, MUS=8, NU=5)
    rng = np.random.default_rng(11)
    W = p.RES_NU * p.RES_MU_S
    c = rng.uniform(0.1, 2.0, (p.RES_R, p.RES_MU, W, 3))
    k = np.arange(W) // p.RES_MU_S
    nu = -1.0 + 2.0 * k / (p.RES_NU - 1.0)
    pR = 3.0 / (16.0 * PI_REF) * (1.0 + nu * nu)
    dS = np.ascontiguousarray(c * pR[None, None, :, None])
    S0 = rng.uniform(0.0, 1.0, (p.RES_R, p.RES_MU, W, 4))
    S1 = O.copy_inscatterN(p, dS, S0)
    rgb = np.max(np.abs((S1[..., :3] - S0[..., :3]) - c) / c)
    alpha = np.max(np.abs(S1[..., 3] - S0[..., 3]))
    return max(float(rgb), float(alpha))


def test_copy_inscatterN_divides_by_rayleigh_phase_of_the_texel_nu(P):
    assert _copy_inscatterN_worst(P) <= 1e-13


# ---------------------------------------------------------------- planted mistakes
_ORACLE_SRC = os.path.join(os.path.dirname(O.__file__), "atmo_oracle.cpp")

# (pin, function whose body is mutated, text replaced, replacement)
_MUTANTS = {
    "segment branches swapped": (_segment_worst, "void T_segment(", "if (mu > 0.0)", "if (!(mu > 0.0))"),
    "segment end radius sign": (_segment_worst, "void T_segment(", "+ 2.0 * r * mu * d", "- 2.0 * r * mu * d"),
    "segment end cosine sign": (_segment_worst, "void T_segment(", "(r * mu + d) / r1", "(r * mu - d) / r1"),
    "copyInscatterN times phase": (_copy_inscatterN_worst, "void oracle_pass_copy_inscatterN(",
                                   "dS[k * 3 + c] / pr", "dS[k * 3 + c] * pr"),
    "copyInscatterN into alpha": (_copy_inscatterN_worst, "void oracle_pass_copy_inscatterN(",
                                  "S[k * 4 + 3] += 0.0", "S[k * 4 + 3] += dS[k * 3] / pr"),
}
_PIN_LIMIT = {_segment_worst: 1.0, _copy_inscatterN_worst: 1e-13}


def _build_mutant(tmp_path, func, old, new):
    src = open(_ORACLE_SRC).read()
    a = src.index(func)
    b = src.index("\n}\n", a)
    body = src[a:b]
    assert old in body, (func, old)
    mutated = src[:a] + body.replace(old, new) + src[b:]
    cpp = os.path.join(str(tmp_path), "atmo_oracle.cpp")
    so = os.path.join(str(tmp_path), "liboracle_mutant.so")
    with open(cpp, "w") as f:
        f.write(mutated)
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
                           "-shared", "-I", os.path.dirname(_ORACLE_SRC), "-o", so, cpp])
    return so


@pytest.mark.parametrize("name", sorted(_MUTANTS))
def test_pins_fail_on_planted_mistakes(P, tmp_path, name):
    """Each pin above rejects an oracle with the mistake planted (a mutated copy
    of oracle/atmo_oracle.cpp built in a temporary directory), while the
    oracle itself passes it."""
    pin, func, old, new = _MUTANTS[name]
    assert pin(P) <= _PIN_LIMIT[pin]
    so = _build_mutant(tmp_path, func, old, new)
    with O.using(so):
        assert pin(P) > _PIN_LIMIT[pin], name
