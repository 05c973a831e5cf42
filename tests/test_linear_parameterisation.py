"""The linear parameterisations of the oracle (the reference built without
TRANSMITTANCE_NON_LINEAR / INSCATTER_NON_LINEAR, common.glsl:105-106, whose
#else branches are at :119-120, :132-133, :162-166 and :194-199).

Pinned by closed forms: the texel -> (r, mu) -> texel round trips are exact at
texel centres, the transmittance table at each texel is exp(-tau) of the
independently pinned optical depth at the closed-form (r, mu) of that texel,
and texture4D at a texel's own (r, mu, muS, nu) returns that texel.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from prolandatmo.params import (FLAG_LINEAR_INSCATTER, FLAG_LINEAR_TRANSMITTANCE, earth, options,
                                with_res)

LT = options(flags=FLAG_LINEAR_TRANSMITTANCE)
L4 = options(flags=FLAG_LINEAR_INSCATTER)


def tiny(R=8, MU=32, MUS=8, NU=4):
    q = with_res(earth(), R, MU, MUS, NU)
    q.TRANSMITTANCE_W, q.TRANSMITTANCE_H, q.SKY_W, q.SKY_H = 64, 16, 16, 8
    return q


def test_linear_transmittance_texels_round_trip():
    p = earth()
    W, H = p.TRANSMITTANCE_W, p.TRANSMITTANCE_H
    for j in range(H):
        for i in range(0, W, 7):
            r, mu = O.transmittance_rmu(p, i, j, LT)
            assert r == pytest.approx(p.Rg + (j + 0.5) / H * (p.Rt - p.Rg), rel=1e-12)
            assert mu == pytest.approx(-0.15 + (i + 0.5) / W * 1.15, abs=1e-12)
            u, v = O.transmittance_uv(p, r, mu, LT)
            assert u == pytest.approx((i + 0.5) / W, abs=1e-12)
            assert v == pytest.approx((j + 0.5) / H, abs=1e-12)


def test_linear_transmittance_table_is_exp_of_the_optical_depth():
    p = tiny()
    T = O.transmittance(p, LT)
    W, H = p.TRANSMITTANCE_W, p.TRANSMITTANCE_H
    bR = np.array([p.betaR.x, p.betaR.y, p.betaR.z])
    bM = np.array([p.betaMEx.x, p.betaMEx.y, p.betaMEx.z])
    for j in range(H):
        for i in range(0, W, 5):
            r = p.Rg + (j + 0.5) / H * (p.Rt - p.Rg)
            mu = -0.15 + (i + 0.5) / W * 1.15
            if mu < -math.sqrt(1.0 - (p.Rg / r) ** 2):
                assert np.all(T[j, i] == 0.0)   # depth 1e9 below the horizon
                continue
            tau = bR * O.optical_depth(p, p.HR, r, mu) + bM * O.optical_depth(p, p.HM, r, mu)
            assert np.allclose(T[j, i], np.exp(-tau), rtol=1e-12, atol=0)
    # the non-linear table differs (its rows are squared in r, its columns tan-spaced)
    assert not np.allclose(T, O.transmittance(p))


@pytest.mark.parametrize("res", [(8, 32, 8, 4), (4, 16, 6, 5)])
def test_linear_inscatter_texels_round_trip(res):
    p = with_res(earth(), *res)
    R, MU, MUS, NU = res
    for l in range(R):
        r, dhdH = O.layer(p, l)
        for y in range(MU):
            for x in range(MUS * NU):
                mu, muS, nu = O.mumusnu(p, r, dhdH, x, y, L4)
                assert mu == pytest.approx(-1.0 + 2.0 * y / (MU - 1), abs=1e-15)
                assert muS == pytest.approx(-0.2 + 1.2 * (x % MUS) / (MUS - 1), abs=1e-15)
                uR, uMu, uMuS, lerp = O.texture4D_coords(p, r, mu, muS, nu, o=L4)
                assert uMu * MU - 0.5 == pytest.approx(y, abs=1e-9)
                assert uMuS * MUS - 0.5 == pytest.approx(x % MUS, abs=1e-9)
                assert lerp == pytest.approx(x // MUS, abs=1e-9)
                # the literal form is the same expression in this parameterisation
                assert np.array_equal(O.texture4D_coords(p, r, mu, muS, nu, literal=True, o=L4),
                                      [uR, uMu, uMuS, lerp])


def test_linear_texture4D_reads_back_its_texels():
    p = with_res(earth(), 6, 16, 6, 4)
    rng = np.random.default_rng(5)
    tab = rng.random((6, 16, 24, 3))
    for l in (1, 3, 4):   # inner layers: uR lands on the layer's texel centre
        r, dhdH = O.layer(p, l)
        for y in range(16):
            for x in range(24):
                mu, muS, nu = O.mumusnu(p, r, dhdH, x, y, L4)
                v = O.texture4D(p, tab, r, mu, muS, nu, o=L4)
                # FP32 layer radii: the r coordinate sits within 5e-3 texel of the centre
                assert np.allclose(v, tab[l, y, x], atol=5e-3 * np.abs(np.diff(tab, axis=0)).max())


def test_linear_chain_is_physical():
    """Both linear parameterisations together: finite, non-negative tables, T in
    [0, 1], the multiple-scattering orders adding to S."""
    p = tiny()
    o = options(flags=FLAG_LINEAR_TRANSMITTANCE | FLAG_LINEAR_INSCATTER)
    T, E, S = O.precompute(p, o)
    T1, E1, S1 = O.precompute(p, options(orders=1, flags=o.flags))
    for a in (T, E, S):
        assert np.all(np.isfinite(a)) and np.all(a >= 0)
    assert np.all(T <= 1)
    assert np.all(S[..., :3] >= S1[..., :3]) and np.array_equal(S[..., 3], S1[..., 3])
    assert np.all(E >= 0) and np.all(E1 == 0)
