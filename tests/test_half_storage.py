"""FP16 table storage (ATMO_FLAG_HALF_STORAGE) in the oracle.

The reference keeps every table in RGB16F / RGBA16F textures
(PreprocessAtmo.cpp:130-143) and adds orders with FP16 ADD blending (:336,
:350).  The oracle's rounding is pinned to NumPy's binary16 conversion (an
independent implementation, round to nearest even with gradual underflow), and
its half-storage chain to the per-pass oracle functions with NumPy doing every
rounding.
"""
import numpy as np
import pytest

from oracle import oracle as O
from prolandatmo.params import FLAG_HALF_STORAGE, draw_sets, earth, options, with_res, SEED_BATCH


def tiny(p=None, R=8, MU=32, MUS=8, NU=4, TW=64, TH=16, SW=16, SH=8):
    q = with_res(p or earth(), R, MU, MUS, NU)
    q.TRANSMITTANCE_W, q.TRANSMITTANCE_H, q.SKY_W, q.SKY_H = TW, TH, SW, SH
    return q


def f16(x):
    with np.errstate(over="ignore"):
        return np.asarray(x, dtype=np.float64).astype(np.float16).astype(np.float64)


def test_q16_matches_numpy_binary16():
    rng = np.random.default_rng(7)
    mags = 10.0 ** rng.uniform(-12, 5, 20000)
    x = mags * rng.choice([-1.0, 1.0], mags.size)
    # exact ties between neighbouring FP16 values (normal and subnormal)
    h = rng.uniform(-60000, 60000, 3000).astype(np.float16)
    hs = (rng.integers(1, 1023, 1000) * 2.0 ** -24).astype(np.float16)
    ties = []
    for a in np.concatenate([h, hs]):
        b = np.nextafter(a, np.float16(np.inf))
        if np.isfinite(b):
            ties.append((float(a) + float(b)) / 2)
    special = [0.0, -0.0, 2.0 ** -14, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -25, 2.0 ** -26,
               65504.0, 65519.99, 65520.0, 1e6, -1e6, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11]
    x = np.concatenate([x, ties, special])
    got = O.q16(x)
    want = f16(x)
    assert np.array_equal(got, want, equal_nan=True)
    assert np.signbit(O.q16(np.array([-0.0])))[0]


def _chain_numpy_rounding(p, o):
    """The half-storage pass chain from the per-pass oracle functions, every
    stored table rounded by NumPy (PreprocessAtmo.cpp:273-354)."""
    T = f16(O.transmittance(p, o))
    dE = f16(O.irradiance1(p, T, o))
    dSR, dSM = O.inscatter1(p, T, o)
    dSR, dSM = f16(dSR), f16(dSM)
    E = np.zeros_like(dE)
    S = O.copy_inscatter1(p, dSR, dSM)
    for n in range(2, o.orders + 1):
        first = n == 2
        dJ = f16(O.inscatterS(p, T, dE, dSR, dSM, first, o))
        dE = f16(O.irradianceN(p, dSR, dSM, first, o))
        dSR = f16(O.inscatterN(p, T, dJ, o))
        E = f16(E + dE)
        S = f16(O.copy_inscatterN(p, dSR, S))
    return T, E, S


@pytest.mark.parametrize("which", ["earth", "seeded"])
def test_half_storage_chain(which):
    p = tiny() if which == "earth" else tiny(draw_sets(SEED_BATCH, 1, res=(8, 32, 8, 4))[0][0])
    o = options(flags=FLAG_HALF_STORAGE)
    T, E, S = O.precompute(p, o)
    for a in (T, E, S):
        assert np.array_equal(a, f16(a))          # every stored value is an FP16 value
    T64, E64, S64 = O.precompute(p, options())
    assert np.array_equal(T, f16(T64))            # T: the FP64 table, rounded once
    rT, rE, rS = _chain_numpy_rounding(p, o)
    assert np.array_equal(T, rT)
    assert np.array_equal(E, rE)
    assert np.array_equal(S, rS)
    # single scattering (S.a) only sees the rounding of T and of its own store
    fl = 1e-4 * np.abs(S64[..., 3]).max()
    rel = np.abs(S[..., 3] - S64[..., 3]) / np.maximum(np.abs(S64[..., 3]), fl)
    assert rel.max() < 4e-3


def test_half_storage_flag_is_ignored_without_it():
    p = tiny()
    a = O.precompute(p, options())
    b = O.precompute(p, options(flags=1))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
