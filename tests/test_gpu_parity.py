"""Parity of the HIP path (libproland_atmo.so through its C ABI) with the FP64
oracle, pass by pass and for the whole precompute.

Bar (floating point): per table, max relative error <= 1e-3 against the FP64
oracle, measured with prolandatmo.metrics.rel_err (relative error, floored at
1e-4 of the table's largest magnitude).  Pass-level checks use tighter bars
where the pass has no amplifying step.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from prolandatmo import lib as L  # noqa: E402
from prolandatmo.metrics import rel_err, summary  # noqa: E402
from prolandatmo.params import (AtmoParams, FLAG_CHECK_NUMERICS, FLAG_HALF_STORAGE, FLAG_NO_AUTOTUNE,  # noqa: E402
                                FLAG_LINEAR_INSCATTER, FLAG_LINEAR_TRANSMITTANCE, SEED_BATCH, draw_sets, earth, options,
                                with_res)

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-3


@pytest.fixture(scope="module")
def eng():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    e = L.Engine(0)
    yield e
    e.close()


def small(p=None, res=(8, 32, 8, 4), T=(64, 16), E=(16, 8)):
    q = with_res(p or earth(), *res)
    q.TRANSMITTANCE_W, q.TRANSMITTANCE_H = T
    q.SKY_W, q.SKY_H = E
    return q


def f4(a):
    return L.to_float4(a)


def rgb(t):
    return t[..., :3].double().cpu().numpy()


def seeded(n, res, seed=SEED_BATCH):
    sets, _ = draw_sets(seed, n, res=res)
    for p in sets:
        p.TRANSMITTANCE_W, p.TRANSMITTANCE_H, p.SKY_W, p.SKY_H = 64, 16, 16, 8
    return sets


CASES = [small(), small(res=(6, 16, 6, 4), T=(32, 8), E=(8, 4))] + seeded(2, (8, 32, 8, 4))


def _assert(name, x, y, tol):
    s = summary(x, y)
    print("%-10s max %.3e mean %.3e (unfloored %.3e at %s)" % (name, s["max"], s["mean"],
                                                               s["max_unfloored"], s["argmax"]),
          flush=True)
    assert s["max"] <= tol, (name, s)


# ---------------------------------------------------------------- per pass
@pytest.mark.parametrize("ci", range(len(CASES)))
def test_pass_transmittance(eng, ci):
    p = CASES[ci]
    o = options()
    T4 = torch.zeros((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 4), device="cuda")
    eng.pass_transmittance(p, o, T4)
    _assert("T", rgb(T4), O.transmittance(p, o), 1e-5)


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_pass_irradiance1_and_inscatter1(eng, ci):
    p = CASES[ci]
    o = options()
    T = O.transmittance(p, o)
    dE4 = torch.zeros((p.SKY_H, p.SKY_W, 4), device="cuda")
    eng.pass_irradiance1(p, o, f4(T), dE4)
    _assert("dE1", rgb(dE4), O.irradiance1(p, T, o), 1e-5)
    shape = (p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 4)
    dSR4, dSM4, S4 = (torch.zeros(shape, device="cuda") for _ in range(3))
    eng.pass_inscatter1(p, o, f4(T), dSR4, dSM4, S4)
    oR, oM = O.inscatter1(p, T, o)
    _assert("dSR1", rgb(dSR4), oR, 1e-4)
    _assert("dSM1", rgb(dSM4), oM, 1e-4)
    S = S4.double().cpu().numpy()
    assert np.array_equal(S[..., :3], dSR4[..., :3].double().cpu().numpy())
    assert np.array_equal(S[..., 3], dSM4[..., 0].double().cpu().numpy())


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("first", [True, False])
def test_pass_inscatterS_irradianceN_inscatterN(eng, ci, first):
    p = CASES[ci]
    o = options()
    st = O.precompute_stages(p, options(orders=3))
    T = st["T"]
    if first:
        dE, dSR, dSM = st["dE1"], st["dSR1"], st["dSM1"]
    else:
        dE, dSR, dSM = st["dE2"], st["dS2"], st["dSM1"]
    dJ4 = torch.zeros((p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 4), device="cuda")
    eng.pass_inscatterS(p, o, f4(T), f4(dE), f4(dSR), f4(dSM), first, dJ4)
    oJ = O.inscatterS(p, T, dE, dSR, dSM, first, o)
    _assert("dJ", rgb(dJ4), oJ, 2e-4)
    dEn4 = torch.zeros((p.SKY_H, p.SKY_W, 4), device="cuda")
    E4 = f4(np.full((p.SKY_H, p.SKY_W, 3), 0.25))
    eng.pass_irradianceN(p, o, f4(dSR), f4(dSM), first, dEn4, E4)
    oE = O.irradianceN(p, dSR, dSM, first, o)
    _assert("dEn", rgb(dEn4), oE, 2e-4)
    assert np.allclose(rgb(E4), 0.25 + rgb(dEn4), rtol=1e-6, atol=0)
    S0 = np.random.default_rng(ci).random((p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 4))
    S4 = torch.from_numpy(S0.astype(np.float32)).cuda()
    dS4 = torch.zeros_like(S4)
    eng.pass_inscatterN(p, o, f4(T), f4(oJ), dS4, S4)
    oS = O.inscatterN(p, T, oJ, o)
    _assert("dSn", rgb(dS4), oS, 2e-4)
    oAcc = O.copy_inscatterN(p, oS, S0.astype(np.float32).astype(np.float64))
    _assert("S+=", S4.double().cpu().numpy(), oAcc, 2e-4)


# ---------------------------------------------------------------- whole precompute
def _chain(eng, p, o=None):
    (T, E, S), = eng.precompute(p, o or options())
    torch.cuda.synchronize()
    return T.double().cpu().numpy(), E.double().cpu().numpy(), S.double().cpu().numpy()


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_chain_matches_oracle(eng, ci):
    p = CASES[ci]
    T, E, S = _chain(eng, p)
    oT, oE, oS = O.precompute(p, options())
    _assert("T", T, oT, TOL)
    _assert("E", E, oE, TOL)
    _assert("S", S, oS, TOL)


@pytest.mark.parametrize("ci", [0, 2, 3])
def test_half_storage_matches_oracle(eng, ci):
    """ATMO_FLAG_HALF_STORAGE (the reference's RGB16F / RGBA16F tables and FP16
    ADD blending) against the oracle's half-storage chain.  Every GPU value is
    an FP16 value; a value can differ where the FP32 result before a rounding
    and the FP64 one fall on either side of a rounding boundary.  Most such
    flips move a table by one FP16 step (up to 9.8e-4 of the value), but most
    dJ values are FP16 subnormals (steps of 6e-8 on values of 1e-7 .. 6e-5),
    and a flipped one moves the ray integrals that gather it by up to a few
    percent.  Bar: T and E bit-equal on at least 99% of the values and within
    4 FP16 steps; S bit-equal on at least 99% of the values and within 1e-2
    (relative, floored at 1e-4 of the table maximum)."""
    p = CASES[ci]
    o = options(flags=FLAG_HALF_STORAGE)
    T, E, S = _chain(eng, p, o)
    oT, oE, oS = O.precompute(p, o)
    for name, x, y in (("T16", T, oT), ("E16", E, oE), ("S16", S, oS)):
        assert np.array_equal(x, x.astype(np.float16).astype(np.float64)), name
        eq = float(np.mean(x == y))
        ulp = np.spacing(np.abs(y).astype(np.float16)).astype(np.float64)
        steps = float(np.max(np.abs(x - y) / ulp))
        sm = summary(x, y)
        print("%s bit-equal %.4f, max %.1f FP16 steps, floored rel err max %.3e mean %.3e"
              % (name, eq, steps, sm["max"], sm["mean"]), flush=True)
        assert eq >= 0.99, (name, eq)
        if name == "S16":
            assert sm["max"] <= 1e-2, (name, sm)
        else:
            assert steps <= 4.0, (name, steps)


@pytest.mark.parametrize("which", ["transmittance", "inscatter", "both"])
def test_linear_parameterisations_match_oracle(eng, which):
    """The reference built without TRANSMITTANCE_NON_LINEAR and/or
    INSCATTER_NON_LINEAR (common.glsl:105-106): ATMO_FLAG_LINEAR_* against the
    oracle's linear branches, whole chain, both k_inscatterN forms."""
    flags = {"transmittance": FLAG_LINEAR_TRANSMITTANCE, "inscatter": FLAG_LINEAR_INSCATTER,
             "both": FLAG_LINEAR_TRANSMITTANCE | FLAG_LINEAR_INSCATTER}[which]
    for ci in (0, 2):
        p = CASES[ci]
        oT, oE, oS = O.precompute(p, options(flags=flags))
        for form in (1, 2):
            T, E, S = _chain(eng, p, options(flags=flags | 8, inscatterN_form=form))
            _assert("T", T, oT, TOL)
            _assert("E", E, oE, TOL)
            _assert("S", S, oS, TOL)


def test_half_storage_is_part_of_the_plan(eng):
    """The same parameters alternately with and without FP16 storage: each call
    gets the tables of its own mode (the cached plan and graph are keyed on it)."""
    p = small()
    a = _chain(eng, p)
    b = _chain(eng, p, options(flags=FLAG_HALF_STORAGE))
    c = _chain(eng, p)
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x, z)
        assert np.array_equal(y, y.astype(np.float16).astype(np.float64))
    assert not np.array_equal(a[2], b[2])


def _extreme_sets():
    """Corners of the batch distribution: thin and thick atmospheres, dense
    grey haze with a sharp Mie peak, weak Rayleigh, a small planet."""
    out = []
    for (Rg, dR, HR, sR, HM, bM, g) in ((3000.0, 40.0, 5.0, 3.0, 0.6, 3e-2, 0.9),
                                        (7000.0, 120.0, 12.0, 0.3, 2.4, 1e-3, 0.6),
                                        (6360.0, 60.0, 8.0, 1.0, 1.2, 3e-2, 0.9)):
        q = small()
        q.Rg, q.Rt, q.RL = Rg, Rg + dR, Rg + dR + 1.0
        q.HR, q.HM, q.mieG = HR, HM, g
        q.betaR.x, q.betaR.y, q.betaR.z = 5.8e-3 * sR, 1.35e-2 * sR, 3.31e-2 * sR
        q.betaMSca.x = q.betaMSca.y = q.betaMSca.z = bM
        q.betaMEx.x = q.betaMEx.y = q.betaMEx.z = bM / 0.9
        out.append(q)
    return out


def test_batch_of_seeded_and_extreme_atmospheres_matches_oracle(eng):
    """Eight sets of the batch distribution (seeded, a stream of their own) and
    three corners of it, in one batch at 8x32x8x4: every table of every set
    within the 1e-3 bar of the oracle."""
    sets = seeded(8, (8, 32, 8, 4), seed=SEED_BATCH + 17) + _extreme_sets()
    outs = eng.precompute(sets, options())
    torch.cuda.synchronize()
    for i, (p, (T, E, S)) in enumerate(zip(sets, outs)):
        oT, oE, oS = O.precompute(p, options())
        _assert("T[%d]" % i, T.double().cpu().numpy(), oT, TOL)
        _assert("E[%d]" % i, E.double().cpu().numpy(), oE, TOL)
        _assert("S[%d]" % i, S.double().cpu().numpy(), oS, TOL)


def _random_case(rng):
    """A seeded atmosphere of the batch distribution at a random small
    resolution (odd sizes included) with random options."""
    R = int(rng.integers(2, 7))
    MU = 2 * int(rng.integers(2, 13))
    MUS = int(rng.integers(2, 9))
    NU = int(rng.integers(2, 9))
    p = seeded(1, (R, MU, MUS, NU), seed=int(rng.integers(1, 2 ** 31)))[0]
    p.TRANSMITTANCE_W, p.TRANSMITTANCE_H = int(rng.integers(2, 80)), int(rng.integers(2, 24))
    p.SKY_W, p.SKY_H = int(rng.integers(2, 24)), int(rng.integers(2, 10))
    o = options(orders=int(rng.integers(1, 6)), transmittance_samples=int(rng.integers(20, 600)),
                inscatter_samples=int(rng.integers(5, 80)),
                irradiance_samples=2 * int(rng.integers(2, 24)),
                spherical_samples=int(rng.integers(4, 24)),
                flags=int(rng.choice([0, FLAG_LINEAR_TRANSMITTANCE, FLAG_LINEAR_INSCATTER])))
    if (o.flags & FLAG_LINEAR_INSCATTER) and (p.RES_MU_S - 1) % 6 == 0:
        # the linear mu_s grid -0.2 + 1.2 m/(RES_MU_S - 1) then has a column at
        # mu_s = 0: the sun on the horizon, where the sun test of a ray's ground
        # sample (inscatter1.glsl:107-110) compares 0 with 0 and rounding decides
        p.RES_MU_S += 1
    return p, o


@pytest.mark.parametrize("seed", range(4))
def test_random_resolutions_and_options_match_oracle(eng, seed):
    """Six random cases per seed (resolutions, table sizes, orders, sample
    counts, parameterisation), each run alone and compared with the oracle."""
    rng = np.random.default_rng(1000 + seed)
    for case in range(6):
        p, o = _random_case(rng)
        print("case", seed, case, (p.RES_R, p.RES_MU, p.RES_MU_S, p.RES_NU),
              (p.TRANSMITTANCE_W, p.TRANSMITTANCE_H, p.SKY_W, p.SKY_H), o.orders,
              o.transmittance_samples, o.inscatter_samples, o.irradiance_samples,
              o.spherical_samples, o.flags, flush=True)
        T, E, S = _chain(eng, p, o)
        oT, oE, oS = O.precompute(p, o)
        _assert("T", T, oT, TOL)
        _assert("E", E, oE, TOL)
        _assert("S", S, oS, TOL)


@pytest.mark.parametrize("seed", range(3))
def test_random_mixed_batches_match_oracle(eng, seed):
    """Batches of five random sets of different resolutions sharing one random
    options draw (seed 2: FP16 storage as well), every set against the oracle."""
    rng = np.random.default_rng(2000 + seed)
    cases = [_random_case(rng) for _ in range(5)]
    o = cases[0][1]
    if seed == 2:
        o.flags |= FLAG_HALF_STORAGE
    sets = [c[0] for c in cases]
    if o.flags & FLAG_LINEAR_INSCATTER:
        for p in sets:
            if (p.RES_MU_S - 1) % 6 == 0:
                p.RES_MU_S += 1   # no exact-horizon column (see _random_case)
    outs = eng.precompute(sets, o)
    torch.cuda.synchronize()
    for i, (p, (T, E, S)) in enumerate(zip(sets, outs)):
        oT, oE, oS = O.precompute(p, o)
        tol = 1e-2 if o.flags & FLAG_HALF_STORAGE else TOL   # FP16 bar of test_half_storage_matches_oracle
        _assert("T[%d]" % i, T.double().cpu().numpy(), oT, tol)
        _assert("E[%d]" % i, E.double().cpu().numpy(), oE, tol)
        _assert("S[%d]" % i, S.double().cpu().numpy(), oS, tol)


def _dense(Rg, dR, dRL, HR, HM, sR, bM, g, alpha=0.1):
    q = small(res=(6, 16, 6, 4), T=(48, 12), E=(12, 6))
    q.Rg, q.Rt, q.RL = Rg, Rg + dR, Rg + dR + dRL
    q.HR, q.HM, q.mieG, q.AVERAGE_GROUND_REFLECTANCE = HR, HM, g, alpha
    q.betaR.x, q.betaR.y, q.betaR.z = 5.8e-3 * sR, 1.35e-2 * sR, 3.31e-2 * sR
    q.betaMSca.x = q.betaMSca.y = q.betaMSca.z = bM
    q.betaMEx.x = q.betaMEx.y = q.betaMEx.z = bM / 0.9
    return q


@pytest.mark.parametrize("which", ["large_planet_dense_haze", "thick_mie", "deep_rayleigh",
                                   "beyond_fp64_range"])
def test_transmittance_far_below_fp32_range(eng, which):
    """Atmospheres whose grazing transmittances fall far below FP32's range
    (e^-300 and beyond): the segment ratios T(x)/T(y) come from the
    optical-depth table in log space, so the tables stay within the 1e-3 bar of
    the FP64 oracle (with plain FP32 quotients these cases were off by 1e2-1e3).
    In the last case both lookups of some ratios fall below FP64's range too
    (optical depths above 745): the oracle's chain also takes its ratios from
    optical depths there, where the quotient of lookups is 0/0."""
    p = {"large_planet_dense_haze": _dense(59594.8, 586.9, 4.6, 48.2, 1.25, 6.3, 1.6e-2, -0.3),
         "thick_mie": _dense(51367.6, 316.8, 5.5, 0.14, 19.75, 0.164, 0.44, -0.33),
         "deep_rayleigh": _dense(8290.8, 1.29, 2.4, 1.97, 23.9, 55.4, 0.11, 0.57),
         "beyond_fp64_range": _dense(46427.34375, 64.5234375, 0.01171875, 42.71, 13.96, 14.77,
                                     3.165e-4, 0.72, alpha=0.39)}[which]
    T, E, S = _chain(eng, p)
    oT, oE, oS = O.precompute(p, options())
    assert oT[oT > 0].min() < 1e-100   # the regime this test is about
    _assert("T", T, oT, TOL)
    _assert("E", E, oE, TOL)
    _assert("S", S, oS, TOL)


def _wide_set(rng):
    """An atmosphere far outside Earth's neighbourhood: Rg 100-70 000, thickness
    1-1000, RL - Rt 1e-3-100 (log-uniform), H_R 0.1-100, H_M 0.1-30, beta_R
    0.01-100x Earth, grey beta_M 1e-4-1 per unit length, g -0.95-0.99."""
    p = small(res=(6, 16, 6, 4), T=(48, 12), E=(12, 6))
    Rg = float(rng.uniform(100, 70000))
    p.Rg, p.Rt = Rg, Rg + float(10 ** rng.uniform(0, 3))
    p.RL = p.Rt + float(10 ** rng.uniform(-3, 2))
    p.HR = float(10 ** rng.uniform(-1, 2))
    p.HM = float(10 ** rng.uniform(-1, 1.5))
    k = float(10 ** rng.uniform(-2, 2))
    p.betaR.x, p.betaR.y, p.betaR.z = 5.8e-3 * k, 1.35e-2 * k, 3.31e-2 * k
    b = float(10 ** rng.uniform(-4, 0))
    p.betaMSca.x = p.betaMSca.y = p.betaMSca.z = b
    p.betaMEx.x = p.betaMEx.y = p.betaMEx.z = b / 0.9
    p.mieG = float(rng.uniform(-0.95, 0.99))
    p.AVERAGE_GROUND_REFLECTANCE = float(rng.uniform(0, 1))
    return p


def test_wide_atmosphere_distribution_in_one_batch(eng):
    """24 seeded atmospheres of a wide distribution (grazing paths through
    dense haze, atmospheres a few km thick on planets of 1e4 km, scale heights
    from 0.1 to 100) in one batch: every table within 2e-4 of the oracle.
    Grazing view rays are the sensitive case there: their segment ratios
    divide two lookups of nearly equal, large optical depth (with FP32 lookup
    coordinates two of these sets were off by 2.4e-3 and 3.7e-3).  A set whose
    top layer's FP32 radius equals Rt makes getMuMuSNu divide 0 by 0 on its
    top layer (PA/PreprocessAtmo.cpp:255-271 with inscatter1.glsl's
    getMuMuSNu): NaN texels in the reference's own tables, reproduced at the
    same texels."""
    rng = np.random.default_rng(77)
    sets = [_wide_set(rng) for _ in range(24)]
    outs = eng.precompute(sets, options())
    torch.cuda.synchronize()
    nan_sets = []
    for i, (p, (T, E, S)) in enumerate(zip(sets, outs)):
        oT, oE, oS = O.precompute(p, options())
        _assert("T[%d]" % i, T.double().cpu().numpy(), oT, 2e-4)
        if np.isnan(oS).any():
            # the NaN texels reach every layer through the later orders: both
            # paths hold them in E and S
            assert np.isnan(S.double().cpu().numpy()).any() and np.isnan(E.double().cpu().numpy()).any()
            nan_sets.append(i)
            continue
        _assert("E[%d]" % i, E.double().cpu().numpy(), oE, 2e-4)
        _assert("S[%d]" % i, S.double().cpu().numpy(), oS, 2e-4)
    assert nan_sets == [2]
    # single scattering of that set: NaN exactly at the oracle's texels (its
    # top layer), everything else within the bar
    p = sets[2]
    (T, E, S), = eng.precompute(p, options(orders=1))
    torch.cuda.synchronize()
    g = S.double().cpu().numpy()
    oS = O.precompute(p, options(orders=1))[2]
    nan = np.isnan(oS)
    assert nan.any() and np.array_equal(np.isnan(g), nan)
    assert np.all(np.argwhere(nan)[:, 0] == p.RES_R - 1)
    _assert("S1[2]", np.where(nan, 0.0, g), np.where(nan, 0.0, oS), 2e-4)


def test_wide_atmosphere_distribution_linear_parameterisations(eng):
    """The 24 sets of the wide distribution with both linear parameterisations
    (common.glsl:105-106 undefined), in one batch: every table of every set
    within 2e-4 of the oracle.  (The linear getMuMuSNu takes mu from the row
    index, :162-166, so the set whose top layer is NaN above is finite here.)"""
    rng = np.random.default_rng(77)
    sets = [_wide_set(rng) for _ in range(24)]
    o = options(flags=FLAG_LINEAR_TRANSMITTANCE | FLAG_LINEAR_INSCATTER)
    outs = eng.precompute(sets, o)
    torch.cuda.synchronize()
    for i, (p, (T, E, S)) in enumerate(zip(sets, outs)):
        oT, oE, oS = O.precompute(p, o)
        _assert("T[%d]" % i, T.double().cpu().numpy(), oT, 2e-4)
        _assert("E[%d]" % i, E.double().cpu().numpy(), oE, 2e-4)
        _assert("S[%d]" % i, S.double().cpu().numpy(), oS, 2e-4)


def _wider_set(rng):
    """A still wider distribution: Rg 100-1e5, thickness 0.3-2000, RL - Rt
    1e-3-300, H_R 0.05-200, H_M 0.05-100, beta_R 1e-3-300x Earth, grey
    beta_M 1e-5-3 with beta_MEx = beta_M / U[0.5, 1], g -0.99-0.99."""
    p = small(res=(8, 32, 8, 4))
    Rg = float(10 ** rng.uniform(2, 5))
    p.Rg, p.Rt = Rg, Rg + float(10 ** rng.uniform(-0.5, 3.3))
    p.RL = p.Rt + float(10 ** rng.uniform(-3, 2.5))
    p.HR = float(10 ** rng.uniform(-1.3, 2.3))
    p.HM = float(10 ** rng.uniform(-1.3, 2))
    k = float(10 ** rng.uniform(-3, 2.5))
    p.betaR.x, p.betaR.y, p.betaR.z = 5.8e-3 * k, 1.35e-2 * k, 3.31e-2 * k
    b = float(10 ** rng.uniform(-5, 0.5))
    p.betaMSca.x = p.betaMSca.y = p.betaMSca.z = b
    p.betaMEx.x = p.betaMEx.y = p.betaMEx.z = b / float(rng.uniform(0.5, 1.0))
    p.mieG = float(rng.uniform(-0.99, 0.99))
    p.AVERAGE_GROUND_REFLECTANCE = float(rng.uniform(0, 1))
    return p


def test_wider_distribution_worst_sets_in_one_batch(eng):
    """16 sets of a still wider seeded distribution at 8x32x8x4, in one batch:
    its first 12 and the 4 worst of its first 96 (haze scale heights of 0.06
    over an atmosphere 1 km thick, Rayleigh at 200x Earth over 0.06 km, Mie
    at 1.3 per unit length on a 550 km planet): within the 1e-3 bar (the worst
    is 4.2e-4)."""
    rng = np.random.default_rng(78)
    every = [_wider_set(rng) for _ in range(96)]
    sets = every[:12] + [every[i] for i in (21, 44, 64, 85)]
    outs = eng.precompute(sets, options())
    torch.cuda.synchronize()
    for i, (p, (T, E, S)) in enumerate(zip(sets, outs)):
        oT, oE, oS = O.precompute(p, options())
        _assert("T[%d]" % i, T.double().cpu().numpy(), oT, TOL)
        _assert("E[%d]" % i, E.double().cpu().numpy(), oE, TOL)
        _assert("S[%d]" % i, S.double().cpu().numpy(), oS, TOL)


def test_chain_orders_1_and_6(eng):
    p = small()
    for n in (1, 6):
        o = options(orders=n)
        T, E, S = _chain(eng, p, o)
        oT, oE, oS = O.precompute(p, o)
        _assert("S(o%d)" % n, S, oS, TOL)
        if n == 1:
            assert np.all(E == 0)
        else:
            _assert("E(o%d)" % n, E, oE, TOL)


def test_chain_finer_sampling_options(eng):
    p = small(res=(6, 16, 6, 4), T=(32, 8), E=(8, 4))
    o = options(transmittance_samples=700, inscatter_samples=64, irradiance_samples=16,
                spherical_samples=24, orders=3)
    T, E, S = _chain(eng, p, o)
    oT, oE, oS = O.precompute(p, o)
    _assert("T", T, oT, TOL)
    _assert("E", E, oE, TOL)
    _assert("S", S, oS, TOL)


def _params(v):
    p = AtmoParams()
    import ctypes
    ctypes.memmove(ctypes.addressof(p), bytes(v), 96)
    return p


@pytest.mark.parametrize("case", ["spherical64", "maximum"])
def test_direction_grids_staged_in_rounds(eng, case):
    """Direction grids whose LDS staging exceeds one round: k_inscatterS and
    k_irradianceN then stage their theta rows in chunks.  spherical64: 64 x 128
    directions with RES_NU 16 (order 2 in two rounds); maximum: the largest
    grids and RES_NU the options allow (128 x 256 spherical directions, 256 x
    512 hemisphere directions, RES_NU 64: up to 13 rounds)."""
    if case == "spherical64":
        p = small(res=(4, 16, 4, 16))
        o = options(spherical_samples=64)
    else:
        p = small(res=(2, 8, 2, 64))
        o = options(spherical_samples=128, irradiance_samples=256)
    T, E, S = _chain(eng, p, o)
    oT, oE, oS = O.precompute(p, o)
    _assert("E", E, oE, TOL)
    _assert("S", S, oS, TOL)


def test_smallest_tables_the_abi_accepts(eng):
    """RES_R 2, RES_MU 4, RES_MU_S 2, RES_NU 2, transmittance and irradiance
    2x2: every (size - 1) divisor of the parameterisations is 1."""
    p = small(res=(2, 4, 2, 2), T=(2, 2), E=(2, 2))
    T, E, S = _chain(eng, p)
    oT, oE, oS = O.precompute(p, options())
    _assert("T", T, oT, TOL)
    _assert("E", E, oE, TOL)
    _assert("S", S, oS, TOL)


def test_largest_tables_the_abi_accepts(eng):
    """RES_R x RES_MU x RES_MU_S x RES_NU = 128 x 1024 x 32 x 32 = 2^27 texels,
    the largest 4D table the ABI accepts (RES_MU_S * RES_NU at its limit of
    1024 too), T 512x128, E 128x32.  Too large for the oracle, so:
    * its corner texels (layers 0 and R-1, rows 0, MU/2-1, MU/2, MU-1, mu_s
      columns 0 and M-1, nu slices 0 and N-1) have the (r, mu, mu_s, nu) of the
      2x4x2x2 table's texels, so their single scattering equals the oracle's
      at that size (1e-4);
    * with 4 orders every table passes the numeric check (finite, T in [0, 1],
      E and S >= 0), T is bit-equal to T of a small run with the same
      transmittance size (the T pass does not depend on the 4D table), and a
      second run gives the same bits."""
    res = (128, 1024, 32, 32)
    p = small(res=res, T=(512, 128), E=(128, 32))
    q = small(res=(2, 4, 2, 2), T=(512, 128), E=(128, 32))
    (T1, E1, S1), = eng.precompute(p, options(orders=1, flags=FLAG_CHECK_NUMERICS))
    torch.cuda.synchronize()
    R, MU, M, N = res
    corners = S1[[0, R - 1]][:, [0, MU // 2 - 1, MU // 2, MU - 1]]
    corners = corners.reshape(2, 4, N, M, 4)[:, :, [0, N - 1]][:, :, :, [0, M - 1]]
    oS1 = O.precompute(q, options(orders=1))[2]
    _assert("S1 corners", corners.reshape(2, 4, 4, 4).double().cpu().numpy(), oS1, 1e-4)
    del S1, corners
    (T, E, S), = eng.precompute(p, options(flags=FLAG_CHECK_NUMERICS))
    torch.cuda.synchronize()
    (Tq, _, _), = eng.precompute(q, options())
    assert torch.equal(T, Tq)
    S0 = S.clone()
    eng.precompute(p, options(), outs=[(T, E, S)])
    torch.cuda.synchronize()
    assert torch.equal(S, S0)


def test_first_and_later_calls_return_the_same_bits():
    """A shape class not tuned before in the process: the first call (which
    runs, tunes, and runs again with the chosen variants), a replay of the
    tuned graph on the same engine, and a new engine (which applies the cached
    tuning before its first launch) return the same tables bit for bit."""
    p = small(res=(6, 24, 6, 4), T=(32, 8), E=(8, 4))
    a = L.Engine(0)
    first = [t.clone() for t in a.precompute(p, options())[0]]
    second = a.precompute(p, options())[0]
    torch.cuda.synchronize()
    assert a.tuning()["tuned"] == 1
    a.close()
    b = L.Engine(0)
    third = b.precompute(p, options())[0]
    torch.cuda.synchronize()
    b.close()
    for x, y, z in zip(first, second, third):
        assert torch.equal(x, y) and torch.equal(x, z)


def test_golden_tiny_fixtures(eng):
    g = np.load(os.path.join(GOLDEN, "tiny_earth.npz"))
    p = _params(g["params"])
    T, E, S = _chain(eng, p)
    _assert("T", T, g["T"], TOL)
    _assert("E", E, g["E"], TOL)
    _assert("S", S, g["S"], TOL)
    g = np.load(os.path.join(GOLDEN, "tiny_seeds.npz"))
    for i in range(3):
        p = _params(g["params%d" % i])
        T, E, S = _chain(eng, p)
        _assert("T%d" % i, T, g["T%d" % i], TOL)
        _assert("E%d" % i, E, g["E%d" % i], TOL)
        _assert("S%d" % i, S, g["S%d" % i], TOL)


def test_reference_configuration_against_golden_samples(eng):
    """The reference configuration (Earth, 32x128x32x8, 4 orders) against the
    FP64 oracle: T and E in full, S at 4096 seeded texels."""
    g = np.load(os.path.join(GOLDEN, "earth_ref.npz"))
    p = _params(g["params"])
    assert bytes(p) == bytes(earth())
    T, E, S = _chain(eng, p)
    _assert("T", T, g["T"], TOL)
    _assert("E", E, g["E"], TOL)
    flat = S.reshape(-1, 4)[g["S_idx"]]
    # floor relative to the whole table's scale
    scale = np.abs(S).max()
    den = np.maximum(np.abs(g["S_val"]), 1e-4 * scale)
    err = np.abs(flat - g["S_val"]) / den
    print("S samples max %.3e mean %.3e" % (err.max(), err.mean()), flush=True)
    assert err.max() <= TOL


def test_reference_configuration_half_storage_against_golden_samples(eng):
    """The reference configuration with ATMO_FLAG_HALF_STORAGE against the
    oracle's FP16-storage tables: T and E in full, S at 4096 seeded texels.
    Bar as in test_half_storage_matches_oracle: at least 99% of the values
    bit-equal, T and E within 4 FP16 steps, S within 1e-2 (floored)."""
    g = np.load(os.path.join(GOLDEN, "earth_ref_half.npz"))
    p = _params(g["params"])
    assert bytes(p) == bytes(earth())
    T, E, S = _chain(eng, p, options(flags=FLAG_HALF_STORAGE))
    flat = S.reshape(-1, 4)[g["S_idx"]]
    for name, x, y in (("T16", T, g["T"].astype(np.float64)), ("E16", E, g["E"].astype(np.float64)),
                       ("S16", flat, g["S_val"].astype(np.float64))):
        eq = float(np.mean(x == y))
        ulp = np.spacing(np.abs(y).astype(np.float16)).astype(np.float64)
        steps = float(np.max(np.abs(x - y) / ulp))
        scale = np.abs(S).max() if name == "S16" else np.abs(y).max()
        err = np.abs(x - y) / np.maximum(np.abs(y), 1e-4 * scale)
        print("%s bit-equal %.4f, max %.1f FP16 steps, floored rel err max %.3e"
              % (name, eq, steps, err.max()), flush=True)
        assert eq >= 0.99, (name, eq)
        if name == "S16":
            assert err.max() <= 1e-2
        else:
            assert steps <= 4.0, (name, steps)


# ---------------------------------------------------------------- batching and graphs
@pytest.mark.parametrize("G,form", [(0, 0), (4, 3)])
def test_mixed_resolution_batch_equals_single_runs(eng, G, form):
    """A set's tables do not depend on the batch it runs in: default variants,
    and the LDS-staged k_inscatterN form with 4 texels per thread."""
    sets = seeded(3, None)
    for p in sets:
        p.RES_R, p.RES_MU, p.RES_MU_S, p.RES_NU = p.RES_R // 2, p.RES_MU // 2, p.RES_MU_S // 2, max(2, p.RES_NU // 2)
    o = options(flags=FLAG_NO_AUTOTUNE, texels_per_thread=G, inscatterN_form=form)
    outs = eng.precompute(sets, o)
    torch.cuda.synchronize()
    for p, (T, E, S) in zip(sets, outs):
        T1, E1, S1 = _chain(eng, p, o)
        assert np.array_equal(T.double().cpu().numpy(), T1)
        assert np.array_equal(E.double().cpu().numpy(), E1)
        assert np.array_equal(S.double().cpu().numpy(), S1)


@pytest.mark.parametrize("flags", [8, 0])
def test_two_engines_on_two_threads(flags):
    """Two engines driven from two host threads at once, each on a stream of its
    own (the C ABI releases the GIL): per-engine plans, capture streams and
    graphs, and the process-wide tuning cache under its lock.  Default variants
    (flags 8) and autotuned ones: bit-identical to a sequential run."""
    import threading
    cases = {"a": small(), "b": small(res=(6, 16, 6, 4), T=(32, 8), E=(8, 4))}
    ref = {}
    for name, p in cases.items():
        e = L.Engine(0)
        ref[name] = [t.double().cpu().numpy() for t in e.precompute(p, options(flags=8))[0]]
        torch.cuda.synchronize()
        e.close()
    res, errs = {}, []

    def work(name, p):
        try:
            e = L.Engine(0)
            st = torch.cuda.Stream()
            outs = L.alloc_outputs(p)
            for _ in range(3):
                e.precompute(p, options(flags=flags), outs=[outs], stream=st)
                st.synchronize()
            res[name] = [t.double().cpu().numpy() for t in outs]
            e.close()
        except Exception as ex:   # reported by the main thread
            errs.append((name, ex))

    th = [threading.Thread(target=work, args=(n, p)) for n, p in cases.items()]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th)
    assert not errs, errs
    for name in cases:
        for x, y in zip(res[name], ref[name]):
            assert np.array_equal(x, y), name


def test_graph_replays_match_eager(eng):
    p = small()
    eager = _chain(eng, p, options(flags=2))
    outs = L.alloc_outputs(p)
    for rep in range(3):
        for t in outs:
            t.fill_(-7.0)
        torch.cuda.synchronize()
        eng.precompute(p, options(), outs=[outs])
        torch.cuda.synchronize()
        x = tuple(t.double().cpu().numpy() for t in outs)
        torch.cuda.synchronize()
        for a, b in zip(x, eager):
            assert np.array_equal(a, b), rep


def test_numeric_check_flag_passes_on_valid_tables(eng):
    p = small()
    eng.precompute(p, options(flags=1))


def test_preprocess_writes_reference_files(eng, tmp_path):
    p = small()
    d = str(tmp_path)
    L.preprocess(p, d)
    # the one-off precompute behind preprocessAtmo runs the default variants
    T, E, S = _chain(eng, p, options(flags=FLAG_NO_AUTOTUNE))
    for name, ref, shape in (("transmittance", T, [p.TRANSMITTANCE_W, p.TRANSMITTANCE_H, 0, 3]),
                             ("irradiance", E, [p.SKY_W, p.SKY_H, 0, 3]),
                             ("inscatter", S, [p.RES_MU_S * p.RES_NU, p.RES_MU * p.RES_R, p.RES_R, 4])):
        data, h = L.read_raw(os.path.join(d, name + ".raw"))
        assert list(h[1:]) == shape
        assert np.array_equal(data, ref.astype(np.float32).ravel())
    m = os.path.getmtime(os.path.join(d, "inscatter.raw"))
    L.preprocess(p, d)  # cached: returns without recomputing
    assert os.path.getmtime(os.path.join(d, "inscatter.raw")) == m


@pytest.mark.parametrize("res", [(8, 32, 16, 8), (4, 16, 32, 16)])
def test_autotuned_variants_agree_with_the_defaults(res):
    """The tuned plan (any k_inscatterS variant, direct or LDS-staged
    k_inscatterN) gives the default plan's tables bit for bit, within the
    oracle bar."""
    p = small(res=res)
    e1 = L.Engine(0)
    ref = e1.precompute(p, options(flags=8))  # default variants, no tuning
    torch.cuda.synchronize()
    e2 = L.Engine(0)
    outs = e2.precompute(p, options())       # first run, then the autotuner
    outs = e2.precompute(p, options(), outs=outs)  # tuned variants, re-captured graph
    torch.cuda.synchronize()
    t = e2.tuning()
    print(t, flush=True)
    assert t["tuned"] == 1 and t["inscatterS_texels_per_thread"] in (0, 2, 3, 4, 8)
    assert t["inscatterN_rows_per_workgroup"] in (1, 2, 4)
    assert t["inscatterN_staged_samples"] in (0, 1, 2)
    for a, b in zip(ref[0], outs[0]):
        assert torch.equal(a, b)
    oT, oE, oS = O.precompute(p, options())
    _assert("S", outs[0][2].double().cpu().numpy(), oS, TOL)
    e1.close()
    e2.close()


def test_per_set_inscatterS_variants_in_one_batch(eng):
    """Untuned, a batch runs k_inscatterS with each set's own texels per thread
    (3 for RES_MU 64 x RES_NU 12, 2 for 16 x 4, 4 for 128 x 8), one launch per
    group of sets; every set's tables equal its single run with G pinned to 4."""
    sets = []
    for res in ((4, 64, 6, 12), (4, 16, 6, 4), (4, 128, 4, 8), (4, 64, 6, 12)):
        q = small(res=res)
        sets.append(q)
    outs = eng.precompute(sets, options(flags=8))
    torch.cuda.synchronize()
    for q, o in zip(sets, outs):
        one = eng.precompute(q, options(flags=8, texels_per_thread=4))[0]
        torch.cuda.synchronize()
        for a, b in zip(o, one):
            assert torch.equal(a, b)


@pytest.mark.parametrize("res", [(8, 32, 16, 8), (6, 16, 6, 5)])
def test_every_inscatterS_variant_gives_identical_tables(eng, res):
    """texels_per_thread 2, 3, 4 and 8 pin the k_inscatterS variant; the tables
    must agree bit for bit (RES_NU = 5 leaves ragged texel groups)."""
    p = small(res=res)
    outs = [eng.precompute(p, options(flags=8, texels_per_thread=g))[0] for g in (2, 3, 4, 8)]
    torch.cuda.synchronize()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert torch.equal(a, b)
    T, E, S = O.precompute(p, options())
    _assert("S", outs[0][2].double().cpu().numpy(), S, TOL)


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("form", [2, 3])
def test_pass_inscatterN_staged_form(eng, ci, form):
    """The LDS-staged k_inscatterN form (2 or 1 samples per LDS round) against
    the oracle, pass level."""
    p = CASES[ci]
    o = options(inscatterN_form=form)
    st = O.precompute_stages(p, options(orders=3))
    T, oJ = st["T"], st["dJ2"]
    S0 = np.random.default_rng(ci).random((p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 4))
    S4 = torch.from_numpy(S0.astype(np.float32)).cuda()
    dS4 = torch.zeros_like(S4)
    eng.pass_inscatterN(p, o, f4(T), f4(oJ), dS4, S4)
    oS = O.inscatterN(p, T, oJ, options())
    _assert("dSn", rgb(dS4), oS, 2e-4)
    _assert("S+=", S4.double().cpu().numpy(), O.copy_inscatterN(p, oS, S0.astype(np.float32).astype(np.float64)),
            2e-4)


def test_inscatterN_forms_agree_on_the_whole_chain(eng):
    p = small(res=(8, 32, 16, 8))
    a = eng.precompute(p, options(flags=8, inscatterN_form=1))[0]
    b = eng.precompute(p, options(flags=8, inscatterN_form=2))[0]
    c = eng.precompute(p, options(flags=8, inscatterN_form=3))[0]
    torch.cuda.synchronize()
    # direct gathers and LDS-staged rows (rounds of 1 or 2 samples): the same
    # column blends in the same order
    for x, y, z in zip(a, b, c):
        assert torch.equal(x, y) and torch.equal(y, z)


@pytest.mark.parametrize("lin", [0, FLAG_LINEAR_INSCATTER])
@pytest.mark.parametrize("res", [(4, 16, 32, 12), (4, 16, 48, 12), (4, 16, 64, 16)])
def test_inscatterN_forms_agree_on_rows_of_more_than_256_texels(eng, res, lin):
    """Rows of 384, 576 and 1024 texels (one mu row per workgroup, 2, 3 and 4
    texels per thread), where the LDS-staged form blends all of a thread's
    texels per sample at once: the same bits as the direct gathers, with the
    non-linear and the linear inscatter parameterisation."""
    p = small(res=res)
    o = dict(flags=8 | lin, orders=4)
    a = eng.precompute(p, options(inscatterN_form=1, **o))[0]
    b = eng.precompute(p, options(inscatterN_form=2, **o))[0]
    c = eng.precompute(p, options(inscatterN_form=3, **o))[0]
    torch.cuda.synchronize()
    for x, y, z in zip(a, b, c):
        assert torch.equal(x, y) and torch.equal(y, z)


def test_precision_study_configuration(eng):
    """The precision-study workload (Earth, T 512x128, E 128x32, inscatter
    64x256x64x16, 32x64 directions, orders 1-8) against the FP64 oracle:
    T and E in full, S at the fixture's 65536 seeded texels; floor 1e-4 of each
    channel's table maximum."""
    g = np.load(os.path.join(GOLDEN, "precision_hires.npz"))
    p = _params(g["params"])
    o = options(orders=8, spherical_samples=32)
    T, E, S = _chain(eng, p, o)
    _assert("T", T, g["T"], TOL)
    _assert("E", E, g["E"], TOL)
    flat = S.reshape(-1, 4)[g["S_idx"]]
    den = np.maximum(np.abs(g["S_val"]), 1e-4 * g["S_max"][None, :])
    err = np.abs(flat - g["S_val"]) / den
    print("S samples max %s mean %.3e" % (np.array2string(err.max(axis=0), precision=3), err.mean()), flush=True)
    assert err.max() <= TOL
