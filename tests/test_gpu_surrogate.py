"""GPU tests of the learned multiple-scattering surrogate: BF16-MFMA forward
and backward against the NumPy MLP oracle, training, and the judge/fallback
contract of atmo_surrogate_precompute."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import mlp_oracle as M  # noqa: E402
from prolandatmo import lib as L  # noqa: E402
from prolandatmo import surrogate as SG  # noqa: E402
from prolandatmo.params import FLAG_HALF_STORAGE, SEED_BATCH, draw_sets, options  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sur():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    s = SG.Surrogate(0, seed=7)
    yield s
    s.close()


def rows(B, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((B, 16), dtype=np.float32)
    X[:, :SG.NIN] = rng.uniform(-1, 1, (B, SG.NIN))
    T = rng.normal(0, 1, (B, 4)).astype(np.float32)
    return X, T


def blocks(v):
    lay = SG.layout()
    return {k: v[o:o + int(np.prod(s))] for k, (o, s) in lay.items()}


def test_forward_matches_oracle(sur):
    X, _ = rows(300, 1)
    Y = sur.predict(torch.from_numpy(X).cuda()).double().cpu().numpy()
    Yo, _ = M.forward(sur.get_params().astype(np.float64), X, emulate_bf16=True)
    err = np.abs(Y - Yo).max() / np.abs(Yo).max()
    print("forward max rel err %.3e" % err, flush=True)
    assert err < 2e-3


@pytest.mark.parametrize("B", [128, 2048])
def test_gradients_match_oracle(sur, B):
    X, T = rows(B, 2 + B)
    loss, g = sur.loss_grad(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda())
    lo, go = M.loss_grad(sur.get_params().astype(np.float64), X, T, emulate_bf16=True)
    assert loss == pytest.approx(lo, rel=2e-3)
    gb, ob = blocks(g.astype(np.float64)), blocks(go)
    for k in gb:
        rel = np.linalg.norm(gb[k] - ob[k]) / max(np.linalg.norm(ob[k]), 1e-30)
        cos = np.dot(gb[k], ob[k]) / max(np.linalg.norm(gb[k]) * np.linalg.norm(ob[k]), 1e-30)
        print("%s rel %.3e cos %.6f" % (k, rel, cos), flush=True)
        assert rel < 3e-2 and cos > 0.999, k


def _log_ray(p, l, y):
    """The mu-row coordinate: the row's ray length (getMuMuSNu with the setLayer
    radius) on a log scale from 1 m to 2 sqrt(Rt^2 - Rg^2)."""
    Rg, Rt = float(p.Rg), float(p.Rt)
    half = p.RES_MU // 2
    r = np.sqrt(Rg ** 2 + (l / (p.RES_R - 1)) ** 2 * (Rt ** 2 - Rg ** 2))
    r += 0.01 if l == 0 else (-0.001 if l == p.RES_R - 1 else 0.0)
    rho, H = np.sqrt(r * r - Rg * Rg), np.sqrt(Rt * Rt - Rg * Rg)
    if y < half:
        dmin, dmax, d = r - Rg, 0.999 * rho, (1 - y / (half - 1)) * rho
    else:
        dmin, dmax, d = Rt - r, 0.999 * (rho + H), (y - half) / (half - 1) * (rho + H)
    d = min(max(d, dmin), dmax)
    return min(max(np.log(d / 1e-3) / np.log(2 * H / 1e-3), 0.0), 1.0)


def test_features_of_texel_codes():
    p = draw_sets(SEED_BATCH, 1)[0][0]
    W4 = p.RES_MU_S * p.RES_NU
    l, y, x = 3, p.RES_MU // 2 + 5, 2 * p.RES_MU_S + 7
    code = (l * p.RES_MU + y) * W4 + x
    e = 0x80000000 | (2 * p.SKY_W + 9)
    # exact-part tables: positive values over the tables' range, and a zero texel
    g = np.random.default_rng(5)
    n4 = p.RES_R * p.RES_MU * W4
    S1 = (10.0 ** g.uniform(-7, 0, (n4, 4))).astype(np.float32)
    E1 = (10.0 ** g.uniform(-7, 0, (p.SKY_W * p.SKY_H, 3))).astype(np.float32)
    S1[code, 1] = 0.0
    S1d, E1d = torch.from_numpy(S1).cuda(), torch.from_numpy(E1).cuda()
    X = SG.features(p, [code, e], S1d, E1d).cpu().numpy()
    half = p.RES_MU // 2
    x1 = 0.25 + 0.75 * _log_ray(p, l, y)           # sky half of the mu rows: [0.25, 1]
    assert X[0, :4] == pytest.approx([2 * l / (p.RES_R - 1) - 1, x1,
                                      2 * 7 / (p.RES_MU_S - 1) - 1, 2 * 2 / (p.RES_NU - 1) - 1], abs=1e-6)
    assert X[1, :4] == pytest.approx([2 * 2 / (p.SKY_H - 1) - 1, -2, 2 * 9 / (p.SKY_W - 1) - 1, -2], abs=1e-6)
    for yg in (0, 3, half - 2, half - 1):          # ground half: [-1, -0.25]
        for lg in (0, l, p.RES_R - 1):
            Xg = SG.features(p, [(lg * p.RES_MU + yg) * W4 + x], S1d, E1d).cpu().numpy()
            assert Xg[0, 1] == pytest.approx(-0.25 - 0.75 * _log_ray(p, lg, yg), abs=1e-5)
    assert X[0, 4] == pytest.approx((p.Rg - 5000) / 2000, abs=1e-6)
    assert X[0, 10] == pytest.approx((p.mieG - 0.75) / 0.15, abs=1e-5)
    # the exact part on a log scale: 0.25 (log(v + 1e-6) + 5)
    logv = lambda v: 0.25 * (np.log(v.astype(np.float64) + 1e-6) + 5.0)
    assert X[0, 11:14] == pytest.approx(logv(S1[code, :3]), abs=2e-6)
    assert X[0, 12] == pytest.approx(0.25 * (np.log(1e-6) + 5.0), abs=2e-6)
    assert X[1, 11:14] == pytest.approx(logv(E1[2 * p.SKY_W + 9]), abs=2e-6)
    assert np.all(X[:, 14:] == 0)


def small_sets(n, seed=SEED_BATCH):
    sets, _ = draw_sets(seed, n, res=(16, 64, 16, 4))
    for p in sets:
        p.TRANSMITTANCE_W, p.TRANSMITTANCE_H, p.SKY_W, p.SKY_H = 64, 16, 16, 8
    return sets


def test_training_reduces_loss():
    s = SG.Surrogate(0, seed=11)
    n = s.build_dataset(small_sets(6), options(), samples_per_set=4096)
    assert n == 6 * (4096 + 16 * 8)
    mu, sd, eps = s.get_norm()
    assert np.all(sd > 0) and eps > 0
    losses = s.train(200, batch=4096, lr=2e-3, lr_end=1e-4)
    print("loss %.4f -> %.4f" % (losses[:5].mean(), losses[-5:].mean()), flush=True)
    assert np.all(np.isfinite(losses))
    # against the untrained network's loss (the first step's; later early
    # steps have already learned most of the mean)
    assert losses[-10:].mean() < 0.3 * losses[0]
    s.close()


def _plain(sets, o):
    e = L.Engine(0)
    outs = e.precompute(sets, o)
    torch.cuda.synchronize()
    res = [tuple(t.double().cpu().numpy() for t in ts) for ts in outs]
    e.close()
    return res


def test_judge_rejects_and_falls_back_to_plain(sur):
    sets = small_sets(5, seed=SEED_BATCH + 1)
    o = options()
    outs, rep = sur.precompute(sets, o, judge_fraction=0.2, tolerance=1e-12)
    torch.cuda.synchronize()
    assert rep["judged"] == 1 and rep["accepted"] == 0 and rep["fallback_sets"] == 5
    ref = _plain(sets, o)
    for (T, E, S), (rT, rE, rS) in zip(outs, ref):
        assert np.array_equal(T.double().cpu().numpy(), rT)
        assert np.array_equal(E.double().cpu().numpy(), rE)
        assert np.array_equal(S.double().cpu().numpy(), rS)


def test_surrogate_rejects_half_storage(sur):
    """The surrogate predicts FP32 tables: FP16 storage is refused, not mixed in."""
    sets = small_sets(1, seed=SEED_BATCH + 3)
    o = options(flags=FLAG_HALF_STORAGE)
    with pytest.raises(L.AtmoError, match="FP16|HALF"):
        sur.precompute(sets, o)
    with pytest.raises(L.AtmoError, match="FP16|HALF"):
        sur.build_dataset(sets, o, samples_per_set=64)


def test_judge_accepts_and_keeps_exact_parts(sur):
    sets = small_sets(4, seed=SEED_BATCH + 2)
    o = options()
    outs, rep = sur.precompute(sets, o, judge_fraction=0.5, tolerance=1e30)
    torch.cuda.synchronize()
    assert rep["judged"] == 2 and rep["accepted"] == 1 and rep["fallback_sets"] == 2
    assert rep["max_rel_err_S"] >= 0 and rep["ms_mlp"] > 0
    ref = _plain(sets, o)
    exact = 0
    for (T, E, S), (rT, rE, rS) in zip(outs, ref):
        S = S.double().cpu().numpy()
        assert np.array_equal(T.double().cpu().numpy(), rT)      # exact transmittance
        assert np.array_equal(S[..., 3], rS[..., 3])             # exact single Mie (alpha)
        assert np.all(np.isfinite(S)) and np.all(S >= 0)
        exact += np.array_equal(S, rS)
    assert exact == 2                                             # the judged sets


def test_fused_inference_matches_the_network(sur):
    """atmo_surrogate_precompute's fused kernel (features, five layers and the
    combination in one launch per set) against the layer-by-layer network
    (atmo_surrogate_predict on atmo_surrogate_features rows) plus the exact
    single scattering, on every texel of two sets."""
    sets = small_sets(2, seed=SEED_BATCH + 3)
    mu0, sd0, eps0 = sur.get_norm()
    sur.set_norm(np.array([-3.0, -2.5, -2.0], dtype=np.float32), np.array([1.0, 0.8, 1.2], dtype=np.float32), 1e-6)
    try:
        o = options()
        outs, rep = sur.precompute(sets, o, judge_fraction=0.0)
        torch.cuda.synchronize()
        assert rep["judged"] == 0 and rep["accepted"] == 1 and rep["fallback_sets"] == 0
        single = _plain(sets, options(orders=1))
        mu, sd, eps = sur.get_norm()
        for p, (T, E, S), (rT, rE, rS) in zip(sets, outs, single):
            n4 = p.RES_R * p.RES_MU * p.RES_MU_S * p.RES_NU
            nE = p.SKY_W * p.SKY_H
            codes = np.concatenate([np.arange(n4, dtype=np.uint32),
                                    np.uint32(0x80000000) | np.arange(nE, dtype=np.uint32)])
            S1d = torch.from_numpy(rS.astype(np.float32)).cuda().reshape(-1, 4).contiguous()
            E1d = torch.from_numpy(rE.astype(np.float32)).cuda().reshape(-1, 3).contiguous()
            Y = sur.predict(SG.features(p, codes, S1d, E1d)).double().cpu().numpy()
            v = np.maximum(np.exp(Y[:, :3] * sd.astype(np.float64) + mu) - eps, 0.0)
            S = S.double().cpu().numpy().reshape(-1, 4)
            vS = S[:, :3] - rS.reshape(-1, 4)[:, :3]
            eS = np.abs(vS - v[:n4]) / np.maximum(v[:n4], 1e-3 * v[:n4].max())
            eE = np.abs(E.double().cpu().numpy().reshape(-1, 3) - v[n4:]) / np.maximum(v[n4:], 1e-3 * v[n4:].max())
            print("fused vs layer-by-layer: S max %.3e mean %.3e, E max %.3e" % (eS.max(), eS.mean(), eE.max()),
                  flush=True)
            assert eS.max() < 2e-2 and eS.mean() < 1e-3 and eE.max() < 2e-2
            assert np.array_equal(S[:, 3], rS.reshape(-1, 4)[:, 3])
            assert np.array_equal(T.double().cpu().numpy(), rT)
    finally:
        sur.set_norm(mu0, sd0, eps0)


def test_trained_network_gains_over_single_scattering_only():
    """What the learned part adds: after a short training on plain-method
    tables, the surrogate inscatter tables (exact single scattering + predicted
    orders >= 2) are closer to the plain multi-order tables, in mean relative
    error over every texel of unseen sets, than the same tables without the
    network (single scattering only); likewise the irradiance tables."""
    s = SG.Surrogate(0, seed=21)
    try:
        s.build_dataset(small_sets(24, seed=SEED_BATCH + 10), options(), samples_per_set=8192)
        s.train(2000, batch=8192, lr=2e-3, lr_end=5e-5)
        test = small_sets(4, seed=SEED_BATCH + 20)
        outs, rep = s.precompute(test, options(), judge_fraction=0.0)
        torch.cuda.synchronize()
        plain = _plain(test, options())
        single = _plain(test, options(orders=1))
        for (T, E, S), (pT, pE, pS), (sT, sE, sS) in zip(outs, plain, single):
            S = S.double().cpu().numpy()
            E = E.double().cpu().numpy()
            fS, fE = 1e-4 * np.abs(pS).max(), 1e-4 * np.abs(pE).max()
            eS = (np.abs(S - pS) / np.maximum(np.abs(pS), fS)).mean()
            eS1 = (np.abs(sS - pS) / np.maximum(np.abs(pS), fS)).mean()
            eE = (np.abs(E - pE) / np.maximum(np.abs(pE), fE)).mean()
            eE1 = (np.abs(sE - pE) / np.maximum(np.abs(pE), fE)).mean()
            print("mean rel err S: surrogate %.4f vs single-only %.4f; E: %.4f vs %.4f" % (eS, eS1, eE, eE1),
                  flush=True)
            assert eS < 0.7 * eS1 and eE < eE1
    finally:
        s.close()
