"""The batched and learned workloads at their full sizes, through the C ABI.

* The batched plain method: 256 parameter sets of the batch distribution
  (seed 0x5EED0003), inscatter resolutions mixed from 16x64x16x4,
  32x128x32x8 and 48x192x48x12, T 256x64, E 64x16, orders 4, in one
  captured graph, with the numeric check on every set.  One set of each
  resolution class is checked against the FP64 oracle (tests/golden/
  batch_classes.npz: T and E in full, S at 4096 seeded texels) at the 1e-3
  bar (relative error floored at 1e-4 of each table's largest magnitude).
* Surrogate training: one forward + backward pass of the MLP at the training
  batch of 65,536 rows against the NumPy oracle (oracle/mlp_oracle.py, BF16
  rounding emulated where the kernels store BF16).
* Surrogate inference with judge and fallback: 1024 unseen sets at
  32x128x32x8 (seed 0x5EED0005).  A rejected batch falls back to tables
  bit-identical to the plain batched path; an accepted one keeps the exact
  transmittance and single-scattering Mie (alpha) of the plain path.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import mlp_oracle as M  # noqa: E402
from prolandatmo import lib as L  # noqa: E402
from prolandatmo import surrogate as SG  # noqa: E402
from prolandatmo.params import (FLAG_CHECK_NUMERICS, RES_CLASSES, SEED_BATCH, SEED_INFER,  # noqa: E402
                                draw_sets, options)

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-3


def _rel(x, y, floor):
    return float((np.abs(x - y) / np.maximum(np.abs(y), floor)).max())


def test_batch_of_256_mixed_sets_against_the_oracle():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    g = np.load(os.path.join(GOLDEN, "batch_classes.npz"))
    sets, _ = draw_sets(SEED_BATCH, 256)
    classes = [(q.RES_R, q.RES_MU, q.RES_MU_S, q.RES_NU) for q in sets]
    assert set(classes) == set(RES_CLASSES)
    eng = L.Engine(0)
    outs = eng.precompute(sets, options())
    torch.cuda.synchronize()
    bad = []
    for i, (T, E, S) in enumerate(outs):
        for name, t, hi in (("T", T, 1.0), ("E", E, None), ("S", S, None)):
            ok = torch.isfinite(t) & (t >= 0)
            if hi is not None:
                ok &= t <= hi
            if not bool(ok.all()):
                k = int(torch.nonzero(~ok.reshape(-1))[0])
                bad.append((i, name, k, float(t.reshape(-1)[k])))
    assert not bad, bad[:8]
    # the library's own check on every set, on the replayed graph of the tuned plan
    o = options(flags=FLAG_CHECK_NUMERICS)
    outs = eng.precompute(sets, o, outs=outs)
    torch.cuda.synchronize()
    print("256 sets, %d 4D texels, tuning %s" % (sum(q.n_texels4d() for q in sets),
                                                 eng.tuning()["inscatterN_staged_samples"]), flush=True)
    for c in range(len(RES_CLASSES)):
        i = int(g["index%d" % c])
        assert bytes(sets[i]) == g["params%d" % c].tobytes()
        T, E, S = outs[i]
        T = T.double().cpu().numpy()
        E = E.double().cpu().numpy()
        flat = S.reshape(-1, 4)[torch.from_numpy(g["S_idx%d" % c].astype(np.int64)).cuda()]
        flat = flat.double().cpu().numpy()
        eT = _rel(T, g["T%d" % c], 1e-4 * np.abs(g["T%d" % c]).max())
        eE = _rel(E, g["E%d" % c], 1e-4 * np.abs(g["E%d" % c]).max())
        eS = _rel(flat, g["S_val%d" % c], 1e-4 * g["S_max%d" % c])
        print("set %3d %s: T %.2e E %.2e S %.2e" % (i, classes[i], eT, eE, eS), flush=True)
        assert max(eT, eE, eS) <= TOL, (i, eT, eE, eS)
    del outs
    eng.close()


def test_training_pass_at_batch_65536_against_the_oracle():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    s = SG.Surrogate(0, seed=5)
    B = 65536
    rng = np.random.default_rng(0x5EED0004)
    X = np.zeros((B, 16), dtype=np.float32)
    X[:, :SG.NIN] = rng.uniform(-1, 1, (B, SG.NIN))
    T = rng.normal(0, 1, (B, 4)).astype(np.float32)
    loss, g = s.loss_grad(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda())
    lo, go = M.loss_grad(s.get_params().astype(np.float64), X, T, emulate_bf16=True)
    print("loss %.6f oracle %.6f" % (loss, lo), flush=True)
    assert loss == pytest.approx(lo, rel=2e-3)
    lay = SG.layout()
    g = g.astype(np.float64)
    for k, (off, shape) in lay.items():
        a, b = g[off:off + int(np.prod(shape))], go[off:off + int(np.prod(shape))]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = np.dot(a, b) / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
        print("%s rel %.3e cos %.6f" % (k, rel, cos), flush=True)
        assert rel < 3e-2 and cos > 0.999, k
    s.close()


def _equal_sets(a, b):
    return all(torch.equal(x, y) for x, y in zip(a, b))


def test_judge_and_fallback_on_1024_unseen_sets():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    sets, _ = draw_sets(SEED_INFER, 1024, res=(32, 128, 32, 8))
    o = options()
    eng = L.Engine(0)
    plain = eng.precompute(sets, o)
    torch.cuda.synchronize()
    eng.close()
    s = SG.Surrogate(0, seed=3)   # an untrained network: the judge must reject it
    s.set_norm(np.array([-12.0, -11.0, -10.0]), np.array([2.0, 2.0, 2.0]), 1e-9)
    outs, rep = s.precompute(sets, o, judge_fraction=0.01, tolerance=1e-3)
    torch.cuda.synchronize()
    print("rejected: %s" % rep, flush=True)
    assert rep["accepted"] == 0 and rep["fallback_sets"] == 1024 and rep["judged"] >= 1
    assert all(_equal_sets(a, b) for a, b in zip(outs, plain))
    # a tolerance every batch meets: accepted, exact parts kept
    outs, rep = s.precompute(sets, o, outs=outs, judge_fraction=0.01, tolerance=1e30)
    torch.cuda.synchronize()
    print("accepted: %s" % rep, flush=True)
    assert rep["accepted"] == 1
    for (T, E, S), (pT, pE, pS) in zip(outs, plain):
        assert torch.equal(T, pT)                      # exact transmittance
        assert torch.equal(S[..., 3], pS[..., 3])      # exact single-scattering Mie (alpha)
        assert bool(torch.isfinite(S).all()) and bool((S >= 0).all())
    del outs, plain
    s.close()
