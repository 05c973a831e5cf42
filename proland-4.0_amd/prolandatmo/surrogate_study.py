"""Training and inference study of the multiple-scattering surrogate.

    python -m prolandatmo.surrogate_study [--train-sets 2048] [--samples 32768]
        [--steps 3000] [--batch 65536] [--infer-sets 1024] [--out result.json]

Training: seeded sets of the batch distribution (seed 0x5EED0004), plain-method
tables (4 orders), `samples` inscatter texels + all irradiance texels per set.
Inference: unseen sets (seed 0x5EED0005) at 32x128x32x8; the plain batched
path is timed on the same sets, then the surrogate path (exact transmittance
and single scattering + network for orders >= 2 + a 1 % spot-check judge).
Prints one JSON object.
"""
import argparse
import json
import sys
import time

import numpy as np


def _log(*x):
    print(*x, file=sys.stderr, flush=True)


def run(a, log=_log):
    """The study with the settings of namespace `a` (the command-line options);
    returns the result dictionary."""
    import torch
    from . import lib as L
    from . import surrogate as SG
    from .params import SEED_INFER, SEED_TRAIN, draw_sets, options

    res = {"config": dict(vars(a))}
    o = options()
    s = SG.Surrogate(0, seed=0x5EED0004)
    s.set_exact_orders(a.exact_orders)
    if a.eps is not None:   # the floor of the log-scaled targets (before standardisation)
        s.set_norm([0.0, 0.0, 0.0], [1.0, 1.0, 1.0], a.eps)
    train_sets, _ = draw_sets(SEED_TRAIN, a.train_sets)
    t = time.perf_counter()
    rows = s.build_dataset(train_sets, o, samples_per_set=a.samples, seed=SEED_TRAIN)
    res["dataset"] = {"rows": rows, "s": round(time.perf_counter() - t, 2)}
    log("dataset", res["dataset"])
    t = time.perf_counter()
    # cosine schedule from lr to lr/100, run in chunks so progress is visible
    chunk = 250
    lr_at = lambda k: a.lr * (1e-2 + (1 - 1e-2) * 0.5 * (1 + np.cos(np.pi * k / max(1, a.steps - 1))))
    parts = []
    for k0 in range(0, a.steps, chunk):
        n = min(chunk, a.steps - k0)
        parts.append(s.train(n, batch=a.batch, lr=lr_at(k0), lr_end=lr_at(k0 + n - 1), seed=1 + k0))
        log("step %d loss %.4f (%.1f s)" % (k0 + n, parts[-1][-10:].mean(), time.perf_counter() - t))
    losses = np.concatenate(parts)
    dt = time.perf_counter() - t
    flop = 3 * 2 * (SG.NIN * 256 + (SG.NL - 1) * 256 * 256 + 256 * 4) * a.batch
    res["training"] = {"steps": a.steps, "s": round(dt, 2), "ms_per_step": round(dt / a.steps * 1e3, 3),
                       "tflops_model": round(flop * a.steps / dt / 1e12, 1),
                       "loss_first": float(losses[:20].mean()), "loss_last": float(losses[-20:].mean()),
                       "loss_curve": [float(losses[i]) for i in np.linspace(0, a.steps - 1, 12).astype(int)]}
    log("training", res["training"])
    if a.save_params:
        np.save(a.save_params, s.get_params())

    if a.infer_sets <= 0:   # training only
        s.close()
        return res
    infer, _ = draw_sets(SEED_INFER, a.infer_sets, res=(32, 128, 32, 8))
    # plain batched baseline
    eng = L.Engine(0)
    plain = eng.precompute(infer, o)
    torch.cuda.synchronize()
    t = time.perf_counter()
    eng.precompute(infer, o, outs=plain)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t
    eng.close()
    res["plain"] = {"sets": len(infer), "s": round(t_plain, 4), "tables_per_s": round(len(infer) / t_plain, 2)}
    log("plain", res["plain"])
    # surrogate, judge disabled for the accuracy census (every table compared)
    outs = [L.alloc_outputs(p) for p in infer]
    s.precompute(infer, o, outs=outs, judge_fraction=0.0, tolerance=1e30)
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, rep_fast = s.precompute(infer, o, outs=outs, judge_fraction=0.0, tolerance=1e30)
    torch.cuda.synchronize()
    t_fast = time.perf_counter() - t
    # the same tables without the learned part: the exactly computed orders only
    eng1 = L.Engine(0)
    single = eng1.precompute(infer, options(orders=a.exact_orders))
    torch.cuda.synchronize()
    eng1.close()
    errS, errE, meanS, meanS1 = [], [], [], []
    for (T, E, S), (pT, pE, pS), (sT, sE, sS) in zip(outs, plain, single):
        fS = 1e-4 * pS.abs().max()
        rS = (S - pS).abs() / torch.clamp(pS.abs(), min=fS)
        errS.append(float(rS.max()))
        meanS.append(float(rS.mean()))
        meanS1.append(float(((sS - pS).abs() / torch.clamp(pS.abs(), min=fS)).mean()))
        fE = 1e-4 * pE.abs().max()
        errE.append(float(((E - pE).abs() / torch.clamp(pE.abs(), min=fE)).max()))
    errS, errE = np.array(errS), np.array(errE)
    res["surrogate_no_judge"] = {
        "s": round(t_fast, 4), "tables_per_s": round(len(infer) / t_fast, 2),
        "speedup_vs_plain": round(t_plain / t_fast, 3), "report": rep_fast,
        "max_rel_err_S": {"median": float(np.median(errS)), "p90": float(np.percentile(errS, 90)),
                          "max": float(errS.max()), "sets_within_tol": int((errS <= a.tol).sum()),
                          "fraction_within": {t: float((errS <= float(t)).mean())
                                              for t in ("1e-3", "1e-2", "1e-1", "1")}},
        "max_rel_err_E": {"median": float(np.median(errE)), "p90": float(np.percentile(errE, 90)),
                          "max": float(errE.max())},
        "mean_rel_err_S": {"surrogate_median_over_sets": float(np.median(meanS)),
                           "single_scattering_only_median_over_sets": float(np.median(meanS1)),
                           "sets_where_surrogate_beats_single_only": int((np.array(meanS) < np.array(meanS1)).sum())}}
    log("surrogate (no judge)", res["surrogate_no_judge"])
    # the production path: 1 % judge, tolerance, fallback
    t = time.perf_counter()
    _, rep = s.precompute(infer, o, outs=outs, judge_fraction=a.judge, tolerance=a.tol)
    torch.cuda.synchronize()
    t_j = time.perf_counter() - t
    res["surrogate_judged"] = {"s": round(t_j, 4), "tables_per_s": round(len(infer) / t_j, 2),
                               "speedup_vs_plain": round(t_plain / t_j, 3), "report": rep}
    log("surrogate (judged)", res["surrogate_judged"])
    # the same judge at the tolerance the judged sets meet (their worst error):
    # the path's time when the batch is accepted
    tol_pass = max(rep["max_rel_err_S"], rep["max_rel_err_E"]) * 1.0001
    t = time.perf_counter()
    _, rep2 = s.precompute(infer, o, outs=outs, judge_fraction=a.judge, tolerance=tol_pass)
    torch.cuda.synchronize()
    t_a = time.perf_counter() - t
    res["surrogate_judged_accepted"] = {"tolerance": tol_pass, "s": round(t_a, 4),
                                        "tables_per_s": round(len(infer) / t_a, 2),
                                        "speedup_vs_plain": round(t_plain / t_a, 3), "report": rep2}
    log("surrogate (judged, passing tolerance)", res["surrogate_judged_accepted"])
    s.close()
    return res


def parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-sets", type=int, default=2048)
    ap.add_argument("--samples", type=int, default=32768)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--infer-sets", type=int, default=1024)
    ap.add_argument("--judge", type=float, default=0.01)
    ap.add_argument("--exact-orders", type=int, default=1,
                    help="orders the plain passes compute (1 or 2); the network predicts the rest")
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--eps", type=float, default=None,
                    help="floor of the log-scaled training targets (default: the library's, 1e-6)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--save-params", default=None)
    return ap


def main():
    a = parser().parse_args()
    txt = json.dumps(run(a))
    print(txt)
    if a.out:
        open(a.out, "w").write(txt)


if __name__ == "__main__":
    main()
