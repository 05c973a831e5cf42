"""Benchmark of the MI355X atmo precompute.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-extra]

One step = one full precompute (transmittance, irradiance and 4D inscatter
tables, all scattering orders) of the high-resolution Earth workload: inscatter
64x256x64x16 (RES_R x RES_MU x RES_MU_S x RES_NU), transmittance 512x128,
irradiance 128x32, 32x64 spherical-integral directions, orders 1-8, 500/50/32
integration samples.  Inputs are parameters only; every table lives in HBM.
Prints one JSON line (rank 0).  Only one GPU is supported.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (REPO, os.path.join(REPO, "proland-4.0_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (packed) peak, AMD public spec
FP32_UNPACKED_PEAK_TFLOPS = 78.6   # the same without packed FP32 (1 FMA per lane per clock)
HBM_PEAK_GBS = 8000.0


def workload_params():
    from prolandatmo.params import earth, options
    p = earth(TRANSMITTANCE_W=512, TRANSMITTANCE_H=128, SKY_W=128, SKY_H=32,
              RES_R=64, RES_MU=256, RES_MU_S=64, RES_NU=16)
    o = options(orders=8, spherical_samples=32)
    return p, o


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join("profiles", "hbm_traffic_earth_64x256x64x16.txt")
COUNTER_PROFILE = os.path.join("profiles", "sq_counters_earth_64x256x64x16.txt")


def profiled(path, tag, kernel):
    """Per-launch figure of `kernel` from a committed rocprofv3 summary (lines
    '<tag> <kernel> <value>'), or None when the summary or the line is absent."""
    full = os.path.join(REPO, path)
    if not os.path.exists(full):
        return None
    with open(full) as f:
        for line in f:
            w = line.split()
            if len(w) == 3 and w[0] == tag and w[1] == kernel:
                return int(w[2])
    return None


def time_steps(eng, sets, o, outs, steps, warmup):
    import torch
    for _ in range(warmup):
        eng.precompute(sets, o, outs=outs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.precompute(sets, o, outs=outs)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def cpu_baseline(p, o, budget_threads=16):
    """FP64 CPU oracle on a bounded sample of the same workload: the full-size
    transmittance and irradiance passes plus the 4D passes at 8x64x16x8
    (1/256 of the texels), extrapolated linearly in texel count."""
    import numpy as np
    from oracle import oracle as O
    from prolandatmo.params import options, with_res
    threads = min(budget_threads, os.cpu_count() or 1)
    O.set_threads(threads)
    q = with_res(p, 8, 64, 16, 8)
    scale = (p.RES_R * p.RES_MU * p.RES_MU_S * p.RES_NU) / float(q.RES_R * q.RES_MU * q.RES_MU_S * q.RES_NU)
    t = {}
    c = time.perf_counter
    a = c(); T = O.transmittance(q, o); t["T"] = c() - a
    a = c(); dE = O.irradiance1(q, T, o); t["E1"] = c() - a
    a = c(); dSR, dSM = O.inscatter1(q, T, o); t["S1"] = c() - a
    t4 = t["S1"]
    tE = t["T"] + t["E1"]
    S = O.copy_inscatter1(q, dSR, dSM)
    for n in range(2, o.orders + 1):
        first = n == 2
        a = c(); dJ = O.inscatterS(q, T, dE, dSR, dSM, first, o); t4 += c() - a
        a = c(); dE = O.irradianceN(q, dSR, dSM, first, o); tE += c() - a
        a = c(); dSR = O.inscatterN(q, T, dJ, o); t4 += c() - a
        a = c(); S = O.copy_inscatterN(q, dSR, S); t4 += c() - a
    est = tE + t4 * scale
    return {"value": 1.0 / est, "unit": "precomputes/s", "cores": threads, "kind": "port",
            "sample": ("FP64 oracle (oracle/, builder's CPU restatement; the reference GL preprocess "
                       "cannot run here) on the same workload with the 4D table reduced to 8x64x16x8 "
                       "(1/%d of the texels), full-size T/E passes; %.1f s measured, %.0f s "
                       "extrapolated per precompute" % (int(scale), tE + t4, est))}


def surrogate_block(steps, exact_orders=1):
    """The learned part on its own workloads: training on 2048 seeded sets of the
    batch distribution (32768 inscatter texels + every irradiance texel per
    set, batch 65536, Adam), then 1024 unseen sets at 32x128x32x8 (4 orders):
    plain batched tables/s, surrogate tables/s with the network trusted and
    with the 1 % judge, the judge's verdict, and the per-set max relative
    error of S against the plain tables (floored at 1e-4 of the table max)."""
    import torch
    from prolandatmo import surrogate_study as SS
    torch.cuda.empty_cache()
    a = SS.parser().parse_args([])
    a.steps = steps
    a.exact_orders = exact_orders
    r = SS.run(a, log=log)
    nj, pl = r["surrogate_no_judge"], r["plain"]
    j, ja = r["surrogate_judged"], r["surrogate_judged_accepted"]
    return {"workload": "train: 2048 sets (seed 0x5EED0004), 32768 texels + E per set, batch 65536, "
                        "%d Adam steps; infer: 1024 unseen sets (seed 0x5EED0005) at 32x128x32x8, "
                        "4 orders; orders 1-%d exact, the network predicts the rest" % (steps, exact_orders),
            "training": {k: r["training"][k] for k in ("ms_per_step", "tflops_model", "loss_first",
                                                       "loss_last")},
            "dataset_s": r["dataset"]["s"],
            "plain_tables_per_s": pl["tables_per_s"],
            "surrogate_trusted_tables_per_s": nj["tables_per_s"],
            "surrogate_trusted_speedup": nj["speedup_vs_plain"],
            "surrogate_judged_tables_per_s": j["tables_per_s"],
            "surrogate_judged_speedup": j["speedup_vs_plain"],
            "judge_passing_tolerance": ja["tolerance"],
            "surrogate_judged_speedup_at_passing_tolerance": ja["speedup_vs_plain"],
            "judge": {"fraction_of_sets": a.judge, "tolerance": a.tol,
                      "judged_sets": j["report"]["judged"], "accepted": j["report"]["accepted"],
                      "fallback_sets": j["report"]["fallback_sets"],
                      "judged_max_rel_err_S": j["report"]["max_rel_err_S"]},
            "per_set_max_rel_err_S": nj["max_rel_err_S"],
            "mean_rel_err_S": nj["mean_rel_err_S"],
            "ms": {"exact": nj["report"]["ms_exact"], "mlp": nj["report"]["ms_mlp"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate block (training on 2048 sets + inference on 1024)")
    ap.add_argument("--sur-steps", type=int, default=3000)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also copy a 256 MiB device buffer (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the precompute shards by parameter set with no "
              "communication, and multi-GPU runs are not part of this build (got --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from prolandatmo import flops as F
    from prolandatmo import lib as L
    from prolandatmo.params import FLAG_TIME_PASSES, earth, options

    torch.cuda.set_device(0)
    eng = L.Engine(0)
    p, o = workload_params()
    outs = [L.alloc_outputs(p)]
    log("workload: 64x256x64x16, orders 8, 32x64 directions")
    sec = time_steps(eng, [p], o, outs, args.steps, args.warmup)
    value = 1.0 / sec
    log("%.2f ms per precompute" % (sec * 1e3))

    # per-pass event timing of one eager run on the same stream
    ot = options(orders=o.orders, spherical_samples=o.spherical_samples, flags=FLAG_TIME_PASSES)
    eng.precompute([p], ot, outs=outs)
    torch.cuda.synchronize()
    times = eng.pass_times()
    by = {}
    for name, order, ms in times:
        by.setdefault(name, []).append(ms)
    pass_ms = {k: sum(v) for k, v in by.items()}
    # dominant kernel: the dJ pass of orders >= 3 (k_inscatterS<false, G, false>,
    # one launch per order; order 2 is its own instantiation with twice the
    # lookups; the third argument is the staging-in-rounds variant)
    later = [ms for name, order, ms in times if name == "inscatterS" and order >= 3]
    avg_ms = sum(later) / len(later)
    kname = "k_inscatterS<false,%d,false>" % eng.tuning().get("inscatterS_texels_per_thread", 4)
    traffic = profiled(TRAFFIC_PROFILE, "traffic", kname)
    # FP32 FLOPs one launch executes (rocprofv3 SQ_INSTS_VALU_FLOPS_FP32 x 64 lanes,
    # FMA = 2) over its live launch time: the roofline figure
    executed = profiled(COUNTER_PROFILE, "executed_fp32_flop", kname)
    # the same launch counted in the reference's formulation (flops.py: its GLSL,
    # each add/mul/transcendental 1, a GL trilinear fetch 75); the kernel stages the
    # separable part of each lookup once per workgroup, so it executes fewer
    formulation = p.n_texels4d() * F.inscatterS_texel(o.spherical_samples, False)
    achieved = executed / (avg_ms * 1e-3) / 1e12 if executed else None
    roof = {"bound": "valu",
            "achieved": round(achieved, 2) if achieved else None, "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4) if achieved else None,
            "traffic": traffic,
            "kernel": "%s (dJ pass, orders 3-%d)" % (kname, o.orders),
            "avg_launch_ms": round(avg_ms, 4), "launches": len(later),
            "flop_per_launch": executed,
            "flop_source": "executed FP32 FLOPs per launch: rocprofv3 SQ_INSTS_VALU_FLOPS_FP32 x 64 "
                           "(%s), over the launch time measured here with hipEvents" % COUNTER_PROFILE,
            # packed FP32 (v_pk_fma_f32) is disabled in this build; one FMA per lane per clock
            "frac_of_unpacked_fp32_peak": (round(achieved / FP32_UNPACKED_PEAK_TFLOPS, 4)
                                           if achieved else None),
            "formulation_flop_per_launch": formulation,
            "formulation_tflops": round(formulation / (avg_ms * 1e-3) / 1e12, 2),
            "formulation_speedup": round(formulation / executed, 2) if executed else None}
    valu = profiled(COUNTER_PROFILE, "valu_insts", kname)
    if valu:
        # VALU wave instructions per texel-direction (SQ_INSTS_VALU x 64 lanes over
        # RES_R RES_MU RES_MU_S RES_NU texels x 2 NS^2 directions) and their issue rate
        dirs = 2 * o.spherical_samples ** 2
        roof["valu_insts_per_texel_direction"] = round(valu * 64 / (p.n_texels4d() * dirs), 2)
        roof["valu_wave_insts_per_s"] = round(valu / (avg_ms * 1e-3), -6)
    if traffic:
        roof["traffic_source"] = TRAFFIC_PROFILE
        roof["hbm_gbs"] = round(traffic / (avg_ms * 1e-3) / 1e9, 1)
        roof["hbm_frac"] = round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    whole_tflops = F.total(p, o) / sec / 1e12

    # second kernel: the dSn pass.  Its HBM traffic is a fraction of peak; what
    # it moves is on-chip: per view ray (r layer, mu row) and sample, the staged
    # form blends the 4 dJ rows around the sample (2 layers x 2 mu rows,
    # RES_NU*RES_MU_S texels of 12 bytes each) into LDS.
    tu = eng.tuning()
    sn_name = "k_inscatterN<%s,false>" % ("true" if tu.get("inscatterN_staged_samples", 0) else "false")
    sn_ms = [ms for name, order, ms in times if name == "inscatterN"]
    sn_avg = sum(sn_ms) / len(sn_ms)
    sn_traffic = profiled(TRAFFIC_PROFILE, "traffic", sn_name)
    row_reads = (p.RES_R * p.RES_MU * (o.inscatter_samples + 1) * 4 * p.RES_NU * p.RES_MU_S * 12)
    second = {"kernel": sn_name + " (dSn pass, orders 2-%d)" % o.orders,
              "bound": "on-chip: dJ rows gathered per view-ray sample (L2 -> LDS)",
              "avg_launch_ms": round(sn_avg, 4), "launches": len(sn_ms),
              "row_gather_bytes": row_reads,
              "row_gather_tbs": round(row_reads / (sn_avg * 1e-3) / 1e12, 2),
              "hbm_traffic": sn_traffic, "hbm_peak_gbs": HBM_PEAK_GBS}
    if sn_traffic:
        second["hbm_gbs"] = round(sn_traffic / (sn_avg * 1e-3) / 1e9, 1)
        second["hbm_frac"] = round(sn_traffic / (sn_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        second["traffic_source"] = TRAFFIC_PROFILE

    extra = {"pass_ms": {k: round(v, 3) for k, v in pass_ms.items()},
             "algorithmic_tflops_whole_precompute": round(whole_tflops, 2),
             "roofline_dsn": second,
             "autotuner": tu}
    if not args.no_extra:
        # the reference configuration (Earth, 32x128x32x8, 4 orders) and its
        # accuracy at the committed FP64-oracle golden texels
        pr = earth()
        orr = options()
        outs1 = [L.alloc_outputs(pr)]
        sec1 = time_steps(eng, [pr], orr, outs1, 20, 2)
        g = np.load(os.path.join(REPO, "tests", "golden", "earth_ref.npz"))
        T, E, S = (t.double().cpu().numpy() for t in outs1[0])
        from prolandatmo.metrics import summary
        scale = np.abs(S).max()
        den = np.maximum(np.abs(g["S_val"]), 1e-4 * scale)
        errS = float((np.abs(S.reshape(-1, 4)[g["S_idx"]] - g["S_val"]) / den).max())
        extra["reference_config"] = {
            "workload": "Earth 32x128x32x8, T 256x64, E 64x16, 16x32 dirs, orders 4",
            "precomputes_per_s": round(1.0 / sec1, 2), "ms_per_precompute": round(sec1 * 1e3, 3),
            "max_rel_err_vs_fp64_oracle": {"T": summary(T, g["T"])["max"],
                                           "E": summary(E, g["E"])["max"],
                                           "S_4096_texels": errS}}
        log("reference config: %.3f ms" % (sec1 * 1e3))
        del outs1
        # batched plain method: 256 seeded atmospheres of mixed inscatter resolution
        from prolandatmo.params import SEED_BATCH, draw_sets
        sets, _ = draw_sets(SEED_BATCH, 256)
        outsb = [L.alloc_outputs(q) for q in sets]
        secb = time_steps(eng, sets, orr, outsb, 2, 1)
        texels = sum(q.n_texels4d() for q in sets)
        extra["batched_256_sets"] = {
            "workload": "256 sets (seed 0x5EED0003), inscatter mixed from 16x64x16x4, "
                        "32x128x32x8, 48x192x48x12; T 256x64, E 64x16, orders 4, one graph",
            "tables_per_s": round(len(sets) / secb, 2), "s_per_batch": round(secb, 4),
            "texels_4d": texels,
            "algorithmic_tflops": round(sum(F.total(q, orr) for q in sets) / secb / 1e12, 2)}
        log("batched 256 sets: %.3f s" % secb)
        del outsb
    eng.close()
    if not args.no_surrogate:
        extra["surrogate"] = surrogate_block(args.sur_steps, 1)
        extra["surrogate_exact_order_2"] = surrogate_block(args.sur_steps, 2)
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(p, o)
    out = {
        "metric": "precomputes/sec at <=1e-3 rel err vs FP64; surrogate vs plain speedup",
        "value": round(value, 4), "unit": "precomputes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (fp64 geometry)",
        "data": "synthetic (atmosphere parameters only)",
        "config": {"workload": "Earth precision-study precompute: inscatter 64x256x64x16, "
                               "transmittance 512x128, irradiance 128x32, 32x64 spherical "
                               "directions, orders 1-8, 500/50/32 samples",
                   "parallelism": "single GPU"},
        "roofline": roof, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(out))
    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copied %d bytes" % (src.numel() * 4))


if __name__ == "__main__":
    main()
