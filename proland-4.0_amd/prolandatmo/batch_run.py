"""The batched plain method on seeded parameter sets, timed.

    python -m prolandatmo.batch_run [--sets 256] [--replays 2] [--seed 0x5EED0003]

Draws the sets from the batch distribution (inscatter resolutions mixed from
16x64x16x4, 32x128x32x8 and 48x192x48x12; T 256x64, E 64x16), runs the
4-order precompute once (graph capture and autotuning of a new batch shape),
then replays it and prints the time of each replay.
"""
import argparse
import time


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sets", type=int, default=256)
    ap.add_argument("--replays", type=int, default=2)
    ap.add_argument("--seed", type=lambda x: int(x, 0), default=0x5EED0003)
    a = ap.parse_args()
    import torch
    from . import lib as L
    from .params import draw_sets, options
    sets, _ = draw_sets(a.seed, a.sets)
    eng = L.Engine(0)
    outs = eng.precompute(sets, options())
    torch.cuda.synchronize()
    for k in range(a.replays):
        t = time.perf_counter()
        eng.precompute(sets, options(), outs=outs)
        torch.cuda.synchronize()
        print("replay %d: %.2f ms (%d sets, %d 4D texels)"
              % (k, (time.perf_counter() - t) * 1e3, len(sets), sum(p.n_texels4d() for p in sets)), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
