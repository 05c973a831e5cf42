"""Writes the golden fixtures of the parity tests from the FP64 oracle.

    python tests/golden/make_golden.py

The reference ships no atmo tables or test vectors (the .raw outputs are
excluded from its distribution, src/atmo/install.sh:17), so these vectors
come from the pinned FP64 oracle (oracle/, checked in tests/test_oracle_pins.py):
(S and T stored as float32: 6e-8 relative rounding, far below the 1e-3 bar)
  tiny_earth.npz      Earth, T 64x16, E 16x8, inscatter 8x32x8x4, orders 4: T, E, S
  tiny_seeds.npz      3 sets of the batch distribution (seed 0x5EED0003 stream),
                      T 32x8, E 8x4, inscatter 6x16x6x4, orders 4
  earth_ref.npz       the reference configuration (Earth, T 256x64, E 64x16,
                      inscatter 32x128x32x8, 500/50/32/16 samples, orders 4):
                      T and E in full, S at 4096 seeded texels
  precision_hires.npz the precision-study configuration (Earth, T 512x128,
                      E 128x32, inscatter 64x256x64x16, 32x64 directions,
                      orders 8; FP64 throughout): T and E in full, S at 65536
                      texels (a seeded subset of a seeded 2 M-texel sample).
                      About 2.5 h of the oracle on 7 threads: run it alone,
                      `python tests/golden/make_golden.py precision_hires`.
  batch_classes.npz   one set of each inscatter resolution class of the 256-set
                      batch (seed 0x5EED0003 stream: 16x64x16x4, 32x128x32x8,
                      48x192x48x12; T 256x64, E 64x16, orders 4): T and E in full,
                      S at 4096 seeded texels, with each set's index in the batch
                      (`... make_golden.py batch_classes`, ~15 min on 7 threads)
  earth_ref_half.npz  the reference configuration with ATMO_FLAG_HALF_STORAGE (every
                      table rounded to FP16 as the reference's RGB16F / RGBA16F
                      textures): T and E in full, S at the 4096 texels of
                      earth_ref.npz (`... make_golden.py half_storage`, which also
                      prints how far the FP16-storage tables are from the FP64 ones)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "proland-4.0_amd"))

from oracle import oracle as O  # noqa: E402
from prolandatmo.params import SEED_BATCH, draw_sets, earth, options, with_res  # noqa: E402


def pvec(p):
    return np.frombuffer(bytes(p), dtype=np.uint8).copy()


def tiny_earth():
    p = with_res(earth(), 8, 32, 8, 4)
    p.TRANSMITTANCE_W, p.TRANSMITTANCE_H, p.SKY_W, p.SKY_H = 64, 16, 16, 8
    T, E, S = O.precompute(p, options())
    np.savez_compressed(os.path.join(HERE, "tiny_earth.npz"), params=pvec(p), T=T, E=E,
                        S=S.astype(np.float32))


def tiny_seeds():
    sets, feats = draw_sets(SEED_BATCH, 3, res=(6, 16, 6, 4))
    out = {}
    for i, p in enumerate(sets):
        p.TRANSMITTANCE_W, p.TRANSMITTANCE_H, p.SKY_W, p.SKY_H = 32, 8, 8, 4
        T, E, S = O.precompute(p, options())
        out["params%d" % i] = pvec(p)
        out["T%d" % i], out["E%d" % i], out["S%d" % i] = T, E, S.astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "tiny_seeds.npz"), **out)


def earth_ref():
    p = earth()
    t = time.time()
    T, E, S = O.precompute(p, options())
    print("earth reference config oracle: %.1f s on %d threads" % (time.time() - t, O.get_threads()))
    rng = np.random.default_rng(0x5EED0001)
    idx = rng.choice(S.shape[0] * S.shape[1] * S.shape[2], size=4096, replace=False)
    idx.sort()
    flat = S.reshape(-1, 4)
    np.savez_compressed(os.path.join(HERE, "earth_ref.npz"), params=pvec(p), T=T.astype(np.float32),
                        E=E, S_idx=idx.astype(np.int32), S_val=flat[idx])


def hires_params():
    p = earth(TRANSMITTANCE_W=512, TRANSMITTANCE_H=128, SKY_W=128, SKY_H=32,
              RES_R=64, RES_MU=256, RES_MU_S=64, RES_NU=16)
    return p, options(orders=8, spherical_samples=32)


def hires_sample(n4):
    """S texels of the precision fixture: 65536 of a seeded 2 M-texel sample."""
    idx = np.sort(np.random.default_rng(0x5EED0002).choice(n4, size=2_000_000, replace=False))
    sub = np.sort(np.random.default_rng(0x5EED0006).choice(idx.size, size=65536, replace=False))
    return idx[sub]


def write_hires(p, T, E, S_idx, S_val, S_max):
    np.savez_compressed(os.path.join(HERE, "precision_hires.npz"), params=pvec(p), T=T, E=E,
                        S_idx=S_idx.astype(np.int64), S_val=S_val, S_max=S_max)


def precision_hires():
    p, o = hires_params()
    t = time.time()
    T, E, S = O.precompute(p, o)
    print("precision-study oracle: %.1f s on %d threads" % (time.time() - t, O.get_threads()))
    flat = S.reshape(-1, 4)
    idx = hires_sample(flat.shape[0])
    write_hires(p, T, E, idx, flat[idx], np.abs(S).max(axis=(0, 1, 2)))


def half_storage():
    p = earth()
    t = time.time()
    T, E, S = O.precompute(p, options(flags=16))   # ATMO_FLAG_HALF_STORAGE
    print("earth reference config oracle, FP16 storage: %.1f s on %d threads"
          % (time.time() - t, O.get_threads()))
    g = np.load(os.path.join(HERE, "earth_ref.npz"))
    idx = g["S_idx"]
    flat = S.reshape(-1, 4)
    np.savez_compressed(os.path.join(HERE, "earth_ref_half.npz"), params=pvec(p), T=T.astype(np.float32),
                        E=E.astype(np.float32), S_idx=idx, S_val=flat[idx].astype(np.float32))
    T64, E64, S64 = O.precompute(p, options())

    def line(name, x, y):
        fl = 1e-4 * np.abs(y).max()
        e = np.abs(x - y) / np.maximum(np.abs(y), fl)
        print("%-4s max %.3e  mean %.3e  >1e-3 %5.1f%%  >1e-2 %5.1f%%"
              % (name, e.max(), e.mean(), 100 * (e > 1e-3).mean(), 100 * (e > 1e-2).mean()))

    print("FP16-storage tables against the FP64 tables (relative error floored at 1e-4 of the max):")
    line("T", T, T64)
    line("E", E, E64)
    for c in range(4):
        line("S." + "rgba"[c], S[..., c], S64[..., c])


def batch_classes():
    """One set of each inscatter resolution class of the 256-set batch (the
    first of its class in the seed 0x5EED0003 stream; T 256x64, E 64x16,
    orders 4): T and E in full, S at 4096 seeded texels each."""
    from prolandatmo.params import RES_CLASSES
    sets, _ = draw_sets(SEED_BATCH, 256)
    out = {}
    for c, res in enumerate(RES_CLASSES):
        i = next(k for k, q in enumerate(sets) if (q.RES_R, q.RES_MU, q.RES_MU_S, q.RES_NU) == res)
        p = sets[i]
        t = time.time()
        T, E, S = O.precompute(p, options())
        print("batch set %d (%dx%dx%dx%d) oracle: %.1f s on %d threads"
              % ((i,) + res + (time.time() - t, O.get_threads())), flush=True)
        flat = S.reshape(-1, 4)
        idx = np.sort(np.random.default_rng(0x5EED0007 + c).choice(flat.shape[0], size=4096, replace=False))
        out["index%d" % c] = np.int32(i)
        out["params%d" % c] = pvec(p)
        out["T%d" % c], out["E%d" % c] = T.astype(np.float32), E
        out["S_idx%d" % c], out["S_val%d" % c] = idx.astype(np.int32), flat[idx]
        out["S_max%d" % c] = np.abs(S).max(axis=(0, 1, 2))
    np.savez_compressed(os.path.join(HERE, "batch_classes.npz"), **out)


if __name__ == "__main__":
    which = sys.argv[1:] or ["tiny_earth", "tiny_seeds", "earth_ref"]
    for w in which:
        globals()[w]()
    print("ok")
