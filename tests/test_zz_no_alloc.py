"""The allocation-free precompute (atmo_precompute_noalloc) against the engine's
batched precompute (atmo_engine_precompute), through the C ABI.

Both run the same kernels; the allocation-free entry point takes its work
tables from the library's static device array instead of hipMalloc, so the
bar is bit-identical tables.  Two calls in a row, the second with a smaller
batch (it runs on the array the first call left behind), then a batch too
large for the array, which is rejected before any device work.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from prolandatmo import lib as L  # noqa: E402
from prolandatmo.params import (FLAG_CHECK_NUMERICS, FLAG_NO_AUTOTUNE, SEED_BATCH,  # noqa: E402
                                draw_sets, earth, options, with_res)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    e = L.Engine(0)
    yield e
    e.close()


def nan_outputs(sets):
    outs = []
    for p in sets:
        outs.append(tuple(t.fill_(float("nan")) for t in L.alloc_outputs(p)))
    return outs


def assert_same(a, b, what):
    for name, x, y in zip("TES", a, b):
        x = x.cpu().numpy()
        y = y.cpu().numpy()
        assert np.isfinite(x).all(), (what, name)
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (what, name)


def test_noalloc_two_calls_match_engine(eng):
    # call 1: Earth at the reference configuration (48.4 MiB of work memory),
    # after an engine call that tuned this shape class: same variants, same bits
    big = [earth()]
    o1 = options(flags=FLAG_CHECK_NUMERICS)
    ref1 = eng.precompute(big, o1)
    got1 = L.precompute_noalloc(big, o1, outs=nan_outputs(big))
    torch.cuda.synchronize()
    print("call 1: Earth 32x128x32x8, tuning", eng.tuning()["tuned"], flush=True)
    assert_same(got1[0], ref1[0], "call 1")

    # call 2: a smaller batch of two seeded sets with the default variants
    small, _ = draw_sets(SEED_BATCH, 2, res=(16, 64, 16, 4))
    for p in small:
        p.TRANSMITTANCE_W, p.TRANSMITTANCE_H, p.SKY_W, p.SKY_H = 128, 32, 32, 8
    o2 = options(flags=FLAG_CHECK_NUMERICS | FLAG_NO_AUTOTUNE)
    ref2 = eng.precompute(small, o2)
    got2 = L.precompute_noalloc(small, o2, outs=nan_outputs(small))
    torch.cuda.synchronize()
    print("call 2: 2 sets at 16x64x16x4", flush=True)
    for i in range(2):
        assert_same(got2[i], ref2[i], "call 2 set %d" % i)


def test_noalloc_rejects_a_batch_beyond_its_array():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    p = with_res(earth(), 48, 192, 48, 12)   # 5.3 M texels: ~230 MiB of work tables
    with pytest.raises(L.AtmoError) as ei:
        L.precompute_noalloc([p], options())
    assert ei.value.status == 1 and "ATMO_NOALLOC_MAX_BYTES" in str(ei.value)
    # two reference-configuration sets do not fit either (2 x 48.4 MiB)
    with pytest.raises(L.AtmoError) as ei:
        L.precompute_noalloc([earth(), earth()], options())
    assert ei.value.status == 1
