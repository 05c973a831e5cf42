"""CPU tests of the C ABI library: it loads, exports every symbol include/*.h
declares, validates its inputs, and handles .raw files (no GPU compute)."""
import ctypes
import os
import re

import numpy as np
import pytest

from prolandatmo import lib as L
from prolandatmo.params import AtmoOptions, AtmoParams, earth, options

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    src = open(os.path.join(REPO, "include", header)).read()
    return set(re.findall(r"ATMO_API\s+[^;(]*?\b(atmo_\w+)\s*\(", src, re.S))


def test_library_exports_every_declared_symbol():
    from prolandatmo import surrogate as SG
    names = declared("atmo.h")
    assert len(names) >= 20
    lib = L.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert names == set(L.SIGNATURES), "binding and header disagree"
    snames = declared("atmo_surrogate.h")
    assert snames == set(SG.SIGNATURES)
    for n in snames:
        assert hasattr(lib, n), n
    assert SG.param_count() == 14 * 256 + 256 + (SG.NL - 1) * (256 * 256 + 256) + 256 * 4 + 4
    from prolandatmo import consumer as C
    cnames = declared("atmo_consumer.h")
    assert cnames == set(C.SIGNATURES) and len(cnames) == 5
    for n in cnames:
        assert hasattr(lib, n), n
    syms = os.popen("nm -D --defined-only %s" % L.LIB_PATH).read()
    for mangled in ("_ZN7proland14preprocessAtmoERKNS_14AtmoParametersEPKc",
                    "_ZN7proland14AtmoParametersC1Ev", "_ZN7proland20preprocessAtmoStatusEv"):
        assert mangled in syms


def test_params_layout_and_defaults():
    assert ctypes.sizeof(AtmoParams) == 96
    p = AtmoParams()
    L.lib().atmo_params_default(p)
    assert bytes(p) == bytes(earth())
    o = AtmoOptions()
    L.lib().atmo_options_default(o)
    assert (o.orders, o.transmittance_samples, o.inscatter_samples, o.irradiance_samples,
            o.spherical_samples) == (4, 500, 50, 32, 16)
    assert L.lib().atmo_version().startswith(b"proland-atmo")


def test_invalid_parameters_are_rejected_before_any_device_call():
    bad = earth()
    bad.Rt = bad.Rg - 1
    with pytest.raises(L.AtmoError) as ei:
        L.precompute_host(bad)
    assert ei.value.status == 1 and "Rg" in str(ei.value)
    odd = earth(RES_MU=127)
    with pytest.raises(L.AtmoError):
        L.precompute_host(odd)
    st = L.lib().atmo_precompute_host(None, None, None, None, None)
    assert st == 1


def test_raw_round_trip_and_trailer(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.random((4, 6, 10, 4)).astype(np.float32)  # depth 4, height 6, width 10, RGBA
    path = str(tmp_path / "inscatter.raw")
    L.write_raw(path, a, 10, 6, 4, 4)
    raw = open(path, "rb").read()
    assert len(raw) == a.size * 4 + 20
    tr = np.frombuffer(raw[-20:], dtype=np.int32)
    assert tr[0] == np.int32(np.uint32(0xCAFEBABE).view(np.int32))
    assert list(tr[1:]) == [10, 24, 4, 4]  # width, height*depth, depth, components
    assert np.array_equal(np.frombuffer(raw[:-20], dtype=np.float32), a.ravel())
    data, h = L.read_raw(path)
    assert np.array_equal(data, a.ravel()) and list(h[1:]) == [10, 24, 4, 4]
    # 2D RGB table (transmittance/irradiance convention: depth 0)
    t = rng.random((16, 64, 3)).astype(np.float32)
    p2 = str(tmp_path / "t.raw")
    L.write_raw(p2, t, 64, 16, 0, 3)
    data, h = L.read_raw(p2)
    assert list(h[1:]) == [64, 16, 0, 3] and np.array_equal(data, t.ravel())
    # truncated file is rejected
    open(p2, "r+b").truncate(100)
    with pytest.raises(L.AtmoError):
        L.read_raw(p2)


def test_workspace_size_counts_work_tables():
    p = earth()
    n4 = 32 * 128 * 32 * 8
    b = L.workspace_size(p, options())
    # dSR, dSM (float4) + dJ (packed rgb) + T, dE, E
    assert b >= 2 * n4 * 16 + n4 * 12 and b < 3 * n4 * 16
    b1 = L.workspace_size(p, options(orders=1))
    assert b1 < b


@pytest.mark.parametrize("bad", [dict(texels_per_thread=5), dict(texels_per_thread=-1),
                                 dict(inscatterN_form=4), dict(orders=0), dict(spherical_samples=129),
                                 dict(irradiance_samples=7)])
def test_invalid_options_are_rejected_before_any_device_call(bad):
    with pytest.raises(L.AtmoError) as ei:
        L.workspace_size(earth(), options(**bad))
    assert ei.value.status == 1


@pytest.mark.parametrize("good", [dict(texels_per_thread=3), dict(inscatterN_form=3),
                                  dict(flags=16), dict(spherical_samples=128, irradiance_samples=256)])
def test_option_ranges_accepted(good):
    assert L.workspace_size(earth(), options(**good)) > 0


def test_engine_create_reports_missing_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    st = L.lib().atmo_engine_create(ctypes.byref(h), 0)
    assert st == 2 and L.lib().atmo_last_error()


def test_consumer_validates_arguments_before_any_device_call():
    from prolandatmo import consumer as C
    lib = C.lib()
    v = C.view([0, 0, 6361.0], [0, 0, 1])
    p = earth()
    # n = 0 is a no-op that needs no device
    assert lib.atmo_sky_radiance(p, None, None, v, None, 0, None, None, None) == 0
    assert lib.atmo_sun_sky_irradiance(p, None, None, None, 0, None, None, None) == 0
    # null tables with work to do, a missing view, bad parameters, bad units
    assert lib.atmo_sky_radiance(p, None, None, v, None, 4, None, None, None) == 1
    assert lib.atmo_in_scattering(p, None, None, None, None, 4, None, None, None) == 1
    bad = earth()
    bad.betaR.y = 0.0
    assert lib.atmo_sky_radiance(bad, None, None, v, None, 0, None, None, None) == 1
    assert b"betaR" in L.lib().atmo_last_error()
    v0 = C.view([0, 0, 6361.0], [0, 0, 1], meters_per_unit=0.0)
    assert lib.atmo_in_scattering(p, None, None, v0, None, 0, None, None, None) == 1
    vc = C.view([0, 0, 0], [0, 0, 1])
    assert lib.atmo_sky_radiance(p, None, None, vc, None, 0, None, None, None) == 1


def test_empty_and_null_batches_are_rejected():
    """An empty batch (nsets 0 or negative) or missing pointers is an
    ATMO_ERR_INVALID_ARGUMENT status, before any device work; so is a batch
    whose second set is invalid (nothing of the first is computed either)."""
    lib = L.lib()
    one = L.params_array([earth()])
    n = ctypes.c_size_t()
    o = AtmoOptions()
    lib.atmo_options_default(o)
    assert lib.atmo_engine_workspace_size(one, 0, ctypes.byref(o), ctypes.byref(n)) == 1
    assert lib.atmo_engine_workspace_size(one, -3, ctypes.byref(o), ctypes.byref(n)) == 1
    assert lib.atmo_engine_workspace_size(None, 1, ctypes.byref(o), ctypes.byref(n)) == 1
    assert lib.atmo_engine_workspace_size(one, 1, ctypes.byref(o), None) == 1
    assert lib.atmo_engine_precompute(None, one, 1, ctypes.byref(o), None, None) == 1
    bad = earth()
    bad.RES_MU_S = 1
    with pytest.raises(L.AtmoError) as ei:
        L.workspace_size([earth(), bad])
    assert ei.value.status == 1
