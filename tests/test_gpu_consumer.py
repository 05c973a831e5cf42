"""Parity of the GPU table consumer (include/atmo_consumer.h: skyRadiance,
inScattering, sunRadiance, skyIrradiance of
src/atmo/examples/atmo/atmosphereShader.glsl) with the FP64 oracle's
restatement, and the end-to-end check that tables produced by the HIP
precompute render like the oracle's tables.

Bars (floating point, prolandatmo.metrics: relative error floored at 1e-4 of
the largest magnitude):
  same tables, GPU consumer vs oracle consumer ............ 1e-4
    (inScattering subtracts two texture values, S(x) - T S(x0); near-equal
    terms leave FP32 cancellation noise of ~1e-7 of the terms, so its floor
    is 1e-2 of the largest value instead of 1e-4)
  GPU tables + GPU consumer vs oracle tables + oracle ..... 1e-3
  .raw files read back vs in-memory tables ................ bit-identical
"""
import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from prolandatmo import consumer as C  # noqa: E402
from prolandatmo import lib as L  # noqa: E402
from prolandatmo.metrics import summary  # noqa: E402
from prolandatmo.params import FLAG_NO_AUTOTUNE, SEED_BATCH, draw_sets, earth, options, with_res  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    e = L.Engine(0)
    yield e
    e.close()


def small(p=None, res=(8, 32, 8, 4)):
    q = with_res(p or earth(), *res)
    q.TRANSMITTANCE_W, q.TRANSMITTANCE_H, q.SKY_W, q.SKY_H = 64, 16, 16, 8
    return q


ATMOS = [small(), small(draw_sets(SEED_BATCH, 1, res=(8, 32, 8, 4))[0][0])]


def panorama(nth=24, nph=32):
    """unit directions on a zenith x azimuth grid (avoiding the exact poles)."""
    th = (np.arange(nth) + 0.5) / nth * math.pi
    ph = np.arange(nph) / nph * 2 * math.pi
    t, f = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(t) * np.cos(f), np.sin(t) * np.sin(f), np.cos(t)], -1).reshape(-1, 3)
    return d.astype(np.float32).astype(np.float64)


def cameras(p):
    Rg, Rt = float(p.Rg), float(p.Rt)
    return [np.array([0.0, 0.0, Rg + 0.05]), np.array([3.0, 0.0, Rg + 0.4 * (Rt - Rg)]),
            np.array([0.0, 20.0, Rt + 80.0])]


SUNS = [np.array([0.0, 0.6, 0.8]), np.array([0.8, 0.0, 0.05])]


def _assert(name, x, y, tol, floor=1e-4):
    s = summary(x, y, floor)
    print("%-12s max %.3e mean %.3e (unfloored %.3e)" % (name, s["max"], s["mean"], s["max_unfloored"]),
          flush=True)
    assert s["max"] <= tol, (name, s)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.parametrize("pi", range(len(ATMOS)))
def test_sky_radiance_matches_oracle_on_the_same_tables(eng, pi):
    p = ATMOS[pi]
    T, E, S = O.precompute(p)
    T32, S32 = T.astype(np.float32).astype(np.float64), S.astype(np.float32).astype(np.float64)
    dT, dS = _dev(T), _dev(S)
    dirs = panorama()
    for ci, cam in enumerate(cameras(p)):
        for si, sun in enumerate(SUNS):
            for shaft in (0.0, 0.5):
                v = C.view(cam, sun, shaft_width=shaft)
                rad, ext = C.sky_radiance(p, dT, dS, v, dirs)
                orad, oext = O.sky_radiance(p, T32, S32, cam, dirs, sun, shaft)
                _assert("sky c%d s%d" % (ci, si), rad[:, :3].double().cpu().numpy(), orad, 1e-4)
                _assert("ext c%d s%d" % (ci, si), ext[:, :3].double().cpu().numpy(), oext, 1e-4)


@pytest.mark.parametrize("pi", range(len(ATMOS)))
def test_in_scattering_matches_oracle_on_the_same_tables(eng, pi):
    p = ATMOS[pi]
    T, E, S = O.precompute(p)
    T32, S32 = T.astype(np.float32).astype(np.float64), S.astype(np.float32).astype(np.float64)
    dT, dS = _dev(T), _dev(S)
    rng = np.random.default_rng(5)
    Rg = float(p.Rg)
    for ci, cam in enumerate(cameras(p)[:2]):
        dirs = panorama(12, 16)
        dist = rng.uniform(0.5, 60.0, size=(len(dirs), 1))
        pts = cam + dirs * dist
        # keep the end points above the ground
        pts = np.where((np.linalg.norm(pts, axis=1) > Rg + 0.01)[:, None], pts, cam + dirs * 0.5)
        pts = pts.astype(np.float32).astype(np.float64)
        for si, sun in enumerate(SUNS):
            v = C.view(cam, sun)
            rad, ext = C.in_scattering(p, dT, dS, v, pts)
            orad, oext = O.in_scattering(p, T32, S32, cam, pts, sun)
            _assert("aerial c%d s%d" % (ci, si), rad[:, :3].double().cpu().numpy(), orad, 1e-4, 1e-2)
            _assert("ext c%d s%d" % (ci, si), ext[:, :3].double().cpu().numpy(), oext, 1e-4)


@pytest.mark.parametrize("pi", range(len(ATMOS)))
def test_sun_and_sky_irradiance_match_oracle(eng, pi):
    p = ATMOS[pi]
    T, E, S = O.precompute(p)
    Rg, Rt = float(p.Rg), float(p.Rt)
    r = np.linspace(Rg, Rt, 7)
    m = np.linspace(-0.3, 1.0, 27)
    rmus = np.stack(np.meshgrid(r, m, indexing="ij"), -1).reshape(-1, 2).astype(np.float32)
    sun, sky = C.sun_sky_irradiance(p, _dev(T), _dev(E), rmus)
    osun, osky = O.sun_sky_irradiance(p, T.astype(np.float32).astype(np.float64),
                                      E.astype(np.float32).astype(np.float64), rmus.astype(np.float64))
    _assert("sun", sun[:, :3].double().cpu().numpy(), osun, 1e-5)
    _assert("sky", sky[:, :3].double().cpu().numpy(), osky, 1e-5)


@pytest.mark.parametrize("pi", range(len(ATMOS)))
def test_gpu_tables_render_like_oracle_tables(eng, pi):
    """End to end: HIP precompute + GPU consumer against FP64 oracle tables +
    oracle consumer."""
    p = ATMOS[pi]
    (dT, dE, dS), = eng.precompute(p)
    torch.cuda.synchronize()
    T, E, S = O.precompute(p)
    dirs = panorama()
    for ci, cam in enumerate(cameras(p)):
        sun = SUNS[0]
        rad, ext = C.sky_radiance(p, dT, dS, C.view(cam, sun), dirs)
        orad, oext = O.sky_radiance(p, T, S, cam, dirs, sun)
        _assert("e2e sky c%d" % ci, rad[:, :3].double().cpu().numpy(), orad, 1e-3)
        _assert("e2e ext c%d" % ci, ext[:, :3].double().cpu().numpy(), oext, 1e-3)


def test_raw_files_render_like_the_in_memory_tables(eng, tmp_path):
    """The .raw files (RGB T, RGBA S with the ork trailer) read back and
    uploaded render bit-identically to the tables the engine produced."""
    p = ATMOS[0]
    d = str(tmp_path)
    L.preprocess(p, d)
    # preprocessAtmo's one-off precompute runs the default kernel variants
    (dT, dE, dS), = eng.precompute(p, options(flags=FLAG_NO_AUTOTUNE))
    torch.cuda.synchronize()
    t, _ = L.read_raw(os.path.join(d, "transmittance.raw"))
    s, _ = L.read_raw(os.path.join(d, "inscatter.raw"))
    fT = torch.from_numpy(t.reshape(dT.shape)).cuda()
    fS = torch.from_numpy(s.reshape(dS.shape)).cuda()
    v = C.view(cameras(p)[1], SUNS[0])
    dirs = panorama()
    a, ea = C.sky_radiance(p, dT, dS, v, dirs)
    b, eb = C.sky_radiance(p, fT, fS, v, dirs)
    torch.cuda.synchronize()
    assert torch.equal(a, b) and torch.equal(ea, eb)


def _descriptor(d, name, kind, source):
    """An ork texture descriptor of the form of src/demo/textures/atmo/*.xml:25."""
    axes = ("wraps", "wrapt", "wrapr")[: (3 if kind == "texture3D" else 2)]
    path = os.path.join(d, name + ".xml")
    with open(path, "w") as f:
        f.write('<?xml version="1.0" ?>\n<%s name="%s" source="%s" internalformat="RGBA16F" min="LINEAR" '
                'mag="LINEAR" %s/>\n' % (kind, name, source, " ".join('%s="CLAMP_TO_EDGE"' % a for a in axes)))
    return path


def test_rgba16f_uploads_through_the_descriptors_render_like_the_oracle(eng, tmp_path):
    """What the demo samples: the .raw files loaded through ork descriptors with
    internalformat RGBA16F (every value an IEEE binary16 value), LINEAR and
    CLAMP_TO_EDGE.  The GPU consumer on those tables against the oracle
    consumer on the same tables quantised by NumPy's binary16 conversion (bar
    as for the same tables above), and the quantised sky within 1e-2 of the
    FP32 tables' sky (the FP16 storage error the reference's textures add)."""
    p = ATMOS[0]
    d = str(tmp_path)
    L.preprocess(p, d, options(flags=FLAG_NO_AUTOTUNE))
    xT = _descriptor(d, "transmittance", "texture2D", "transmittance.raw")
    xE = _descriptor(d, "irradiance", "texture2D", "irradiance.raw")
    xS = _descriptor(d, "inscatter", "texture3D", "inscatter.raw")
    T16, dT = C.load_texture(xT)
    E16, dE = C.load_texture(xE)
    S16, dS = C.load_texture(xS)
    assert (dT["width"], dT["height"], dT["components"]) == (p.TRANSMITTANCE_W, p.TRANSMITTANCE_H, 3)
    assert (dS["width"], dS["height"], dS["depth"], dS["components"]) == (
        p.RES_MU_S * p.RES_NU, p.RES_MU, p.RES_R, 4)
    t32, _ = L.read_raw(os.path.join(d, "transmittance.raw"))
    s32, _ = L.read_raw(os.path.join(d, "inscatter.raw"))
    e32, _ = L.read_raw(os.path.join(d, "irradiance.raw"))
    # the loader's quantisation is binary16's
    for a, b in ((T16, t32), (E16, e32), (S16, s32)):
        assert np.array_equal(a.ravel().view(np.uint32), b.astype(np.float16).astype(np.float32).view(np.uint32))
    S16 = S16.reshape(p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 4)
    gT, gE, gS = _dev(T16), _dev(E16), _dev(S16)
    fT, fS = _dev(t32.reshape(T16.shape)), _dev(s32.reshape(S16.shape))
    T64, E64, S64 = (a.astype(np.float64) for a in (T16, E16, S16))
    dirs = panorama()
    for ci, cam in enumerate(cameras(p)):
        for si, sun in enumerate(SUNS):
            v = C.view(cam, sun)
            rad, ext = C.sky_radiance(p, gT, gS, v, dirs)
            orad, oext = O.sky_radiance(p, T64, S64, cam, dirs, sun)
            _assert("16F sky c%d s%d" % (ci, si), rad[:, :3].double().cpu().numpy(), orad, 1e-4)
            _assert("16F ext c%d s%d" % (ci, si), ext[:, :3].double().cpu().numpy(), oext, 1e-4)
            frad, _ = C.sky_radiance(p, fT, fS, v, dirs)
            _assert("16F vs 32F c%d s%d" % (ci, si), rad[:, :3].double().cpu().numpy(),
                    frad[:, :3].double().cpu().numpy(), 1e-2)
    cam = cameras(p)[1]
    pts = (cam + panorama(12, 16) * 5.0).astype(np.float32).astype(np.float64)
    v = C.view(cam, SUNS[0])
    rad, ext = C.in_scattering(p, gT, gS, v, pts)
    orad, oext = O.in_scattering(p, T64, S64, cam, pts, SUNS[0])
    _assert("16F aerial", rad[:, :3].double().cpu().numpy(), orad, 1e-4, 1e-2)
    r = np.linspace(float(p.Rg), float(p.Rt), 5)
    m = np.linspace(-0.3, 1.0, 13)
    rmus = np.stack(np.meshgrid(r, m, indexing="ij"), -1).reshape(-1, 2).astype(np.float32)
    sun, sky = C.sun_sky_irradiance(p, gT, gE, rmus)
    osun, osky = O.sun_sky_irradiance(p, T64, E64, rmus.astype(np.float64))
    _assert("16F sun", sun[:, :3].double().cpu().numpy(), osun, 1e-5)
    _assert("16F skyE", sky[:, :3].double().cpu().numpy(), osky, 1e-5)
