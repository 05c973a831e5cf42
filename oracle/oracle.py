"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle_atmo.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product library.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

from prolandatmo.params import AtmoOptions, AtmoParams, options

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_atmo.so")

_dp = ctypes.POINTER(ctypes.c_double)
_pp = ctypes.POINTER(AtmoParams)
_po = ctypes.POINTER(AtmoOptions)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def load(path):
    """Loads an oracle library built from oracle/atmo_oracle.cpp (the default
    one, or a mutated copy a test builds) with every signature declared."""
    L = ctypes.CDLL(path)
    sig = {
        "oracle_layer": (None, [_pp, ctypes.c_int, _dp, _dp]),
        "oracle_limit": (ctypes.c_double, [_pp, ctypes.c_double, ctypes.c_double]),
        "oracle_phaseR": (ctypes.c_double, [ctypes.c_double]),
        "oracle_phaseM": (ctypes.c_double, [_pp, ctypes.c_double]),
        "oracle_transmittance_uv": (None, [_pp, _po, ctypes.c_double, ctypes.c_double, _dp]),
        "oracle_transmittance_rmu": (None, [_pp, _po, ctypes.c_int, ctypes.c_int, _dp]),
        "oracle_irradiance_rmus": (None, [_pp, ctypes.c_int, ctypes.c_int, _dp]),
        "oracle_mumusnu": (None, [_pp, _po, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_int, _dp]),
        "oracle_texture4D_coords": (None, [_pp, _po] + [ctypes.c_double] * 4 + [_dp]),
        "oracle_texture4D_coords_literal": (None, [_pp, _po] + [ctypes.c_double] * 4 + [_dp]),
        "oracle_sample2D": (None, [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double, _dp]),
        "oracle_sample3D": (None, [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]),
        "oracle_texture4D": (None, [_pp, _po, _dp, ctypes.c_int] + [ctypes.c_double] * 4 + [_dp]),
        "oracle_transmittance_lookup": (None, [_pp, _dp, ctypes.c_double, ctypes.c_double, _dp]),
        "oracle_transmittance_segment": (None, [_pp, _dp, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, _dp]),
        "oracle_transmittance_segment_tau": (None, [_pp, _po, _dp, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_double, _dp]),
        "oracle_optical_depth_table": (None, [_pp, _po, _dp]),
        "oracle_optical_depth": (ctypes.c_double, [_pp, _po, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_double]),
        "oracle_pass_transmittance": (None, [_pp, _po, _dp]),
        "oracle_pass_irradiance1": (None, [_pp, _po, _dp, _dp]),
        "oracle_pass_inscatter1": (None, [_pp, _po, _dp, _dp, _dp]),
        "oracle_pass_copy_inscatter1": (None, [_pp, _dp, _dp, _dp]),
        "oracle_pass_inscatterS": (None, [_pp, _po, _dp, _dp, _dp, _dp, ctypes.c_int, _dp]),
        "oracle_pass_irradianceN": (None, [_pp, _po, _dp, _dp, ctypes.c_int, _dp]),
        "oracle_pass_inscatterN": (None, [_pp, _po, _dp, _dp, _dp]),
        "oracle_pass_copy_inscatterN": (None, [_pp, _dp, _dp]),
        "oracle_pass_copy_irradiance": (None, [_pp, ctypes.c_double, _dp, _dp]),
        "oracle_precompute": (None, [_pp, _po, _dp, _dp, _dp]),
        "oracle_q16": (ctypes.c_double, [ctypes.c_double]),
        "oracle_sky_radiance": (None, [_pp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp]),
        "oracle_in_scattering": (None, [_pp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double,
                                        ctypes.c_double, _dp, _dp]),
        "oracle_sun_sky_irradiance": (None, [_pp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                             _dp, _dp]),
        "oracle_set_threads": (None, [ctypes.c_int]),
        "oracle_get_threads": (ctypes.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = load(LIB)
    return _lib


@contextlib.contextmanager
def using(path):
    """Routes every function of this module through the library at path
    (a mutated oracle built by a test) for the duration of the block."""
    global _lib
    saved = lib()
    _lib = load(path)
    try:
        yield
    finally:
        _lib = saved


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def set_threads(n):
    lib().oracle_set_threads(int(n))


def get_threads():
    return lib().oracle_get_threads()


# ---- tables -------------------------------------------------------------
def zeros_T(p):
    return np.zeros((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3))


def zeros_E(p):
    return np.zeros((p.SKY_H, p.SKY_W, 3))


def zeros_4D(p, c=3):
    return np.zeros((p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, c))


def transmittance(p, o=None):
    o = o or options()
    T = zeros_T(p)
    lib().oracle_pass_transmittance(p, o, _d(T))
    return T


def irradiance1(p, T, o=None):
    o = o or options()
    dE = zeros_E(p)
    lib().oracle_pass_irradiance1(p, o, _d(T), _d(dE))
    return dE


def inscatter1(p, T, o=None):
    o = o or options()
    dSR, dSM = zeros_4D(p), zeros_4D(p)
    lib().oracle_pass_inscatter1(p, o, _d(T), _d(dSR), _d(dSM))
    return dSR, dSM


def copy_inscatter1(p, dSR, dSM):
    S = zeros_4D(p, 4)
    lib().oracle_pass_copy_inscatter1(p, _d(dSR), _d(dSM), _d(S))
    return S


def inscatterS(p, T, dE, dSR, dSM, first, o=None):
    o = o or options()
    dJ = zeros_4D(p)
    lib().oracle_pass_inscatterS(p, o, _d(T), _d(dE), _d(dSR), _d(dSM), int(first), _d(dJ))
    return dJ


def irradianceN(p, dSR, dSM, first, o=None):
    o = o or options()
    dE = zeros_E(p)
    lib().oracle_pass_irradianceN(p, o, _d(dSR), _d(dSM), int(first), _d(dE))
    return dE


def inscatterN(p, T, dJ, o=None):
    o = o or options()
    dS = zeros_4D(p)
    lib().oracle_pass_inscatterN(p, o, _d(T), _d(dJ), _d(dS))
    return dS


def copy_inscatterN(p, dS, S):
    S = np.ascontiguousarray(S.copy())
    lib().oracle_pass_copy_inscatterN(p, _d(dS), _d(S))
    return S


def precompute(p, o=None):
    """Whole FP64 precompute -> (T, E, S) in .raw layouts."""
    o = o or options()
    T, E, S = zeros_T(p), zeros_E(p), zeros_4D(p, 4)
    lib().oracle_precompute(p, o, _d(T), _d(E), _d(S))
    return T, E, S


def q16(x):
    """Nearest FP16 value of each element (oracle_q16), as float64."""
    a = np.asarray(x, dtype=np.float64)
    f = lib().oracle_q16
    return np.array([f(float(v)) for v in a.ravel()]).reshape(a.shape)


def precompute_stages(p, o=None):
    """Whole precompute returning every intermediate table:
    dict with T, dE1, dSR1, dSM1, and per order n>=2: dJ{n}, dE{n}, dS{n}; plus E, S."""
    o = o or options()
    out = {}
    T = transmittance(p, o)
    dE = irradiance1(p, T, o)
    dSR, dSM = inscatter1(p, T, o)
    out.update(T=T, dE1=dE, dSR1=dSR, dSM1=dSM)
    E = zeros_E(p)
    S = copy_inscatter1(p, dSR, dSM)
    for n in range(2, o.orders + 1):
        first = n == 2
        dJ = inscatterS(p, T, dE, dSR, dSM, first, o)
        dE = irradianceN(p, dSR, dSM, first, o)
        dSR = inscatterN(p, T, dJ, o)
        E = E + dE
        S = copy_inscatterN(p, dSR, S)
        out["dJ%d" % n] = dJ
        out["dE%d" % n] = dE
        out["dS%d" % n] = dSR
    out["E"] = E
    out["S"] = S
    return out


# ---- primitives ----------------------------------------------------------
def layer(p, l):
    r = ctypes.c_double()
    d = np.zeros(4)
    lib().oracle_layer(p, l, ctypes.byref(r), _d(d))
    return r.value, d


def limit(p, r, mu):
    return lib().oracle_limit(p, r, mu)


def phaseR(mu):
    return lib().oracle_phaseR(mu)


def phaseM(p, mu):
    return lib().oracle_phaseM(p, mu)


def _opt(o):
    return ctypes.byref(o) if o is not None else None


def transmittance_uv(p, r, mu, o=None):
    out = np.zeros(2)
    lib().oracle_transmittance_uv(p, _opt(o), r, mu, _d(out))
    return out


def transmittance_rmu(p, i, j, o=None):
    out = np.zeros(2)
    lib().oracle_transmittance_rmu(p, _opt(o), i, j, _d(out))
    return out


def irradiance_rmus(p, i, j):
    o = np.zeros(2)
    lib().oracle_irradiance_rmus(p, i, j, _d(o))
    return o


def mumusnu(p, r, dhdH, x, y, o=None):
    out = np.zeros(3)
    lib().oracle_mumusnu(p, _opt(o), r, _d(np.ascontiguousarray(dhdH, dtype=np.float64)), x, y,
                         _d(out))
    return out


def texture4D_coords(p, r, mu, muS, nu, literal=False, o=None):
    out = np.zeros(4)
    f = lib().oracle_texture4D_coords_literal if literal else lib().oracle_texture4D_coords
    f(p, _opt(o), r, mu, muS, nu, _d(out))
    return out


def sample2D(tex, s, t):
    tex = np.ascontiguousarray(tex, dtype=np.float64)
    H, W, C = tex.shape
    o = np.zeros(C)
    lib().oracle_sample2D(_d(tex), W, H, C, s, t, _d(o))
    return o


def sample3D(tex, s, t, u):
    tex = np.ascontiguousarray(tex, dtype=np.float64)
    D, H, W, C = tex.shape
    o = np.zeros(C)
    lib().oracle_sample3D(_d(tex), W, H, D, C, s, t, u, _d(o))
    return o


def texture4D(p, tab, r, mu, muS, nu, o=None):
    tab = np.ascontiguousarray(tab, dtype=np.float64)
    C = tab.shape[-1]
    out = np.zeros(C)
    lib().oracle_texture4D(p, _opt(o), _d(tab), C, r, mu, muS, nu, _d(out))
    return out


def transmittance_lookup(p, T, r, mu):
    o = np.zeros(3)
    lib().oracle_transmittance_lookup(p, _d(T), r, mu, _d(o))
    return o


def transmittance_segment(p, T, r, mu, d):
    o = np.zeros(3)
    lib().oracle_transmittance_segment(p, _d(T), r, mu, d, _d(o))
    return o


def optical_depth_table(p, o=None):
    """-ln T of every transmittance texel, (TH, TW, 3); finite where T underflows"""
    tau = np.zeros((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3))
    lib().oracle_optical_depth_table(p, _opt(o), _d(tau))
    return tau


def transmittance_segment_tau(p, tau, r, mu, d, o=None):
    """the segment ratio as oracle precompute evaluates it, from optical depths"""
    out = np.zeros(3)
    lib().oracle_transmittance_segment_tau(p, _opt(o), _d(tau), r, mu, d, _d(out))
    return out


def optical_depth(p, H, r, mu, o=None):
    o = o or options()
    return lib().oracle_optical_depth(p, o, H, r, mu)


# ---------------------------------------------------------------- table consumer
def _v3(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(3))


def _tabs(*t):
    return [np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in t]


def sky_radiance(p, T, S, camera, dirs, sun, shaft_width=0.0):
    """atmosphereShader.glsl:371-418 skyRadiance for each row of dirs (n, 3)
    -> (radiance (n, 3), extinction (n, 3))."""
    T, S = _tabs(T, S)
    dirs = np.asarray(dirs, dtype=np.float64).reshape(-1, 3)
    rad, ext = np.zeros((len(dirs), 3)), np.zeros((len(dirs), 3))
    cam, sn = _v3(camera), _v3(sun)
    for i, d in enumerate(dirs):
        ro, eo = np.zeros(3), np.zeros(3)
        lib().oracle_sky_radiance(p, _d(T), _d(S), _d(cam), _d(_v3(d)), _d(sn), shaft_width,
                                  _d(ro), _d(eo))
        rad[i], ext[i] = ro, eo
    return rad, ext


def in_scattering(p, T, S, camera, points, sun, shaft_width=0.0, meters_per_unit=1000.0):
    """atmosphereShader.glsl:425-533 inScattering from camera to each row of
    points (n, 3) -> (radiance (n, 3), extinction (n, 3))."""
    T, S = _tabs(T, S)
    points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    rad, ext = np.zeros((len(points), 3)), np.zeros((len(points), 3))
    cam, sn = _v3(camera), _v3(sun)
    for i, q in enumerate(points):
        ro, eo = np.zeros(3), np.zeros(3)
        lib().oracle_in_scattering(p, _d(T), _d(S), _d(cam), _d(_v3(q)), _d(sn), shaft_width,
                                   meters_per_unit, _d(ro), _d(eo))
        rad[i], ext[i] = ro, eo
    return rad, ext


def sun_sky_irradiance(p, T, E, rmus):
    """sunRadiance (:342-350) and skyIrradiance (:355-361) for rows (r, muS)."""
    T, E = _tabs(T, E)
    rmus = np.asarray(rmus, dtype=np.float64).reshape(-1, 2)
    sun, sky = np.zeros((len(rmus), 3)), np.zeros((len(rmus), 3))
    for i, (r, m) in enumerate(rmus):
        a, b = np.zeros(3), np.zeros(3)
        lib().oracle_sun_sky_irradiance(p, _d(T), _d(E), r, m, _d(a), _d(b))
        sun[i], sky[i] = a, b
    return sun, sky
