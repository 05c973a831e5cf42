"""ctypes mirror of the C ABI's parameter structs (include/atmo.h) and the
seeded parameter-set draws used by the batched and surrogate workloads.

``AtmoParams`` is byte-compatible with proland::AtmoParameters
(src/atmo/sources/proland/preprocess/atmo/PreprocessAtmo.h:44-67, 96 bytes);
``earth()`` returns the reference defaults (PreprocessAtmo.cpp:90-98).
"""
import ctypes
import math

import numpy as np


class Vec3f(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


class AtmoParams(ctypes.Structure):
    _fields_ = [
        ("Rg", ctypes.c_float), ("Rt", ctypes.c_float), ("RL", ctypes.c_float),
        ("TRANSMITTANCE_W", ctypes.c_int), ("TRANSMITTANCE_H", ctypes.c_int),
        ("SKY_W", ctypes.c_int), ("SKY_H", ctypes.c_int),
        ("RES_R", ctypes.c_int), ("RES_MU", ctypes.c_int),
        ("RES_MU_S", ctypes.c_int), ("RES_NU", ctypes.c_int),
        ("AVERAGE_GROUND_REFLECTANCE", ctypes.c_float), ("HR", ctypes.c_float),
        ("betaR", Vec3f), ("HM", ctypes.c_float),
        ("betaMSca", Vec3f), ("betaMEx", Vec3f), ("mieG", ctypes.c_float),
    ]

    def copy(self):
        c = AtmoParams()
        ctypes.pointer(c)[0] = self
        return c

    # table sizes (floats) in the .raw layouts
    def n_transmittance(self):
        return self.TRANSMITTANCE_W * self.TRANSMITTANCE_H * 3

    def n_irradiance(self):
        return self.SKY_W * self.SKY_H * 3

    def n_inscatter(self):
        return self.RES_R * self.RES_MU * self.RES_MU_S * self.RES_NU * 4

    def n_texels4d(self):
        return self.RES_R * self.RES_MU * self.RES_MU_S * self.RES_NU

    def shape_transmittance(self):
        return (self.TRANSMITTANCE_H, self.TRANSMITTANCE_W)

    def shape_irradiance(self):
        return (self.SKY_H, self.SKY_W)

    def shape_inscatter(self):
        return (self.RES_R, self.RES_MU, self.RES_NU * self.RES_MU_S)


class AtmoOptions(ctypes.Structure):
    _fields_ = [
        ("orders", ctypes.c_int), ("transmittance_samples", ctypes.c_int),
        ("inscatter_samples", ctypes.c_int), ("irradiance_samples", ctypes.c_int),
        ("spherical_samples", ctypes.c_int), ("flags", ctypes.c_int),
        ("texels_per_thread", ctypes.c_int), ("inscatterN_form", ctypes.c_int),
        ("reserved", ctypes.c_int * 8),
    ]


assert ctypes.sizeof(AtmoParams) == 96
assert ctypes.sizeof(AtmoOptions) == 64

FLAG_CHECK_NUMERICS = 1
FLAG_NO_GRAPH = 2
FLAG_TIME_PASSES = 4
FLAG_NO_AUTOTUNE = 8
FLAG_HALF_STORAGE = 16
FLAG_LINEAR_TRANSMITTANCE = 32
FLAG_LINEAR_INSCATTER = 64


def earth(**over):
    """Reference defaults, PreprocessAtmo.cpp:90-98 (betaMEx = 4.44e-3)."""
    p = AtmoParams()
    p.Rg, p.Rt, p.RL = 6360.0, 6420.0, 6421.0
    p.TRANSMITTANCE_W, p.TRANSMITTANCE_H = 256, 64
    p.SKY_W, p.SKY_H = 64, 16
    p.RES_R, p.RES_MU, p.RES_MU_S, p.RES_NU = 32, 128, 32, 8
    p.AVERAGE_GROUND_REFLECTANCE = 0.1
    p.HR = 8.0
    p.betaR = Vec3f(5.8e-3, 1.35e-2, 3.31e-2)
    p.HM = 1.2
    p.betaMSca = Vec3f(4e-3, 4e-3, 4e-3)
    p.betaMEx = Vec3f(4.44e-3, 4.44e-3, 4.44e-3)
    p.mieG = 0.8
    for k, v in over.items():
        if k in ("betaR", "betaMSca", "betaMEx") and not isinstance(v, Vec3f):
            v = Vec3f(*v)
        setattr(p, k, v)
    return p


def options(orders=4, transmittance_samples=500, inscatter_samples=50,
            irradiance_samples=32, spherical_samples=16, flags=0, texels_per_thread=0,
            inscatterN_form=0):
    """Reference sample counts (common.glsl:94-97) and order count (PreprocessAtmo.cpp:351);
    texels_per_thread and inscatterN_form pin kernel variants (0 = default / autotuned)."""
    o = AtmoOptions()
    o.texels_per_thread = texels_per_thread
    o.inscatterN_form = inscatterN_form
    o.orders = orders
    o.transmittance_samples = transmittance_samples
    o.inscatter_samples = inscatter_samples
    o.irradiance_samples = irradiance_samples
    o.spherical_samples = spherical_samples
    o.flags = flags
    return o


def with_res(p, R, MU, MUS, NU):
    q = p.copy()
    q.RES_R, q.RES_MU, q.RES_MU_S, q.RES_NU = R, MU, MUS, NU
    return q


# ---------------------------------------------------------------------------
# Seeded parameter-set draws (counter-based splitmix64 stream).
# ---------------------------------------------------------------------------
_M64 = (1 << 64) - 1
EARTH_BETA_R = (5.8e-3, 1.35e-2, 3.31e-2)
RES_CLASSES = ((16, 64, 16, 4), (32, 128, 32, 8), (48, 192, 48, 12))

SEED_BATCH = 0x5EED0003
SEED_TRAIN = 0x5EED0004
SEED_INFER = 0x5EED0005


def splitmix64(state):
    """Returns (next_state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & _M64

    def u01(self):
        self.s, z = splitmix64(self.s)
        return (z >> 11) * (1.0 / (1 << 53))

    def uniform(self, a, b):
        return a + (b - a) * self.u01()

    def loguniform(self, a, b):
        return math.exp(self.uniform(math.log(a), math.log(b)))

    def choice(self, n):
        return min(int(self.u01() * n), n - 1)


def draw_set(rng, res=None):
    """One atmosphere drawn from the batch distribution: Rg~U[3000,7000] km,
    Rt-Rg~U[40,120], RL=Rt+1, HR~U[5,12], betaR=Earth*U[0.3,3], HM~U[0.6,2.4],
    betaMSca~logU[1e-3,3e-2] (grey), betaMEx=betaMSca/0.9, mieG~U[0.6,0.9],
    ground reflectance 0.1, inscatter resolution from RES_CLASSES unless given.
    Returns (params, features) with the 7 raw features
    (Rg, Rt-Rg, HR, betaR scale, HM, log betaMSca, mieG)."""
    Rg = rng.uniform(3000.0, 7000.0)
    dR = rng.uniform(40.0, 120.0)
    HR = rng.uniform(5.0, 12.0)
    sR = rng.uniform(0.3, 3.0)
    HM = rng.uniform(0.6, 2.4)
    bM = rng.loguniform(1e-3, 3e-2)
    g = rng.uniform(0.6, 0.9)
    rc = rng.choice(len(RES_CLASSES))
    p = earth()
    p.Rg = Rg
    p.Rt = Rg + dR
    p.RL = p.Rt + 1.0
    p.HR = HR
    p.betaR = Vec3f(*(b * sR for b in EARTH_BETA_R))
    p.HM = HM
    p.betaMSca = Vec3f(bM, bM, bM)
    p.betaMEx = Vec3f(bM / 0.9, bM / 0.9, bM / 0.9)
    p.mieG = g
    R, MU, MUS, NU = res if res is not None else RES_CLASSES[rc]
    p.RES_R, p.RES_MU, p.RES_MU_S, p.RES_NU = R, MU, MUS, NU
    feats = np.array([Rg, dR, HR, sR, HM, math.log(bM), g], dtype=np.float64)
    return p, feats


def draw_sets(seed, n, res=None):
    rng = SplitMix64(seed)
    out = [draw_set(rng, res) for _ in range(n)]
    return [o[0] for o in out], np.stack([o[1] for o in out])


def params_array(plist):
    arr = (AtmoParams * len(plist))()
    for i, p in enumerate(plist):
        arr[i] = p
    return arr
