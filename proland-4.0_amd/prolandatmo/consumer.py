"""ctypes binding of the table consumer (include/atmo_consumer.h): skyRadiance,
inScattering, sunRadiance and skyIrradiance of
src/atmo/examples/atmo/atmosphereShader.glsl evaluated on the GPU from the
.raw-layout tables this library writes.  Arrays are torch tensors on the HIP
device."""
import ctypes

from .lib import _ptr, _stream, check, lib as _lib
from .params import AtmoParams

_vp = ctypes.c_void_p
_pp = ctypes.POINTER(AtmoParams)


class AtmoView(ctypes.Structure):
    _fields_ = [("camera", ctypes.c_double * 3), ("sun", ctypes.c_double * 3),
                ("shaft_width", ctypes.c_double), ("meters_per_unit", ctypes.c_double)]


assert ctypes.sizeof(AtmoView) == 64


class AtmoTextureDesc(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("source", ctypes.c_char * 512), ("format", ctypes.c_char * 16),
                ("kind", ctypes.c_int), ("format_channels", ctypes.c_int), ("half", ctypes.c_int),
                ("min_linear", ctypes.c_int), ("mag_linear", ctypes.c_int), ("clamp_to_edge", ctypes.c_int * 3),
                ("width", ctypes.c_int), ("height", ctypes.c_int), ("depth", ctypes.c_int),
                ("components", ctypes.c_int)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        for k in ("name", "source", "format"):
            d[k] = d[k].decode()
        d["clamp_to_edge"] = list(self.clamp_to_edge)
        return d

# every symbol include/atmo_consumer.h declares, with its signature
SIGNATURES = {
    "atmo_sky_radiance": (ctypes.c_int, [_pp, _vp, _vp, ctypes.POINTER(AtmoView), _vp, ctypes.c_int,
                                         _vp, _vp, _vp]),
    "atmo_in_scattering": (ctypes.c_int, [_pp, _vp, _vp, ctypes.POINTER(AtmoView), _vp, ctypes.c_int,
                                          _vp, _vp, _vp]),
    "atmo_sun_sky_irradiance": (ctypes.c_int, [_pp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp]),
    "atmo_texture_load_desc": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(AtmoTextureDesc)]),
    "atmo_texture_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t,
                                         ctypes.POINTER(AtmoTextureDesc)]),
}

_bound = False


def lib():
    global _bound
    L = _lib()
    if not _bound:
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _bound = True
    return L


def view(camera, sun, shaft_width=0.0, meters_per_unit=1000.0):
    v = AtmoView()
    for i in range(3):
        v.camera[i] = float(camera[i])
        v.sun[i] = float(sun[i])
    v.shaft_width = float(shaft_width)
    v.meters_per_unit = float(meters_per_unit)
    return v


def _vec4(x):
    """(n, 3) or (n, 4) array/tensor -> contiguous float32 (n, 4) device tensor."""
    import torch
    t = torch.as_tensor(x, dtype=torch.float32)
    if t.shape[-1] == 3:
        t = torch.nn.functional.pad(t, (0, 1))
    return t.reshape(-1, 4).contiguous().to("cuda")


def sky_radiance(p, T, S, v, dirs, stream=None):
    """skyRadiance along dirs (n, 3) -> (radiance (n, 4), extinction (n, 4)) device tensors."""
    import torch
    d = _vec4(dirs)
    n = d.shape[0]
    rad = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    ext = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    check(lib().atmo_sky_radiance(p, _ptr(T), _ptr(S), ctypes.byref(v), _ptr(d), n, _ptr(rad),
                                  _ptr(ext), _stream(stream)))
    return rad, ext


def in_scattering(p, T, S, v, points, stream=None):
    """inScattering from the view's camera to points (n, 3) -> (radiance, extinction)."""
    import torch
    q = _vec4(points)
    n = q.shape[0]
    rad = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    ext = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    check(lib().atmo_in_scattering(p, _ptr(T), _ptr(S), ctypes.byref(v), _ptr(q), n, _ptr(rad),
                                   _ptr(ext), _stream(stream)))
    return rad, ext


def sun_sky_irradiance(p, T, E, rmus, stream=None):
    """sunRadiance and skyIrradiance at rows (r, muS) -> (sun (n, 4), sky (n, 4))."""
    import torch
    a = torch.as_tensor(rmus, dtype=torch.float32).reshape(-1, 2).contiguous().to("cuda")
    n = a.shape[0]
    sun = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    sky = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    check(lib().atmo_sun_sky_irradiance(p, _ptr(T), _ptr(E), _ptr(a), n, _ptr(sun), _ptr(sky),
                                        _stream(stream)))
    return sun, sky


def texture_desc(xml_path):
    """atmo_texture_load_desc: the descriptor's attributes and its source's sizes."""
    d = AtmoTextureDesc()
    check(lib().atmo_texture_load_desc(xml_path.encode(), ctypes.byref(d)))
    return d.as_dict()


def load_texture(xml_path):
    """atmo_texture_load: (payload as the GPU texture stores it, numpy float32 in
    the .raw layout [depth][height][width][components] or [height][width][components],
    descriptor dict)."""
    import numpy as np
    d = texture_desc(xml_path)
    shape = ((d["depth"],) if d["depth"] > 0 else ()) + (d["height"], d["width"], d["components"])
    a = np.zeros(int(np.prod(shape)), dtype=np.float32)
    out = AtmoTextureDesc()
    check(lib().atmo_texture_load(xml_path.encode(), a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                  a.size, ctypes.byref(out)))
    return a.reshape(shape), out.as_dict()
