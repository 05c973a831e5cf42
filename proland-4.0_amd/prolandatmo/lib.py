"""ctypes binding of libproland_atmo.so (the C ABI declared in include/atmo.h).

Device tables are torch tensors on the HIP device (torch is only used for
device memory and streams).  The product path is the HIP library: if it is
missing this module raises instead of falling back to anything else.
"""
import ctypes
import os

from .params import AtmoOptions, AtmoParams, options as default_options, params_array

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("PROLAND_ATMO_LIB") or os.path.join(PKG_DIR, "libproland_atmo.so")

ATMO_OK = 0
STATUS = {0: "OK", 1: "INVALID_ARGUMENT", 2: "DEVICE", 3: "OUT_OF_MEMORY", 4: "IO",
          5: "NUMERIC", 6: "UNSUPPORTED"}


class AtmoTables(ctypes.Structure):
    _fields_ = [("transmittance", ctypes.c_void_p), ("irradiance", ctypes.c_void_p),
                ("inscatter", ctypes.c_void_p)]


class AtmoPassTime(ctypes.Structure):
    _fields_ = [("pass_", ctypes.c_int), ("order", ctypes.c_int), ("ms", ctypes.c_float)]


PASS_NAMES = ("transmittance", "irradiance1", "inscatter1", "inscatterS", "irradianceN",
              "inscatterN")


class AtmoTuning(ctypes.Structure):
    _fields_ = [("tuned", ctypes.c_int), ("inscatterS_texels_per_thread", ctypes.c_int),
                ("inscatterN_rows_per_workgroup", ctypes.c_int), ("inscatterN_xcd_order", ctypes.c_int),
                ("inscatterS_ms", ctypes.c_float * 5), ("inscatterN_ms", ctypes.c_float * 27),
                ("inscatterN_staged_samples", ctypes.c_int)]

    def as_dict(self):
        return {"tuned": self.tuned, "inscatterS_texels_per_thread": self.inscatterS_texels_per_thread,
                "inscatterN_rows_per_workgroup": self.inscatterN_rows_per_workgroup,
                "inscatterN_xcd_order": self.inscatterN_xcd_order,
                "inscatterS_ms": [round(x, 3) for x in self.inscatterS_ms],
                "inscatterN_staged_samples": self.inscatterN_staged_samples,
                "inscatterN_ms": [round(x, 3) for x in self.inscatterN_ms]}


class AtmoError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


_vp = ctypes.c_void_p
_pp = ctypes.POINTER(AtmoParams)
_po = ctypes.POINTER(AtmoOptions)
_i32p = ctypes.POINTER(ctypes.c_int32)
_fp = ctypes.POINTER(ctypes.c_float)

# every symbol include/atmo.h declares, with its signature
SIGNATURES = {
    "atmo_params_default": (None, [_pp]),
    "atmo_options_default": (None, [_po]),
    "atmo_last_error": (ctypes.c_char_p, []),
    "atmo_version": (ctypes.c_char_p, []),
    "atmo_preprocess": (ctypes.c_int, [_pp, ctypes.c_char_p]),
    "atmo_preprocess_ex": (ctypes.c_int, [_pp, _po, ctypes.c_char_p, ctypes.c_int]),
    "atmo_precompute_host": (ctypes.c_int, [_pp, _po, _fp, _fp, _fp]),
    "atmo_engine_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int]),
    "atmo_engine_destroy": (ctypes.c_int, [_vp]),
    "atmo_engine_precompute": (ctypes.c_int, [_vp, _pp, ctypes.c_int, _po,
                                              ctypes.POINTER(AtmoTables), _vp]),
    "atmo_engine_pass_times": (ctypes.c_int, [_vp, ctypes.POINTER(AtmoPassTime), ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int)]),
    "atmo_engine_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(AtmoTuning)]),
    "atmo_engine_workspace_size": (ctypes.c_int, [_pp, ctypes.c_int, _po,
                                                  ctypes.POINTER(ctypes.c_size_t)]),
    "atmo_precompute_noalloc": (ctypes.c_int, [_pp, ctypes.c_int, _po, ctypes.POINTER(AtmoTables),
                                               _vp]),
    "atmo_pass_transmittance": (ctypes.c_int, [_vp, _pp, _po, _vp, _vp]),
    "atmo_pass_irradiance1": (ctypes.c_int, [_vp, _pp, _po, _vp, _vp, _vp]),
    "atmo_pass_inscatter1": (ctypes.c_int, [_vp, _pp, _po, _vp, _vp, _vp, _vp, _vp]),
    "atmo_pass_inscatterS": (ctypes.c_int, [_vp, _pp, _po, _vp, _vp, _vp, _vp, ctypes.c_int,
                                            _vp, _vp]),
    "atmo_pass_irradianceN": (ctypes.c_int, [_vp, _pp, _po, _vp, _vp, ctypes.c_int, _vp, _vp,
                                             _vp]),
    "atmo_pass_inscatterN": (ctypes.c_int, [_vp, _pp, _po, _vp, _vp, _vp, _vp, _vp]),
    "atmo_write_raw": (ctypes.c_int, [ctypes.c_char_p, _fp, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int]),
    "atmo_read_raw_header": (ctypes.c_int, [ctypes.c_char_p, _i32p]),
    "atmo_read_raw": (ctypes.c_int, [ctypes.c_char_p, _fp, ctypes.c_size_t, _i32p]),
}

_lib = None


def lib():
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libproland_atmo.so not built: run `make -C proland-4.0_amd` "
                               "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != ATMO_OK:
        raise AtmoError(status, lib().atmo_last_error().decode())


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


# ---- device table allocation --------------------------------------------
def alloc_outputs(p, device="cuda"):
    """Output tables of one set in the .raw layouts (torch float32)."""
    import torch
    T = torch.empty((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3), dtype=torch.float32, device=device)
    E = torch.empty((p.SKY_H, p.SKY_W, 3), dtype=torch.float32, device=device)
    S = torch.empty((p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 4), dtype=torch.float32,
                    device=device)
    return T, E, S


def to_float4(a, device="cuda"):
    """(…, 3) or (…, 4) array -> contiguous float32 (…, 4) device tensor."""
    import numpy as np
    import torch
    a = np.asarray(a)
    out = np.zeros(a.shape[:-1] + (4,), dtype=np.float32)
    out[..., :a.shape[-1]] = a
    return torch.from_numpy(out).to(device)


class Engine:
    """An atmo_engine: owns device work tables and the captured pass graph."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().atmo_engine_create(ctypes.byref(h), int(device)))
        self.h = h

    def close(self):
        if self.h:
            lib().atmo_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def precompute(self, sets, opts=None, outs=None, stream=None):
        """Batched precompute.  sets: AtmoParams or list; outs: list of (T, E, S)
        tensors (allocated if None).  Returns the list of (T, E, S)."""
        if isinstance(sets, AtmoParams):
            sets = [sets]
        opts = opts or default_options()
        if outs is None:
            outs = [alloc_outputs(p) for p in sets]
        tabs = (AtmoTables * len(sets))()
        for i, (T, E, S) in enumerate(outs):
            tabs[i] = AtmoTables(T.data_ptr(), E.data_ptr(), S.data_ptr())
        arr = params_array(sets)
        check(lib().atmo_engine_precompute(self.h, arr, len(sets), ctypes.byref(opts), tabs,
                                           _stream(stream)))
        return outs

    def pass_times(self):
        """[(pass name, order, ms)] of the last run made with FLAG_TIME_PASSES."""
        n = ctypes.c_int()
        check(lib().atmo_engine_pass_times(self.h, None, 0, ctypes.byref(n)))
        buf = (AtmoPassTime * max(n.value, 1))()
        check(lib().atmo_engine_pass_times(self.h, buf, n.value, ctypes.byref(n)))
        return [(PASS_NAMES[buf[i].pass_], buf[i].order, buf[i].ms) for i in range(n.value)]

    def tuning(self):
        t = AtmoTuning()
        check(lib().atmo_engine_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    # ---- pass level (float4 device tables) ----
    def pass_transmittance(self, p, opts, T4, stream=None):
        check(lib().atmo_pass_transmittance(self.h, p, opts, _ptr(T4), _stream(stream)))

    def pass_irradiance1(self, p, opts, T4, dE4, stream=None):
        check(lib().atmo_pass_irradiance1(self.h, p, opts, _ptr(T4), _ptr(dE4), _stream(stream)))

    def pass_inscatter1(self, p, opts, T4, dSR4, dSM4, S4=None, stream=None):
        check(lib().atmo_pass_inscatter1(self.h, p, opts, _ptr(T4), _ptr(dSR4), _ptr(dSM4),
                                         _ptr(S4), _stream(stream)))

    def pass_inscatterS(self, p, opts, T4, dE4, dSR4, dSM4, first, dJ4, stream=None):
        check(lib().atmo_pass_inscatterS(self.h, p, opts, _ptr(T4), _ptr(dE4), _ptr(dSR4),
                                         _ptr(dSM4), int(first), _ptr(dJ4), _stream(stream)))

    def pass_irradianceN(self, p, opts, dSR4, dSM4, first, dE4, E4=None, stream=None):
        check(lib().atmo_pass_irradianceN(self.h, p, opts, _ptr(dSR4), _ptr(dSM4), int(first),
                                          _ptr(dE4), _ptr(E4), _stream(stream)))

    def pass_inscatterN(self, p, opts, T4, dJ4, dSR4, S4=None, stream=None):
        check(lib().atmo_pass_inscatterN(self.h, p, opts, _ptr(T4), _ptr(dJ4), _ptr(dSR4),
                                         _ptr(S4), _stream(stream)))


# bytes of the allocation-free precompute's static device array (ATMO_NOALLOC_MAX_BYTES)
NOALLOC_MAX_BYTES = 64 * 1024 * 1024


def precompute_noalloc(sets, opts=None, outs=None, stream=None):
    """atmo_precompute_noalloc: the batched precompute with its work tables in
    the library's static device array (no device allocation by the library).
    outs: list of (T, E, S) tensors (allocated here if None).  Returns them."""
    if isinstance(sets, AtmoParams):
        sets = [sets]
    opts = opts or default_options()
    if outs is None:
        outs = [alloc_outputs(p) for p in sets]
    tabs = (AtmoTables * len(sets))()
    for i, (T, E, S) in enumerate(outs):
        tabs[i] = AtmoTables(T.data_ptr(), E.data_ptr(), S.data_ptr())
    check(lib().atmo_precompute_noalloc(params_array(sets), len(sets), ctypes.byref(opts), tabs,
                                        _stream(stream)))
    return outs


def workspace_size(sets, opts=None):
    if isinstance(sets, AtmoParams):
        sets = [sets]
    n = ctypes.c_size_t()
    check(lib().atmo_engine_workspace_size(params_array(sets), len(sets),
                                           ctypes.byref(opts or default_options()),
                                           ctypes.byref(n)))
    return n.value


def precompute_host(p, opts=None):
    """One set, host numpy outputs (T, E, S) in the .raw layouts."""
    import numpy as np
    T = np.zeros((p.TRANSMITTANCE_H, p.TRANSMITTANCE_W, 3), dtype=np.float32)
    E = np.zeros((p.SKY_H, p.SKY_W, 3), dtype=np.float32)
    S = np.zeros((p.RES_R, p.RES_MU, p.RES_NU * p.RES_MU_S, 4), dtype=np.float32)
    fp = lambda a: a.ctypes.data_as(_fp)
    check(lib().atmo_precompute_host(p, opts or default_options(), fp(T), fp(E), fp(S)))
    return T, E, S


def preprocess(p, outdir, opts=None, skip_if_exists=True):
    """proland::preprocessAtmo: writes transmittance/irradiance/inscatter.raw."""
    check(lib().atmo_preprocess_ex(p, opts or default_options(), outdir.encode(),
                                   int(bool(skip_if_exists))))


def write_raw(path, data, width, height, depth, components):
    import numpy as np
    a = np.ascontiguousarray(data, dtype=np.float32)
    check(lib().atmo_write_raw(path.encode(), a.ctypes.data_as(_fp), width, height, depth,
                               components))


def read_raw(path):
    """Returns (data float32 1-D, trailer int32[5])."""
    import numpy as np
    h = (ctypes.c_int32 * 5)()
    check(lib().atmo_read_raw_header(path.encode(), h))
    n = h[1] * h[2] * h[4]
    a = np.zeros(n, dtype=np.float32)
    check(lib().atmo_read_raw(path.encode(), a.ctypes.data_as(_fp), n, h))
    return a, np.array(list(h), dtype=np.int32)
