"""ctypes binding of the learned multiple-scattering surrogate
(include/atmo_surrogate.h, part of libproland_atmo.so)."""
import ctypes

import numpy as np

from . import lib as L
from .params import AtmoOptions, AtmoParams, options as default_options, params_array

NIN, XS, NH, NOUT = 14, 16, 256, 4
NL = 3   # hidden layers (H1 .. H_NL); layers 2 .. NL are 256x256


class Report(ctypes.Structure):
    _fields_ = [("nsets", ctypes.c_int), ("judged", ctypes.c_int), ("accepted", ctypes.c_int),
                ("fallback_sets", ctypes.c_int), ("max_rel_err_S", ctypes.c_float),
                ("max_rel_err_E", ctypes.c_float), ("ms_exact", ctypes.c_float),
                ("ms_mlp", ctypes.c_float), ("ms_judge", ctypes.c_float),
                ("ms_fallback", ctypes.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_vp = ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)
SIGNATURES = {
    "atmo_surrogate_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_uint64]),
    "atmo_surrogate_destroy": (ctypes.c_int, [_vp]),
    "atmo_surrogate_param_count": (ctypes.c_size_t, []),
    "atmo_surrogate_get_params": (ctypes.c_int, [_vp, _fp, ctypes.c_size_t]),
    "atmo_surrogate_set_params": (ctypes.c_int, [_vp, _fp, ctypes.c_size_t]),
    "atmo_surrogate_set_exact_orders": (ctypes.c_int, [_vp, ctypes.c_int]),
    "atmo_surrogate_get_norm": (ctypes.c_int, [_vp, _fp, _fp, _fp]),
    "atmo_surrogate_set_norm": (ctypes.c_int, [_vp, _fp, _fp, ctypes.c_float]),
    "atmo_surrogate_build_dataset": (ctypes.c_int, [_vp, ctypes.POINTER(AtmoParams), ctypes.c_int,
                                                    ctypes.POINTER(AtmoOptions), ctypes.c_int,
                                                    ctypes.c_uint64,
                                                    ctypes.POINTER(ctypes.c_size_t)]),
    "atmo_surrogate_train": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                            ctypes.c_float, ctypes.c_uint64, _fp]),
    "atmo_surrogate_loss_grad": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _fp, _fp]),
    "atmo_surrogate_predict": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    "atmo_surrogate_features": (ctypes.c_int, [ctypes.POINTER(AtmoParams), _vp, _vp, _vp, ctypes.c_int, _vp]),
    "atmo_surrogate_precompute": (ctypes.c_int, [_vp, ctypes.POINTER(AtmoParams), ctypes.c_int,
                                                 ctypes.POINTER(AtmoOptions),
                                                 ctypes.POINTER(L.AtmoTables), ctypes.c_float,
                                                 ctypes.c_float, ctypes.c_uint64,
                                                 ctypes.POINTER(Report), _vp]),
}

_bound = False


def lib():
    global _bound
    lb = L.lib()
    if not _bound:
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lb, name)
            f.restype = res
            f.argtypes = args
        _bound = True
    return lb


def param_count():
    return lib().atmo_surrogate_param_count()


def layout():
    """name -> (offset, shape) of the FP32 parameter vector."""
    out, off = {}, 0
    blocks = [("W1", (NIN, NH)), ("b1", (NH,))]
    for L in range(2, NL + 1):   # the hidden 256x256 layers
        blocks += [("W%d" % L, (NH, NH)), ("b%d" % L, (NH,))]
    blocks += [("W5", (NH, NOUT)), ("b5", (NOUT,))]   # the output layer ("layer 5" for any NL)
    for name, shape in blocks:
        out[name] = (off, shape)
        off += int(np.prod(shape))
    assert off == param_count()
    return out


class Surrogate:
    def __init__(self, device=0, seed=1):
        h = ctypes.c_void_p()
        L.check(lib().atmo_surrogate_create(ctypes.byref(h), int(device), int(seed)))
        self.h = h

    def close(self):
        if self.h:
            lib().atmo_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        n = param_count()
        a = np.zeros(n, dtype=np.float32)
        L.check(lib().atmo_surrogate_get_params(self.h, a.ctypes.data_as(_fp), n))
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        L.check(lib().atmo_surrogate_set_params(self.h, a.ctypes.data_as(_fp), a.size))

    def set_exact_orders(self, n):
        """Orders computed by the plain passes (1 or 2); the network predicts the rest."""
        L.check(lib().atmo_surrogate_set_exact_orders(self.h, int(n)))

    def get_norm(self):
        mu, sd, eps = (ctypes.c_float * 3)(), (ctypes.c_float * 3)(), ctypes.c_float()
        L.check(lib().atmo_surrogate_get_norm(self.h, mu, sd, ctypes.byref(eps)))
        return np.array(mu[:]), np.array(sd[:]), eps.value

    def set_norm(self, mu, sd, eps):
        m = (ctypes.c_float * 3)(*mu)
        s = (ctypes.c_float * 3)(*sd)
        L.check(lib().atmo_surrogate_set_norm(self.h, m, s, float(eps)))

    def build_dataset(self, sets, opts=None, samples_per_set=32768, seed=0x5EED0004):
        n = ctypes.c_size_t()
        L.check(lib().atmo_surrogate_build_dataset(self.h, params_array(sets), len(sets),
                                                   ctypes.byref(opts or default_options()),
                                                   int(samples_per_set), int(seed), ctypes.byref(n)))
        return n.value

    def train(self, steps, batch=65536, lr=1e-3, lr_end=1e-5, seed=1):
        losses = np.zeros(steps, dtype=np.float32)
        L.check(lib().atmo_surrogate_train(self.h, int(steps), int(batch), float(lr), float(lr_end),
                                           int(seed), losses.ctypes.data_as(_fp)))
        return losses

    def loss_grad(self, X, T):
        """X: (B, 16) device float32 tensor, T: (B, 4); returns (loss, grad ndarray)."""
        g = np.zeros(param_count(), dtype=np.float32)
        loss = ctypes.c_float()
        L.check(lib().atmo_surrogate_loss_grad(self.h, ctypes.c_void_p(X.data_ptr()),
                                               ctypes.c_void_p(T.data_ptr()), X.shape[0],
                                               ctypes.byref(loss), g.ctypes.data_as(_fp)))
        return loss.value, g

    def predict(self, X):
        import torch
        Y = torch.empty((X.shape[0], NOUT), dtype=torch.float32, device=X.device)
        L.check(lib().atmo_surrogate_predict(self.h, ctypes.c_void_p(X.data_ptr()), X.shape[0],
                                             ctypes.c_void_p(Y.data_ptr())))
        return Y

    def precompute(self, sets, opts=None, outs=None, judge_fraction=0.01, tolerance=1e-3,
                   seed=0x5EED0005, stream=None):
        if isinstance(sets, AtmoParams):
            sets = [sets]
        if outs is None:
            outs = [L.alloc_outputs(p) for p in sets]
        tabs = (L.AtmoTables * len(sets))()
        for i, (T, E, S) in enumerate(outs):
            tabs[i] = L.AtmoTables(T.data_ptr(), E.data_ptr(), S.data_ptr())
        rep = Report()
        L.check(lib().atmo_surrogate_precompute(self.h, params_array(sets), len(sets),
                                                ctypes.byref(opts or default_options()), tabs,
                                                float(judge_fraction), float(tolerance), int(seed),
                                                ctypes.byref(rep), L._stream(stream)))
        return outs, rep.as_dict()


def features(p, codes, S_exact, E_exact):
    """Network input rows (n, 16) for texel codes of set p (device tensors).
    S_exact, E_exact: the set's device tables of the exactly computed orders
    (float32, inscatter with 4 and irradiance with 3 components per texel)."""
    import torch
    for t, c in ((S_exact, 4), (E_exact, 3)):
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.shape[-1] == c):
            raise ValueError("exact tables: contiguous float32 device tensors with %d components" % c)
    codes = torch.as_tensor(np.asarray(codes, dtype=np.uint32).view(np.int32)).cuda()
    X = torch.empty((codes.shape[0], XS), dtype=torch.float32, device="cuda")
    L.check(lib().atmo_surrogate_features(p, ctypes.c_void_p(S_exact.data_ptr()), ctypes.c_void_p(E_exact.data_ptr()),
                                          ctypes.c_void_p(codes.data_ptr()), codes.shape[0],
                                          ctypes.c_void_p(X.data_ptr())))
    return X
