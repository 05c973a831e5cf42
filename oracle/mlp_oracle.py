"""NumPy restatement of the surrogate MLP (include/atmo_surrogate.h):
forward, mean-square loss and backward.

TEST INFRASTRUCTURE ONLY: the reference the GPU forward/backward kernels are
tested against.
With emulate_bf16=True the activations, the hidden-layer weights and the
hidden-layer gradients are rounded to bfloat16 at the points where the GPU
kernels store them, so the two agree up to accumulation order.
"""
import numpy as np

NIN, XS, NH, NOUT = 14, 16, 256, 4
NL = 3   # hidden layers (H1 .. H_NL); layers 2 .. NL are 256x256


def layout():
    out, off = {}, 0
    blocks = [("W1", (NIN, NH)), ("b1", (NH,))]
    for L in range(2, NL + 1):   # the hidden 256x256 layers
        blocks += [("W%d" % L, (NH, NH)), ("b%d" % L, (NH,))]
    blocks += [("W5", (NH, NOUT)), ("b5", (NOUT,))]   # the output layer ("layer 5" for any NL)
    for name, shape in blocks:
        out[name] = (off, shape)
        off += int(np.prod(shape))
    return out, off


def unpack(p):
    lay, n = layout()
    assert p.size == n
    return {k: np.asarray(p[o:o + int(np.prod(s))], dtype=np.float64).reshape(s)
            for k, (o, s) in lay.items()}


def pack(d):
    lay, n = layout()
    out = np.zeros(n)
    for k, (o, s) in lay.items():
        out[o:o + int(np.prod(s))] = np.asarray(d[k]).ravel()
    return out


def bf16(x):
    """Round to bfloat16 (round to nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def act(z):
    """Hidden activation: z times the smoothstep of t = clip(z / 6 + 1/2, 0, 1)."""
    t = np.clip(z / 6.0 + 0.5, 0.0, 1.0)
    return z * t * t * (3.0 - 2.0 * t)


def dact(z):
    t = np.clip(z / 6.0 + 0.5, 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t) + z * t * (1.0 - t)


def forward(p, X, emulate_bf16=True):
    """Returns (Y, cache).  X: (B, >=14)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    w = unpack(p)
    X = np.asarray(X, dtype=np.float64)[:, :NIN]
    Z = [X @ w["W1"] + w["b1"]]
    H = [r(act(Z[0]))]
    for L in range(2, NL + 1):
        Z.append(H[-1] @ r(w["W%d" % L]) + w["b%d" % L])
        H.append(r(act(Z[-1])))
    Y = H[-1] @ w["W5"] + w["b5"]
    return Y, (X, Z, H, w)


def loss_grad(p, X, T, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    Y, (X, Z, H, w) = forward(p, X, emulate_bf16)
    T = np.asarray(T, dtype=np.float64)
    B = X.shape[0]
    d = Y - T
    loss = float(np.sum(d * d) / (B * NOUT))
    dY = 2.0 * d / (B * NOUT)
    g = {"W5": H[NL - 1].T @ dY, "b5": dY.sum(0)}
    dZ = r((dY @ w["W5"].T) * dact(Z[NL - 1]))
    for L in range(NL, 1, -1):
        g["W%d" % L] = H[L - 2].T @ dZ
        g["b%d" % L] = dZ.sum(0)
        dH = dZ @ r(w["W%d" % L]).T
        dZ = dH * dact(Z[L - 2])
        if L > 2:
            dZ = r(dZ)
    g["W1"] = X.T @ dZ
    g["b1"] = dZ.sum(0)
    return loss, pack(g)
