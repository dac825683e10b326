"""ORACLE -- test infrastructure only: FP64 numpy restatement of the C_l emulator's network
(include/crime_mlp.h), used to check the GPU forward pass, gradients and Adam steps.

Network: h1 = silu(x W0 + b0), h_{k+1} = silu(h_k W_k + b_k) for the four square layers,
y = h5 . w5 + b5; loss (1/B) sum (y - t)^2.  With ``bf16=True`` the values the GPU rounds to BF16
(the square layers' weights, every activation h_k that feeds a BF16 GEMM or the output layer, and every
back-propagated dz) are rounded here at the same points, round-to-nearest-even, so that the GPU and the
oracle differ only by FP32 accumulation order.  The backward pass is pinned by central finite
differences of the loss in FP64 (``bf16=False``).
"""
import numpy as np

NIN, WD, NH = 8, 256, 5
NMID = NH - 1


def layout():
    """(name, offset, shape) of every parameter block in the flat FP32 vector."""
    out, off = [], 0
    for name, shape in [("W0", (NIN, WD)), ("b0", (WD,))] + \
            [(f"{p}{k + 1}", s) for k in range(NMID) for p, s in (("W", (WD, WD)), ("b", (WD,)))] + \
            [("w5", (WD,)), ("b5", (1,))]:
        out.append((name, off, shape))
        off += int(np.prod(shape))
    return out, off


def unpack(p):
    lay, n = layout()
    assert p.size == n
    return {name: p[o:o + int(np.prod(s))].reshape(s).astype(np.float64) for name, o, s in lay}


def pack(d):
    lay, n = layout()
    out = np.empty(n)
    for name, o, s in lay:
        out[o:o + int(np.prod(s))] = d[name].ravel()
    return out


def to_bf16(a):
    """Round to the nearest BF16 (ties to even), through FP32 like the GPU's values."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed):
    """He-uniform weights, zero biases (the starting point used by the emulator)."""
    rng = np.random.default_rng(seed)
    d = {}
    d["W0"] = rng.uniform(-1, 1, (NIN, WD)) * np.sqrt(6.0 / NIN)
    d["b0"] = np.zeros(WD)
    for k in range(1, NMID + 1):
        d[f"W{k}"] = rng.uniform(-1, 1, (WD, WD)) * np.sqrt(6.0 / WD)
        d[f"b{k}"] = np.zeros(WD)
    d["w5"] = rng.uniform(-1, 1, WD) * np.sqrt(3.0 / WD)
    d["b5"] = np.zeros(1)
    return pack(d).astype(np.float32)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, bf16=True):
    d = unpack(np.asarray(p, np.float64))
    r = to_bf16 if bf16 else (lambda a: a)
    x = np.asarray(x, np.float64)
    zs, hs = [], []
    z = x @ d["W0"] + d["b0"]
    zs.append(z)
    hs.append(r(_silu(z)))
    for k in range(1, NMID + 1):
        z = hs[-1] @ r(d[f"W{k}"]) + d[f"b{k}"]
        zs.append(z)
        hs.append(r(_silu(z)))
    y = hs[-1] @ d["w5"] + d["b5"][0]
    return y, zs, hs


def loss_grad(p, x, t, bf16=True, w=None):
    """(loss, flat gradient) of (1/B) sum w (y - t)^2 (w = 1 by default)."""
    d = unpack(np.asarray(p, np.float64))
    r = to_bf16 if bf16 else (lambda a: a)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    n = x.shape[0]
    y, zs, hs = forward(p, x, bf16)
    e = y - t
    w = np.ones(n) if w is None else np.asarray(w, np.float64)
    loss = float((w * e * e).sum() / n)
    dy = 2.0 * w * e / n
    g = {"w5": hs[-1].T @ dy, "b5": np.array([dy.sum()])}
    dz = r(dy[:, None] * d["w5"][None, :] * _dsilu(zs[-1]))
    for k in range(NMID, 0, -1):
        g[f"W{k}"] = hs[k - 1].T @ dz
        g[f"b{k}"] = dz.sum(axis=0)
        dz = r((dz @ r(d[f"W{k}"]).T) * _dsilu(zs[k - 1]))
    g["W0"] = x.T @ dz
    g["b0"] = dz.sum(axis=0)
    return loss, pack(g)


def adam_step(p, m, v, g, step, lr0, total, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update with the cosine-decayed learning rate lr0 (1 + cos(pi step/total)) / 2."""
    lr = lr0 * 0.5 * (1 + np.cos(np.pi * min(step / total, 1.0)))
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** (step + 1))
    vh = v / (1 - b2 ** (step + 1))
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
