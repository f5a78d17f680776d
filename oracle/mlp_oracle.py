"""Reference arithmetic of the scattering-source surrogate (test
infrastructure: the checker of the GPU training and inference kernels).

The network is 24 -> 4 x 256 (SiLU) -> 1 with the parameter layout of the
library (W1 [32][256] with rows 24..31 zero, b1, W2..W4 [256][256], b2..b4,
W5 [256], b5).  `bf16=True` mirrors the GPU's rounding points: inputs,
weights, activations and the backward dz are rounded to BF16 (round to
nearest even) where the GPU stores them in BF16; products accumulate in
float64 here (FP32 on the matrix cores), so the two agree to accumulation
order.  `bf16=False` is the exact float64 network, pinned by finite
differences in the tests.
"""
import numpy as np

IN, INP, H = 24, 32, 256
OFF = {}
_o = 0
for name, size in [("W1", INP * H), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * H), ("b3", H),
                   ("W4", H * H), ("b4", H), ("W5", H), ("b5", 1)]:
    OFF[name] = (_o, size)
    _o += size
NPARAM = _o


def bf16(a):
    """Round to BF16 (nearest even), returned as float32 values."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return r.astype(np.uint32).view(np.float32)


def unpack(p):
    d = {}
    for k, (o, n) in OFF.items():
        v = p[o:o + n]
        if k == "W1":
            v = v.reshape(INP, H)
        elif k in ("W2", "W3", "W4"):
            v = v.reshape(H, H)
        d[k] = v
    return d


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, bf=True):
    """x [B][24] normalised features -> (y [B], cache)."""
    q = unpack(np.asarray(p, dtype=np.float32 if bf else np.float64))
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    xb = np.zeros((x.shape[0], INP))
    xb[:, :IN] = x
    a = r(xb).astype(np.float64)
    zs, acts = [], [a]
    for l in range(1, 5):
        W = r(q["W%d" % l]).astype(np.float64)
        z = a @ W
        if bf:
            z = (z.astype(np.float32) + q["b%d" % l]).astype(np.float32)
        else:
            z = z + q["b%d" % l]
        a = r(silu(z.astype(np.float64))).astype(np.float64)
        zs.append(z.astype(np.float64))
        acts.append(a)
    w5 = r(q["W5"]).astype(np.float64)
    y = a @ w5 + float(q["b5"][0])
    return y, (acts, zs, q, r)


def grads(p, x, t, bf=True):
    """Gradients of mean((y - t)^2) in the parameter layout, and the loss."""
    y, (acts, zs, q, r) = forward(p, x, bf)
    B = x.shape[0]
    d = y - t
    loss = float(np.mean(d * d))
    dy = 2.0 * d / B
    g = np.zeros(NPARAM)
    w5 = r(q["W5"]).astype(np.float64)
    o, n = OFF["W5"]
    g[o:o + n] = acts[4].T @ dy
    g[OFF["b5"][0]] = dy.sum()
    dz = r(dy[:, None] * w5[None, :] * dsilu(zs[3])).astype(np.float64)
    for l in range(4, 0, -1):
        o, n = OFF["W%d" % l]
        g[o:o + n] = (acts[l - 1].T @ dz).ravel()
        o, n = OFF["b%d" % l]
        g[o:o + n] = dz.sum(axis=0)
        if l > 1:
            W = r(q["W%d" % l]).astype(np.float64)
            da = dz @ W.T
            dz = r(da * dsilu(zs[l - 2])).astype(np.float64)
    return g, loss


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step = number of completed steps); float32 like the GPU."""
    p, g, m, v = (np.asarray(a, dtype=np.float32) for a in (p, g, m, v))
    t = np.float32(step + 1)
    m = np.float32(b1) * m + np.float32(1 - b1) * g
    v = np.float32(b2) * v + np.float32(1 - b2) * g * g
    mh = m / (np.float32(1) - np.float32(b1) ** t)
    vh = v / (np.float32(1) - np.float32(b2) ** t)
    return p - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps)), m, v


def init_params(seed=1):
    """He-uniform initial parameters in the library's layout (any seed)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    for k, fan in [("W1", IN), ("W2", H), ("W3", H), ("W4", H), ("W5", 2 * H)]:
        o, n = OFF[k]
        a = np.sqrt(6.0 / fan)
        w = rng.uniform(-a, a, n).astype(np.float32)
        if k == "W1":
            w = w.reshape(INP, H)
            w[IN:] = 0
            w = w.ravel()
        p[o:o + n] = w
    return p
