"""NumPy reference of the learned con2prim guess network (4 -> 16 -> 16 -> 2,
ReLU), TEST INFRASTRUCTURE ONLY: the checker of the HIP training and
inference kernels.  Nothing in the product imports it.

Parameter packing (386 floats): W1[16][4], b1[16], W2[16][16], b2[16],
W3[2][16], b3[2], row = output unit.

Loss: mean over samples of s (y0 - t0)^2 + w1 (y1 - t1)^2, s = 1 + (q/q_ref)^2
with q = X[:, 2]^2 exp(t0) when q_ref > 0, else s = 1.

`qat=True` evaluates the network the way the BF16 inference kernel does:
inputs, W1, W2, W3 and both hidden layers rounded to BF16 (round to nearest
even); b1, b2, b3 (the kernel's FP32 accumulator initialisation) and all
sums in full precision.  The backward pass treats the rounding as the
identity (straight-through), with the rounded values in every product.

How it is pinned: `loss_grad` against central finite differences of
`forward` in float64 (tests/test_c2p_oracle.py), which a dropped term, a
transposed weight or a wrong ReLU mask fails.
"""
import numpy as np

NIN, NH, NOUT = 4, 16, 2
SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NOUT, NH)),
          ("b3", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
assert NPARAM == 386


def unpack(params):
    params = np.asarray(params)
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = params[o:o + k].reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float32).ravel() for name, _ in SHAPES])


def bf16_round(a):
    """float32 -> nearest-even bfloat16, returned as float32 (NaN kept)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).copy()
    out[np.isnan(a)] = np.nan
    return out


def _q(a, qat):
    return bf16_round(a) if qat else a


def forward(params, X, qat=False, dtype=np.float64):
    p = unpack(params)
    W1, W2, W3 = (_q(p[k], qat).astype(dtype) for k in ("W1", "W2", "W3"))
    b1, b2, b3 = (p[k].astype(dtype) for k in ("b1", "b2", "b3"))
    x = _q(np.asarray(X, dtype=np.float32), qat).astype(dtype)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0)
    if qat:
        h1 = bf16_round(h1.astype(np.float32)).astype(dtype)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    if qat:
        h2 = bf16_round(h2.astype(np.float32)).astype(dtype)
    y = h2 @ W3.T + b3
    return dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, y=y, W1=W1, W2=W2, W3=W3)


def _weights(X, T, w1, q_ref, dtype):
    """Per-sample loss weights [n][2]."""
    n = len(X)
    s = np.ones(n, dtype=dtype)
    if q_ref > 0:
        q = np.asarray(X, dtype=dtype)[:, 2] ** 2 * np.exp(np.asarray(T, dtype=dtype)[:, 0])
        s = 1.0 + (q / q_ref) ** 2
    return np.stack([s, np.full(n, w1, dtype=dtype)], axis=1)


def loss_grad(params, X, T, w1=1.0, qat=False, dtype=np.float64, q_ref=0.0):
    """(loss, gradient[NPARAM]) of the mean weighted squared error."""
    f = forward(params, X, qat, dtype)
    n = len(X)
    e = f["y"] - np.asarray(T, dtype=dtype)
    wv = _weights(X, T, w1, q_ref, dtype)
    loss = float(((e * e) * wv).sum() / n)
    dy = 2.0 * e * wv / n
    gW3 = dy.T @ f["h2"]
    gb3 = dy.sum(0)
    dz2 = (dy @ f["W3"]) * (f["h2"] > 0)
    gW2 = dz2.T @ f["h1"]
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ f["W2"]) * (f["h1"] > 0)
    gW1 = dz1.T @ f["x"]
    gb1 = dz1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g


def loss_only(params, X, T, w1=1.0, dtype=np.float64, q_ref=0.0):
    f = forward(params, X, False, dtype)
    e = f["y"] - np.asarray(T, dtype=dtype)
    return float(((e * e) * _weights(X, T, w1, q_ref, dtype)).sum() / len(X))


def he_init(seed):
    rng = np.random.default_rng(seed)
    d = {"W1": rng.standard_normal((NH, NIN)) * np.sqrt(2.0 / NIN), "b1": np.zeros(NH),
         "W2": rng.standard_normal((NH, NH)) * np.sqrt(2.0 / NH), "b2": np.zeros(NH),
         "W3": rng.standard_normal((NOUT, NH)) * np.sqrt(1.0 / NH), "b3": np.zeros(NOUT)}
    return pack(d)
