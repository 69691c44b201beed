"""FP64 oracle of the learned WENO5 weights (the MLP of include/mfc_rhs.h).

TEST INFRASTRUCTURE ONLY: the checker for the GPU training and inference
kernels.  With ``bf16=True`` it rounds to bfloat16 (round to nearest even)
at the points where the kernels feed MFMA operands: the features, the
weights, the two hidden activations, dz, da2 and da1; everything else is
FP64, so the remaining difference from the GPU is FP32 accumulation and
exp rounding.  With ``bf16=False`` it is the exact model, whose gradient is
pinned by central finite differences in tests/test_mlp_oracle.py.

The reference has no learned component; this oracle restates the model this
build defines (see the kernels' header), not reference code.
"""
import numpy as np

NPARAM = 2128
OFF = dict(W1=0, b1=512, W2=544, b2=1568, W3=1600, b3=2112)


def bf16(a):
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    W1 = p[OFF["W1"]:OFF["W1"] + 512].reshape(16, 32)
    b1 = p[OFF["b1"]:OFF["b1"] + 32]
    W2 = p[OFF["W2"]:OFF["W2"] + 1024].reshape(32, 32)
    b2 = p[OFF["b2"]:OFF["b2"] + 32]
    W3 = p[OFF["W3"]:OFF["W3"] + 512].reshape(32, 16)
    b3 = p[OFF["b3"]:OFF["b3"] + 16]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    out = np.zeros(NPARAM)
    out[OFF["W1"]:OFF["W1"] + 512] = np.asarray(W1).reshape(-1)
    out[OFF["b1"]:OFF["b1"] + 32] = b1
    out[OFF["W2"]:OFF["W2"] + 1024] = np.asarray(W2).reshape(-1)
    out[OFF["b2"]:OFF["b2"] + 32] = b2
    out[OFF["W3"]:OFF["W3"] + 512] = np.asarray(W3).reshape(-1)
    out[OFF["b3"]:OFF["b3"] + 16] = b3
    return out


def _r(a, on):
    return bf16(a) if on else np.asarray(a, dtype=np.float64)


GATE_IDEAL = np.array([0.3, 0.6, 0.1])
GATE_OFF, GATE_DEV = 0.005, 0.03


def gate(wp):
    """Logit gate: min(1, max(0, sum_k |w_plain_k - d_k| - 0.005) / 0.03), d
    the uniform-grid ideal weights of the oriented face; the logits act in
    full away from smooth, resolved flow, fade out as the plain weights
    approach d, and where g = 0 the plain weights are kept exactly."""
    dev = np.abs(np.asarray(wp) - GATE_IDEAL).sum(axis=1)
    return np.minimum(1.0, np.maximum(0.0, dev - GATE_OFF) / GATE_DEV)


def forward(params, S, bf16_points=True):
    """S: (12, n) samples.  Returns a dict with the intermediate values."""
    S = np.asarray(S, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(params)
    n = S.shape[1]
    x = np.zeros((n, 16))
    x[:, :5] = S[0:5].T
    xr = _r(x, bf16_points)
    W1r, W2r, W3r = _r(W1, bf16_points), _r(W2, bf16_points), _r(W3, bf16_points)
    a1 = np.maximum(xr @ W1r + b1, 0.0)
    h1 = _r(a1, bf16_points)
    a2 = np.maximum(h1 @ W2r + b2, 0.0)
    h2 = _r(a2, bf16_points)
    z = (h2 @ W3r + b3)[:, :3]
    wp = S[8:11].T
    p = S[5:8].T
    t = S[11]
    g = gate(wp)
    zg = g[:, None] * z
    u = wp * np.exp(zg - zg.max(axis=1, keepdims=True))
    w = np.where((g == 0.0)[:, None], wp, u / u.sum(axis=1, keepdims=True))
    y = (w * p).sum(axis=1)
    loss = np.mean((y - t) ** 2)
    return dict(x=xr, a1=a1, h1=h1, a2=a2, h2=h2, z=z, g=g, w=w, y=y, p=p, t=t, loss=loss,
                W1r=W1r, W2r=W2r, W3r=W3r, n=n)


def gradient(params, S, bf16_points=True):
    f = forward(params, S, bf16_points)
    n = f["n"]
    dy = 2.0 * (f["y"] - f["t"]) / n
    dw = dy[:, None] * f["p"]
    sdw = (f["w"] * dw).sum(axis=1, keepdims=True)
    dz3 = f["g"][:, None] * f["w"] * (dw - sdw)
    db3 = np.zeros(16)
    db3[:3] = dz3.sum(axis=0)
    dz = np.zeros((n, 16))
    dz[:, :3] = dz3
    dzr = _r(dz, bf16_points)
    dW3 = f["h2"].T @ dzr
    da2 = (dzr @ f["W3r"].T) * (f["a2"] > 0)
    db2 = da2.sum(axis=0)
    da2r = _r(da2, bf16_points)
    dW2 = f["h1"].T @ da2r
    da1 = (da2r @ f["W2r"].T) * (f["a1"] > 0)
    db1 = da1.sum(axis=0)
    da1r = _r(da1, bf16_points)
    dW1 = f["x"].T @ da1r
    return pack(dW1, db1, dW2, db2, dW3, db3), f["loss"]


def infer(params, S, bf16_points=True):
    f = forward(params, S, bf16_points)
    return np.vstack([f["w"].T, f["y"][None, :]])


def samples_1d(coarse, fine, coef27, eps=1e-16):
    """Training pairs of a 1D variable, restating the sample kernel's
    definition: coarse (m+1) and fine (2m+2) interior values with 4 ghost
    cells on each side.  Returns (13, 2(m+1))."""
    b = 4
    C = lambda i: coarse[i + b]  # noqa: E731
    Fv = lambda i: fine[i + b]  # noqa: E731
    cf = np.asarray(coef27, dtype=np.float64)
    m = len(coarse) - 2 * b - 1

    def parts(v):
        d1, d0, dm1, dm2 = v[4] - v[3], v[3] - v[2], v[2] - v[1], v[1] - v[0]
        p = [v[2] + cf[6] * d1 + cf[7] * d0, v[2] + cf[8] * d0 + cf[9] * dm1, v[2] + cf[10] * dm1 + cf[11] * dm2]
        bb = [cf[18] * d1 * d1 + cf[19] * d1 * d0 + cf[20] * d0 * d0 + eps,
              cf[21] * d0 * d0 + cf[22] * d0 * dm1 + cf[23] * dm1 * dm1 + eps,
              cf[24] * dm1 * dm1 + cf[25] * dm1 * dm2 + cf[26] * dm2 * dm2 + eps]
        a = [cf[15 + k] / (bb[k] * bb[k]) for k in range(3)]
        s = sum(a)
        return p, [ak / s for ak in a]

    def value(v):
        p, w = parts(v)
        return sum(wk * pk for wk, pk in zip(w, p))

    out = np.zeros((13, 2 * (m + 1)))
    for i in range(m + 1):
        for side in (0, 1):
            o = [k - 2 for k in range(5)] if side == 0 else [2 - k for k in range(5)]
            v = [C(i + oo) for oo in o]
            fa = 2 * i + 1 if side == 0 else 2 * i
            sg = 1 if side == 0 else -1
            a = [Fv(fa + sg * (k - 2)) for k in range(5)]
            bb = [Fv(fa + sg * (3 - k)) for k in range(5)]
            tface = 0.5 * (value(a) + value(bb))
            p, w = parts(v)
            s = max(abs(v[k + 1] - v[k]) for k in range(4))
            vmax = max(abs(x) for x in v)
            valid = s > 1e-6 * vmax and s > 1e-300
            inv = 1.0 / s if valid else 0.0
            col = 2 * i + side
            out[0:5, col] = [(vk - v[2]) * inv for vk in v]
            out[5:8, col] = [(pk - v[2]) * inv for pk in p]
            out[8:11, col] = w
            out[11, col] = (tface - v[2]) * inv
            out[12, col] = 1.0 if valid else 0.0
    return out
