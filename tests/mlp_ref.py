"""Reference (numpy, float64) of the surrogate MLP kernels, for tests only.

The model: x^ = (B - mu) / sigma over nin <= 16 descriptor components, f = w3 . s(W2^T s(W1^T x^ + b1) + b2)
+ b3 with s = SiLU, E_atom = e_shift + e_scale f.  Parameters are the flat FP32 layout of the kernels:
W1[16][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.

``quant=True`` rounds to bfloat16 (round to nearest even) at exactly the points where the GPU kernels
feed the matrix cores: weights, the standardized inputs, the layer activations and the back-propagated
vectors; everything else stays float64 here (float32 on the GPU).
"""
import numpy as np

NPARAM = 5313
OFF = dict(W1=0, b1=1024, W2=1088, b2=5184, w3=5248, b3=5312)


def bf16(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[0:1024].reshape(16, 64), p[1024:1088], p[1088:5184].reshape(64, 64), p[5184:5248], p[5248:5312],
            p[5312])


def init_params(seed=0, nin=14):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    W1 = np.zeros((16, 64))
    W1[:nin] = rng.normal(0, 1.0 / np.sqrt(nin), (nin, 64))
    p[0:1024] = W1.ravel()
    p[1088:5184] = rng.normal(0, 1.0 / 8.0, 64 * 64)
    p[5248:5312] = rng.normal(0, 1.0 / 8.0, 64)
    return p


def _s(z):
    return 1.0 / (1.0 + np.exp(-z))


def _act(z):
    s = _s(z)
    return z * s, s * (1 + z * (1 - s)), s * (1 - s) * (2 + z * (1 - 2 * s))


def _q(a, quant):
    return bf16(a) if quant else np.asarray(a, np.float64)


def _inputs(aff, nin, X):
    mu, isg = np.asarray(aff[:16]), np.asarray(aff[16:32])
    xh = np.zeros((len(X), 16))
    xh[:, :nin] = (np.asarray(X)[:, :nin] - mu[:nin]) * isg[:nin]
    return xh, isg


def forward(p, aff, nin, X, quant=False):
    """(e (n,), g (n, nin)) for descriptors X (n, nin)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    W1q, W2q = _q(W1, quant), _q(W2, quant)
    xh, isg = _inputs(aff, nin, X)
    z1 = _q(xh, quant) @ W1q + b1
    h1, d1, _ = _act(z1)
    z2 = _q(h1, quant) @ W2q + b2
    h2, d2, _ = _act(z2)
    f = h2 @ w3 + b3
    e = aff[33] + aff[32] * f
    dz2 = _q(w3 * d2, quant)
    dz1 = _q((dz2 @ W2q.T) * d1, quant)
    gx = dz1 @ W1q.T
    g = aff[32] * gx[:, :nin] * isg[:nin]
    return e, g


def train_grad(p, aff, nin, X, U, alpha, quant=False):
    """Gradient of sum_a (alpha_a E_a + U_a . g_a) with respect to the flat parameters."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    W1q, W2q = _q(W1, quant), _q(W2, quant)
    xh, isg = _inputs(aff, nin, X)
    es = aff[32]
    uh = np.zeros_like(xh)
    uh[:, :nin] = np.asarray(U)[:, :nin] * isg[:nin] * es
    ap = es * np.asarray(alpha)
    xq, uq = _q(xh, quant), _q(uh, quant)
    z1 = xq @ W1q + b1
    t1 = uq @ W1q
    h1, s1p, s1pp = _act(z1)
    ht1 = s1p * t1
    h1q, ht1q = _q(h1, quant), _q(ht1, quant)
    z2 = h1q @ W2q + b2
    t2 = ht1q @ W2q
    h2, s2p, s2pp = _act(z2)
    D2 = w3 * s2p
    P2 = ap[:, None] * w3 * s2p + w3 * s2pp * t2
    P2q, D2q = _q(P2, quant), _q(D2, quant)
    aH = P2q @ W2q.T
    aHt = D2q @ W2q.T
    D1 = s1p * aHt
    P1 = s1p * aH + s1pp * t1 * aHt
    P1q, D1q = _q(P1, quant), _q(D1, quant)
    g = np.zeros(NPARAM)
    g[0:1024] = (xq.T @ P1q + uq.T @ D1q).ravel()
    g[1024:1088] = P1.sum(0)
    g[1088:5184] = (h1q.T @ P2q + ht1q.T @ D2q).ravel()
    g[5184:5248] = P2.sum(0)
    g[5248:5312] = (ap[:, None] * h2 + s2p * t2).sum(0)
    g[5312] = ap.sum()
    return g


def adam(p, g, m, v, lr, b1, b2, eps, t, wd):
    """One Adam step (Kingma & Ba 2015) with decoupled weight decay on the weight matrices."""
    p, m, v = (np.asarray(a, np.float64).copy() for a in (p, m, v))
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mask = np.zeros(NPARAM, bool)
    mask[0:1024] = mask[1088:5184] = mask[5248:5312] = True
    p = np.where(mask, p - lr * wd * p, p)
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
