"""CPU oracle of the truncation-closure network (test infrastructure only).

The network maps an ADO rho_n of the last tier (multi-index n, tier L) to
the ADOs of its D missing next-tier neighbours:

    x  = [Re rho_n / s, Im rho_n / s, n / L]              (2 N^2 + D inputs)
    h1 = GELU(W1 x + b1),  h2 = GELU(W2 h1 + b2),  y = W3 h2 + b3
    rho_{n+e_j} ~ s * (y_j[:N^2] + i y_j[N^2:]),  y = [y_0, ..., y_{D-1}]

(ADOs column-major), with s = the largest |Re| or |Im| entry of rho_n
(floored) and GELU(z) = z/2 (1 + erf(z/sqrt 2)).
There is no counterpart in PHI, which truncates by dropping the tier
(hierarchy_integrator.cc:1204-1338 links no neighbour beyond the last tier);
this module restates the arithmetic of the project's HIP kernels so their
outputs and gradients can be checked:

  * `exact=True`: float64 throughout (pinned by finite differences in
    tests/test_mlp_oracle.py);
  * `exact=False`: the kernels' arithmetic: GEMM operands rounded to
    bfloat16 (round to nearest even), products accumulated in float32,
    bias/GELU/loss in float32.

Matrices are stored as the kernels store them: W_l is (out, in) row-major,
activations are (batch, features) row-major.
"""

import numpy as np
from scipy.special import erf

SQRT1_2 = 0.7071067811865476
INV_SQRT_2PI = 0.3989422804014327


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).copy()
    out[nan] = np.nan
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z * SQRT1_2)) + z * INV_SQRT_2PI * np.exp(-0.5 * z * z)


def init_params(n_in, n_hidden, n_out, seed=0):
    """He-style initialisation (float32), layer list [(W, b), ...]."""
    rng = np.random.default_rng(seed)
    dims = [n_in, n_hidden, n_hidden, n_out]
    params = []
    for i in range(3):
        fan_in = dims[i]
        W = (rng.standard_normal((dims[i + 1], fan_in)) * np.sqrt(2.0 / fan_in)).astype(np.float32)
        if i == 2:
            W *= 0.1
        b = np.zeros(dims[i + 1], dtype=np.float32)
        params.append((W, b))
    return params


def _mm(a, b, exact):
    """a @ b.T for a (M, K), b (N, K)."""
    if exact:
        return np.asarray(a, np.float64) @ np.asarray(b, np.float64).T
    return (bf16(a).astype(np.float64) @ bf16(b).astype(np.float64).T).astype(np.float32)


def forward(params, x, exact=False):
    """Returns (y, cache) for x (batch, n_in)."""
    dt = np.float64 if exact else np.float32
    h = np.asarray(x, dt)
    cache = []
    for i, (W, b) in enumerate(params):
        z = _mm(h, W, exact).astype(dt) + np.asarray(b, dt)
        cache.append((h, z))
        h = gelu(z).astype(dt) if i < 2 else z
    return h, cache


def mse(y, t):
    d = np.asarray(y, np.float64) - np.asarray(t, np.float64)
    return float(np.mean(d * d))


def backward(params, cache, y, t, exact=False):
    """Gradients of mean((y - t)^2) over all entries: [(dW, db), ...]."""
    dt = np.float64 if exact else np.float32
    g = (2.0 / y.size) * (np.asarray(y, dt) - np.asarray(t, dt))
    grads = [None] * 3
    for i in (2, 1, 0):
        W, _ = params[i]
        h, z = cache[i]
        if i < 2:
            g = (g * gelu_grad(z)).astype(dt)
        # dW = g^T h  (out, in), db = sum_batch g, dh = g W
        if exact:
            dW = g.T @ h
            dh = g @ np.asarray(W, np.float64)
        else:
            dW = (bf16(g).astype(np.float64).T @ bf16(h).astype(np.float64)).astype(np.float32)
            dh = (bf16(g).astype(np.float64) @ bf16(W).astype(np.float64)).astype(np.float32)
        db = g.sum(axis=0, dtype=np.float64).astype(dt)
        grads[i] = (dW, db)
        g = dh
    return grads


def adam_step(params, grads, state, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update in float32 (state: dict with 't' and moment lists)."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    if "m" not in state:
        state["m"] = [(np.zeros_like(W), np.zeros_like(b)) for W, b in params]
        state["v"] = [(np.zeros_like(W), np.zeros_like(b)) for W, b in params]
    out = []
    for i, ((W, b), (dW, db)) in enumerate(zip(params, grads)):
        new = []
        for j, (p, g) in enumerate(((W, dW), (b, db))):
            m = state["m"][i][j]
            v = state["v"][i][j]
            m[...] = beta1 * m + (1 - beta1) * g
            v[...] = beta2 * v + (1 - beta2) * g * g
            mh = m / (1 - beta1 ** t)
            vh = v / (1 - beta2 ** t)
            new.append((p - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32))
        out.append(tuple(new))
    return out


def scale(rho, floor=1e-30):
    """s = the largest |Re| or |Im| entry of rho (floored), in float32."""
    rho = np.asarray(rho, complex)
    m = max(float(np.abs(rho.real).max()), float(np.abs(rho.imag).max()))
    return float(np.float32(max(np.float32(m), np.float32(floor))))


def encode(rho, index, depth, floor=1e-30):
    """Network input of a last-tier ADO rho (N, N) with multi-index `index`
    in a hierarchy of last tier `depth`; returns (x, s)."""
    rho = np.asarray(rho, complex)
    s = scale(rho, floor)
    cm = rho.T.reshape(-1) / s            # column-major, as stored on the device
    x = np.concatenate([cm.real, cm.imag, np.asarray(index, float) / depth])
    return x.astype(np.float32), s


def encode_children(children, s):
    """Target row of the D next-tier ADOs (list of (N, N)) at scale s."""
    parts = []
    for c in children:
        cm = np.asarray(c, complex).T.reshape(-1) / s
        parts += [cm.real, cm.imag]
    return np.concatenate(parts).astype(np.float32)


def decode_children(y, s, N, D):
    """Network output row -> the D next-tier ADOs, each (N, N)."""
    ne = N * N
    y = np.asarray(y, float)
    out = []
    for j in range(D):
        blk = y[2 * ne * j:2 * ne * (j + 1)]
        cm = (blk[:ne] + 1j * blk[ne:]) * s
        out.append(cm.reshape(N, N).T)
    return out
