"""Pins of the closure-network oracle (CPU only).

* bfloat16 rounding equals torch's float32 -> bfloat16 conversion;
* the exact (float64) backward pass equals central finite differences of
  the loss for every parameter tensor;
* the kernel-arithmetic mode (bfloat16 GEMM operands, float32 sums) stays
  within bfloat16 rounding of the exact mode;
* GELU and its derivative against their closed forms at known points;
* Adam matches a hand-computed first step.
"""

import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-8, 8, 20000),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -0.0, 65504.0, 1e-40],
                                 dtype=np.float32)])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16(x), ref)
    # ties go to even
    assert M.bf16(np.float32(1.0 + 2 ** -8))[()] == 1.0


def test_gelu_closed_form():
    assert M.gelu(0.0) == 0.0
    assert abs(M.gelu(1.0) - 0.8413447460685429) < 1e-15
    assert abs(M.gelu_grad(0.0) - 0.5) < 1e-15
    z = np.linspace(-4, 4, 41)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.abs(fd - M.gelu_grad(z)).max() < 1e-8


def _problem(seed=0, n_in=13, n_hid=16, n_out=9, batch=7):
    rng = np.random.default_rng(seed)
    params = M.init_params(n_in, n_hid, n_out, seed=seed)
    params = [(W.astype(np.float64), (rng.standard_normal(b.shape) * 0.1)) for W, b in params]
    x = rng.standard_normal((batch, n_in))
    t = rng.standard_normal((batch, n_out))
    return params, x, t


def test_exact_backward_equals_finite_differences():
    params, x, t = _problem()
    y, cache = M.forward(params, x, exact=True)
    grads = M.backward(params, cache, y, t, exact=True)
    h = 1e-6
    for li in range(3):
        for pi in range(2):
            p = params[li][pi]
            g = grads[li][pi]
            it = np.nditer(p, flags=["multi_index"])
            checked = 0
            for _ in it:
                idx = it.multi_index
                old = p[idx]
                p[idx] = old + h
                lp = M.mse(M.forward(params, x, exact=True)[0], t)
                p[idx] = old - h
                lm = M.mse(M.forward(params, x, exact=True)[0], t)
                p[idx] = old
                fd = (lp - lm) / (2 * h)
                assert abs(fd - g[idx]) <= 1e-7 * max(1.0, abs(fd)), (li, pi, idx)
                checked += 1
            assert checked == p.size


def test_kernel_mode_within_bf16_rounding_of_exact():
    params, x, t = _problem(seed=3, n_in=105, n_hid=256, n_out=98, batch=64)
    p32 = [(W.astype(np.float32), b.astype(np.float32)) for W, b in params]
    ye, ce = M.forward(params, x, exact=True)
    yk, ck = M.forward(p32, x.astype(np.float32), exact=False)
    scale = np.abs(ye).max()
    # three layers of bf16 operands: a few 2^-8 relative
    assert np.abs(yk - ye).max() < 3e-2 * scale
    assert np.abs(yk - ye).max() > 0  # the rounding is really applied
    ge = M.backward(params, ce, ye, t, exact=True)
    gk = M.backward(p32, ck, yk, t.astype(np.float32), exact=False)
    for (dWe, dbe), (dWk, dbk) in zip(ge, gk):
        assert np.abs(dWk - dWe).max() < 5e-2 * np.abs(dWe).max()
        assert np.abs(dbk - dbe).max() < 5e-2 * np.abs(dbe).max()


def test_adam_first_step():
    W = np.array([[1.0, -2.0]], np.float32)
    b = np.array([0.5], np.float32)
    g = [(np.array([[0.1, -0.3]], np.float32), np.array([2.0], np.float32))]
    st = {}
    (W1, b1), = M.adam_step([(W, b)], g, st, lr=0.01)
    # first step: m_hat = g, v_hat = g^2 -> update = lr * sign(g) (eps aside)
    assert np.allclose(W1, W - 0.01 * np.sign(g[0][0]), atol=1e-6)
    assert np.allclose(b1, b - 0.01, atol=1e-6)


def test_encode_decode_roundtrip():
    rng = np.random.default_rng(5)
    rho = rng.standard_normal((7, 7)) + 1j * rng.standard_normal((7, 7))
    x, s = M.encode(rho, [1, 0, 2, 0, 0, 1, 0], 4)
    assert x.shape == (105,)
    assert s == np.float32(max(np.abs(rho.real).max(), np.abs(rho.imag).max()))
    assert np.allclose(x[98:], np.array([1, 0, 2, 0, 0, 1, 0]) / 4)
    kids = [rng.standard_normal((7, 7)) + 1j * rng.standard_normal((7, 7)) for _ in range(7)]
    y = M.encode_children(kids, s)
    assert y.shape == (686,)
    for a, b in zip(M.decode_children(y, s, 7, 7), kids):
        assert np.allclose(a, b, atol=1e-6 * np.abs(b).max())
    # the first block of a child row is its column-major real part
    assert np.allclose(y[:49] * s, kids[0].T.reshape(-1).real, atol=1e-5)
