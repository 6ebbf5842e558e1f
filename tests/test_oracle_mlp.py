"""CPU: the FP64 oracle of the learned-h network, pinned by finite differences, closed forms and bit patterns
(the network is new code: there is no reference vector to pin it), plus the host-side weight handling."""
import math

import numpy as np
import pytest

from botheadstudios_amd import mlp
from oracle import mlp_oracle as mo


def _rand(seed=0, n=257):
    rng = np.random.default_rng(seed)
    w = mlp.init_weights(seed).astype(np.float64)
    w[mlp.NPARAM - 33:mlp.NPARAM - 1] = rng.normal(0, 0.3, 32)  # a non-trivial output layer
    x = rng.normal(0, 1, (n, 6))
    t = rng.normal(0, 1, n)
    return w, x, t


def test_gradient_matches_central_differences():
    w, x, t = _rand()
    _, g = mo.loss_and_grad(w, x, t)
    rng = np.random.default_rng(5)
    # every bias and output weight, plus a sample of the weight matrices
    idx = np.unique(np.concatenate([np.arange(192, 224), np.arange(1248, 1313), rng.choice(1313, 120, False)]))
    eps = 1e-6
    fd = np.empty(len(idx))
    for k, i in enumerate(idx):
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        fd[k] = (mo.loss_and_grad(wp, x, t)[0] - mo.loss_and_grad(wm, x, t)[0]) / (2 * eps)
    err = np.abs(fd - g[idx])
    assert np.max(err / (np.abs(g[idx]) + 1e-6 * np.abs(g).max())) < 1e-5


def test_one_path_network_closed_form():
    """Only W1[0,0], W2[0,0], W3[0] and b3 non-zero: y = v gelu(u gelu(a x0)) + c (gelu(0) = 0 elsewhere)."""
    w = np.zeros(mo.NPARAM)
    a, u, v, c = 0.7, -1.3, 2.1, 0.25
    w[0] = a
    w[224] = u
    w[1280] = v
    w[1312] = c
    x = np.linspace(-3, 3, 13)[:, None] * np.array([[1, 0, 0, 0, 0, 0]]) + np.array([[0, 9, -9, 4, 1, 5]])

    def g(z):
        return 0.5 * z * (1 + math.tanh(math.sqrt(2 / math.pi) * (z + 0.044715 * z ** 3)))
    want = [v * g(u * g(a * x0)) + c for x0 in x[:, 0]]
    assert np.allclose(mo.forward(w, x), want, rtol=1e-14, atol=1e-14)


def test_gelu_known_values():
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    # Large |z|: identity and zero.
    assert mo.gelu(np.array([10.0]))[0] == pytest.approx(10.0, rel=1e-15)
    assert abs(mo.gelu(np.array([-10.0]))[0]) < 1e-15
    # The tanh form is within 1e-3 of the exact erf GELU (Hendrycks & Gimpel 2016).
    z = np.linspace(-5, 5, 101)
    exact = 0.5 * z * (1 + np.array([math.erf(v / math.sqrt(2)) for v in z]))
    assert np.max(np.abs(mo.gelu(z) - exact)) < 1e-3
    # derivative against differences
    h = 1e-6
    assert np.allclose(mo.gelu_grad(z), (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h), atol=1e-8)


def test_bf16_rounding_bit_patterns():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159265, 65504.0, 1e-3], np.float32)
    r = mo.bf16_round(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0          # exactly half an ulp above 1.0: ties to even (mantissa 0)
    assert r[2] == 1.0078125    # 0.75 ulp above 1.0 (ulp = 2^-7): rounds up to 1 + 2^-7
    assert r[3] == 1.0078125    # representable
    assert (r.view(np.uint32) & 0xFFFF == 0).all()
    assert np.all(np.abs(r - vals) <= np.abs(vals) * 2.0 ** -8)


def test_weights_roundtrip(tmp_path):
    w = mlp.init_weights(3)
    p = str(tmp_path / "w.npz")
    mlp.save_weights(p, w, {"a": 1})
    assert np.array_equal(mlp.load_weights(p), w)
    assert mlp.load_meta(p) == {"a": 1}
    W1, b1, W2, b2, W3, b3 = mlp.unpack(w)
    assert W1.shape == (32, 6) and W2.shape == (32, 32) and W3.shape == (32,) and b3.shape == (1,)
    assert np.array_equal(mo.unpack(w)[0], W1)


def test_seed_hit_rate_definition():
    rows = np.zeros((4, 8), np.float32)
    rows[:, 6] = [0.0, 0.001, 0.01, 1.0]
    pred = np.zeros(4, np.float32)
    # |err| * 0.01 < 1e-4 / 3  <=>  |err| < 3.33e-3
    assert mlp.seed_hit_rate(rows, pred) == 0.5
