"""Pins of the CPU MLP restatement (oracle/mlp.py) that checks the HIP
training and inference kernels: finite differences, explicit loops, closed
forms and BF16 rounding on known bit patterns."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp as M


def _init(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in M.SHAPES:
        fan_in = shape[1] if len(shape) == 2 else 1
        d[name] = rng.uniform(-1, 1, shape) * scale * (np.sqrt(6.0 / fan_in) if len(shape) == 2 else 0.1)
    return M.pack(d)


def _batch(n, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 2, (n, 24))
    e = (rng.random(n) < 0.4).astype(np.float64)
    t = rng.random(n) * e
    return x, e, t


def test_parameter_layout():
    assert M.N_PARAMS == 24 * 128 + 128 + 128 * 128 + 128 + 2 * 128 + 2 == 19970
    w = np.arange(M.N_PARAMS, dtype=np.float64)
    p = M.unpack(w)
    assert p["W1"][1, 0] == 24 and p["b1"][0] == 3072 and p["W2"][0, 1] == 3201
    assert p["b3"][1] == M.N_PARAMS - 1
    assert np.array_equal(M.pack(p), w)


def test_forward_matches_explicit_loops():
    w = _init()
    x, _, _ = _batch(3)
    np.testing.assert_allclose(M.forward(w, x), M.forward_loops(w, x), rtol=1e-12, atol=1e-12)


def test_gradient_matches_central_differences():
    w = _init(scale=1.0)
    x, e, t = _batch(16)
    loss, g = M.loss_and_grad(w, x, e, t)
    rng = np.random.default_rng(5)
    # a sample of coordinates from every layer
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SHAPES])
    idx = np.concatenate([rng.choice(np.arange(a, b), size=min(12, b - a), replace=False)
                          for a, b in zip(offs[:-1], offs[1:])])
    h = 1e-6
    fds = {}
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (M.loss_and_grad(wp, x, e, t)[0] - M.loss_and_grad(wm, x, e, t)[0]) / (2 * h)
        fds[i] = fd
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    assert np.isfinite(loss)
    # planted mistakes in the gradient fail the same comparison
    for plant in ("no_sigmoid_deriv", "w2_untransposed"):
        _, gp = M.loss_and_grad(w, x, e, t, _plant=plant)
        bad = sum(abs(fds[i] - gp[i]) > 1e-6 * max(1.0, abs(fds[i])) for i in idx)
        assert bad > 0, plant


def test_loss_closed_forms():
    w = _init()
    x, _, _ = _batch(8)
    out = M.forward(w, x)
    z, s = out[:, 0], out[:, 1]
    e0 = np.zeros(8)
    l0, _ = M.loss_and_grad(w, x, e0, np.zeros(8))
    assert np.isclose(l0, np.mean(np.log1p(np.exp(z))))
    e1 = np.ones(8)
    t = np.full(8, 0.25)
    l1, _ = M.loss_and_grad(w, x, e1, t)
    assert np.isclose(l1, np.mean(np.log1p(np.exp(-z)) + (1 / (1 + np.exp(-s)) - 0.25) ** 2))


def test_bf16_rounding():
    def f(bits):
        return np.array([bits], np.uint32).view(np.float32)
    # exact values stay
    assert M.bf16(np.float32(1.0))[()] == 1.0
    # halfway: ties to even (0x3f808000 -> 0x3f80, 0x3f818000 -> 0x3f82)
    assert M.bf16(f(0x3f808000)).view(np.uint32)[0] == 0x3f800000
    assert M.bf16(f(0x3f818000)).view(np.uint32)[0] == 0x3f820000
    # above halfway rounds up, below rounds down
    assert M.bf16(f(0x3f808001)).view(np.uint32)[0] == 0x3f810000
    assert M.bf16(f(0x3f807fff)).view(np.uint32)[0] == 0x3f800000
    # NaN stays NaN, the largest float rounds to infinity
    assert np.isnan(M.bf16(f(0x7fc00001))[0])
    assert np.isinf(M.bf16(f(0x7f7fffff))[0])


@pytest.mark.parametrize("seed", range(5))
def test_cubic_coefficients_are_the_determinant(seed):
    rng = np.random.default_rng(seed)
    xo, v = rng.normal(size=(4, 3)), rng.normal(size=(4, 3))
    a, b, c, d = M.cubic_coeffs(xo, v)
    for t in (-1.3, 0.0, 0.4, 2.0):
        p = xo + t * v
        det = np.linalg.det(np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]]))
        assert np.isclose(((a * t + b) * t + c) * t + d, det, rtol=1e-10, atol=1e-12)


def test_features_are_translation_invariant():
    rng = np.random.default_rng(3)
    xo, v = rng.normal(size=(4, 3)), rng.normal(size=(4, 3))
    f1 = M.features(xo, v, 0.01, 0.5)
    f2 = M.features(xo + 7.0, v, 0.01, 0.5)
    np.testing.assert_allclose(f1, f2, atol=1e-5)
    assert np.allclose(f1[12:], (v * 0.01 / 0.5).ravel())
    assert np.allclose(f1[:12].reshape(4, 3).sum(0), 0, atol=1e-5)
