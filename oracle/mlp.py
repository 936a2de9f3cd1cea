"""CPU restatement of the time-of-impact proposer's network, its loss and
its gradient, in FP64 NumPy; and the CCD sub-test quantities it is
trained on (the 24 features, the coplanarity cubic's coefficients).

TEST INFRASTRUCTURE ONLY: the checker for the HIP training and inference
kernels and for the learned form's labels.  Loaded by tests/ and never by
the product package.  The network has no counterpart in the reference; the
labels follow isCoplanar (dcollid3d.cpp:371-482), which the C oracle
restates and pins.

Parity pins (tests/test_mlp_oracle.py): the analytic gradient matches
central finite differences of the FP64 loss; the forward pass matches a
layer-by-layer evaluation with explicit loops on a few samples; the loss
reduces to softplus(z) - y z for samples without a root; the BF16 rounding
helper matches round-to-nearest-even on hand-picked bit patterns; the cubic's
coefficients reproduce the determinant det[x_i(t) - x_0(t)] at sample times;
the features are translation invariant and sum to zero over the points.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 24, 128, 2
SHAPES = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
          ("W3", (N_OUT, N_H)), ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)
assert N_PARAMS == 19970


def unpack(w):
    """flat parameter vector -> dict of arrays (views, row = output unit)"""
    w = np.asarray(w)
    out, off = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = w[off:off + n].reshape(shape)
        off += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(a):
    """round float32 values to bfloat16 (nearest even), returned as float32"""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7fffffff) > 0x7f800000
    r = ((u + 0x7fff + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, 0x7fc00000, r).astype(np.uint32)
    return r.view(np.float32)


def softplus(z):
    return np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z)))


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(w, x, emulate_bf16=False):
    """(logit, raw seed) of each row of x [n][24]; with emulate_bf16 the
    inputs, weights and hidden activations are rounded to BF16 as on the
    device (sums stay exact-ish in FP64, the device accumulates in FP32)"""
    p = unpack(np.asarray(w, np.float64))
    x = np.asarray(x, np.float64)
    q = (lambda a: bf16(a).astype(np.float64)) if emulate_bf16 else (lambda a: a)
    h1 = q(np.maximum(q(x) @ q(p["W1"]).T + p["b1"], 0))
    h2 = q(np.maximum(h1 @ q(p["W2"]).T + p["b2"], 0))
    return h2 @ q(p["W3"]).T + p["b3"]


def loss_and_grad(w, x, exists, toi, _plant=None):
    """mean over the batch of softplus(z) - y z + y (sigmoid(s) - t)^2 and its
    gradient with respect to the flat parameters, FP64.  _plant (tests only)
    plants a known mistake in the gradient: "no_sigmoid_deriv" drops the
    seed's sigmoid derivative, "w2_untransposed" back-propagates through W2
    instead of its transpose"""
    p = unpack(np.asarray(w, np.float64))
    x = np.asarray(x, np.float64)
    y = np.asarray(exists, np.float64)
    t = np.asarray(toi, np.float64)
    n = x.shape[0]
    a1 = x @ p["W1"].T + p["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(a2, 0)
    out = h2 @ p["W3"].T + p["b3"]
    z, s = out[:, 0], out[:, 1]
    q = sigmoid(s)
    loss = np.mean(softplus(z) - y * z + y * (q - t) ** 2)
    dq = 1.0 if _plant == "no_sigmoid_deriv" else q * (1 - q)
    d3 = np.stack([sigmoid(z) - y, y * 2 * (q - t) * dq], axis=1) / n
    g = {"W3": d3.T @ h2, "b3": d3.sum(0)}
    d2 = (d3 @ p["W3"]) * (a2 > 0)
    g["W2"], g["b2"] = d2.T @ h1, d2.sum(0)
    d1 = (d2 @ (p["W2"].T if _plant == "w2_untransposed" else p["W2"])) * (a1 > 0)
    g["W1"], g["b1"] = d1.T @ x, d1.sum(0)
    return loss, pack(g)


def forward_loops(w, x):
    """the same network with explicit loops (for a handful of samples)"""
    p = unpack(np.asarray(w, np.float64))
    res = []
    for row in np.asarray(x, np.float64):
        h1 = [max(0.0, sum(p["W1"][o, i] * row[i] for i in range(N_IN)) + p["b1"][o])
              for o in range(N_H)]
        h2 = [max(0.0, sum(p["W2"][o, i] * h1[i] for i in range(N_H)) + p["b2"][o])
              for o in range(N_H)]
        res.append([sum(p["W3"][o, i] * h2[i] for i in range(N_H)) + p["b3"][o]
                    for o in range(N_OUT)])
    return np.array(res)


def features(xo, v, dt, h):
    """the 24 features of a CCD sub-test: x_old relative to the four points'
    centroid, then avgVel * dt, both over the thickness h"""
    xo = np.asarray(xo, np.float64)
    v = np.asarray(v, np.float64)
    c = 0.25 * (((xo[0] + xo[1]) + xo[2]) + xo[3])
    return np.concatenate([((xo - c) / h).ravel(), (v * dt / h).ravel()]).astype(np.float32)


def cubic_coeffs(xo, v):
    """a, b, c, d of det[x_i + t v_i] relative to point 0 (dcollid3d.cpp:389-427)"""
    xo = np.asarray(xo, np.float64)
    v = np.asarray(v, np.float64)
    x = xo[1:] - xo[0]
    w = v[1:] - v[0]
    # det(x1 + t w1, x2 + t w2, x3 + t w3) expanded by powers of t
    def det(a, b, c):
        return float(np.dot(c, np.cross(a, b)))
    a = det(w[0], w[1], w[2])
    b = det(x[0], w[1], w[2]) + det(w[0], x[1], w[2]) + det(w[0], w[1], x[2])
    c = det(x[0], x[1], w[2]) + det(x[0], w[1], x[2]) + det(w[0], x[1], x[2])
    d = det(x[0], x[1], x[2])
    return a, b, c, d
