"""Pins the numpy MLP oracle (oracle/mlp_oracle.py) that the device MLP of the
learned Mesher proposer is checked against: gradients against torch autograd
(float64) and central finite differences, GELU against its definition."""
import math

import numpy as np
import torch

from oracle import mlp_oracle as M


def _setup(seed=0, nin=7, width=5, depth=4, n=11):
    rng = np.random.default_rng(seed)
    npar = sum(o * i + o for o, i in M.shapes(nin, width, depth))
    p = rng.normal(0, 0.5, npar)
    x = rng.normal(0, 1, (n, nin))
    y = (rng.random(n) < 0.4).astype(float)
    w = rng.uniform(0.5, 2.0, n)
    mean = rng.normal(0, 0.1, nin)
    inv = rng.uniform(0.5, 2.0, nin)
    return p, x, y, w, mean, inv, (nin, width, depth)


def test_gradient_matches_torch_autograd():
    p, x, y, w, mean, inv, (nin, width, depth) = _setup()
    loss, g = M.loss_grad(p, x, y, w, nin, width, depth, mean, inv)
    P = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    a = (torch.tensor(x) - torch.tensor(mean)) * torch.tensor(inv)
    o = 0
    for l, (dn, di) in enumerate(M.shapes(nin, width, depth)):
        W = P[o:o + dn * di].reshape(dn, di)
        o += dn * di
        b = P[o:o + dn]
        o += dn
        a = a @ W.T + b
        if l + 1 < depth:
            a = torch.nn.functional.gelu(a)        # exact (erf) GELU
    z = a[:, 0]
    L = torch.nn.functional.binary_cross_entropy_with_logits(
        z, torch.tensor(y), weight=torch.tensor(w), reduction="sum") / float(w.sum())
    L.backward()
    assert abs(L.item() - loss) < 1e-12
    np.testing.assert_allclose(g, P.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_gradient_matches_finite_differences():
    p, x, y, w, mean, inv, (nin, width, depth) = _setup(seed=1, nin=4, width=3, depth=3, n=6)
    _, g = M.loss_grad(p, x, y, w, nin, width, depth, mean, inv)
    h = 1e-6
    for k in range(0, p.size, 3):
        e = np.zeros_like(p)
        e[k] = h
        lp, _ = M.loss_grad(p + e, x, y, w, nin, width, depth, mean, inv)
        lm, _ = M.loss_grad(p - e, x, y, w, nin, width, depth, mean, inv)
        assert abs((lp - lm) / (2 * h) - g[k]) < 1e-7


def test_gelu_definition():
    z = np.linspace(-4, 4, 17)
    phi = 0.5 * (1 + np.array([math.erf(v / math.sqrt(2)) for v in z]))
    np.testing.assert_allclose(M.gelu(z), z * phi, rtol=1e-15)
    assert M.gelu(np.array([0.0]))[0] == 0.0
