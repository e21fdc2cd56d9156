"""CPU restatement (numpy, float64) of the device MLP of the learned Mesher
proposer (bsalib_amd/csrc/mlp.hip) -- test infrastructure, not product code.
This network has no counterpart in the reference; the oracle is pinned by
torch autograd and by central finite differences in tests/test_mlp_oracle.py.

Layout: params = concat over layers of W_l (out x in, row-major) then b_l;
hidden layers GELU(x) = x Phi(x); loss = weighted binary cross-entropy on the
logit, normalised by the weight sum."""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf


def shapes(nin, width, depth):
    return [((1 if l + 1 == depth else width), (nin if l == 0 else width)) for l in range(depth)]


def unpack(params, nin, width, depth):
    out, o = [], 0
    for (dn, di) in shapes(nin, width, depth):
        W = np.asarray(params[o:o + dn * di], dtype=np.float64).reshape(dn, di)
        o += dn * di
        b = np.asarray(params[o:o + dn], dtype=np.float64)
        o += dn
        out.append((W, b))
    assert o == len(params)
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def dgelu(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def forward(params, x, nin, width, depth, mean=None, inv=None):
    a = np.asarray(x, dtype=np.float64)
    if mean is not None:
        a = (a - mean) * inv
    cache = [a]
    layers = unpack(params, nin, width, depth)
    for l, (W, b) in enumerate(layers):
        z = a @ W.T + b
        if l + 1 < depth:
            cache.append(z)
            a = gelu(z)
            cache.append(a)
        else:
            a = z
    return a[:, 0], cache


def loss_grad(params, x, y, w, nin, width, depth, mean=None, inv=None):
    z, cache = forward(params, x, nin, width, depth, mean, inv)
    y = np.asarray(y, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    ws = w.sum()
    sp = np.logaddexp(0.0, z)                  # softplus(z)
    loss = float(np.sum(w * (y * (sp - z) + (1 - y) * sp)) / ws)
    dz = (w * (1.0 / (1.0 + np.exp(-z)) - y) / ws)[:, None]
    layers = unpack(params, nin, width, depth)
    grads = []
    d = dz
    for l in range(depth - 1, -1, -1):
        W, _ = layers[l]
        a_in = cache[0] if l == 0 else cache[2 * l]
        gW = d.T @ a_in
        gb = d.sum(axis=0)
        grads.append((gW, gb))
        if l > 0:
            zprev = cache[2 * l - 1]
            d = (d @ W) * dgelu(zprev)
    flat = []
    for gW, gb in reversed(grads):
        flat.append(gW.ravel())
        flat.append(gb)
    return loss, np.concatenate(flat)


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    g = g + wd * p
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
