"""GPU parity of the device MLP (bsahip_mlp_*) with the numpy oracle
(oracle/mlp_oracle.py, itself pinned against torch autograd): forward logits,
loss and gradients, one Adam step, and a training run that separates a
two-class data set.  FP32 on the device against FP64: tolerance 1e-4 normwise
relative (accumulations over at most a few hundred terms)."""
import numpy as np
import pytest

from bsalib_amd.mlp import MLP
from oracle import mlp_oracle as MO
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _data(seed, n, nin):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, nin)).astype(np.float32)
    y = (rng.random(n) < 0.3).astype(np.float32)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    mean = rng.normal(0, 0.2, nin).astype(np.float32)
    inv = rng.uniform(0.5, 1.5, nin).astype(np.float32)
    return x, y, w, mean, inv


def test_forward_and_gradient_match_oracle(gpu_lib):
    nin, width, depth = 37, 128, 4
    x, y, w, mean, inv = _data(0, 300, nin)
    net = MLP(nin, width, depth, seed=3)
    net.set(mean=mean, inv_scale=inv)
    p = net.params().astype(np.float64)
    z = net.forward(x)
    zo, _ = MO.forward(p, x, nin, width, depth, mean, inv)
    assert rel_err(z, zo) < 1e-4
    loss, g = net.loss_grad(x, y, w)
    lo, go = MO.loss_grad(p, x, y, w, nin, width, depth, mean, inv)
    assert abs(loss - lo) < 1e-4 * abs(lo)
    o = 0
    for (dn, di) in MO.shapes(nin, width, depth):        # per layer, weights and biases
        for n in (dn * di, dn):
            assert rel_err(g[o:o + n], go[o:o + n]) < 1e-4
            o += n
    net.close()


def test_one_adam_step_matches_oracle(gpu_lib):
    nin, width, depth = 9, 32, 3
    x, y, w, mean, inv = _data(1, 64, nin)
    net = MLP(nin, width, depth, seed=5)
    net.set(mean=mean, inv_scale=inv)
    p0 = net.params().astype(np.float64)
    net.train(x, y, w, epochs=1, batch=64, lr=1e-2, weight_decay=1e-3)
    p1 = net.params()
    _, g = MO.loss_grad(p0, x, y, w, nin, width, depth, mean, inv)
    pe, _, _ = MO.adam_step(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1, 1e-2, wd=1e-3)
    assert rel_err(p1 - p0, pe - p0) < 1e-3
    net.close()


def test_training_separates_two_classes(gpu_lib):
    rng = np.random.default_rng(7)
    n, nin = 4096, 16
    x = rng.normal(0, 1, (n, nin)).astype(np.float32)
    y = ((x[:, 0] * x[:, 1] + 0.5 * x[:, 2]) > 0.2).astype(np.float32)   # not linearly separable
    net = MLP(nin, 64, 4, seed=1)
    losses = net.train(x, y, epochs=30, batch=256, lr=3e-3)
    acc = np.mean((net.forward(x) > 0) == (y > 0.5))
    print(f"loss {losses[0]:.3f} -> {losses[-1]:.3f}, accuracy {acc:.3f}")
    assert losses[-1] < 0.3 * losses[0] and acc > 0.95
    net.close()
