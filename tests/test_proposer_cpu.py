"""CPU tests of the proposer's host logic: the greedy cost-aware level
selection that makes the training targets, the ordinal decoding of the logit,
tile costs, and the committed weights file (loaded without pickle)."""
import numpy as np

from bsalib_amd import proposer as PR


def test_greedy_levels_by_hand():
    # tile 1 buys 9e-4 for 10 evaluations (best ratio) -> level 2; the bound is then
    # 2.2e-4 > 2e-4, next best is tile 0 -> level 2 (5e-5 for 10) -> 1.7e-4 <= budget
    err = np.array([[1e-4, 5e-5, 1e-5, 0], [1e-3, 1e-4, 1e-5, 0], [2e-5, 1e-5, 0, 0]])
    cost = np.array([[0, 10, 50, 200]] * 3, dtype=float)
    np.testing.assert_array_equal(PR.greedy_levels(err, cost, 2e-4), [1, 1, 0])
    np.testing.assert_array_equal(PR.greedy_levels(err, cost, 1.0), [0, 0, 0])
    np.testing.assert_array_equal(PR.greedy_levels(err, cost, 0.0), [3, 3, 2])


def test_ordinal_decoding_midpoints():
    # p = sigmoid(z): level 1 below 1/6, 2 below 1/2, 4 below 5/6, else 8
    z = np.log(np.array([0.1, 0.2, 0.49, 0.51, 0.8, 0.9]) / (1 - np.array([0.1, 0.2, 0.49, 0.51, 0.8, 0.9])))
    np.testing.assert_array_equal(PR.levels_from_logits(z, 0.0), [1, 2, 2, 4, 4, 8])
    np.testing.assert_array_equal(PR.levels_from_logits(z + 10, 10.0), [1, 2, 2, 4, 4, 8])


def test_tile_costs_count_new_points():
    geom = np.zeros((1, 10))
    geom[0, 4], geom[0, 5] = 2, 3
    # level l grid (2l+1)(3l+1) minus the 12 level-0 points
    np.testing.assert_array_equal(PR.tile_costs(geom)[0], [0, 35 - 12, 117 - 12, 425 - 12])


def test_committed_weights_load_without_pickle():
    with np.load(PR.WEIGHTS, allow_pickle=False) as z:
        nin, width, depth = int(z["nin"]), int(z["width"]), int(z["depth"])
        assert (nin, width, depth) == (PR.NFEAT, 128, 4)
        npar = sum(o * i + o for o, i in [(width, nin), (width, width), (width, width), (1, width)])
        assert z["params"].shape == (npar,) and np.all(np.isfinite(z["params"]))
        assert z["mean"].shape == (nin,) and z["inv"].shape == (nin,)
