"""The stored training set of the spin-up network and the error measure that
compares the warm start with the prior (host code, no GPU)."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def test_training_set_covers_2000_basins():
    from cmf import surrogate as sg
    X, Y, w, bid = sg.load_training_set(os.path.join(HERE, "golden", "surrogate_train.npz"))
    assert X.shape == (Y.shape[0], sg.N_IN) and Y.shape[1] == sg.N_OUT
    assert np.isfinite(X).all() and np.isfinite(Y).all()
    assert np.unique(bid).size == 2000
    assert np.bincount(bid).max() <= 120
    assert sg.fast_cells(X).all()                      # only cells the network corrects
    assert (w >= 0.05).all() and (w <= 50.0).all()
    # the target is log(equilibrium / prior); the prior is within a few percent
    assert 0.005 < np.mean(np.abs(Y[:, 0])) < 0.1


def test_correction_error_identities():
    from cmf import surrogate as sg
    rng = np.random.default_rng(0)
    n = 1000
    X = np.zeros((n, sg.N_IN), np.float32)
    X[:, 6] = np.log(10.0)                             # 10 days of memory: every cell corrected
    prior = rng.uniform(1e3, 1e6, n)
    truth = rng.normal(0.0, 0.05, n)
    eq = prior * np.exp(truth)
    exact = sg.correction_error(truth[:, None], X, eq, prior)
    assert exact["warm_log_err"] < 1e-12 and exact["warm_l1"] < 1e-12
    none = sg.correction_error(np.zeros((n, 1)), X, eq, prior)
    assert abs(none["warm_log_err"] - none["prior_log_err"]) < 1e-15
    assert abs(none["warm_l1"] - none["prior_l1"]) < 1e-15
    # a correction of the right sign and half the size halves the log error
    half = sg.correction_error(0.5 * truth[:, None], X, eq, prior)
    assert abs(half["warm_log_err"] - 0.5 * half["prior_log_err"]) < 1e-12
    # cells with long memory keep the prior whatever the network says
    X[:, 6] = np.log(1000.0)
    slow = sg.correction_error(truth[:, None], X, eq, prior)
    assert abs(slow["warm_l1"] - slow["prior_l1"]) < 1e-15
