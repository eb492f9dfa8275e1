"""Pins of the oracle's step-size predictor (no GPU): a known-answer forward,
bfloat16 round-to-nearest-even ties, central finite differences of the FP64
backward, and the censored loss."""
import numpy as np

from oracle import oracle as O

NI, NH = 8, 32
O_MU, O_IS = 0, 8
O_W1 = 16
O_B1 = O_W1 + NH * NI
O_W2 = O_B1 + NH
O_B2 = O_W2 + NH * NH
O_W3 = O_B2 + NH
O_B3 = O_W3 + NH


def _identity_params(b3=0.25):
    P = np.zeros(O.mlp_n_params())
    P[O_IS:O_IS + NI] = 1.0
    W1 = np.zeros((NH, NI))
    W1[:NI, :NI] = np.eye(NI)
    P[O_W1:O_B1] = W1.ravel()
    P[O_W2:O_B2] = np.eye(NH).ravel()
    P[O_W3:O_B3] = 1.0
    P[O_B3] = b3
    return P


def test_layout_size():
    assert O.mlp_n_params() == 16 + NH * NI + NH + NH * NH + NH + NH + 1


def test_forward_known_answer_identity_network():
    # W1 = [I; 0], W2 = I, W3 = 1: out = sum_i relu((x_i - mu_i) * s_i) + b3
    P = _identity_params()
    P[O_MU:O_MU + NI] = np.arange(NI) * 0.5
    P[O_IS:O_IS + NI] = 2.0
    X = np.array([[1.0, -2.0, 3.0, 0.25, 8.0, 1.5, -0.5, 4.0],
                  [0.0, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 3.5]], np.float32)
    want = np.maximum((X - P[O_MU:O_MU + NI]) * 2.0, 0).sum(axis=1) + 0.25
    for quant in (0, 1):   # every value above is exact in bfloat16
        np.testing.assert_allclose(O.mlp_forward(P, X, quant), want, rtol=0, atol=1e-12)


def test_forward_bfloat16_rounding_ties():
    # one active path: out = q(x) for quant=1; 1 + 2^-8 is a tie between
    # 1 and 1 + 2^-7 (round to even -> 1), 1 + 3*2^-8 -> 1 + 2^-6
    P = _identity_params(b3=0.0)
    P[O_W3:O_B3] = 0.0
    P[O_W3] = 1.0
    X = np.zeros((3, NI), np.float32)
    X[:, 0] = [1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -9]
    np.testing.assert_array_equal(O.mlp_forward(P, X, 1), [1.0, 1 + 2 ** -6, 1.0])
    np.testing.assert_allclose(O.mlp_forward(P, X, 0), X[:, 0], rtol=1e-7)


def _random_problem(seed, n=64):
    rng = np.random.default_rng(seed)
    P = np.zeros(O.mlp_n_params())
    P[O_MU:O_MU + NI] = rng.normal(0, 1, NI)
    P[O_IS:O_IS + NI] = rng.uniform(0.5, 2.0, NI)
    P[O_W1:O_B3 + 1] = rng.normal(0, 0.4, O_B3 + 1 - O_W1)
    X = rng.normal(0, 1.5, (n, NI)).astype(np.float32)
    T = rng.normal(0, 1, n).astype(np.float32)
    C = (rng.random(n) < 0.3).astype(np.uint8)
    return P, X, T, C


def test_backward_matches_central_differences():
    P, X, T, C = _random_problem(3)
    loss, g = O.mlp_loss_grad(P, X, T, C, quant=0)
    assert loss > 0
    assert np.all(g[:16] == 0)
    h = 1e-6
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(np.arange(16, P.size), 300, replace=False), [O_B3]])
    for k in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        fd = (O.mlp_loss_grad(Pp, X, T, C, 0)[0] - O.mlp_loss_grad(Pm, X, T, C, 0)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_quantized_gradient_is_straight_through():
    # quant=1 differentiates the rounded network with respect to the
    # rounded weights: with weights and inputs already on the bfloat16 grid
    # and no activation rounding effect it equals the real-valued gradient
    P = _identity_params(b3=0.0)
    X = np.array([[1.0, 2.0, 0.5, 0.25, 1.5, 3.0, 0.75, 2.5]], np.float32)
    T = np.array([1.0], np.float32)
    C = np.zeros(1, np.uint8)
    l0, g0 = O.mlp_loss_grad(P, X, T, C, 0)
    l1, g1 = O.mlp_loss_grad(P, X, T, C, 1)
    assert l0 == l1
    np.testing.assert_array_equal(g0, g1)


def test_censored_samples():
    P, X, T, C = _random_problem(5, n=8)
    out = O.mlp_forward(P, X, 0)
    # targets below the output: censored -> zero loss and gradient
    T1 = (out - 1.0).astype(np.float32)
    loss, g = O.mlp_loss_grad(P, X, T1, np.ones(8, np.uint8), 0)
    assert loss == 0 and np.all(g == 0)
    # targets above the output: censoring does not matter
    T2 = (out + 1.0).astype(np.float32)
    la, ga = O.mlp_loss_grad(P, X, T2, np.ones(8, np.uint8), 0)
    lb, gb = O.mlp_loss_grad(P, X, T2, np.zeros(8, np.uint8), 0)
    assert la == lb and np.array_equal(ga, gb)
    np.testing.assert_allclose(la, np.mean((out - T2.astype(np.float64)) ** 2), rtol=1e-12)
