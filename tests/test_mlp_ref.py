"""CPU checks of the surrogate MLP reference formulas (tests/mlp_ref.py) against finite differences:
the descriptor gradient g = dE/dB, and the parameter gradient of sum_a (alpha_a E_a + u_a . g_a),
which includes the second-order (tangent + reverse) term the force loss needs.  These pin the formulas
that the GPU kernels are then compared with."""
import numpy as np

import mlp_ref as R


def _case(seed=3, n=7, nin=14):
    rng = np.random.default_rng(seed)
    p = R.init_params(seed, nin).astype(np.float64)
    p[1024:1088] = rng.normal(0, 0.3, 64)
    p[5184:5248] = rng.normal(0, 0.3, 64)
    p[5312] = 0.2
    aff = np.zeros(34)
    aff[:nin] = rng.normal(0, 1, nin)
    aff[16:16 + nin] = rng.uniform(0.5, 2.0, nin)
    aff[32], aff[33] = 1.7, -3.0
    X = rng.normal(0, 1.5, (n, nin))
    return p, aff, nin, X, rng


def test_descriptor_gradient_matches_finite_differences():
    p, aff, nin, X, _ = _case()
    e, g = R.forward(p, aff, nin, X)
    h = 1e-6
    for k in range(nin):
        d = np.zeros_like(X)
        d[:, k] = h
        fd = (R.forward(p, aff, nin, X + d)[0] - R.forward(p, aff, nin, X - d)[0]) / (2 * h)
        assert np.max(np.abs(fd - g[:, k])) < 1e-7 * max(1.0, np.max(np.abs(g)))


def test_parameter_gradient_with_force_term_matches_finite_differences():
    p, aff, nin, X, rng = _case(5)
    U = rng.normal(0, 1, X.shape)
    alpha = rng.normal(0, 1, len(X))

    def loss(q):
        e, g = R.forward(q, aff, nin, X)
        return np.sum(alpha * e) + np.sum(U * g)

    grad = R.train_grad(p, aff, nin, X, U, alpha)
    h = 1e-6
    idx = list(range(0, 14 * 64, 37)) + list(range(1024, 1088, 5)) + list(range(1088, 5184, 211)) + \
        list(range(5184, 5248, 7)) + list(range(5248, 5312, 9)) + [5312]
    for j in idx:
        d = np.zeros_like(p)
        d[j] = h
        fd = (loss(p + d) - loss(p - d)) / (2 * h)
        assert abs(fd - grad[j]) < 1e-6 * max(1.0, abs(fd)), j


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float32)
    y = R.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to the even mantissa (1); 1 + 3*2^-8 rounds up
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == 1.0 + 2 ** -6 and y[3] == -2.5
    assert abs(y[4] - 3.0e-3) < 3.0e-3 * 2 ** -8
