"""Pins the CPU reference of the cost model: analytic gradient against
central finite differences (float64)."""
import numpy as np

import mlp_ref as R


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, R.NPARAM)
    x = rng.normal(0, 1, (40, 8))
    t = rng.normal(0, 1, (40, 3))
    m = rng.random((40, 3)) < 0.7
    _, g = R.loss_grad(p, x, t, m)
    for i in list(range(0, R.NPARAM, 37)) + [R.NPARAM - 1]:
        h = 1e-6
        q1, q2 = p.copy(), p.copy()
        q1[i] += h
        q2[i] -= h
        fd = (R.loss(q1, x, t, m) - R.loss(q2, x, t, m)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_layout_and_masking():
    assert R.NPARAM == 1443
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, R.NPARAM)
    x = rng.normal(0, 1, (10, 8))
    t = rng.normal(0, 1, (10, 3))
    l0, g0 = R.loss_grad(p, x, t, np.zeros((10, 3), bool))
    assert l0 == 0 and not g0.any()
    # a model with only biases outputs b3 everywhere
    q = np.zeros(R.NPARAM)
    q[-3:] = [1.0, 2.0, 3.0]
    y, _ = R.forward(q, x)
    assert np.allclose(y, [1.0, 2.0, 3.0])
