"""Pins of the surrogate network's CPU restatement (tests/mlp_oracle.py):
its backward pass against central finite differences of its own forward
pass, in FP64 without BF16 rounding."""
import numpy as np

import mlp_oracle as M


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    dims = [6, 12, 10, 9]
    npar = sum(k * n + n for k, n in zip(dims[:-1], dims[1:]))
    p = rng.standard_normal(npar) * 0.4
    x = rng.standard_normal((5, dims[0]))
    t = rng.standard_normal((5, dims[-1]))
    loss, g = M.loss_and_grad(p, dims, x, t, bf=False)
    h = 1e-6
    fd = np.empty(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        fd[i] = (M.loss_and_grad(p + e, dims, x, t, bf=False)[0] - M.loss_and_grad(p - e, dims, x, t, bf=False)[0]) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-30], dtype=np.float32)
    got = M.bf16(x)
    want = np.array([1.0, 1.0, 1.0 + 4 * 2 ** -8, 1.0, -2.5, got[-1]])
    assert np.array_equal(got[:5], want[:5])
    assert abs(got[-1] - 3.0e-30) / 3.0e-30 < 2 ** -8
