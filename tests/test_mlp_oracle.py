"""CPU pins of the MLP oracle (oracle/mlp_oracle.py) that the device network's
gradients and Adam steps are checked against on the GPU."""
import numpy as np

from oracle import mlp_oracle as mo


def tiny(seed=0, dims=(5, 7, 6, 3), batch=9):
    rng = np.random.default_rng(seed)
    n = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    p = rng.standard_normal(n) * 0.4
    X = rng.standard_normal((batch, dims[0]))
    T = rng.standard_normal((batch, dims[-1]))
    M = (rng.random((batch, dims[-1])) > 0.3).astype(float)
    return p, dims, X, T, M


def test_gradient_matches_central_differences():
    p, dims, X, T, M = tiny()
    _, g = mo.loss_grad(p, dims, X, T, M, bf16=False)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (mo.loss_grad(p + e, dims, X, T, M, bf16=False)[0] - mo.loss_grad(p - e, dims, X, T, M, bf16=False)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)


def test_masked_outputs_get_no_gradient_and_loss_is_masked_mean():
    p, dims, X, T, M = tiny(1)
    M[:, 0] = 0.0
    Y = mo.forward(p, dims, X, bf16=False)
    loss, _ = mo.loss_grad(p, dims, X, T, M, bf16=False)
    assert abs(loss - np.sum(M * (Y - T) ** 2) / M.sum()) < 1e-14
    T2 = T.copy()
    T2[:, 0] += 100.0            # targets under a zero mask change nothing
    assert mo.loss_grad(p, dims, X, T2, M, bf16=False)[0] == loss


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5e-3, 3.0e5], dtype=np.float32)
    r = mo.bf16_round(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                     # exactly half way between 1 and 1+2^-7: ties to even (1.0)
    assert r[2] == 1.0 + 2 ** -6           # half way between 1+2^-7 and 1+2^-6: ties to even (1+2^-6)
    assert r[3] == 1.0                     # below half way
    assert abs(r[4] / -2.5e-3 - 1) < 2 ** -8 and abs(r[5] / 3.0e5 - 1) < 2 ** -8


def test_first_adam_step_is_sign_step():
    rng = np.random.default_rng(2)
    p, g = rng.standard_normal(20), rng.standard_normal(20)
    z = np.zeros(20)
    p1, m, v = mo.adam(p, g, z, z, 1, 1e-3, eps=0.0)
    np.testing.assert_allclose(p1, p - 1e-3 * np.sign(g), rtol=0, atol=1e-15)
