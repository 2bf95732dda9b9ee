"""Pins of the learned-closure oracle (oracle/closure_oracle.py) on the CPU.

* its gradient equals central finite differences of its loss (roundings off, FP64);
* BF16 rounding is round-to-nearest-even on known values, and the host packing of the device block uses
  the same rounding;
* the features' invariants equal their matrix definitions (tr(q^3) via a matrix power).
A planted mistake (a dropped tanh derivative) is caught by the finite-difference pin.
"""
import numpy as np
import pytest

from oracle import closure_oracle as co


def _setup(seed=0, n=200):
    rng = np.random.default_rng(seed)
    master = rng.normal(0, 0.3, co.MASTER_SIZE)
    master[co.OFF_W1T:co.OFF_W2T].reshape(co.H, co.IN_PAD)[:, co.IN:] = 0.0
    x = rng.normal(0, 1, (n, co.IN))
    y = rng.uniform(0.1, 2.0, n)
    return master, np.zeros(co.IN), np.ones(co.IN), x, y


def _fd_check(grad_fn, master, shift, scale, x, y, idx):
    _, g = grad_fn(master, shift, scale, x, y, round_bf16=False)
    h = 1e-6
    for i in idx:
        mp, mm = master.copy(), master.copy()
        mp[i] += h
        mm[i] -= h
        lp, _ = co.loss_and_grad(mp, shift, scale, x, y, round_bf16=False)
        lm, _ = co.loss_and_grad(mm, shift, scale, x, y, round_bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(g[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, g[i], fd)


def test_gradient_matches_finite_differences():
    master, shift, scale, x, y = _setup()
    rng = np.random.default_rng(1)
    idx = list(rng.choice(np.arange(co.MASTER_SIZE), 40, replace=False))
    idx = [i for i in idx if not (i < co.OFF_W2T and i % co.IN_PAD >= co.IN)] + [co.OFF_B3, co.OFF_W3, co.OFF_B1]
    _fd_check(co.loss_and_grad, master, shift, scale, x, y, idx)


def test_planted_mistake_is_caught(monkeypatch):
    master, shift, scale, x, y = _setup(seed=3)

    def wrong(master, shift, scale, x, y, round_bf16=True):
        o, (a0, a1, h1, h2, w2t, w3) = co.forward(master, shift, scale, x, round_bf16)
        n = len(x)
        diff = co.softplus(o) - y
        g = 2.0 / n * diff * co.sigmoid(o)
        d2 = g[:, None] * w3[None, :]                   # tanh derivative dropped
        d1 = (d2 @ w2t) * (1 - h1 ** 2)
        grad = np.zeros(co.MASTER_SIZE)
        grad[co.OFF_W2T:co.OFF_W3] = (d2.T @ a1).ravel()
        grad[co.OFF_W1T:co.OFF_W2T] = (d1.T @ a0).ravel()
        return np.mean(diff ** 2), grad
    with pytest.raises(AssertionError):
        _fd_check(wrong, master, shift, scale, x, y, [co.OFF_W2T + 5, co.OFF_W2T + 77])


def test_bf16_rounding():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)                     # BF16 spacing at 1
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 0.0], dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 * ulp, 1.0 + ulp, -2.5, 0.0])
    np.testing.assert_array_equal(co.bf16(vals), want)
    assert one == 1.0


def test_host_packing_uses_the_same_rounding():
    from lbmpy_amd.closure import _bf16_bits
    v = np.random.default_rng(4).normal(0, 3, 1000).astype(np.float32)
    bits = _bf16_bits(v).astype(np.uint32) << 16
    np.testing.assert_array_equal(bits.view(np.float32).astype(np.float64), co.bf16(v))


def test_feature_invariants():
    rng = np.random.default_rng(5)
    p = rng.normal(0, 1e-3, (50, 6))
    rho = rng.uniform(0.9, 1.1, 50)
    f = co.features(p, rho)
    for k in range(50):
        q = p[k] / rho[k]
        Q = np.array([[q[0], q[3], q[4]], [q[3], q[1], q[5]], [q[4], q[5], q[2]]])
        assert np.isclose(f[k, 6], np.sqrt(2 * np.sum(Q * Q)))
        assert np.isclose(f[k, 7], np.trace(Q))
        assert np.isclose(f[k, 8], np.sum(Q * Q))
        assert np.isclose(f[k, 9], np.trace(Q @ Q @ Q))
