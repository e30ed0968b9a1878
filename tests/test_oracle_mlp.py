"""Pins the FP64 MLP restatement (oracle/mlp.py): with the BF16 rounding
switched off, its gradients equal central finite differences of its loss."""
import numpy as np

from oracle import mlp


def test_mlp_oracle_gradients_finite_differences(monkeypatch):
    monkeypatch.setattr(mlp, "bf16", lambda a: np.asarray(a, dtype=np.float64))
    rng = np.random.default_rng(0)
    dims = (7, 6, 5, 4)
    params = [(rng.standard_normal((dims[l], dims[l + 1])), rng.standard_normal(dims[l + 1])) for l in range(3)]
    X = rng.standard_normal((3, 7))
    T = rng.standard_normal((3, 4))
    _, _, g = mlp.forward_backward(params, X, T)
    eps = 1e-6
    for l in range(3):
        for which in (0, 1):
            arr = params[l][which]
            G = g[l][which]
            for idx in [tuple(rng.integers(0, s) for s in arr.shape) for _ in range(4)]:
                old = arr[idx]
                arr[idx] = old + eps
                lp, _, _ = mlp.forward_backward(params, X, T)
                arr[idx] = old - eps
                lm, _, _ = mlp.forward_backward(params, X, T)
                arr[idx] = old
                fd = (lp - lm) / (2 * eps)
                assert abs(fd - G[idx]) < 1e-6 * max(1.0, abs(fd))


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38], dtype=np.float64)
    r = mlp.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie -> even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
