"""The CPU restatement of the far-field surrogate network (oracle/mlp.py) is
the checker of the GPU training kernels; its analytic gradient is pinned
here by central finite differences in FP64, and its BF16 rounding by exact
BF16 values."""
import numpy as np

from oracle import mlp


def _setup(seed=5, n=37):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.2, mlp.NPARAM)
    xn = rng.normal(size=(n, mlp.NIN))
    yn = rng.normal(size=(n, mlp.NOUT))
    return p, xn, yn


def test_gradient_matches_finite_differences():
    p, xn, yn = _setup()
    loss, g = mlp.loss_and_grad(p, xn, yn, bf=False)
    rng = np.random.default_rng(1)
    # a spread of parameters from every tensor
    picks = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False) for a, b in
                            [(mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2),
                             (mlp.OFF_B2, mlp.OFF_W3), (mlp.OFF_W3, mlp.OFF_B3)]] + [np.arange(mlp.OFF_B3, mlp.NPARAM)])
    h = 1e-6
    for k in picks:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mlp.loss_and_grad(pp, xn, yn)[0] - mlp.loss_and_grad(pm, xn, yn)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_bf16_rounding_is_exact_on_representable_values_and_ties_to_even():
    vals = np.array([1.0, -2.5, 0.15625, 3.0e-3], dtype=np.float32)
    r = mlp.bf16(vals)
    assert np.array_equal(r[:3], vals[:3].astype(np.float64))
    # 1 + 2^-8 lies halfway between BF16 neighbours 1 and 1 + 2^-7: ties to even (1)
    assert mlp.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    # 1 + 3 * 2^-8 lies halfway between 1 + 2^-7 and 1 + 2^-6: ties to even (1 + 2^-6)
    assert mlp.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
