"""The MLP oracle's backpropagation against central finite differences (FP64),
and its bfloat16 rounding against the definition."""
import numpy as np

from oracle import mlp


def _net(rng, dims):
    return [(rng.normal(0, 1 / np.sqrt(a), (a, b)), rng.normal(0, 0.1, b)) for a, b in zip(dims[:-1], dims[1:])]


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    layers = _net(rng, [5, 7, 6, 4])
    x = rng.normal(size=(9, 5))
    y = rng.integers(0, 4, 9)
    loss, grads = mlp.loss_and_gradients(layers, x, y, bf=False)
    h = 1e-6
    for li, (W, b) in enumerate(layers):
        for arr, g in ((W, grads[li][0]), (b, grads[li][1])):
            for idx in [tuple(rng.integers(0, s) for s in arr.shape) for _ in range(6)]:
                old = arr[idx]
                arr[idx] = old + h
                up, _ = mlp.loss_and_gradients(layers, x, y, bf=False)
                arr[idx] = old - h
                dn, _ = mlp.loss_and_gradients(layers, x, y, bf=False)
                arr[idx] = old
                assert abs((up - dn) / (2 * h) - g[idx]) < 1e-7 * max(1.0, abs(g[idx]))


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159, 65504.0, 1e-3], dtype=np.float32)
    got = mlp.bf16(vals)
    # 8 significand bits: 1 + k/128 grid near 1; ties to even
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0078125 and got[3] == 1.0078125
    assert abs(got[4] + 3.140625) < 1e-12
    for v, g in zip(vals, got):
        assert abs(g - v) <= abs(v) * 2 ** -8
