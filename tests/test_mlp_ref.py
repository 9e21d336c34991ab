"""CPU pins of the MLP checker (tests/mlp_ref.py) before it judges the device:

* BF16 rounding against hand-worked bit patterns (nearest, ties to even);
* forward_backward's gradients against central finite differences of the
  plain FP64 loss, on a network whose weights and inputs are already
  BF16-representable (so the checker's rounding is the identity there);
* the BF16-emulating mode differs from the exact mode only by rounding-sized
  amounts, and equals it when every intermediate is BF16-representable.
"""
import numpy as np

import mlp_ref


def test_bf16_rounding_bit_patterns():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # BF16 spacing at 1.0
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5, 0.0],
                    dtype=np.float32)
    got = mlp_ref.bf16(vals)
    want = np.array([one, one, one + 2 * ulp, one + ulp, -2.5, 0.0], dtype=np.float32)
    assert np.array_equal(got, want)


def _net(rng, dims):
    return [(mlp_ref.bf16(rng.standard_normal((a, b)) / np.sqrt(a)).astype(np.float64),
             mlp_ref.bf16(rng.standard_normal(b) * 0.1).astype(np.float64)) for a, b in zip(dims, dims[1:])]


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(7)
    layers = _net(rng, [8, 16, 16, 4])
    x = mlp_ref.bf16(rng.standard_normal((5, 8))).astype(np.float64)
    t = rng.standard_normal((5, 4))
    loss, grads = mlp_ref.forward_backward(layers, x, t)
    assert abs(loss - mlp_ref.loss_exact(layers, x, t)) <= 1e-14 * loss
    h = 1e-6
    for k, (W, b) in enumerate(layers):
        for arr, g in ((W, grads[k][0]), (b, grads[k][1])):
            for idx in list(np.ndindex(arr.shape))[:: max(1, arr.size // 12)]:
                keep = arr[idx]
                arr[idx] = keep + h
                up = mlp_ref.loss_exact(layers, x, t)
                arr[idx] = keep - h
                dn = mlp_ref.loss_exact(layers, x, t)
                arr[idx] = keep
                fd = (up - dn) / (2 * h)
                assert abs(fd - g[idx]) <= 1e-6 * (1.0 + abs(fd)), (k, idx, fd, g[idx])


def test_emulated_rounding_is_rounding_sized():
    rng = np.random.default_rng(8)
    layers = _net(rng, [32, 64, 64, 32])
    x = rng.standard_normal((16, 32))
    t = rng.standard_normal((16, 32))
    l0, g0 = mlp_ref.forward_backward(layers, x, t)
    l1, g1 = mlp_ref.forward_backward(layers, x, t, emulate_bf16=True)
    assert abs(l0 - l1) <= 2e-2 * l0
    for (w0, b0), (w1, b1) in zip(g0, g1):
        rel = np.linalg.norm(w0 - w1) / np.linalg.norm(w0)
        assert 0.0 < rel <= 3e-2
    # A zero-output network with a zero target has exactly-zero gradients in
    # both modes: nothing for the rounding to change.
    zl = [(np.zeros_like(W), np.zeros_like(b)) for W, b in layers]
    _, gz = mlp_ref.forward_backward(zl, x, np.zeros_like(t), emulate_bf16=True)
    assert all(not np.any(w) and not np.any(b) for w, b in gz)
