"""GPU tests of the learned warm start's MFMA network (csrc/mlp.hip).

* the BF16 MFMA GEMM against an exact product (integer-valued operands, so
  BF16 rounding and FP32 accumulation are exact) for all four transposes --
  this pins the MFMA lane layout;
* the device backward pass against an FP64 CPU restatement (tests/mlp_ref.py,
  itself pinned by finite differences in test_mlp_ref.py) with the same
  BF16-rounded weights and inputs: against the restatement that also rounds
  the BF16-stored intermediates, loss within 1e-3 and every gradient tensor
  within 5e-3 relative L2 (FP32 accumulation is the only difference left);
  against the exact-intermediate restatement, within 2e-2 / 8e-2 (the size
  of BF16 storage error through three layers);
* training on the device lowers the loss of a fixed teacher mapping 5x.
"""
import numpy as np
import pytest

import mcpdlp as mp
import mlp_ref

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_bf16_exact(ta, tb):
    rng = np.random.default_rng(3)
    M, N, K = 64, 96, 48
    A = rng.integers(-8, 9, (M, K)).astype(np.float32)
    B = rng.integers(-8, 9, (K, N)).astype(np.float32)
    got = mp.gemm_bf16(A.T.copy() if ta else A, B.T.copy() if tb else B, transpose_a=ta, transpose_b=tb)
    assert np.array_equal(got, A.astype(np.float64) @ B.astype(np.float64))


def test_mlp_gradients_match_fp64_reference():
    net = mp.MLP(256, 256, 3, 64, capacity=64, seed=11)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((64, 256)).astype(np.float32)
    t = rng.standard_normal((64, 64)).astype(np.float32)
    loss, g = net.loss_grad(x, t)
    layers = net.unpack(net.params())
    x64, t64 = x.astype(np.float64), t.astype(np.float64)
    for emulate, tol_loss, tol_grad in ((True, 1e-3, 5e-3), (False, 2e-2, 8e-2)):
        ref_loss, ref = mlp_ref.forward_backward(layers, x64, t64, emulate_bf16=emulate)
        assert abs(loss - ref_loss) <= tol_loss * ref_loss, (emulate, loss, ref_loss)
        for k, ((dW, db), (rW, rb)) in enumerate(zip(net.unpack(g), ref)):
            ew = np.linalg.norm(dW - rW) / np.linalg.norm(rW)
            eb = np.linalg.norm(db - rb) / (np.linalg.norm(rb) + 1e-12)
            print(f"layer {k} emulate={emulate}: rel err dW {ew:.2e} db {eb:.2e}")
            assert ew <= tol_grad and eb <= tol_grad, (emulate, k, ew, eb)
    net.close()


def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(5)
    R = rng.standard_normal((256, 64)) / 16.0
    x = rng.standard_normal((512, 256)).astype(np.float32)
    t = np.tanh(x @ R).astype(np.float32)
    net = mp.MLP(256, 256, 3, 64, capacity=128, seed=12)
    before = net.loss_grad(x[:128], t[:128])[0]
    loss, ms = net.train(x, t, batch=128, epochs=60, lr=2e-3)
    after = float(np.mean((net.forward(x[:128]) - t[:128]) ** 2))
    assert after < before / 5, (before, after)
    assert ms > 0
    net.close()
