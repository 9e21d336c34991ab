"""FP64 CPU restatement of the MLP forward/backward (test infrastructure):
the checker for the device MLP's gradients.

Inputs and weights are always rounded to BF16 the way the device rounds
them.  With ``emulate_bf16=True`` the reference also rounds what the device
stores in BF16 between layers -- the ReLU activations that feed the next
product and the pre-activation gradients dZ that feed dW and dA -- so the
remaining differences come from FP32 accumulation alone.  With the default
``False`` the intermediates stay exact FP64 (the mathematically intended
network)."""
import numpy as np


def bf16(a):
    """Round to BF16 (nearest even) and back to float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def forward_backward(layers, x, t, emulate_bf16=False):
    """layers: [(W, b)], ReLU on hidden layers, loss = mean((y - t)^2).
    Returns loss and [(dW, db)] in float64."""
    rnd = (lambda a: bf16(a).astype(np.float64)) if emulate_bf16 else (lambda a: a)
    acts, zs = [bf16(x).astype(np.float64)], []
    for k, (W, b) in enumerate(layers):
        z = acts[-1] @ bf16(W).astype(np.float64) + b.astype(np.float64)
        zs.append(z)
        if k < len(layers) - 1:
            acts.append(rnd(np.maximum(z, 0.0)))
    y = zs[-1]
    n, o = y.shape
    loss = np.mean((y - t) ** 2)
    d = 2.0 * (y - t) / (n * o)
    grads = [None] * len(layers)
    for k in range(len(layers) - 1, -1, -1):
        W, _ = layers[k]
        grads[k] = (acts[k].T @ rnd(d), d.sum(axis=0))
        if k:
            d = (rnd(d) @ bf16(W).astype(np.float64).T) * (zs[k - 1] > 0)
    return loss, grads


def loss_exact(layers, x, t):
    """Loss of the network in plain FP64 (no BF16 anywhere): the function
    whose finite differences pin forward_backward's gradient formulas."""
    a = x
    for k, (W, b) in enumerate(layers):
        a = a @ W + b
        if k < len(layers) - 1:
            a = np.maximum(a, 0.0)
    return np.mean((a - t) ** 2)
