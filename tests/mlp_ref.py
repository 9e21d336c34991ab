"""FP64 CPU restatement of the MLP forward/backward (test infrastructure):
the checker for the device MLP's gradients.  Inputs and weights are rounded
to BF16 the way the device rounds them; intermediate activations are not."""
import numpy as np


def bf16(a):
    """Round to BF16 (nearest even) and back to float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def forward_backward(layers, x, t):
    """layers: [(W, b)], ReLU on hidden layers, loss = mean((y - t)^2).
    Returns loss and [(dW, db)] in float64."""
    acts, zs = [bf16(x).astype(np.float64)], []
    for k, (W, b) in enumerate(layers):
        z = acts[-1] @ bf16(W).astype(np.float64) + b.astype(np.float64)
        zs.append(z)
        if k < len(layers) - 1:
            acts.append(np.maximum(z, 0.0))
    y = zs[-1]
    n, o = y.shape
    loss = np.mean((y - t) ** 2)
    d = 2.0 * (y - t) / (n * o)
    grads = [None] * len(layers)
    for k in range(len(layers) - 1, -1, -1):
        W, _ = layers[k]
        grads[k] = (acts[k].T @ d, d.sum(axis=0))
        if k:
            d = (d @ bf16(W).astype(np.float64).T) * (zs[k - 1] > 0)
    return loss, grads
