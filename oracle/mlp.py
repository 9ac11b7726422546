"""ORACLE -- test infrastructure only. NumPy restatement of the learned
shortlist's MLP (ReLU hidden layers, softmax cross-entropy, mean over the
batch) with its gradients by backpropagation.

bf16=True rounds to bfloat16 (nearest even) exactly where the GPU path stores
bf16: the inputs, the weights used in products, the hidden activations, the
logit gradient and the hidden gradients; products accumulate in FP64. With
bf16=False everything is FP64, and finite differences pin the gradients.
"""
import numpy as np


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(layers, x, bf=True):
    """layers: [(W [rows, cols], b [cols])...]; returns (activations, logits)."""
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    h = r(x)
    acts = [h]
    for i, (W, b) in enumerate(layers):
        z = h @ r(W) + np.asarray(b, np.float64)
        if i < len(layers) - 1:
            h = r(np.maximum(z, 0.0).astype(np.float32) if bf else np.maximum(z, 0.0))
            acts.append(h)
        else:
            return acts, z
    raise ValueError("no layers")


def loss_and_gradients(layers, x, y, bf=True):
    """Mean softmax cross-entropy and its gradients [(dW, db)...]."""
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    acts, z = forward(layers, x, bf)
    B = z.shape[0]
    z = z - z.max(axis=1, keepdims=True)
    p = np.exp(z)
    p /= p.sum(axis=1, keepdims=True)
    loss = float(np.mean(-np.log(p[np.arange(B), y])))
    d = p.copy()
    d[np.arange(B), y] -= 1.0
    d = r((d / B).astype(np.float32) if bf else d / B)
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        h = acts[i]
        grads[i] = (h.T @ d, d.sum(axis=0))
        if i == 0:
            break
        g = d @ r(layers[i][0]).T
        g = np.where(acts[i] > 0, g, 0.0)
        d = r(g.astype(np.float32)) if bf else g
    return loss, grads
