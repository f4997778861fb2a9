"""TEST INFRASTRUCTURE ONLY: numpy restatement of the per-orbital MLP
(quack_amd/csrc/qk_mlp.hip) used to check its forward pass and gradients.

Parameters are the flat FP32 buffer of the HIP code: per layer W (out x in,
row-major) then b (out); hidden layers use SiLU, the last layer is linear;
loss = mean((y - t)^2).  Computed in float64 so that it is the reference for
the FP32 device kernels.
"""
import numpy as np


def layers(nin, width, depth):
    sizes = []
    prev = nin
    for l in range(depth + 1):
        out = width if l < depth else 1
        sizes.append((prev, out))
        prev = out
    return sizes


def unpack(p, nin, width, depth):
    off = 0
    out = []
    for (i, o) in layers(nin, width, depth):
        W = p[off:off + o * i].reshape(o, i)
        off += o * i
        b = p[off:off + o]
        off += o
        out.append((W, b))
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def forward(p, X, nin, width, depth):
    a = X.astype(np.float64)
    cache = []
    L = unpack(np.asarray(p, np.float64), nin, width, depth)
    for l, (W, b) in enumerate(L):
        z = a @ W.T + b
        cache.append((a, z))
        a = silu(z) if l < len(L) - 1 else z
    return a[:, 0], cache, L


def loss_grad(p, X, T, nin, width, depth):
    y, cache, L = forward(p, X, nin, width, depth)
    B = X.shape[0]
    loss = np.mean((y - T) ** 2)
    dz = (2.0 * (y - T) / B)[:, None]
    grads = []
    for l in range(len(L) - 1, -1, -1):
        a_in, z = cache[l]
        W, b = L[l]
        grads.append((dz.T @ a_in, dz.sum(axis=0)))
        if l > 0:
            da = dz @ W
            zp = cache[l - 1][1]
            s = 1.0 / (1.0 + np.exp(-zp))
            dz = da * s * (1.0 + zp * (1.0 - s))
    flat = []
    for (dW, db) in reversed(grads):
        flat.append(dW.ravel())
        flat.append(db)
    return loss, np.concatenate(flat)
