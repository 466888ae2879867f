"""CPU restatement of the per-body MLP of the learned warm start (colb_amd/csrc/abd_mlp.hip).

TEST INFRASTRUCTURE ONLY: tests compare the GPU forward, backward and training kernels with
this file; the product never imports it.

Model: y = W_{L-1}^T relu(... relu(W_0^T x + b_0) ...) + b_{L-1} in the row convention
Z_l = A_l W_l + b_l, A_{l+1} = relu(Z_l).  With ``bf16=True`` the values the GPU stores in BF16
(inputs, weights read by the products, activations, loss gradients) are rounded to BF16 (nearest
even) and products accumulate in FP32; with ``bf16=False`` everything is FP64 (the exact model the
finite-difference test checks).
"""
import numpy as np


def to_bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unflatten(flat, dims):
    Ws, bs, off = [], [], 0
    for l in range(len(dims) - 1):
        n = dims[l] * dims[l + 1]
        Ws.append(np.asarray(flat[off:off + n]).reshape(dims[l], dims[l + 1]))
        off += n
        bs.append(np.asarray(flat[off:off + dims[l + 1]]))
        off += dims[l + 1]
    return Ws, bs


def flatten(Ws, bs):
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in zip(Ws, bs)])


def forward(flat, dims, X, bf16=True):
    Ws, bs = unflatten(flat, dims)
    if bf16:
        A = to_bf16(X)
        Ws = [to_bf16(W) for W in Ws]
        bs = [b.astype(np.float32) for b in bs]
    else:
        A = np.asarray(X, dtype=np.float64)
        Ws = [W.astype(np.float64) for W in Ws]
        bs = [b.astype(np.float64) for b in bs]
    As, Zs = [A], []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        Z = As[-1] @ W + b
        Zs.append(Z)
        if l + 1 < len(Ws):
            R = np.maximum(Z, 0)
            As.append(to_bf16(R) if bf16 else R)
    return As, Zs, Ws


def loss_and_grad(flat, dims, X, T, w, bf16=True):
    """Loss sum_ij w_j (Y_ij - T_ij)^2 / n and its gradient (flat layout)."""
    n = X.shape[0]
    As, Zs, Ws = forward(flat, dims, X, bf16)
    Y = Zs[-1]
    D = Y - T
    loss = float(np.sum(w * D * D) / n)
    G = w * 2.0 * D / n
    if bf16:
        G = to_bf16(G.astype(np.float32))
    gWs, gbs = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        gWs[l] = As[l].T @ G
        gbs[l] = G.sum(axis=0)
        if l > 0:
            dA = G @ Ws[l].T
            G = dA * (Zs[l - 1] > 0)
            if bf16:
                G = to_bf16(G.astype(np.float32))
    return loss, flatten(gWs, gbs)
