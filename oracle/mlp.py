"""FP64 CPU restatement of the chemistry-surrogate MLP (TEST INFRASTRUCTURE ONLY).

The oracle of the HIP network: forward with GELU (tanh form) and hand-written backward in
float64 numpy, reading the same parameter vector layout as eu3d_mlp_get_params (per layer
W^T [dout x din] then bias [dout]; padded inputs 32, hidden 256, padded outputs 16). The
surrogate has no counterpart in the reference, so its pins are this restatement, finite
differences of it, and the BF16 rounding model below.
"""
import numpy as np

DIN = [32, 256, 256, 256, 256]
DOUT = [256, 256, 256, 256, 16]
NOUT_REAL = 9


def bf16_round(x):
    """Round float64/float32 values to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    Ws, bs, off = [], [], 0
    for din, dout in zip(DIN, DOUT):
        Wt = params[off:off + din * dout].reshape(dout, din).astype(np.float64)
        off += din * dout
        b = params[off:off + dout].astype(np.float64)
        off += dout
        Ws.append(Wt)
        bs.append(b)
    return Ws, bs


C_GELU = np.sqrt(2.0 / np.pi)


def gelu(z):
    """GELU, tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))) (the kernels evaluate the
    equal z * sigmoid(2 sqrt(2/pi) (z + 0.044715 z^3)))."""
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def dgelu(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3 * 0.044715 * z * z)


def forward(params, X, bf16=False):
    """X [n x 32] -> Y [n x 16]. bf16=True rounds weights and every activation to BF16 as the
    GPU stores them (accumulation stays in FP64 here)."""
    Ws, bs = unpack(params)
    r = bf16_round if bf16 else (lambda v: v)
    A = r(X)
    cache = [(A, None)]
    for l in range(4):
        Z = A @ r(Ws[l]).T + bs[l]
        if bf16:
            Z = bf16_round(Z)
        A = r(gelu(Z))
        cache.append((A, Z))
    Y = A @ r(Ws[4]).T + bs[4]
    return Y, cache


def loss_and_grads(params, X, T):
    """MSE over the 9 real outputs and the gradient of every parameter (FP64)."""
    Ws, bs = unpack(params)
    Y, cache = forward(params, X)
    n = X.shape[0]
    d = np.zeros_like(Y)
    d[:, :NOUT_REAL] = Y[:, :NOUT_REAL] - T[:, :NOUT_REAL]
    loss = float((d ** 2).sum() / (n * NOUT_REAL))
    dZ = 2.0 * d / (n * NOUT_REAL)
    gW, gb = [None] * 5, [None] * 5
    for l in range(4, -1, -1):
        A_prev = cache[l][0]
        gW[l] = dZ.T @ A_prev
        gb[l] = dZ.sum(0)
        if l > 0:
            dA = dZ @ Ws[l]
            dZ = dA * dgelu(cache[l][1])
    g = np.concatenate([np.concatenate([gW[l].ravel(), gb[l]]) for l in range(5)])
    return loss, g
