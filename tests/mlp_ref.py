"""CPU oracle for the MLP kernels (test infrastructure).

FP64 numpy forward/backward of a dense SiLU stack.  With quantize=True it
rounds to BF16 at exactly the points where the HIP kernels hold BF16
operands (inputs, weights, hidden activations, and dZ before the backward
GEMMs), so GPU and oracle differ only by FP32 vs FP64 accumulation.
"""
import numpy as np


def bf16(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(X, Ws, bs, quantize=True):
    q = bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))
    x = q(X)
    acts, zs = [x], []
    for i, (W, b) in enumerate(zip(Ws, bs)):
        z = x @ q(W) + b
        zs.append(z)
        if i + 1 < len(Ws):
            x = q(silu(z))
            acts.append(x)
        else:
            out = z
    return out, acts, zs


def backward(dout, Ws, acts, zs, quantize=True):
    q = bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))
    gWs, gbs = [None] * len(Ws), [None] * len(Ws)
    dY = dout
    for i in range(len(Ws) - 1, -1, -1):
        dz = dY * (silu_grad(zs[i]) if i + 1 < len(Ws) else 1.0)
        dzq = q(dz)
        gWs[i] = acts[i].T @ dzq
        gbs[i] = dz.sum(axis=0)
        dY = dzq @ q(Ws[i]).T
    return gWs, gbs


def h0_loss(out, tgt, mask):
    """Squared error on column 0, BCE on logit column 1; mean over rows."""
    M = out.shape[0]
    e = out[:, 0] - tgt[:, 0]
    p = 1.0 / (1.0 + np.exp(-out[:, 1]))
    t = tgt[:, 1]
    loss = np.sum(e * e * mask[:, 0]) - np.sum(
        (t * np.log(np.maximum(p, 1e-7)) + (1 - t) * np.log(np.maximum(1 - p, 1e-7))) * mask[:, 1])
    d = np.zeros_like(out)
    d[:, 0] = 2 * e * mask[:, 0] / M
    d[:, 1] = (p - t) * mask[:, 1] / M
    return loss, d


def mse_heads(out, tgt, heads, width, mask=None):
    """Squared error of `heads` column blocks of out against tgt [M x live],
    weighted per row and head by mask [M x heads]; returns (sum, dout) with
    dout = d(sum / (M*live))/d(out)."""
    M, live = tgt.shape
    d = np.zeros_like(out)
    total = 0.0
    for h in range(heads):
        w = np.ones(M) if mask is None else mask[:, h]
        e = out[:, h * width:h * width + live] - tgt
        total += np.sum(e * e * w[:, None])
        d[:, h * width:h * width + live] = 2 * e * w[:, None] / (M * live)
    return total, d
