"""CPU oracle for the learned boundary's training kernels (test infrastructure only).

float64 numpy restatement of a zero-padded 1D convolution along a strip with one weight set
per boundary side, its two backward products and the weighted MSE loss:

    y[b, p, co] = sum_k sum_ci w[side[b], co, k, ci] * x[b, p + k - K//2, ci]

Pinned in tests/test_learned.py by the adjoint identities
<conv(x, w), dy> = <x, bwd_data(dy, w)> = <w, bwd_weight(dy, x)>, by central finite
differences of the loss, and by a dense-matrix brute force of the convolution.  adam_step is
pinned against torch.optim.Adam over three steps and against its first-step closed form
w - lr*g/(abs(g) + eps); to_bf16_bits against torch's float32 to bfloat16 cast on values where
round-to-nearest-even and truncation differ.  Planted mistakes (no second-moment bias
correction, a truncating conversion) fail the same comparisons.

Also: bf16 bit conversion (round to nearest even) for feeding exact bf16 values to the device.
"""
import numpy as np

__all__ = ['conv_fwd', 'conv_bwd_data', 'conv_bwd_weight', 'mse', 'adam_step', 'to_bf16_bits',
           'from_bf16_bits', 'round_bf16']


def conv_fwd(x, w, side):
    B, P, Ci = x.shape
    S, Co, K, _ = w.shape
    R = K // 2
    y = np.zeros((B, P, Co))
    xp = np.zeros((B, P + 2 * R, Ci))
    xp[:, R:R + P] = x
    for b in range(B):
        ws = w[side[b]]
        for k in range(K):
            y[b] += xp[b, k:k + P] @ ws[:, k, :].T
    return y


def conv_bwd_data(dy, w, side):
    B, P, Co = dy.shape
    S, _, K, Ci = w.shape
    R = K // 2
    dx = np.zeros((B, P, Ci))
    for b in range(B):
        ws = w[side[b]]
        for k in range(K):
            # x position q receives dy at p = q - k + R
            lo, hi = max(0, k - R), min(P, P + k - R)
            dx[b, lo:hi] += dy[b, lo - k + R:hi - k + R] @ ws[:, k, :]
    return dx


def conv_bwd_weight(dy, x, side, S, K):
    B, P, Co = dy.shape
    Ci = x.shape[2]
    R = K // 2
    dw = np.zeros((S, Co, K, Ci))
    xp = np.zeros((B, P + 2 * R, Ci))
    xp[:, R:R + P] = x
    for b in range(B):
        for k in range(K):
            dw[side[b], :, k, :] += dy[b].T @ xp[b, k:k + P]
    return dw


def mse(y, t, lo, hi, wch, norm):
    """Loss and gradient of norm * sum_{b, p in [lo_b, hi_b), c} wch[c] (y - t)^2."""
    B, P, C = y.shape
    mask = np.zeros((B, P, 1))
    for b in range(B):
        mask[b, lo[b]:hi[b]] = 1.0
    e = (y - t) * mask
    loss = norm * np.sum(wch[None, None, :] * e * e)
    grad = 2.0 * wch[None, None, :] * e * norm
    return loss, grad


def adam_step(w, m1, m2, g, lr, b1, b2, eps, step):
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    w = w - lr * (m1 / (1 - b1 ** step)) / (np.sqrt(m2 / (1 - b2 ** step)) + eps)
    return w, m1, m2


def to_bf16_bits(a):
    """float32 -> bf16 bit patterns (uint16), round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = u + 0x7FFF + ((u >> 16) & 1)
    return (u >> 16).astype(np.uint16)


def from_bf16_bits(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def round_bf16(a):
    return from_bf16_bits(to_bf16_bits(a)).astype(np.float64)
