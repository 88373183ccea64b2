"""FP64 numpy restatement of the orientation-proposal network (test infrastructure).

Only tests, the graft entry's smoke check and the bench's CPU baseline leg may
import this module; the library never does.  The network has no counterpart
in Frealign: it is pinned by central finite differences of its own loss
(tests/test_cnn_oracle.py), and the GPU kernels are checked against it.

forward_bf16_act (activations rounded to BF16, the reference for
fr_cnn_forward_bf16) is pinned by tests/test_cnn_oracle.py: it equals forward
where every activation is exact in BF16 and differs from it by rounding only.

Shapes follow include/frealign_hip.h: nconv 3x3 / stride 2 / pad 1
convolutions with ReLU (1 -> C -> ... -> C channels), then one fully
connected layer to K logits.  Flat parameters: per convolution W[Co][Ci][3][3]
then b[Co]; then W_fc[K][F] and b_fc[K], F = C * (S >> nconv)^2.  The FC
input flattens each image's last activation as [channel][y][x].
"""
from __future__ import annotations

import numpy as np


def layout(S: int, C: int, L: int, K: int):
    offs = []
    o = 0
    for l in range(L):
        ci = 1 if l == 0 else C
        w = (o, o + C * ci * 9)
        o = w[1]
        b = (o, o + C)
        o = b[1]
        offs.append((w, b, ci))
    F = C * (S >> L) ** 2
    fw = (o, o + K * F)
    o = fw[1]
    fb = (o, o + K)
    o = fb[1]
    return offs, fw, fb, F, o


def im2col(x: np.ndarray) -> np.ndarray:
    """x (B, Ci, H, H) -> cols (B, Ci*9, Ho*Ho) for 3x3 / stride 2 / pad 1."""
    B, Ci, H, _ = x.shape
    Ho = H // 2
    xp = np.zeros((B, Ci, H + 2, H + 2), x.dtype)
    xp[:, :, 1:H + 1, 1:H + 1] = x
    cols = np.empty((B, Ci, 9, Ho, Ho), x.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, ky * 3 + kx] = xp[:, :, ky:ky + 2 * Ho:2, kx:kx + 2 * Ho:2]
    return cols.reshape(B, Ci * 9, Ho * Ho)


def col2im(cols: np.ndarray, Ci: int, H: int) -> np.ndarray:
    """Adjoint of im2col: (B, Ci*9, Ho*Ho) -> (B, Ci, H, H)."""
    B = cols.shape[0]
    Ho = H // 2
    c = cols.reshape(B, Ci, 9, Ho, Ho)
    xp = np.zeros((B, Ci, H + 2, H + 2), cols.dtype)
    for ky in range(3):
        for kx in range(3):
            xp[:, :, ky:ky + 2 * Ho:2, kx:kx + 2 * Ho:2] += c[:, :, ky * 3 + kx]
    return xp[:, :, 1:H + 1, 1:H + 1]


def _ops(bf16_ops: bool):
    """The operand rounding of a GEMM: BF16 (via FP32) for the mixed-precision
    step, none otherwise."""
    return (lambda a: bf16_round(a).astype(np.float64)) if bf16_ops else (lambda a: a)


def forward(p: np.ndarray, shape, img: np.ndarray, bf16_ops: bool = False):
    """Logits (B, K) and the cache for backward.  bf16_ops: every GEMM's
    operands rounded to BF16 (the activations where a product reads them),
    as fr_cnn_gradients_bf16 stages them; the arithmetic stays FP64."""
    S, C, L, K = shape
    r = _ops(bf16_ops)
    offs, fw, fb, F, _ = layout(S, C, L, K)
    p = np.asarray(p, np.float64)
    x = np.asarray(img, np.float64).reshape(-1, 1, S, S)
    B = x.shape[0]
    cache = []
    for (w, b, ci) in offs:
        W = p[w[0]:w[1]].reshape(C, ci * 9)
        cols = im2col(r(x))
        z = np.einsum("ok,bkn->bon", r(W), cols) + p[b[0]:b[1]][None, :, None]
        a = np.maximum(z, 0.0)
        Ho = x.shape[2] // 2
        cache.append((x, cols, a))
        x = a.reshape(B, C, Ho, Ho)
    X = x.reshape(B, F)
    Wf = p[fw[0]:fw[1]].reshape(K, F)
    logits = r(X) @ r(Wf).T + p[fb[0]:fb[1]][None, :]
    return logits, (cache, X)


def forward_bf16_act(p: np.ndarray, shape, img: np.ndarray):
    """Logits (B, K) with every convolution's output (after bias and ReLU)
    rounded to BF16, as the BF16-activation inference stores them; the
    arithmetic itself in FP64."""
    S, C, L, K = shape
    offs, fw, fb, F, _ = layout(S, C, L, K)
    p = np.asarray(p, np.float64)
    x = np.asarray(img, np.float64).reshape(-1, 1, S, S)
    B = x.shape[0]
    for (w, b, ci) in offs:
        W = p[w[0]:w[1]].reshape(C, ci * 9)
        z = np.einsum("ok,bkn->bon", W, im2col(x)) + p[b[0]:b[1]][None, :, None]
        a = bf16_round(np.maximum(z, 0.0))
        Ho = x.shape[2] // 2
        x = a.reshape(B, C, Ho, Ho)
    Wf = p[fw[0]:fw[1]].reshape(K, F)
    return x.reshape(B, F) @ Wf.T + p[fb[0]:fb[1]][None, :]


def loss_grad(p: np.ndarray, shape, img: np.ndarray, label: np.ndarray, bf16_ops: bool = False):
    """Per-image softmax cross-entropy and the gradient of their mean;
    bf16_ops as in forward (the backward products too: the loss gradient,
    the layer gradients and the weights where a product reads them; the bias
    sums and the ReLU masks stay exact)."""
    S, C, L, K = shape
    r = _ops(bf16_ops)
    offs, fw, fb, F, n = layout(S, C, L, K)
    p = np.asarray(p, np.float64)
    logits, (cache, X) = forward(p, shape, img, bf16_ops)
    B = logits.shape[0]
    z = logits - logits.max(axis=1, keepdims=True)
    e = np.exp(z)
    s = e.sum(axis=1, keepdims=True)
    prob = e / s
    lab = np.asarray(label, np.int64)
    loss = np.log(s[:, 0]) - z[np.arange(B), lab]
    gL = prob.copy()
    gL[np.arange(B), lab] -= 1.0
    gL /= B
    g = np.zeros(n, np.float64)
    Wf = p[fw[0]:fw[1]].reshape(K, F)
    g[fw[0]:fw[1]] = (r(gL).T @ r(X)).reshape(-1)
    g[fb[0]:fb[1]] = gL.sum(axis=0)
    gX = r(gL) @ r(Wf)
    ga = gX.reshape(cache[-1][2].shape)  # (B, C, P)
    for l in range(L - 1, -1, -1):
        (w, b, ci) = offs[l]
        x, cols, a = cache[l]
        gz = ga * (a > 0)
        W = p[w[0]:w[1]].reshape(C, ci * 9)
        g[w[0]:w[1]] = np.einsum("bon,bkn->ok", r(gz), cols).reshape(-1)
        g[b[0]:b[1]] = gz.sum(axis=(0, 2))
        if l > 0:
            gcols = np.einsum("ok,bon->bkn", r(W), r(gz))
            gx = col2im(gcols, ci, x.shape[2])
            ga = gx.reshape(x.shape[0], ci, -1)
    return loss, g


def adam(master, grad, m, v, step: int, lr: float, is_weight: np.ndarray):
    """One bias-corrected Adam step (beta 0.9/0.999, eps 1e-8); returns new
    (master, wq, m, v) with wq the BF16-rounded weights and exact biases."""
    b1, b2, eps = 0.9, 0.999, 1e-8
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    c1 = 1 - b1 ** step
    c2 = 1 - b2 ** step
    master = master - lr * (m / c1) / (np.sqrt(v / c2) + eps)
    wq = np.where(is_weight, bf16_round(master), master)
    return master, wq, m, v


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round FP32 values to the nearest BF16 (ties to even), kept as FP32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def weight_mask(shape) -> np.ndarray:
    S, C, L, K = shape
    offs, fw, fb, F, n = layout(S, C, L, K)
    m = np.zeros(n, bool)
    for (w, b, ci) in offs:
        m[w[0]:w[1]] = True
    m[fw[0]:fw[1]] = True
    return m
