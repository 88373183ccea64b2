"""Pins the FP64 network oracle (oracle/cnn_oracle.py) by mathematics:
central finite differences of its loss, the adjointness of its im2col/col2im
pair, a closed-form loss, and BF16 rounding on hand-checked values."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import cnn_oracle as CO  # noqa: E402


def _case(seed=3, shape=(8, 3, 2, 5), B=3):
    rng = np.random.default_rng(seed)
    n = CO.layout(*shape)[-1]
    p = rng.standard_normal(n) * 0.5
    img = rng.standard_normal((B, shape[0], shape[0]))
    lab = rng.integers(0, shape[3], B)
    return p, img, lab


def test_finite_differences():
    shape = (8, 3, 2, 5)
    p, img, lab = _case(shape=shape)
    loss, g = CO.loss_grad(p, shape, img, lab)
    rng = np.random.default_rng(11)
    idx = rng.choice(len(p), 60, replace=False)
    h = 1e-6
    worst = 0.0
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (CO.loss_grad(pp, shape, img, lab)[0].mean() -
              CO.loss_grad(pm, shape, img, lab)[0].mean()) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / max(1e-6, abs(g[i])))
    assert worst < 1e-5, worst


def test_planted_mistake_is_caught():
    # a gradient with a dropped ReLU mask would fail the same check
    shape = (8, 3, 2, 5)
    p, img, lab = _case(shape=shape, seed=5)
    _, g = CO.loss_grad(p, shape, img, lab)
    offs = CO.layout(*shape)[0]
    w0 = offs[0][0]
    i = w0[0] + 4
    h = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[i] += h
    pm[i] -= h
    fd = (CO.loss_grad(pp, shape, img, lab)[0].mean() -
          CO.loss_grad(pm, shape, img, lab)[0].mean()) / (2 * h)
    assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(g[i]))
    assert abs(fd - 1.5 * g[i]) > 1e-4 * abs(g[i]) or abs(g[i]) < 1e-9


def test_im2col_adjoint():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 3, 8, 8))
    y = rng.standard_normal((2, 27, 16))
    lhs = np.sum(CO.im2col(x) * y)
    rhs = np.sum(x * CO.col2im(y, 3, 8))
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


def test_im2col_matches_direct_convolution():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((1, 2, 6, 6))
    w = rng.standard_normal((3, 2, 3, 3))
    cols = CO.im2col(x)
    out = np.einsum("ok,bkn->bon", w.reshape(3, 18), cols).reshape(1, 3, 3, 3)
    ref = np.zeros((1, 3, 3, 3))
    for co in range(3):
        for oy in range(3):
            for ox in range(3):
                s = 0.0
                for ci in range(2):
                    for ky in range(3):
                        for kx in range(3):
                            yy, xx = 2 * oy + ky - 1, 2 * ox + kx - 1
                            if 0 <= yy < 6 and 0 <= xx < 6:
                                s += w[co, ci, ky, kx] * x[0, ci, yy, xx]
                ref[0, co, oy, ox] = s
    assert np.allclose(out, ref, atol=1e-12)


def test_zero_network_loss_is_log_k():
    shape = (8, 2, 2, 7)
    n = CO.layout(*shape)[-1]
    img = np.random.default_rng(0).standard_normal((4, 8, 8))
    lab = np.array([0, 3, 6, 2])
    loss, g = CO.loss_grad(np.zeros(n), shape, img, lab)
    assert np.allclose(loss, np.log(7.0), atol=1e-15)
    fb = CO.layout(*shape)[2]
    want = np.full(7, 4 / 7.0)
    for y in lab:
        want[y] -= 1.0
    assert np.allclose(g[fb[0]:fb[1]], want / 4, atol=1e-15)


def test_bf16_round():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 0.0], np.float32)
    assert np.array_equal(CO.bf16_round(v), want)


def test_forward_bf16_act_rounds_only_the_activations():
    """With integer weights, biases and images small enough that every
    activation is an integer below 256 (exact in BF16), the BF16-activation
    forward is the forward; on random values it differs from it by BF16
    rounding, not more."""
    shape = (16, 4, 2, 5)
    offs, fw, fb, F, n = CO.layout(*shape)
    rng = np.random.default_rng(3)
    p = np.zeros(n)
    for (w, b, ci) in offs:
        p[w[0]:w[1]] = rng.integers(-1, 2, w[1] - w[0])
        p[b[0]:b[1]] = rng.integers(0, 2, b[1] - b[0])
    p[fw[0]:fw[1]] = rng.integers(-2, 3, fw[1] - fw[0])
    img = rng.integers(0, 3, (3, 16, 16)).astype(np.float64)
    exact, _ = CO.forward(p, shape, img)
    act = CO.forward_bf16_act(p, shape, img)
    assert np.array_equal(exact, act)
    p2 = rng.standard_normal(n) * 0.3
    img2 = rng.standard_normal((3, 16, 16))
    a, _ = CO.forward(p2, shape, img2)
    b = CO.forward_bf16_act(p2, shape, img2)
    rel = np.linalg.norm(a - b) / np.linalg.norm(a)
    assert 1e-6 < rel < 2e-2, rel


def test_bf16_operand_step_rounds_only_the_products():
    """The mixed-precision step (loss_grad with bf16_ops): with integer
    operands exact in BF16 its forward is the exact one, and the gradients
    differ from the exact ones by BF16 rounding of the backward operands
    only (the FC bias gradient, a plain sum of the loss gradient, is
    unchanged); on random values the gradient differs by BF16 rounding (a
    segment by more only where a rounded operand moves a pre-activation
    across a ReLU's zero)."""
    shape = (16, 4, 2, 5)
    offs, fw, fb, F, n = CO.layout(*shape)
    rng = np.random.default_rng(5)
    p = np.zeros(n)
    for (w, b, ci) in offs:
        p[w[0]:w[1]] = rng.integers(-1, 2, w[1] - w[0])
        p[b[0]:b[1]] = rng.integers(0, 2, b[1] - b[0])
    p[fw[0]:fw[1]] = rng.integers(-2, 3, fw[1] - fw[0]) * 0.0625
    img = rng.integers(0, 3, (3, 16, 16)).astype(np.float64)
    lab = np.array([0, 3, 4])
    exact, _ = CO.forward(p, shape, img)
    mixed, _ = CO.forward(p, shape, img, bf16_ops=True)
    assert np.array_equal(exact, mixed)
    l0, g0 = CO.loss_grad(p, shape, img, lab)
    l1, g1 = CO.loss_grad(p, shape, img, lab, bf16_ops=True)
    assert np.array_equal(l0, l1)
    assert np.array_equal(g0[fb[0]:fb[1]], g1[fb[0]:fb[1]])
    assert 0 < np.linalg.norm(g1 - g0) <= 2e-2 * np.linalg.norm(g0)
    p2 = rng.standard_normal(n) * 0.3
    img2 = rng.standard_normal((3, 16, 16))
    _, a = CO.loss_grad(p2, shape, img2, lab)
    _, b = CO.loss_grad(p2, shape, img2, lab, bf16_ops=True)
    # three images of a 4-channel network: a single flipped ReLU moves a
    # segment by a few per cent, so the bar here is loose
    for (w, bb, ci) in offs + [(fw, fb, 0)]:
        rel = np.linalg.norm(a[w[0]:w[1]] - b[w[0]:w[1]]) / np.linalg.norm(a[w[0]:w[1]])
        assert 1e-6 < rel < 0.2, rel
