"""Pins of the NumPy decoder oracle: float64 gradients against central finite differences,
im2col_up / col2im_up adjointness, and the bfloat16 rounding rule."""
import numpy as np

from oracle import nn_oracle as no


def test_bf16_rounding_rule():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40])
    y = no.bf16(x)
    # ties to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 → 1; 1 + 3·2^-8 → 1 + 2^-6
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == 1.0 + 2 ** -7 and y[3] == 1.0 + 2 ** -6
    assert y[4] == -2.5
    r = np.random.default_rng(0).normal(size=1000)
    assert np.max(np.abs(no.bf16(r) - r) / np.abs(r)) <= 2 ** -8


def test_im2col_col2im_adjoint():
    r = np.random.default_rng(1)
    h = r.normal(size=(2, 3, 3, 4))
    d = r.normal(size=(2 * 36, 36))
    lhs = (no.im2col_up(h) * d).sum()
    rhs = (h * no.col2im_up(d, 2, 3, 4)).sum()
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_im2col_up_is_nearest_upsample_then_conv_patch():
    h = np.arange(2 * 2 * 2 * 1, dtype=float).reshape(2, 2, 2, 1)
    cols = no.im2col_up(h).reshape(2, 4, 4, 9)
    # centre tap (ky = kx = 1) of output pixel (y, x) reads input pixel (y // 2, x // 2)
    for y in range(4):
        for x in range(4):
            assert cols[1, y, x, 4] == h[1, y // 2, x // 2, 0]
    assert cols[0, 0, 0, 0] == 0.0  # zero padding


def test_gradients_match_finite_differences():
    arch = no.Arch(in_dim=2, hidden=5, latent_hw=2, latent_c=3, c1=3, c2=2, out_c=2)
    r = np.random.default_rng(2)
    flat = r.normal(scale=0.5, size=arch.nparam)
    x = r.normal(size=(2, 2))
    t = r.normal(size=(2, 16, 16, 2))
    loss, g, _ = no.loss_grad(arch, flat, x, t, bf=False)
    idx = r.choice(arch.nparam, size=60, replace=False)
    eps = 1e-6
    for i in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += eps
        fm[i] -= eps
        lp = no.loss_grad(arch, fp, x, t, bf=False)[0]
        lm = no.loss_grad(arch, fm, x, t, bf=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
