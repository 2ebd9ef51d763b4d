"""NumPy oracle of the warm-start conv decoder (forward, MSE loss, backward).

TEST INFRASTRUCTURE ONLY (tests/ use it as the checker of the GPU network).  Float64
throughout; with bf16=True the operands of every matrix product are rounded to bfloat16
(round to nearest even) at the same points as the GPU kernels, so the GPU differs from it
only by FP32 accumulation order.  Pinned by central finite differences in float64
(tests/test_oracle_nn.py).
"""
from __future__ import annotations

import numpy as np


def bf16(a):
    """Round float64/float32 values to bfloat16 (nearest even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = np.where(finite, (u + 0x7FFF + ((u >> 16) & 1)) >> 16, u >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(a):
    return np.asarray(a, dtype=np.float64)


class Arch:
    def __init__(self, in_dim=2, hidden=256, latent_hw=16, latent_c=64, c1=64, c2=32, out_c=3):
        self.in_dim, self.hidden, self.hw, self.c0 = in_dim, hidden, latent_hw, latent_c
        self.c1, self.c2, self.out_c = c1, c2, out_c
        self.L = latent_hw * latent_hw * latent_c
        shapes = [("W1", (in_dim, hidden)), ("b1", (hidden,)), ("W2", (hidden, self.L)), ("b2", (self.L,)),
                  ("C1", (9 * latent_c, c1)), ("bc1", (c1,)), ("C2", (9 * c1, c2)), ("bc2", (c2,)),
                  ("C3", (9 * c2, out_c)), ("bc3", (out_c,))]
        self.shapes = shapes
        self.nparam = sum(int(np.prod(s)) for _, s in shapes)

    def unpack(self, flat):
        out, o = {}, 0
        for name, s in self.shapes:
            n = int(np.prod(s))
            out[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(s)
            o += n
        return out

    def pack(self, d):
        return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in self.shapes])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def im2col_up(h):
    """[B, Hi, Hi, C] → nearest ×2 → 3×3 patches with zero padding → [B·Ho·Ho, 9C], k = (ky, kx, c)."""
    B, Hi, _, C = h.shape
    Ho = 2 * Hi
    up = h.repeat(2, axis=1).repeat(2, axis=2)
    pad = np.zeros((B, Ho + 2, Ho + 2, C))
    pad[:, 1:-1, 1:-1] = up
    cols = np.empty((B, Ho, Ho, 9, C))
    for t in range(9):
        ky, kx = divmod(t, 3)
        cols[:, :, :, t, :] = pad[:, ky:ky + Ho, kx:kx + Ho, :]
    return cols.reshape(B * Ho * Ho, 9 * C)


def col2im_up(dcols, B, Hi, C):
    """Adjoint of im2col_up: [B·Ho·Ho, 9C] → [B, Hi, Hi, C]."""
    Ho = 2 * Hi
    d = dcols.reshape(B, Ho, Ho, 9, C)
    dpad = np.zeros((B, Ho + 2, Ho + 2, C))
    for t in range(9):
        ky, kx = divmod(t, 3)
        dpad[:, ky:ky + Ho, kx:kx + Ho, :] += d[:, :, :, t, :]
    dup = dpad[:, 1:-1, 1:-1]
    return dup.reshape(B, Hi, 2, Hi, 2, C).sum(axis=(2, 4))


def forward(arch, flat, x, bf=True):
    r = bf16 if bf else _ident
    p = arch.unpack(flat)
    B = x.shape[0]
    c = {"xb": r(x)}
    c["z1"] = c["xb"] @ r(p["W1"]) + p["b1"]
    c["h1b"] = r(silu(c["z1"]))
    c["z2"] = c["h1b"] @ r(p["W2"]) + p["b2"]
    c["h2b"] = r(silu(c["z2"])).reshape(B, arch.hw, arch.hw, arch.c0)
    c["cols1"] = im2col_up(c["h2b"])
    c["z3"] = c["cols1"] @ r(p["C1"]) + p["bc1"]
    H1 = 2 * arch.hw
    c["h3b"] = r(silu(c["z3"])).reshape(B, H1, H1, arch.c1)
    c["cols2"] = im2col_up(c["h3b"])
    c["z4"] = c["cols2"] @ r(p["C2"]) + p["bc2"]
    H2 = 4 * arch.hw
    c["h4b"] = r(silu(c["z4"])).reshape(B, H2, H2, arch.c2)
    c["cols3"] = im2col_up(c["h4b"])
    y = c["cols3"] @ r(p["C3"]) + p["bc3"]
    H3 = 8 * arch.hw
    return y.reshape(B, H3, H3, arch.out_c), c


def loss_grad(arch, flat, x, t, bf=True):
    """MSE loss and its gradient w.r.t. the flat parameters (same rounding points as the GPU)."""
    r = bf16 if bf else _ident
    p = arch.unpack(flat)
    B = x.shape[0]
    y, c = forward(arch, flat, x, bf)
    n = y.size
    loss = float(((y - t) ** 2).mean())
    dy = (2.0 / n) * (y - t).reshape(-1, arch.out_c)
    g = {}
    stages = [("C1", "bc1", "cols1", arch.hw, arch.c0, arch.c1, c["z2"]),
              ("C2", "bc2", "cols2", 2 * arch.hw, arch.c1, arch.c2, c["z3"]),
              ("C3", "bc3", "cols3", 4 * arch.hw, arch.c2, arch.out_c, c["z4"])]
    dz = dy
    for name, bname, cols, Hi, Ci, Co, zin in reversed(stages):
        dzb = r(dz)
        g[bname] = dzb.sum(axis=0)  # the GPU sums the BF16 copy (a row of ones in the product)
        g[name] = c[cols].T @ dzb
        dcols = dzb @ r(p[name]).T
        dh = col2im_up(dcols, B, Hi, Ci).reshape(-1, Ci) if name != "C1" else col2im_up(dcols, B, Hi, Ci).reshape(B, -1)
        dz = dh * dsilu(zin.reshape(dh.shape))
    # dense 2 (dz = dZ2, [B, L])
    dzb = r(dz)
    g["b2"] = dzb.sum(axis=0)
    g["W2"] = c["h1b"].T @ dzb
    dh1 = dzb @ r(p["W2"]).T
    dz1 = dh1 * dsilu(c["z1"])
    g["b1"] = r(dz1).sum(axis=0)  # a row of ones in the BF16 product, as for b2
    g["W1"] = c["xb"].T @ r(dz1)
    return loss, arch.pack(g), y
