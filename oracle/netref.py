"""CPU restatement of the learned pressure-preconditioner network.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product
path).  The network has no counterpart in the reference; this module states
the same computation as pawsim_amd/csrc/pg_net.hip in numpy so that the GPU
forward pass, loss and parameter gradients can be checked against it, and its
own backward pass is pinned by finite differences (float64, no rounding).

Layout: NHWC arrays (B, H, W, C); parameters as one flat float32 vector,
per layer w[Cout][9][Cin] (tap = 3*(dy+1) + (dx+1)) followed by b[Cout].
With bf16=True every tensor the GPU stores in BF16 (weights, activations,
activation gradients) is rounded to BF16 (round to nearest even) at the same
points; sums are float32 (float64 when bf16=False).
"""
from __future__ import annotations

import numpy as np

CIN = (4, 16, 32, 64, 64, 32, 16)
COUT = (16, 32, 64, 64, 32, 16, 1)


def param_count() -> int:
    return sum(co * 9 * ci + co for ci, co in zip(CIN, COUT))


def split_params(p):
    out, off = [], 0
    for ci, co in zip(CIN, COUT):
        w = p[off:off + co * 9 * ci].reshape(co, 9, ci)
        off += co * 9 * ci
        b = p[off:off + co]
        off += co
        out.append((w, b))
    return out


def init_params(seed: int) -> np.ndarray:
    """He-style initialisation, float32, deterministic."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in zip(CIN, COUT):
        std = np.sqrt(2.0 / (9 * ci))
        parts.append((rng.standard_normal(co * 9 * ci) * std).astype(np.float32))
        parts.append(np.zeros(co, np.float32))
    return np.concatenate(parts)


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(np.shape(x))


class Net:
    def __init__(self, params, bf16_storage=True):
        self.bf = bf16_storage
        self.dt = np.float32 if bf16_storage else np.float64
        self.layers = [(w.astype(self.dt), b.astype(self.dt)) for w, b in split_params(np.asarray(params))]
        if self.bf:
            self.wq = [(bf16(w), b) for w, b in self.layers]
        else:
            self.wq = self.layers

    def r(self, x):
        return bf16(x) if self.bf else x

    # -- primitives --------------------------------------------------------
    @staticmethod
    def conv(x, w):
        """3x3 zero-padded convolution, x (B,H,W,Ci), w (Co,9,Ci)."""
        B, H, W, _ = x.shape
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
        out = np.zeros((B, H, W, w.shape[0]), dtype=x.dtype)
        for tap in range(9):
            ty, tx = tap // 3, tap % 3
            out += xp[:, ty:ty + H, tx:tx + W, :] @ w[:, tap, :].T
        return out

    @staticmethod
    def conv_wgrad(x, dz):
        """d/dw of sum(dz * conv(x, w)): (Co, 9, Ci)."""
        B, H, W, Ci = x.shape
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
        g = np.zeros((dz.shape[3], 9, Ci), dtype=x.dtype)
        dzf = dz.reshape(-1, dz.shape[3])
        for tap in range(9):
            ty, tx = tap // 3, tap % 3
            g[:, tap, :] = dzf.T @ xp[:, ty:ty + H, tx:tx + W, :].reshape(-1, Ci)
        return g

    @staticmethod
    def conv_dgrad(dz, w):
        """d/dx of sum(dz * conv(x, w)) = conv(dz, rotated transpose of w)."""
        wt = np.transpose(w[:, ::-1, :], (2, 1, 0))  # (Ci, 9, Co), tap reversed
        return Net.conv(dz, wt)

    @staticmethod
    def pool(x, scale):
        return scale * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2])

    @staticmethod
    def up(x):
        return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)

    # -- network -----------------------------------------------------------
    def forward(self, x0):
        """x0 (B,H,W,4) -> y (B,H,W); keeps the activations for backward."""
        r, L = self.r, self.wq
        relu = lambda z: np.maximum(z, 0)  # noqa: E731
        x0 = r(x0.astype(self.dt))
        a0 = r(relu(self.conv(x0, L[0][0]) + L[0][1]))
        p0 = r(self.pool(a0, 0.25))
        a1 = r(relu(self.conv(p0, L[1][0]) + L[1][1]))
        p1 = r(self.pool(a1, 0.25))
        a2 = r(relu(self.conv(p1, L[2][0]) + L[2][1]))
        p2 = r(self.pool(a2, 0.25))
        a3 = r(relu(self.conv(p2, L[3][0]) + L[3][1]))
        s2 = r(a2 + self.up(a3))
        a4 = r(relu(self.conv(s2, L[4][0]) + L[4][1]))
        s1 = r(a1 + self.up(a4))
        a5 = r(relu(self.conv(s1, L[5][0]) + L[5][1]))
        s0 = r(a0 + self.up(a5))
        y = (self.conv(s0, L[6][0]) + L[6][1])[..., 0]
        self.cache = dict(x0=x0, a0=a0, p0=p0, a1=a1, p1=p1, a2=a2, p2=p2, a3=a3, s2=s2, a4=a4, s1=s1, a5=a5, s0=s0)
        return y

    def loss_grad(self, x0, t):
        """Mean squared error against t (B,H,W) and the flat parameter gradient."""
        y = self.forward(x0)
        c, r, L = self.cache, self.r, self.wq
        n = y.size
        d = y - t.astype(self.dt)
        loss = float(np.mean(d.astype(np.float64) ** 2))
        dy = r((2.0 / n) * d)[..., None]
        grads = [None] * 7

        def lay(i, x, dz):
            grads[i] = (self.conv_wgrad(x, dz), dz.sum(axis=(0, 1, 2)))

        lay(6, c["s0"], dy)
        ds0 = r(self.conv_dgrad(dy, L[6][0]))
        dz5 = r(self.pool(ds0, 1.0)) * (c["a5"] > 0)
        lay(5, c["s1"], dz5)
        ds1 = r(self.conv_dgrad(dz5, L[5][0]))
        dz4 = r(self.pool(ds1, 1.0)) * (c["a4"] > 0)
        lay(4, c["s2"], dz4)
        ds2 = r(self.conv_dgrad(dz4, L[4][0]))
        dz3 = r(self.pool(ds2, 1.0)) * (c["a3"] > 0)
        lay(3, c["p2"], dz3)
        dp2 = r(self.conv_dgrad(dz3, L[3][0]))
        dz2 = r(ds2 + 0.25 * self.up(dp2)) * (c["a2"] > 0)
        lay(2, c["p1"], dz2)
        dp1 = r(self.conv_dgrad(dz2, L[2][0]))
        dz1 = r(ds1 + 0.25 * self.up(dp1)) * (c["a1"] > 0)
        lay(1, c["p0"], dz1)
        dp0 = r(self.conv_dgrad(dz1, L[1][0]))
        dz0 = r(ds0 + 0.25 * self.up(dp0)) * (c["a0"] > 0)
        lay(0, c["x0"], dz0)
        flat = np.concatenate([np.concatenate([g.ravel(), gb.ravel()]) for g, gb in grads])
        return loss, flat, y


def pack_input(c0, c1, mask):
    """(B,H,W) channels + (H,W) mask -> (B,H,W,4) with a zero fourth channel."""
    B = c0.shape[0]
    x = np.zeros(c0.shape + (4,), np.float32)
    x[..., 0] = c0
    x[..., 1] = c1
    x[..., 2] = np.broadcast_to(mask, c0.shape)
    return x
