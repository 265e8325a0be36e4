"""CPU oracle of the shift-predictor MLP (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker of the GPU kernels in stab_amd/csrc/mlp.hip; the
product path never calls it.

The network has no counterpart in the reference (sscollis/stab takes the
eigenvalue guess from its user, thesis/TStest/run.sh:13-17), so the oracle is
pinned by mathematics instead of goldens: its gradient against central finite
differences in FP64, and Adam against the closed form it takes for a constant
gradient.

Restated here (numpy, FP64 arithmetic):
  * the flat parameter layout of include/stab_hip.h (stab_mlp section),
  * forward: h_0 = x' (padded to nin_p), h_{l+1} = tanh(h_l W_l^T + b_l),
    y' = h_depth W_o^T + b_o,
  * loss: mean over rows and the nout real outputs of (y' - t')^2,
  * backward: the chain rule of the above,
  * Adam (Kingma & Ba 2015), beta1 0.9, beta2 0.999, eps 1e-8,
  * optional rounding of every matrix-product operand to BF16 (nearest even),
    at the places the GPU kernels round (both operands of each product),
    with the products and sums then exact/FP64, as the GPU's FP32
    accumulation of exact BF16 products approximates.
"""
from __future__ import annotations

import dataclasses

import numpy as np


def pad16(k: int) -> int:
    return (k + 15) // 16 * 16


@dataclasses.dataclass
class Layout:
    nin: int
    width: int
    depth: int
    nout: int

    @property
    def nin_p(self):
        return pad16(self.nin)

    @property
    def nout_p(self):
        return 16

    def layers(self):
        """(offset_W, out, in, offset_b) per linear layer."""
        out, off = [], 0
        for l in range(self.depth + 1):
            i = self.nin_p if l == 0 else self.width
            o = self.nout_p if l == self.depth else self.width
            out.append((off, o, i, off + o * i))
            off += o * i + o
        return out

    @property
    def nparams(self):
        off, o, i, ob = self.layers()[-1]
        return ob + o

    def unpack(self, p):
        Ws, bs = [], []
        for off, o, i, ob in self.layers():
            Ws.append(np.asarray(p[off:off + o * i], dtype=np.float64).reshape(o, i))
            bs.append(np.asarray(p[ob:ob + o], dtype=np.float64))
        return Ws, bs

    def random_params(self, rng, scale=0.3):
        """Random parameters with zero padding, as the GPU layout requires."""
        p = np.zeros(self.nparams)
        for l, (off, o, i, ob) in enumerate(self.layers()):
            ro = self.nout if l == self.depth else o
            ri = self.nin if l == 0 else i
            W = np.zeros((o, i))
            W[:ro, :ri] = rng.standard_normal((ro, ri)) * scale
            p[off:off + o * i] = W.ravel()
            b = np.zeros(o)
            b[:ro] = rng.standard_normal(ro) * scale
            p[ob:ob + o] = b
        return p


def bf16(a):
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _mm(a, b, rnd):
    if rnd:
        a, b = bf16(a), bf16(b)
    return a @ b


def forward(lay: Layout, p, xn, rnd=False):
    """xn: (B, nin) normalised inputs.  Returns (activations, y' (B, nout_p))."""
    Ws, bs = lay.unpack(p)
    B = xn.shape[0]
    h = np.zeros((B, lay.nin_p))
    h[:, :lay.nin] = xn
    acts = [h]
    for l in range(lay.depth):
        z = _mm(acts[-1], Ws[l].T, rnd) + bs[l]
        if rnd:
            z = z.astype(np.float32).astype(np.float64)
        acts.append(np.tanh(z))
    y = _mm(acts[-1], Ws[-1].T, rnd) + bs[-1]
    return acts, y


def loss_grad(lay: Layout, p, xn, tn, rnd=False):
    """Loss and flat gradient for normalised inputs xn (B, nin) and targets tn (B, nout)."""
    Ws, bs = lay.unpack(p)
    acts, y = forward(lay, p, xn, rnd)
    B = xn.shape[0]
    d = y[:, :lay.nout] - tn
    loss = float(np.mean(d * d))
    dY = np.zeros_like(y)
    dY[:, :lay.nout] = 2.0 * d / (B * lay.nout)
    g = np.zeros(lay.nparams)
    L = lay.layers()
    # output layer
    off, o, i, ob = L[-1]
    g[off:off + o * i] = _mm(dY.T, acts[-1], rnd).ravel()
    g[ob:ob + o] = dY.sum(0)
    dz = _mm(dY, Ws[-1], rnd) * (1.0 - acts[-1] ** 2)
    for l in range(lay.depth - 1, -1, -1):
        off, o, i, ob = L[l]
        g[off:off + o * i] = _mm(dz.T, acts[l], rnd).ravel()
        g[ob:ob + o] = dz.sum(0)
        if l > 0:
            dz = _mm(dz, Ws[l], rnd) * (1.0 - acts[l] ** 2)
    return loss, g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t >= 1); returns new (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def finite_difference_grad(lay: Layout, p, xn, tn, idx, h=1e-6):
    """Central differences of the FP64 loss at parameter indices idx."""
    out = []
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        out.append((loss_grad(lay, pp, xn, tn)[0] - loss_grad(lay, pm, xn, tn)[0]) / (2 * h))
    return np.array(out)
