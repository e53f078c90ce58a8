"""PyTorch fp32 reference of the RealNVP proposal (test infrastructure).

Restates the flow computed by the bf16-MFMA kernels in
3photons-rust_amd/csrc/tp_flow.hip from its definition (csrc/tp_flow.h):
z0 = logit(v); six masked affine couplings z <- z exp(s) + t with
(s_raw, t) = MLP_l(z * cond_l), s = 1.5 tanh(s_raw), MLP 12(16) -> 64 -> 64 ->
32 with hard-swish (PyTorch's hardswish); u = sigmoid(z_6);
log q(u) = sum log v(1 - v) - sum s - sum log u(1 - u).

With ``emulate=True`` every operand of a matrix product (weights, biases and
activations) is rounded to bf16 as the kernels round them, with f32
arithmetic elsewhere; the rounding passes gradients straight through, so
autograd gives the gradient the kernels' backward pass approximates.  The
kernels evaluate hard-swish as c min(max(c / 6 + 1/2, 0), 1), which the
emulation follows.
"""
from __future__ import annotations

import numpy as np
import torch

D, DP, H, O, L = 12, 16, 64, 32, 6
LP = 7328
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 1024, 1088, 5184, 5248, 7296
MASKS = (0x03F, 0xFC0, 0x555, 0xAAA, 0x1C7, 0xE38)
SCALE = 1.5


def bf(x: torch.Tensor) -> torch.Tensor:
    """Round to bf16 (nearest even), gradient straight through."""
    return x + (x.to(torch.bfloat16).to(torch.float32) - x).detach()


def layer_params(w: torch.Tensor, l: int):
    p = w[l * LP:(l + 1) * LP]
    return (p[OFF_W1:OFF_B1].view(H, DP), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].view(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].view(O, H), p[OFF_B3:LP])


def cond_mask(l: int) -> torch.Tensor:
    return torch.tensor([(MASKS[l] >> j) & 1 == 1 for j in range(D)])


def mlp(a16: torch.Tensor, prm, emulate: bool):
    W1, b1, W2, b2, W3, b3 = prm
    if not emulate:
        h1 = torch.nn.functional.hardswish(a16 @ W1.T + b1)
        h2 = torch.nn.functional.hardswish(h1 @ W2.T + b2)
        o = h2 @ W3.T + b3
        return o[:, :D], o[:, 16:16 + D]
    c1 = bf(a16) @ bf(W1).T + bf(b1)
    h1 = c1 * torch.clamp(c1 * (1. / 6.) + 0.5, 0., 1.)
    c2 = bf(h1) @ bf(W2).T + bf(b2)
    h2 = c2 * torch.clamp(c2 * (1. / 6.) + 0.5, 0., 1.)
    o = bf(h2) @ bf(W3).T + bf(b3)
    return o[:, :D], o[:, 16:16 + D]


def forward(v: torch.Tensor, w: torch.Tensor, emulate: bool = True, z0: torch.Tensor | None = None):
    """(z_out [n,12], log q [n], sum of s [n], z0 [n,12]) for base points v [n,12]."""
    if z0 is None:
        z0 = torch.log(v) - torch.log1p(-v)
    z = z0
    ld = (torch.log(v) + torch.log1p(-v)).sum(1)
    ssum = torch.zeros(len(v))
    n = len(v)
    for l in range(L):
        c = cond_mask(l)
        a = torch.where(c, z, torch.zeros_like(z))
        a16 = torch.cat([a, torch.zeros(n, DP - D)], 1)
        sraw, t = mlp(a16, layer_params(w, l), emulate)
        s = SCALE * torch.tanh(sraw)
        s = torch.where(c, torch.zeros_like(s), s)
        z = torch.where(c, z, z * torch.exp(s) + t)
        ssum = ssum + s.sum(1)
    logq = ld - ssum + (torch.nn.functional.softplus(-z) + torch.nn.functional.softplus(z)).sum(1)
    return z, logq, ssum, z0


def inverse_logq(u: torch.Tensor, w: torch.Tensor, emulate: bool = True) -> torch.Tensor:
    """log q(u) by inverting the couplings (layers 5 .. 0)."""
    z = torch.log(u) - torch.log1p(-u)
    ld = -(torch.log(u) + torch.log1p(-u)).sum(1)
    n = len(u)
    for l in reversed(range(L)):
        c = cond_mask(l)
        a = torch.where(c, z, torch.zeros_like(z))
        a16 = torch.cat([a, torch.zeros(n, DP - D)], 1)
        sraw, t = mlp(a16, layer_params(w, l), emulate)
        s = torch.where(c, torch.zeros_like(sraw), SCALE * torch.tanh(sraw))
        z = torch.where(c, z, (z - t) * torch.exp(-s))
        ld = ld - s.sum(1)
    return ld - (torch.nn.functional.softplus(-z) + torch.nn.functional.softplus(z)).sum(1)


def grad(v: np.ndarray, gz: np.ndarray, w: np.ndarray, emulate: bool = True):
    """Gradient of (1/n) sum_i [gz_i . z_i - sum s_i] w.r.t. the parameters,
    and per sample w.r.t. z0 (without the 1/n), as tp_flow_grad defines."""
    wt = torch.tensor(w, dtype=torch.float32, requires_grad=True)
    vt = torch.tensor(v[:, :D], dtype=torch.float32)
    z0 = (torch.log(vt) - torch.log1p(-vt)).requires_grad_(True)
    z, _, ssum, _ = forward(vt, wt, emulate, z0=z0)
    g = torch.tensor(gz[:, :D], dtype=torch.float32)
    total = (g * z).sum() - ssum.sum()
    gw, gz0 = torch.autograd.grad(total, [wt, z0])
    return (gw / len(v)).numpy(), gz0.numpy()


def random_params(seed: int, out_scale: float = 0.3) -> np.ndarray:
    """Random weights with a non-trivial output layer (padding entries zero)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(L * LP, np.float32)
    for l in range(L):
        p = w[l * LP:(l + 1) * LP]
        W1 = rng.uniform(-1, 1, (H, DP)) / np.sqrt(D)
        W1[:, D:] = 0
        p[OFF_W1:OFF_B1] = W1.ravel()
        p[OFF_B1:OFF_W2] = rng.uniform(-1, 1, H) / np.sqrt(D)
        p[OFF_W2:OFF_B2] = (rng.uniform(-1, 1, (H, H)) / np.sqrt(H)).ravel()
        p[OFF_B2:OFF_W3] = rng.uniform(-1, 1, H) / np.sqrt(H)
        W3 = rng.uniform(-1, 1, (O, H)) * out_scale / np.sqrt(H)
        b3 = rng.uniform(-1, 1, O) * out_scale
        for r in range(O):
            if (D <= r < 16) or r >= 16 + D:
                W3[r] = 0
                b3[r] = 0
        p[OFF_W3:OFF_B3] = W3.ravel()
        p[OFF_B3:LP] = b3
    return w
