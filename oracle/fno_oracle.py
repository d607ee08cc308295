"""FP64 reference of the FNO warm-start network (TEST INFRASTRUCTURE ONLY:
imported by tests/ to check the device forward and backward passes).

The same network as generalizedgrosspitaevskii.jl_amd/fno.py, written with
complex einsums over the explicit truncated Fourier factors and
differentiated by torch autograd on the CPU in float64.
"""
import math

import numpy as np
import torch


def gelu_tanh(x):
    return 0.5 * x * (1 + torch.tanh(math.sqrt(2 / math.pi) * (x + 0.044715 * x ** 3)))


def forward(params, x, N, m, layers):
    """params: dict of float64 tensors (W0, b0, Wr, Wi, Wc, bc, Wp, bp); x: (B, cin, N, N)."""
    W0, b0, Wr, Wi, Wc, bc, Wp, bp = (params[k] for k in ("W0", "b0", "Wr", "Wi", "Wc", "bc", "Wp", "bp"))
    Wd = W0.shape[0]
    kx = torch.tensor(np.r_[np.arange(m), np.arange(N - m, N)], dtype=torch.float64)
    ky = torch.arange(m, dtype=torch.float64)
    xs = torch.arange(N, dtype=torch.float64)
    Fx = torch.exp(-2j * math.pi * kx[:, None] * xs[None, :] / N)          # (2m, N)
    Fy = torch.exp(-2j * math.pi * xs[:, None] * ky[None, :] / N)          # (N, m)
    Gx = torch.exp(2j * math.pi * xs[:, None] * kx[None, :] / N)           # (N, 2m)
    cw = torch.where(ky == 0, 1.0, 2.0).to(torch.float64) / (N * N)
    Gy = cw[:, None] * torch.exp(2j * math.pi * ky[:, None] * xs[None, :] / N)  # (m, N)
    h = torch.einsum("oc,bcxy->boxy", W0, x) + b0[None, :, None, None]
    for l in range(layers):
        hh = torch.einsum("kx,bcxy,yj->bckj", Fx, h.to(torch.complex128), Fy)
        R = (Wr[l] + 1j * Wi[l]).reshape(2 * m, m, Wd, Wd)
        oh = torch.einsum("bikj,kjio->bokj", hh, R)
        spec = torch.einsum("xk,bokj,jy->boxy", Gx, oh, Gy).real
        pre = spec + torch.einsum("oi,bixy->boxy", Wc[l], h) + bc[l][None, :, None, None]
        h = gelu_tanh(pre) if l < layers - 1 else pre
    return torch.einsum("oc,bcxy->boxy", Wp, h) + bp[None, :, None, None]


def loss_and_grads(params_np, x_np, t_np, N, m, layers):
    params = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in params_np.items()}
    x = torch.tensor(x_np, dtype=torch.float64)
    t = torch.tensor(t_np, dtype=torch.float64)
    out = forward(params, x, N, m, layers)
    loss = torch.mean((out - t) ** 2)
    loss.backward()
    return out.detach().numpy(), float(loss.detach()), {k: v.grad.numpy() for k, v in params.items()}
