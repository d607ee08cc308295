"""Host composition of the FNO warm start, checked on the CPU.

The network in fno.py is a sequence of strided GEMM / bias / activation /
axpby calls.  Here those five calls are emulated in float64 on CPU tensors
with exactly the stride-and-offset contract of gpe_fno_gemm and friends
(include/gpe_hip.h), and the composed forward pass and every parameter
gradient are compared with the float64 autograd reference
(oracle/fno_oracle.py).  The device kernels themselves are checked in
test_gpu_fno.py.
"""
import numpy as np
import pytest
import torch

from oracle import fno_oracle


class EmulatedOps:
    """float64 CPU stand-in for the device kernels (test-only)."""

    @staticmethod
    def _view(t, off, shape, strides):
        return torch.as_strided(t, shape, strides, t.storage_offset() + int(off))

    def mm(self, A, B, C, m, n, k, a, b, c, batch=1, alpha=1.0, beta=0.0, offs=(0, 0, 0), fp32=False):
        Av = self._view(A, offs[0], (batch, m, k), (a[2], a[0], a[1]))
        Bv = self._view(B, offs[1], (batch, k, n), (b[2], b[0], b[1]))
        Cv = self._view(C, offs[2], (batch, m, n), (c[2], c[0], c[1]))
        prev = Cv.clone() if beta != 0 else 0.0
        Cv.copy_(alpha * torch.bmm(Av, Bv) + beta * prev)

    def bias_act(self, x, bias, pre, y, rows, cols, C, act):
        v = x.reshape(rows, cols) + (0 if bias is None else bias.reshape(-1)[torch.arange(rows) % C][:, None])
        if pre is not None:
            pre.reshape(rows, cols).copy_(v)
        y.reshape(rows, cols).copy_(fno_oracle.gelu_tanh(v) if act else v)

    def act_backward(self, pre, dy, dx, n):
        p = pre.reshape(-1).clone().requires_grad_(True)
        g, = torch.autograd.grad(fno_oracle.gelu_tanh(p).sum(), p)
        dx.reshape(-1).copy_(dy.reshape(-1) * g)

    def axpby(self, x, y, a, b, n):
        y.reshape(-1).copy_(a * x.reshape(-1) + b * y.reshape(-1))


def _cpu_net(ggp, N, m, W, layers, cin, cout):
    from gpe_hip_amd import fno

    net = fno.FNO.__new__(fno.FNO)
    net.torch = torch
    net.N, net.cin, net.cout, net.W, net.m, net.nl = N, cin, cout, W, m, layers
    net.Q = 2 * m * m
    net.dev = "cpu"
    rng = np.random.default_rng(5)
    shapes = dict(W0=(W, cin), b0=(W,), Wr=(layers, net.Q, W, W), Wi=(layers, net.Q, W, W), Wc=(layers, W, W),
                  bc=(layers, W), Wp=(cout, W), bp=(cout,))
    net.p = {k: torch.tensor(rng.standard_normal(s) * 0.3) for k, s in shapes.items()}
    net.g = {k: torch.zeros_like(v) for k, v in net.p.items()}
    net.c = {k: torch.tensor(v) for k, v in fno.dft_matrices(N, m).items()}
    net.ops = EmulatedOps()
    net.ones = None
    net._buf = lambda *s: torch.empty(s, dtype=torch.float64)
    net._ensure_ones = lambda n: setattr(net, "ones", torch.ones(max(n, 1), dtype=torch.float64))
    return net


def test_dft_factors_are_c_contiguous(ggp):
    from gpe_hip_amd import fno

    for k, v in fno.dft_matrices(16, 4).items():
        assert v.flags["C_CONTIGUOUS"], k


@pytest.mark.parametrize("N,m,W,layers,B", [(16, 4, 8, 2, 2), (32, 6, 5, 3, 1), (64, 8, 4, 1, 3)])
def test_fno_composition_matches_autograd(ggp, N, m, W, layers, B):
    net = _cpu_net(ggp, N, m, W, layers, 4, 4)
    rng = np.random.default_rng(N + B)
    x = rng.standard_normal((B, 4, N, N))
    t = rng.standard_normal((B, 4, N, N))
    params = {k: v.numpy().copy() for k, v in net.p.items()}
    out_ref, _, g_ref = fno_oracle.loss_and_grads(params, x, t, N, m, layers)
    out, _ = net.loss_and_grad(torch.tensor(x), torch.tensor(t))
    assert np.linalg.norm(out.numpy() - out_ref) <= 1e-12 * np.linalg.norm(out_ref)
    for k, gr in g_ref.items():
        assert np.linalg.norm(net.g[k].numpy() - gr) <= 1e-11 * np.linalg.norm(gr), k


def test_oracle_spectral_layer_matches_fft():
    """The reference's spectral layer equals the truncated-mode FFT definition."""
    N, m, W = 16, 4, 3
    rng = np.random.default_rng(0)
    h = rng.standard_normal((1, W, N, N))
    Wr = rng.standard_normal((2 * m * m, W, W))
    Wi = rng.standard_normal((2 * m * m, W, W))
    H = np.fft.fft2(h, axes=(2, 3))
    kx = np.r_[np.arange(m), np.arange(N - m, N)]
    R = (Wr + 1j * Wi).reshape(2 * m, m, W, W)
    xs = np.arange(N)
    spec = np.zeros_like(h)
    for a, k in enumerate(kx):
        for j in range(m):
            O = np.einsum("bi,io->bo", H[:, :, k, j], R[a, j])
            ph = np.exp(2j * np.pi * (xs[:, None] * k + xs[None, :] * j) / N)
            spec += (1 if j == 0 else 2) * np.real(O[:, :, None, None] * ph) / N ** 2
    params = dict(W0=np.eye(W), b0=np.zeros(W), Wr=Wr[None], Wi=Wi[None], Wc=np.zeros((1, W, W)),
                  bc=np.zeros((1, W)), Wp=np.eye(W), bp=np.zeros(W))
    o = fno_oracle.forward({k: torch.tensor(v) for k, v in params.items()}, torch.tensor(h), N, m, 1).numpy()
    assert np.abs(o - spec).max() <= 1e-12 * np.abs(spec).max()


def test_band_limit_keeps_retained_modes_and_removes_the_rest(ggp):
    N, m = 32, 6
    net = _cpu_net(ggp, N, m, 4, 1, 4, 4)
    x = np.arange(N)
    X, Y = np.meshgrid(x, x, indexing="ij")
    rng = np.random.default_rng(3)
    inband = np.zeros((N, N))
    for kx in range(-m + 1, m):
        for ky in range(0, m):
            a, b = rng.standard_normal(2)
            inband += a * np.cos(2 * np.pi * (kx * X + ky * Y) / N) + b * np.sin(2 * np.pi * (kx * X + ky * Y) / N)
    outband = np.cos(2 * np.pi * (3 * X + (m + 2) * Y) / N) + np.sin(2 * np.pi * ((m + 4) * X + 1 * Y) / N)
    planes = torch.tensor(np.stack([inband, inband + outband])[None])
    out = net.band_limit(planes).numpy()[0]
    assert np.abs(out[0] - inband).max() <= 1e-12 * np.abs(inband).max()
    assert np.abs(out[1] - inband).max() <= 1e-12 * np.abs(inband).max()
