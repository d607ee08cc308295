"""Fourier-neural-operator warm start for steady states (no reference
counterpart: GeneralizedGrossPitaevskii.jl has no learned component).

The network maps input feature maps on an N x N grid (pump profile,
detuning, x, y) to the steady-state field of a two-component problem
(Re/Im of both components): a lifting layer, L spectral layers and a
projection,

    h_0 = W_0 f + b_0
    h_{l+1} = act(K_l(h_l) + W_l h_l + b_l),   act = GELU except after the last
    out = W_p h_L + b_p
    K(h)(x, y) = (1/N^2) sum_{kx in Kx, ky < m} c_ky Re[ (R(kx, ky) h^(kx, ky)) e^{2 pi i (x kx + y ky)/N} ]

with h^ the DFT of h restricted to kx in {0..m-1, N-m..N-1}, ky in {0..m-1},
R a complex width x width matrix per retained mode, c_0 = 1, c_ky = 2.

Every product -- the truncated DFTs, the per-mode channel mixing, the 1x1
convolutions, lifting, projection, and all their backward products -- runs
through gpe_fno_gemm (BF16 operands on the matrix cores, FP32 accumulation);
GELU, bias and Adam through the elementwise kernels of fno_kernels.hip.
Parameters and activations live in torch tensors on the device (storage
only; no torch arithmetic on the forward/backward path).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib as L


def _ptr(t, off=0):
    return ctypes.c_void_p(t.data_ptr() + 4 * int(off))


class _Ops:
    """Thin wrappers over the FNO kernels on one stream."""

    def __init__(self, stream=None):
        self.lib = L.load()
        self.stream = stream
        self.flops = 0  # multiply-adds x 2 issued through mm (roofline accounting)
        self.bytes = 0  # operand bytes: each operand read once, C written (and read when beta != 0)

    def mm(self, A, B, C, m, n, k, a, b, c, batch=1, alpha=1.0, beta=0.0, offs=(0, 0, 0), fp32=False):
        """C = alpha A B + beta C; a/b/c = (row stride, col stride, batch stride);
        BF16 operands unless fp32."""
        st = (ctypes.c_int64 * 9)(a[0], a[1], b[0], b[1], c[0], c[1], a[2], b[2], c[2])
        self.flops += 2 * m * n * k * batch
        self.bytes += 4 * (m * k * (batch if a[2] else 1) + k * n * (batch if b[2] else 1)
                           + m * n * batch * (2 if beta != 0.0 else 1))
        fn = self.lib.gpe_fno_gemm_f32 if fp32 else self.lib.gpe_fno_gemm
        L.check(fn(_ptr(A, offs[0]), _ptr(B, offs[1]), _ptr(C, offs[2]), m, n, k, st, batch, float(alpha),
                   float(beta), self.stream))

    def bias_act(self, x, bias, pre, y, rows, cols, C, act):
        self.bytes += 4 * rows * cols * (3 if pre is not None else 2)
        L.check(self.lib.gpe_fno_bias_act(_ptr(x), None if bias is None else _ptr(bias),
                                          None if pre is None else _ptr(pre), _ptr(y), rows, cols, C, 0, int(act),
                                          self.stream))

    def act_backward(self, pre, dy, dx, n):
        self.bytes += 12 * n
        L.check(self.lib.gpe_fno_act_backward(_ptr(pre), _ptr(dy), _ptr(dx), n, self.stream))

    def axpby(self, x, y, a, b, n):
        self.bytes += 12 * n
        L.check(self.lib.gpe_fno_axpby(_ptr(x), _ptr(y), float(a), float(b), n, self.stream))

    def adam(self, p, g, m, v, n, lr, step, b1=0.9, b2=0.999, eps=1e-8):
        L.check(self.lib.gpe_fno_adam(_ptr(p), _ptr(g), _ptr(m), _ptr(v), n, float(lr), float(b1), float(b2),
                                      float(eps), int(step), self.stream))


def dft_matrices(N, m):
    """Truncated DFT / inverse-DFT factors (FP64 -> FP32)."""
    y = np.arange(N)
    ky = np.arange(m)
    kx = np.r_[np.arange(m), np.arange(N - m, N)]
    ay = 2 * np.pi * np.outer(y, ky) / N                     # (N, m)
    Fy = np.concatenate([np.cos(ay), -np.sin(ay)], axis=1)   # (N, 2m): [Re | Im] of e^{-i ay}
    ax = 2 * np.pi * np.outer(kx, np.arange(N)) / N          # (2m, N)
    Fxr, Fxi = np.cos(ax), -np.sin(ax)
    Gxr, Gxi = np.cos(ax.T), np.sin(ax.T)                    # (N, 2m): e^{+i ax}
    c = np.where(ky == 0, 1.0, 2.0)[:, None] / (N * N)
    Gc = c * np.cos(ay.T)                                    # (m, N)
    Gs = -c * np.sin(ay.T)
    # C order: the products address these by raw pointer and row stride
    return {k: np.ascontiguousarray(v) for k, v in
            dict(Fy=Fy, Fxr=Fxr, Fxi=Fxi, Gxr=Gxr, Gxi=Gxi, Gc=Gc, Gs=Gs).items()}


class FNO:
    """The surrogate: parameters, forward, backward and Adam on the device."""

    PARAM_NAMES = ("W0", "b0", "Wr", "Wi", "Wc", "bc", "Wp", "bp")

    def __init__(self, N, in_ch=4, out_ch=4, width=32, modes=16, layers=4, seed=0, device="cuda:0"):
        import torch

        self.torch = torch
        self.N, self.cin, self.cout, self.W, self.m, self.nl = N, in_ch, out_ch, width, modes, layers
        self.Q = 2 * modes * modes
        self.dev = device
        rng = np.random.default_rng(seed)
        f32 = lambda a: torch.tensor(np.ascontiguousarray(a, dtype=np.float32), device=device)
        W, Q = width, self.Q
        self.p = {
            "W0": f32(rng.standard_normal((W, in_ch)) / math.sqrt(in_ch)),
            "b0": f32(np.zeros(W)),
            "Wr": f32(rng.standard_normal((layers, Q, W, W)) / W),
            "Wi": f32(rng.standard_normal((layers, Q, W, W)) / W),
            "Wc": f32(rng.standard_normal((layers, W, W)) / math.sqrt(W)),
            "bc": f32(np.zeros((layers, W))),
            "Wp": f32(rng.standard_normal((out_ch, W)) / math.sqrt(W)),
            "bp": f32(np.zeros(out_ch)),
        }
        self.g = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.adam_m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.adam_v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.step_count = 0
        self.c = {k: f32(v) for k, v in dft_matrices(N, modes).items()}
        self.ops = _Ops()
        self.ones = None

    # ----------------------------------------------------------- helpers
    def _buf(self, *shape):
        return self.torch.empty(shape, dtype=self.torch.float32, device=self.dev)

    def _zeros(self, *shape):
        return self.torch.zeros(shape, dtype=self.torch.float32, device=self.dev)

    def _ensure_ones(self, n):
        if self.ones is None or self.ones.numel() < n:
            self.ones = self.torch.ones(n, dtype=self.torch.float32, device=self.dev)

    # ------------------------------------------------------- spectral op
    def _dft_truncated(self, h, P, fp32=False):
        """(Yr, Yi), each (P, Q): the DFT of P real N x N planes at the
        retained modes (kx in {0..m-1, N-m..N-1}, ky in {0..m-1})."""
        o, N, m, Q = self.ops, self.N, self.m, self.Q
        c = self.c
        Z = self._buf(P, N, 2 * m)
        o.mm(h, c["Fy"], Z, P * N, 2 * m, N, (N, 1, 0), (2 * m, 1, 0), (2 * m, 1, 0), fp32=fp32)
        Yr, Yi = self._buf(P, Q), self._buf(P, Q)
        # Yr = Fxr Zr - Fxi Zi ; Yi = Fxr Zi + Fxi Zr   (per plane)
        zr, zi = (2 * m, 1, N * 2 * m), (2 * m, 1, N * 2 * m)
        fx = (N, 1, 0)
        yo = (m, 1, Q)
        o.mm(c["Fxr"], Z, Yr, 2 * m, m, N, fx, zr, yo, batch=P, fp32=fp32)
        o.mm(c["Fxi"], Z, Yr, 2 * m, m, N, fx, zi, yo, batch=P, alpha=-1.0, beta=1.0, offs=(0, m, 0), fp32=fp32)
        o.mm(c["Fxr"], Z, Yi, 2 * m, m, N, fx, zi, yo, batch=P, offs=(0, m, 0), fp32=fp32)
        o.mm(c["Fxi"], Z, Yi, 2 * m, m, N, fx, zr, yo, batch=P, beta=1.0, fp32=fp32)
        return Yr, Yi

    def _idft_truncated(self, Or, Oi, P, out, beta, fp32=False):
        """out (P, N, N) = beta out + (1/N^2) sum_modes c_ky Re[O e^{+i(kx x + ky y)}]."""
        o, N, m, Q = self.ops, self.N, self.m, self.Q
        c = self.c
        Vr, Vi = self._buf(P, N, m), self._buf(P, N, m)
        gx = (2 * m, 1, 0)
        ob = (m, 1, Q)
        vc = (m, 1, N * m)
        o.mm(c["Gxr"], Or, Vr, N, m, 2 * m, gx, ob, vc, batch=P, fp32=fp32)
        o.mm(c["Gxi"], Oi, Vr, N, m, 2 * m, gx, ob, vc, batch=P, alpha=-1.0, beta=1.0, fp32=fp32)
        o.mm(c["Gxr"], Oi, Vi, N, m, 2 * m, gx, ob, vc, batch=P, fp32=fp32)
        o.mm(c["Gxi"], Or, Vi, N, m, 2 * m, gx, ob, vc, batch=P, beta=1.0, fp32=fp32)
        o.mm(Vr, c["Gc"], out, P * N, N, m, (m, 1, 0), (N, 1, 0), (N, 1, 0), beta=beta, fp32=fp32)
        o.mm(Vi, c["Gs"], out, P * N, N, m, (m, 1, 0), (N, 1, 0), (N, 1, 0), beta=1.0, fp32=fp32)

    def band_limit(self, y):
        """Low-pass of real planes y (..., N, N): the inverse transform of
        their retained modes (the spectral layers' band), as a new tensor.
        FP32 products: rounding the basis functions to BF16 would put error
        back at every wavenumber."""
        N = self.N
        P = y.numel() // (N * N)
        Yr, Yi = self._dft_truncated(y, P, fp32=True)
        out = self._buf(*y.shape)
        self._idft_truncated(Yr, Yi, P, out, beta=0.0, fp32=True)
        return out

    def _spectral(self, h, l, out, B, save):
        """out (B*W, N, N) += K_l(h)."""
        o, N, m, W, Q = self.ops, self.N, self.m, self.W, self.Q
        P = B * W
        c = self.c
        Yr, Yi = self._dft_truncated(h, P)
        # per-mode mixing: O = Y R (complex), Or = Yr Wr - Yi Wi, Oi = Yr Wi + Yi Wr
        Or, Oi = self._buf(P, Q), self._buf(P, Q)
        ya = (W * Q, Q, 1)
        wb = (W, 1, W * W)
        oc = (W * Q, Q, 1)
        woff = l * Q * W * W
        Wr, Wi = self.p["Wr"], self.p["Wi"]
        o.mm(Yr, Wr, Or, B, W, W, ya, wb, oc, batch=Q, offs=(0, woff, 0))
        o.mm(Yi, Wi, Or, B, W, W, ya, wb, oc, batch=Q, alpha=-1.0, beta=1.0, offs=(0, woff, 0))
        o.mm(Yr, Wi, Oi, B, W, W, ya, wb, oc, batch=Q, offs=(0, woff, 0))
        o.mm(Yi, Wr, Oi, B, W, W, ya, wb, oc, batch=Q, beta=1.0, offs=(0, woff, 0))
        # inverse transform (x then y) to a real field, accumulated into out
        self._idft_truncated(Or, Oi, P, out, beta=1.0)
        if save:
            return dict(Yr=Yr, Yi=Yi)
        return None

    def _spectral_backward(self, dS, l, cache, dh, B):
        """dh (B*W, N, N) += K_l^T(dS); accumulates dR into self.g."""
        o, N, m, W, Q = self.ops, self.N, self.m, self.W, self.Q
        P = B * W
        c = self.c
        dVr, dVi = self._buf(P, N, m), self._buf(P, N, m)
        # dV = dS G^T  (G (m, N) -> G^T element (y, ky) at ky*N + y)
        o.mm(dS, c["Gc"], dVr, P * N, m, N, (N, 1, 0), (1, N, 0), (m, 1, 0))
        o.mm(dS, c["Gs"], dVi, P * N, m, N, (N, 1, 0), (1, N, 0), (m, 1, 0))
        # dOr = Gxr^T dVr + Gxi^T dVi ; dOi = -Gxi^T dVr + Gxr^T dVi
        dOr, dOi = self._buf(P, Q), self._buf(P, Q)
        gT = (1, 2 * m, 0)
        vb = (m, 1, N * m)
        oc = (m, 1, Q)
        o.mm(c["Gxr"], dVr, dOr, 2 * m, m, N, gT, vb, oc, batch=P)
        o.mm(c["Gxi"], dVi, dOr, 2 * m, m, N, gT, vb, oc, batch=P, beta=1.0)
        o.mm(c["Gxi"], dVr, dOi, 2 * m, m, N, gT, vb, oc, batch=P, alpha=-1.0)
        o.mm(c["Gxr"], dVi, dOi, 2 * m, m, N, gT, vb, oc, batch=P, beta=1.0)
        # weights: dWr = Yr^T dOr + Yi^T dOi ; dWi = -Yi^T dOr + Yr^T dOi   (per mode)
        Yr, Yi = cache["Yr"], cache["Yi"]
        yT = (Q, W * Q, 1)
        db = (W * Q, Q, 1)
        wc = (W, 1, W * W)
        woff = l * Q * W * W
        gWr, gWi = self.g["Wr"], self.g["Wi"]
        o.mm(Yr, dOr, gWr, W, W, B, yT, db, wc, batch=Q, beta=1.0, offs=(0, 0, woff))
        o.mm(Yi, dOi, gWr, W, W, B, yT, db, wc, batch=Q, beta=1.0, offs=(0, 0, woff))
        o.mm(Yi, dOr, gWi, W, W, B, yT, db, wc, batch=Q, alpha=-1.0, beta=1.0, offs=(0, 0, woff))
        o.mm(Yr, dOi, gWi, W, W, B, yT, db, wc, batch=Q, beta=1.0, offs=(0, 0, woff))
        # dYr = dOr Wr^T + dOi Wi^T ; dYi = -dOr Wi^T + dOi Wr^T
        dYr, dYi = self._buf(P, Q), self._buf(P, Q)
        wT = (1, W, W * W)
        ya = (W * Q, Q, 1)
        Wr, Wi = self.p["Wr"], self.p["Wi"]
        o.mm(dOr, Wr, dYr, B, W, W, ya, wT, ya, batch=Q, offs=(0, woff, 0))
        o.mm(dOi, Wi, dYr, B, W, W, ya, wT, ya, batch=Q, beta=1.0, offs=(0, woff, 0))
        o.mm(dOr, Wi, dYi, B, W, W, ya, wT, ya, batch=Q, alpha=-1.0, offs=(0, woff, 0))
        o.mm(dOi, Wr, dYi, B, W, W, ya, wT, ya, batch=Q, beta=1.0, offs=(0, woff, 0))
        # dZr = Fxr^T dYr + Fxi^T dYi ; dZi = -Fxi^T dYr + Fxr^T dYi  into dZ (P, N, 2m)
        dZ = self._buf(P, N, 2 * m)
        fT = (1, N, 0)
        yb = (m, 1, Q)
        zc = (2 * m, 1, N * 2 * m)
        o.mm(c["Fxr"], dYr, dZ, N, m, 2 * m, fT, yb, zc, batch=P)
        o.mm(c["Fxi"], dYi, dZ, N, m, 2 * m, fT, yb, zc, batch=P, beta=1.0)
        o.mm(c["Fxi"], dYr, dZ, N, m, 2 * m, fT, yb, zc, batch=P, alpha=-1.0, offs=(0, 0, m))
        o.mm(c["Fxr"], dYi, dZ, N, m, 2 * m, fT, yb, zc, batch=P, beta=1.0, offs=(0, 0, m))
        # dh += dZ Fy^T
        o.mm(dZ, c["Fy"], dh, P * N, N, 2 * m, (2 * m, 1, 0), (1, 2 * m, 0), (N, 1, 0), beta=1.0)

    # ------------------------------------------------------------ forward
    def forward(self, x, save=False):
        """x: (B, in_ch, N, N) float32 device tensor -> (B, out_ch, N, N)."""
        o, N, W = self.ops, self.N, self.W
        B = x.shape[0]
        NN = N * N
        self._ensure_ones(NN)
        cache = dict(x=x, B=B, layers=[])
        h = self._buf(B, W, N, N)
        o.mm(self.p["W0"], x, h, W, NN, self.cin, (self.cin, 1, 0), (NN, 1, self.cin * NN), (NN, 1, W * NN), batch=B)
        o.bias_act(h, self.p["b0"], None, h, B * W, NN, W, act=False)
        for l in range(self.nl):
            pre = self._buf(B, W, N, N)
            o.mm(self.p["Wc"], h, pre, W, NN, W, (W, 1, 0), (NN, 1, W * NN), (NN, 1, W * NN), batch=B,
                 offs=(l * W * W, 0, 0))
            sc = self._spectral(h, l, pre, B, save)
            hn = self._buf(B, W, N, N)
            last = l == self.nl - 1
            o.bias_act(pre, self.p["bc"][l], pre if (save and not last) else None, hn, B * W, NN, W, act=not last)
            cache["layers"].append(dict(h=h, pre=pre, spec=sc))
            h = hn
        out = self._buf(B, self.cout, N, N)
        o.mm(self.p["Wp"], h, out, self.cout, NN, W, (W, 1, 0), (NN, 1, W * NN), (NN, 1, self.cout * NN), batch=B)
        o.bias_act(out, self.p["bp"], None, out, B * self.cout, NN, self.cout, act=False)
        cache["hL"] = h
        return out, cache

    def _wgrad(self, A, Bm, gname, m, n, B, a_bstride, b_bstride, a_rc, b_rc, off=0):
        """g[gname][off: off + m*n] (contiguous m x n) += sum_b A_b Bm_b with
        A_b (m x K) at b * a_bstride, Bm_b (K x n) at b * b_bstride and
        K = N*N.  K is split into chunks computed as one batched product per
        b (partials), then reduced by one product with a ones vector."""
        o = self.ops
        K = self.N * self.N
        # short chunks: m, n <= width, so the partial products need many
        # workgroups (K / kc per case) to fill the chip
        kc = min(K, 128)
        S = K // kc
        part = self._buf(B * S, m, n)
        for b in range(B):
            o.mm(A, Bm, part, m, n, kc, (a_rc[0], a_rc[1], kc * a_rc[1]), (b_rc[0], b_rc[1], kc * b_rc[0]),
                 (n, 1, m * n), batch=S, offs=(b * a_bstride, b * b_bstride, b * S * m * n))
        self._reduce_rows(part, B * S, m * n, self.g[gname], off)

    def _reduce_rows(self, part, rows, cols, out, off, group=64):
        """out[off:off+cols] += sum over the rows of part (rows x cols, row-major).
        Long sums are taken in two stages (groups of `group` rows batched over all
        groups, then the group sums), so the product never runs thousands of
        k-tiles on a handful of workgroups; the order of the sums is fixed."""
        o = self.ops
        self._ensure_ones(max(rows, group))
        if rows > 4 * group and rows % group == 0:
            ng = rows // group
            mid = self._buf(ng, cols)
            o.mm(part, self.ones, mid, cols, 1, group, (1, cols, group * cols), (1, 1, 0), (1, 1, cols), batch=ng)
            part, rows = mid, ng
        o.mm(part, self.ones, out, cols, 1, rows, (1, cols, 0), (1, 1, 0), (1, 1, 0), beta=1.0, offs=(0, 0, off))

    def _bias_grad(self, d, gname, C, B, NN, off=0):
        """g[gname][off:off+C] += sum over batch and pixels of d (B, C, NN)."""
        self._ensure_ones(NN)
        self._wgrad(d, self.ones, gname, C, 1, B, C * NN, 0, (NN, 1), (1, 1), off=off)

    def backward(self, dout, cache):
        """Accumulate parameter gradients of sum(dout * out) into self.g."""
        o, N, W = self.ops, self.N, self.W
        B = cache["B"]
        NN = N * N
        # projection
        hL = cache["hL"]
        self._wgrad(dout, hL, "Wp", self.cout, W, B, self.cout * NN, W * NN, (NN, 1), (1, NN))
        self._bias_grad(dout, "bp", self.cout, B, NN)
        dh = self._buf(B, W, N, N)
        o.mm(self.p["Wp"], dout, dh, W, NN, self.cout, (1, W, 0), (NN, 1, self.cout * NN), (NN, 1, W * NN), batch=B)
        for l in reversed(range(self.nl)):
            lay = cache["layers"][l]
            last = l == self.nl - 1
            dpre = self._buf(B, W, N, N)
            if last:
                dpre.copy_(dh)
            else:
                o.act_backward(lay["pre"], dh, dpre, B * W * NN)
            self._bias_grad(dpre, "bc", W, B, NN, off=l * W)
            h = lay["h"]
            self._wgrad(dpre, h, "Wc", W, W, B, W * NN, W * NN, (NN, 1), (1, NN), off=l * W * W)
            dh = self._buf(B, W, N, N)
            o.mm(self.p["Wc"], dpre, dh, W, NN, W, (1, W, 0), (NN, 1, W * NN), (NN, 1, W * NN), batch=B,
                 offs=(l * W * W, 0, 0))
            self._spectral_backward(dpre, l, lay["spec"], dh, B)
        # lifting
        x = cache["x"]
        self._wgrad(dh, x, "W0", W, self.cin, B, W * NN, self.cin * NN, (NN, 1), (1, NN))
        self._bias_grad(dh, "b0", W, B, NN)

    def zero_grad(self):
        for g in self.g.values():
            g.zero_()

    def loss_and_grad(self, x, target):
        """MSE loss (mean over all values) and its gradient accumulated into g."""
        out, cache = self.forward(x, save=True)
        n = out.numel()
        d = self._buf(*out.shape)
        d.copy_(target)
        self.ops.axpby(out, d, 2.0 / n, -2.0 / n, n)  # d = 2 (out - target) / n
        self.backward(d, cache)
        return out, d

    def adam_step(self, lr):
        self.step_count += 1
        for k in self.PARAM_NAMES:
            self.ops.adam(self.p[k], self.g[k], self.adam_m[k], self.adam_v[k], self.p[k].numel(), lr,
                          self.step_count)
