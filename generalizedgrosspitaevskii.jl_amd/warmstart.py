"""Learned warm start for steady states (no reference counterpart: the
reference integrates a fixed tspan and has no learned component).

    generate_steady_states  steady states of one problem family over a list
                            of parameter sets; several handles run at once,
                            each on its own HIP stream, one host thread each
    WarmStart               the FNO surrogate (fno.FNO) with per-channel
                            normalisation: fit, predict, save, load
    refine                  steady_state started from a predicted field

The solver work runs in libgpe_hip.so (gpe_steady_state), the network's
dense products and elementwise work in its FNO kernels; torch tensors hold
device storage and do the data shuffling (batch gathers, normalisation).
"""
from __future__ import annotations

import math
import threading
import time

import numpy as np

from . import problem as P
from .fno import FNO


def generate_steady_states(make_problem, params, *, dt, tol, check_every=200, max_steps=200000, workers=4,
                           terms=("dispersion", "pump"), progress=None):
    """Steady states of make_problem(param) from its u0 for every param.

    make_problem(param) -> GrossPitaevskiiProblem with host (numpy) u0; all
    problems share shape, dtype and every term not listed in `terms` (the
    ones re-evaluated per parameter set on a reused handle).  Returns
    (fields, info): fields complex (n, M, *spatial) in the problem's dtype,
    info a list of dict(steps, residual, ms) per parameter set.
    """
    n = len(params)
    if n == 0:
        raise ValueError("generate_steady_states: no parameter sets")
    first = make_problem(params[0])
    M = first.M
    shape = tuple(first.u0[0].shape)
    fields = np.empty((n, M) + shape, dtype=first.u0[0].dtype)
    info = [None] * n
    lock = threading.Lock()
    nxt = [0]
    errors = []
    done = [0]
    t0 = time.time()

    def worker():
        it = None
        try:
            it = P.init(make_problem(params[0]), P.StrangSplitting(), (0.0, dt), dt=dt, nsaves=1)
            while True:
                with lock:
                    i = nxt[0]
                    nxt[0] += 1
                if i >= n or errors:
                    return
                it.set_param(params[i], terms)
                u, inf = it.steady_state(tol=tol, check_every=check_every, max_steps=max_steps)
                for c in range(M):
                    fields[i, c] = u[c]
                info[i] = dict(steps=inf["steps"], residual=inf["residual"], ms=inf["ms"])
                with lock:
                    done[0] += 1
                    k = done[0]
                if progress and (k % progress == 0 or k == n):
                    print(f"  steady states {k}/{n}  ({time.time() - t0:.1f} s)", flush=True)
        except BaseException as e:  # surfaced in the calling thread
            with lock:
                errors.append(e)
        finally:
            if it is not None:
                it.close()

    threads = [threading.Thread(target=worker, daemon=True) for _ in range(max(1, min(workers, n)))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return fields, info


def fields_to_channels(fields):
    """(n, M, *spatial) complex -> (n, 2M, *spatial) float32: Re/Im per component."""
    n, M = fields.shape[:2]
    out = np.empty((n, 2 * M) + fields.shape[2:], dtype=np.float32)
    out[:, 0::2] = fields.real
    out[:, 1::2] = fields.imag
    return out


class WarmStart:
    """FNO surrogate from feature maps (n, C_in, N, N) to the steady-state
    field channels (n, 2M, N, N), with per-channel input standardisation
    and per-channel output scaling fitted on the training set."""

    def __init__(self, N, in_ch, out_ch, width=32, modes=16, layers=4, seed=0, device="cuda:0"):
        import torch

        self.torch = torch
        self.cfg = dict(N=N, in_ch=in_ch, out_ch=out_ch, width=width, modes=modes, layers=layers, seed=seed)
        self.net = FNO(N, in_ch=in_ch, out_ch=out_ch, width=width, modes=modes, layers=layers, seed=seed,
                       device=device)
        self.dev = device
        self.x_mean = self.x_std = self.y_scale = None

    # ----------------------------------------------------- normalisation
    def _norm_x(self, X):
        return (X - self.x_mean) / self.x_std

    def fit(self, X, Y, *, epochs=50, batch=16, lr=1e-3, lr_min=1e-5, time_budget=None, seed=0, log_every=1):
        """Adam on the mean squared error of the scaled outputs, cosine
        learning-rate decay over `epochs`; stops early when time_budget
        seconds are spent.  X, Y: float32 device tensors.  Returns the
        per-epoch mean training loss."""
        torch = self.torch
        n = X.shape[0]
        self.x_mean = X.mean(dim=(0, 2, 3), keepdim=True)[0:1]
        self.x_std = X.std(dim=(0, 2, 3), keepdim=True)[0:1].clamp_min(1e-6)
        self.y_scale = Y.pow(2).mean(dim=(0, 2, 3), keepdim=True).sqrt()[0:1].clamp_min(1e-12)
        g = torch.Generator(device="cpu")
        g.manual_seed(seed)
        nb = n // batch
        total = epochs * nb
        losses = []
        t0 = time.time()
        step = 0
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g)[: nb * batch].to(self.dev)
            acc = torch.zeros((), dtype=torch.float32, device=self.dev)
            for b in range(nb):
                idx = perm[b * batch:(b + 1) * batch]
                xb = self._norm_x(X.index_select(0, idx)).contiguous()
                yb = (Y.index_select(0, idx) / self.y_scale).contiguous()
                self.net.zero_grad()
                out, _ = self.net.loss_and_grad(xb, yb)
                acc += (out - yb).pow(2).mean()
                frac = step / max(1, total - 1)
                self.net.adam_step(lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * frac)))
                step += 1
            losses.append(float(acc) / nb)
            if log_every and (ep % log_every == 0 or ep == epochs - 1):
                print(f"  epoch {ep + 1}/{epochs}  loss {losses[-1]:.3e}  ({time.time() - t0:.1f} s)", flush=True)
            if time_budget is not None and time.time() - t0 > time_budget:
                print(f"  time budget reached after epoch {ep + 1}", flush=True)
                break
        return losses

    def predict(self, X, batch=16, band_limit=True):
        """Output channels in physical units for feature maps X (device).

        band_limit: keep only the spectral layers' retained modes of the
        output (FP32 transforms).  A smooth steady state has no content
        beyond them, while the network's error spreads over all wavenumbers;
        error left at high k sits in the slowest-decaying modes of the
        refinement (for the polariton model, exciton-like modes that decay
        at gamma_x), which a start from zero never populates."""
        torch = self.torch
        outs = []
        for s in range(0, X.shape[0], batch):
            xb = self._norm_x(X[s:s + batch]).contiguous()
            out, _ = self.net.forward(xb)
            if band_limit:
                out = self.net.band_limit(out)
            outs.append(out * self.y_scale)
        return torch.cat(outs, 0)

    def predict_fields(self, X, dtype=None, band_limit=True):
        """Predicted fields as a tuple of M complex (n, N, N) device tensors."""
        torch = self.torch
        Yp = self.predict(X, band_limit=band_limit)
        M = Yp.shape[1] // 2
        cdt = dtype or torch.complex64
        return tuple(torch.complex(Yp[:, 2 * c], Yp[:, 2 * c + 1]).to(cdt) for c in range(M))

    # ---------------------------------------------------------- storage
    def save(self, path):
        arrs = {f"p_{k}": v.detach().cpu().numpy() for k, v in self.net.p.items()}
        arrs.update(x_mean=self.x_mean.cpu().numpy(), x_std=self.x_std.cpu().numpy(),
                    y_scale=self.y_scale.cpu().numpy(), cfg=np.array([self.cfg[k] for k in
                                                                       ("N", "in_ch", "out_ch", "width", "modes",
                                                                        "layers", "seed")]))
        np.savez(path, **arrs)

    @classmethod
    def load(cls, path, device="cuda:0"):
        import torch

        z = np.load(path, allow_pickle=False)
        N, in_ch, out_ch, width, modes, layers, seed = (int(v) for v in z["cfg"])
        ws = cls(N, in_ch, out_ch, width=width, modes=modes, layers=layers, seed=seed, device=device)
        for k in ws.net.p:
            ws.net.p[k].copy_(torch.tensor(z[f"p_{k}"], device=device))
        ws.x_mean = torch.tensor(z["x_mean"], device=device)
        ws.x_std = torch.tensor(z["x_std"], device=device)
        ws.y_scale = torch.tensor(z["y_scale"], device=device)
        return ws


def refine(it, u0, *, tol=1e-8, check_every=50, max_steps=400000):
    """Plain split-step from u0 (e.g. a prediction) to the steady state on an
    existing iterator.  Returns (u, info) as StrangSplittingIterator.steady_state."""
    return it.steady_state(u0=u0, tol=tol, check_every=check_every, max_steps=max_steps)


def solve_warm(it, ws, features, *, tol=1e-8, check_every=50, max_steps=400000, fallback=True):
    """Steady state of the iterator's problem started from the surrogate's
    band-limited prediction for `features` (one case, (1, C_in, N, N) on the
    device), refined by plain split-step.  If the refinement does not reach
    `tol` within max_steps and `fallback` is set, the case is rerun cold from
    the problem's own u0.  Returns (u, info); info["start"] is "warm" or
    "cold-fallback", info["ms"] the device time of everything run (inference
    and, on fallback, both refinements)."""
    torch = ws.torch
    dt = torch.complex128 if "128" in str(it.prob.u0[0].dtype) else torch.complex64
    ea, eb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ea.record()
    pred = ws.predict_fields(features, dtype=dt)
    eb.record()
    torch.cuda.synchronize()
    t_inf = ea.elapsed_time(eb)
    u, info = it.steady_state(u0=tuple(x[0] for x in pred), tol=tol, check_every=check_every, max_steps=max_steps)
    info = dict(info, start="warm", inference_ms=t_inf, ms=info["ms"] + t_inf)
    if info["residual"] <= tol or not fallback:
        return u, info
    spent = info["ms"]
    u, info = it.steady_state(tol=tol, check_every=check_every, max_steps=max_steps)
    return u, dict(info, start="cold-fallback", inference_ms=t_inf, ms=info["ms"] + spent)
