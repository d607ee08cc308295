"""FP64 CPU oracle for the Strang-splitting path of GeneralizedGrossPitaevskii.jl.

TEST INFRASTRUCTURE ONLY.  This module is the checker the HIP library is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path never
does.

It restates, in numpy complex128 arithmetic, what the reference computes for
``solve(prob, StrangSplitting(), tspan; dt, nsaves)``:

* ``resolve_fixed_timestepping``  <- src/fixed_time_stepping.jl:14-24
* ``solve`` (save loop, t += dt)   <- src/fixed_time_stepping.jl:26-54
* ``init`` (tables, pump buffers)  <- src/strang_splitting.jl:32-67
* ``_diffusion_step``              <- src/strang_splitting.jl:69-76
* ``_potential_pump_step``         <- src/strang_splitting.jl:78-84
* ``_step``                        <- src/strang_splitting.jl:86-90
* ``_muladd``                      <- src/kernels.jl:37-54 (muladd_kernel!)
* ``_mul``/``_add``/``_cis``       <- src/kernels.jl:9-27
* ``direct_grid``/``reciprocal_grid`` <- src/problem.jl:129-139
* ``get_exponential``              <- src/misc.jl:12-20
* ``evaluate_pump``                <- src/misc.jl:22-42
* FFT over the spatial dims        <- src/misc.jl:53-64 (FFTW plan_fft / inv)
* ``kspace_g2`` (n(k), G2(k,k'))    <- examples/truncated_wigner.jl:110-165
* ``windowed_correlation``, ``analytic_commutation`` <- test/windowed_ft.jl:31-59
* ``density_g2_x`` (g2(r) along x)  <- examples/truncated_wigner.jl:140-157 (ordering correction)

User closures follow the reference's calling conventions, evaluated
vectorised over the grid instead of point by point:

* ``dispersion(k, param)``, ``potential(r, param)``: ``k``/``r`` is a tuple of
  arrays broadcastable to the spatial grid.
* ``nonlinearity(u, param)``: ``u`` is a tuple of the M field arrays.
* ``pump(r, param, t)``.
* ``position_noise_func(u, r, param)``.

A closure's return value selects the reference's static-array form: a numpy
array or Python number is the scalar form, a Python ``list``/``tuple`` of
values is an ``SVector`` (elementwise products), and a list of lists is an
``SMatrix`` (matrix products, matrix exponential).

Noise: the reference draws ``randn!`` from Julia's RNG (src/misc.jl:44-51),
whose stream cannot be reproduced outside Julia.  The oracle instead draws
xi from the counter-based Philox4x32-10 generator with the same counter
layout as the HIP library (``philox_xi``), so both sides see identical
Wiener increments.  E|xi|^2 = 1 as in src/problem.jl:38-41.

Arrays use Julia's index order: axis 0 is the first spatial dimension; the
trailing axes beyond the spatial ones are batch (trajectory) axes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Callable, Optional, Sequence

import numpy as np

# --------------------------------------------------------------------------
# time stepping: src/fixed_time_stepping.jl:14-24
# --------------------------------------------------------------------------


def resolve_fixed_timestepping(dt, tspan, nsaves):
    """Return (dt_resolved, steps_per_save, t0) like fixed_time_stepping.jl:14-24.

    dT = (t1 - t0) / nsaves ; steps = ceil(dT / dt) ; dt' = dT / steps.
    ``t1 - t0`` is evaluated in the input type (integers exactly) before the
    conversion to Float64, as ``T(last(tspan) - first(tspan))`` does.
    """
    t0, t1 = tspan
    dT = float(t1 - t0) / nsaves
    steps = int(math.ceil(dT / float(dt)))
    return dT / steps, steps, float(t0)


# --------------------------------------------------------------------------
# grids: src/problem.jl:129-139
# --------------------------------------------------------------------------


def direct_grid_1d(L, N):
    """x_j = j * (L/N), j = 0..N-1 (StepRangeLen(zero(L), L/N, N))."""
    step = float(L) / N
    return np.arange(N, dtype=np.float64) * step


def reciprocal_grid_1d(L, N):
    """fftfreq(N, 2pi*N/L): k_j = m_j * ((2pi*N/L)/N), FFT order."""
    fs = (2.0 * math.pi) * N / float(L)
    mult = fs / N
    nneg = ((N - 1) >> 1) + 1
    m = np.arange(N, dtype=np.int64)
    m = np.where(m < nneg, m, m - N)
    return m.astype(np.float64) * mult


# --------------------------------------------------------------------------
# static-array forms and the identity algebra: src/kernels.jl:1-27
# --------------------------------------------------------------------------

AI = "AdditiveIdentity"
MI = "MultiplicativeIdentity"


@dataclass
class Form:
    kind: str  # "S" scalar, "V" vector (axis 0 = component), "M" matrix (axes 0,1)
    a: np.ndarray


def as_form(x, shape):
    """Interpret a closure result (see module docstring) broadcast to ``shape``."""
    if isinstance(x, (list, tuple)):
        if len(x) > 0 and all(isinstance(r, (list, tuple)) for r in x):
            M = len(x)
            arr = np.empty((M, len(x[0])) + tuple(shape), dtype=np.complex128)
            for i, row in enumerate(x):
                for j, v in enumerate(row):
                    arr[i, j] = np.broadcast_to(np.asarray(v, dtype=np.complex128), shape)
            return Form("M", arr)
        arr = np.empty((len(x),) + tuple(shape), dtype=np.complex128)
        for i, v in enumerate(x):
            arr[i] = np.broadcast_to(np.asarray(v, dtype=np.complex128), shape)
        return Form("V", arr)
    return Form("S", np.broadcast_to(np.asarray(x, dtype=np.complex128), shape).copy())


def _matvec(A, v):
    return np.einsum("ij...,j...->i...", A, v)


def _matmat(A, B):
    return np.einsum("ij...,jk...->ik...", A, B)


def _mul(x, y):
    """kernels.jl:9-15."""
    if isinstance(x, str) or isinstance(y, str):
        if x == MI:
            return y
        if y == MI:
            return x
        if y == AI:
            return AI
        if x == AI:  # generic x*y with an AdditiveIdentity: reference has no method
            raise TypeError("reference has no method _mul(AdditiveIdentity, x)")
    if not isinstance(x, Form):
        x = Form("S", np.asarray(x, dtype=np.complex128))
    if not isinstance(y, Form):
        y = Form("S", np.asarray(y, dtype=np.complex128))
    if x.kind == "M" and y.kind == "M":
        return Form("M", _matmat(x.a, y.a))
    if x.kind == "M" and y.kind == "V":
        return Form("V", _matvec(x.a, y.a))
    if x.kind == "V" and y.kind == "M":
        raise TypeError("SVector * SMatrix is not defined in the reference")
    kind = "M" if "M" in (x.kind, y.kind) else ("V" if "V" in (x.kind, y.kind) else "S")
    return Form(kind, x.a * y.a)


def _add(x, y):
    """kernels.jl:17-21 (x .+ y)."""
    if isinstance(x, str) and x == AI:
        return y
    if isinstance(y, str) and y == AI:
        return x
    kind = "V" if "V" in (x.kind, y.kind) else x.kind
    if "M" in (x.kind, y.kind):
        kind = "M"
    return Form(kind, x.a + y.a)


def expm_2x2(A):
    """exp(A) for a batch of 2x2 complex matrices, A shape (2,2,...).

    Closed form exp(A) = e^{tau} [cosh(s) I + sinh(s)/s (A - tau I)],
    tau = (a+d)/2, s = sqrt(((a-d)/2)^2 + b c), with a Taylor series of
    cosh and sinh(s)/s for small |s|.
    """
    a, b, c, d = A[0, 0], A[0, 1], A[1, 0], A[1, 1]
    tau = 0.5 * (a + d)
    h = 0.5 * (a - d)
    s2 = h * h + b * c
    s = np.sqrt(s2)
    small = np.abs(s) < 1e-3
    with np.errstate(divide="ignore", invalid="ignore"):
        ch = np.cosh(s)
        shs = np.where(small, 1.0, np.sinh(s) / np.where(small, 1.0, s))
    # series: cosh = 1 + s2/2 + s2^2/24 + s2^3/720 + s2^4/40320
    #         sinh(s)/s = 1 + s2/6 + s2^2/120 + s2^3/5040 + s2^4/362880
    ch_s = 1 + s2 * (1 / 2 + s2 * (1 / 24 + s2 * (1 / 720 + s2 * (1 / 40320 + s2 / 3628800))))
    sh_s = 1 + s2 * (1 / 6 + s2 * (1 / 120 + s2 * (1 / 5040 + s2 * (1 / 362880 + s2 / 39916800))))
    ch = np.where(small, ch_s, ch)
    shs = np.where(small, sh_s, shs)
    et = np.exp(tau)
    out = np.empty_like(A)
    out[0, 0] = et * (ch + shs * h)
    out[1, 1] = et * (ch - shs * h)
    out[0, 1] = et * shs * b
    out[1, 0] = et * shs * c
    return out


def expm_general(A):
    """exp(A) for a batch of MxM matrices (scipy's Pade scaling-and-squaring)."""
    import scipy.linalg

    M = A.shape[0]
    if M == 1:
        return np.exp(A)
    if M == 2:
        return expm_2x2(A)
    flat = A.reshape(M, M, -1)
    out = np.empty_like(flat)
    for i in range(flat.shape[2]):
        out[:, :, i] = scipy.linalg.expm(flat[:, :, i])
    return out.reshape(A.shape)


def _cis(x):
    """kernels.jl:23-25: cis(z) = exp(i z); elementwise for SVector, matrix
    exponential exp(iZ) for SMatrix (LinearAlgebra.cis)."""
    if isinstance(x, str):
        if x == AI:
            return MI
        raise TypeError(x)
    if x.kind == "M":
        return Form("M", expm_general(1j * x.a))
    return Form(x.kind, np.exp(1j * x.a))


# --------------------------------------------------------------------------
# Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the xi layout
# --------------------------------------------------------------------------

_PHILOX_M0 = np.uint64(0xD2511F53)
_PHILOX_M1 = np.uint64(0xCD9E8D57)
_PHILOX_W0 = 0x9E3779B9
_PHILOX_W1 = 0xBB67AE85
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32 with 10 rounds.  Inputs are uint32-valued arrays
    (any integer dtype) or ints; returns four uint32 arrays."""
    c0 = np.asarray(c0, dtype=np.uint64) & _MASK32
    c1 = np.asarray(c1, dtype=np.uint64) & _MASK32
    c2 = np.asarray(c2, dtype=np.uint64) & _MASK32
    c3 = np.asarray(c3, dtype=np.uint64) & _MASK32
    k0 = int(k0) & 0xFFFFFFFF
    k1 = int(k1) & 0xFFFFFFFF
    for _ in range(10):
        p0 = _PHILOX_M0 * c0
        p1 = _PHILOX_M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK32
        n0 = hi1 ^ c1 ^ np.uint64(k0)
        n2 = hi0 ^ c3 ^ np.uint64(k1)
        c0, c1, c2, c3 = n0, lo1, n2, lo0
        k0 = (k0 + _PHILOX_W0) & 0xFFFFFFFF
        k1 = (k1 + _PHILOX_W1) & 0xFFFFFFFF
    return (c0.astype(np.uint32), c1.astype(np.uint32), c2.astype(np.uint32), c3.astype(np.uint32))


def philox_xi(shape, comp, seed, call, complex_noise=True, pair_base=0):
    """xi for one component of one half-step, FP64.

    Counter layout (shared with the HIP library): the flat column-major index
    p over ``shape`` splits into line l = p // N0 and x = p % N0 along the
    first axis; pair index P = pair_base + l*(N0/2) + (x mod N0/2) and half
    h = x // (N0/2) (pair_base: an instance's offset / 2 inside a sweep).
    ctr = (P mod 2^32, comp + 16*(P >> 32), call mod 2^32, call >> 32),
    key = (seed mod 2^32, seed >> 32).  Words (x0,x1) feed h=0, (x2,x3) h=1.
    Box-Muller: u = (w + 0.5) 2^-32, z = sqrt(-log u_a) e^{2 pi i u_b};
    complex xi = z (E|xi|^2 = 1), real xi = sqrt(2) Re z.
    """
    shape = tuple(shape)
    N0 = shape[0]
    if N0 % 2:
        raise ValueError("first axis must be even for the paired Philox layout")
    half = N0 // 2
    n = int(np.prod(shape))
    p = np.arange(n, dtype=np.int64)
    line = p // N0
    x = p % N0
    h = x // half
    P = int(pair_base) + line * half + (x % half)
    call = int(call)
    w0, w1, w2, w3 = philox4x32_10(P & 0xFFFFFFFF, comp + 16 * (P >> 32), call & 0xFFFFFFFF,
                                   (call >> 32) & 0xFFFFFFFF, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    wa = np.where(h == 0, w0, w2).astype(np.float64)
    wb = np.where(h == 0, w1, w3).astype(np.float64)
    ua = (wa + 0.5) * 2.0 ** -32
    ub = (wb + 0.5) * 2.0 ** -32
    r = np.sqrt(-np.log(ua))
    th = 2.0 * math.pi * ub
    z = r * np.cos(th) + 1j * (r * np.sin(th))
    if not complex_noise:
        z = (math.sqrt(2.0) * z.real).astype(np.complex128)
    return z.reshape(shape, order="F")


# --------------------------------------------------------------------------
# the problem: src/problem.jl:97-123
# --------------------------------------------------------------------------


@dataclass
class Problem:
    u0: tuple
    lengths: tuple
    dispersion: Any = None
    potential: Any = None
    nonlinearity: Any = None
    pump: Any = None
    position_noise_func: Any = None
    noise_prototype: Any = None  # None, "complex" or "real"
    param: Any = None
    noise_seed: int = 0
    noise_call0: int = 0
    noise_pair_base: int = 0  # Philox pair index of the first point (an instance's offset / 2 in a sweep)

    def __post_init__(self):
        # problem.jl:112-115
        assert all(np.ndim(x) >= len(self.lengths) for x in self.u0)
        assert all(np.shape(x) == np.shape(self.u0[0]) for x in self.u0)
        self.u0 = tuple(np.asarray(x, dtype=np.complex128) for x in self.u0)


def _spatial_bcast(arr_1d, d, ndim_total):
    shape = [1] * ndim_total
    shape[d] = arr_1d.shape[0]
    return arr_1d.reshape(shape)


def _table(f, grid_axes, spatial_shape, ndim_total, *args):
    """grid_map! of misc.jl:1-10 over the spatial grid, reshaped to broadcast over
    trailing batch axes."""
    res = f(grid_axes, *args)
    form = as_form(res, spatial_shape)
    pad = (1,) * (ndim_total - len(spatial_shape))
    lead = form.a.shape[: form.a.ndim - len(spatial_shape)]
    return Form(form.kind, form.a.reshape(lead + tuple(spatial_shape) + pad))


class Iterator:
    """State of init(prob, StrangSplitting(), tspan; ...) (strang_splitting.jl:32-67)."""

    def __init__(self, prob: Problem, tspan, dt, nsaves, save_start=True):
        self.prob = prob
        self.save_start = bool(save_start)
        self.dt, self.steps_per_save, t0 = resolve_fixed_timestepping(dt, tspan, nsaves)
        self.nsaves = nsaves
        self.t0 = t0
        self.u = [x.copy() for x in prob.u0]
        shape = prob.u0[0].shape
        self.shape = shape
        N = len(prob.lengths)
        self.N = N
        self.spatial = shape[:N]
        nd = len(shape)
        self.nd = nd
        rg = [reciprocal_grid_1d(L, n) for L, n in zip(prob.lengths, self.spatial)]
        dg = [direct_grid_1d(L, n) for L, n in zip(prob.lengths, self.spatial)]
        self.k_axes = tuple(_spatial_bcast(g, d, N) for d, g in enumerate(rg))
        self.r_axes = tuple(_spatial_bcast(g, d, N) for d, g in enumerate(dg))
        self.rg, self.dg = rg, dg
        # get_exponential: misc.jl:12-20
        if prob.dispersion is None:
            self.exp_D = MI
        else:
            self.exp_D = _cis(_mul(-self.dt, _table(prob.dispersion, self.k_axes, self.spatial, nd, prob.param)))
        if prob.potential is None:
            self.exp_V = MI
        else:
            self.exp_V = _cis(_mul(-self.dt / 2, _table(prob.potential, self.r_axes, self.spatial, nd, prob.param)))
        # pump buffers: strang_splitting.jl:56-58
        self.F_next = AI
        self.F_now = AI
        if prob.pump is not None:
            self.F_next = self._eval_pump(t0)
        self.noise_call = int(prob.noise_call0)

    def _eval_pump(self, t):
        return _table(lambda r, p: self.prob.pump(r, p, t), self.r_axes, self.spatial, self.nd, self.prob.param)

    # misc.jl:39-42
    def _evaluate_pump(self, t):
        if self.prob.pump is None:
            return
        self.F_now = self.F_next
        self.F_next = self._eval_pump(t)

    def _sample_noise(self):
        """misc.jl:44-51 with Philox in place of randn!."""
        p = self.prob
        if p.position_noise_func is None or p.noise_prototype is None:
            return AI
        cplx = p.noise_prototype != "real"
        M = len(self.u)
        xi = np.stack([philox_xi(self.shape, c, p.noise_seed, self.noise_call, cplx, p.noise_pair_base)
                       for c in range(M)])
        self.noise_call += 1
        return Form("V", xi)

    def _muladd(self, dest, exp_dt, F_next, F_now, dt_, nonlinearity, noise_func, xi, point):
        """kernels.jl:37-54 evaluated on the whole array at once."""
        M = len(dest)
        fields = Form("V", np.stack(dest))
        field_tuple = tuple(dest)
        param = self.prob.param
        if noise_func is None or isinstance(xi, str):
            noise = AI
        else:
            eta = as_form(noise_func(field_tuple, point, param), self.shape)
            noise = _mul(-1j * math.sqrt(dt_), _mul(eta, xi))
        if nonlinearity is None:
            Gd = AI
        else:
            G = as_form(nonlinearity(field_tuple, param), self.shape)
            Gd = _mul(-dt_, G)
        exp_val = _mul(_cis(Gd), exp_dt)
        Fn = AI if isinstance(F_next, str) else _mul(dt_ / 2, F_next)
        Fo = AI if isinstance(F_now, str) else _mul(dt_ / 2, F_now)
        result = _add(_mul(exp_val, _add(fields, Fo)), Fn)
        result = _add(result, noise)
        a = result.a
        if result.kind == "S":
            a = np.broadcast_to(a, (M,) + a.shape)
        elif result.kind == "M":
            raise TypeError("matrix-valued update is not a field")
        for n in range(M):
            dest[n][...] = a[n]

    # strang_splitting.jl:69-76
    def _diffusion_step(self):
        axes = tuple(range(self.N))
        ft = [np.fft.fftn(x, axes=axes) for x in self.u]
        if not isinstance(self.exp_D, str):
            # The reference passes position_noise_func with dt = false (sqrt(0) = 0)
            # in k-space, so its noise term is exactly zero (strang_splitting.jl:73-74).
            self._muladd(ft, self.exp_D, AI, AI, 0.0, None, None, AI, None)
        self.u = [np.fft.ifftn(x, axes=axes) for x in ft]

    # strang_splitting.jl:78-84
    def _potential_pump_step(self, t, dt_):
        xi = self._sample_noise()
        self._evaluate_pump(t)
        # build_field_at(grid, K) indexes every grid axis with K[1] (kernels.jl:27,31-35):
        # the noise closure sees r = (x_i, y_i) for point (i, j).
        point = tuple(_spatial_bcast(g, 0, self.nd) for g in self.dg)
        self._muladd(self.u, self.exp_V, self.F_next, self.F_now, dt_, self.prob.nonlinearity,
                     self.prob.position_noise_func, xi, point)

    # strang_splitting.jl:86-90
    def step(self, t, dt):
        self._potential_pump_step(t + dt / 2, dt / 2)
        self._diffusion_step()
        self._potential_pump_step(t + dt, dt / 2)


def solve(prob: Problem, tspan, dt, nsaves, save_start=True, callback: Optional[Callable] = None):
    """solve! of fixed_time_stepping.jl:26-54.  Returns (ts, result) where
    result[c] has shape size(u0) + (nsaves + save_start,)."""
    it = Iterator(prob, tspan, dt, nsaves, save_start)
    nfr = nsaves + int(save_start)
    result = [np.empty(it.shape + (nfr,), dtype=np.complex128) for _ in it.u]
    if save_start:
        for c, x in enumerate(prob.u0):
            result[c][..., 0] = x
    ts = np.empty(nsaves + 1)
    ts[0] = it.t0
    t = it.t0
    for n in range(nsaves):
        for _ in range(it.steps_per_save):
            t += it.dt
            it.step(t, it.dt)
        for c in range(len(it.u)):
            result[c][..., n + int(save_start)] = it.u[c]
        ts[n + 1] = t
        if callback is not None:
            callback(n, it)
    return (ts if save_start else ts[1:]), result, it


# -------------------------------------------- momentum-space correlations
def kspace_g2(frames, commutator, pairwise=True):
    """n(k) and G2(k, k') of a 1D truncated-Wigner ensemble frame
    (examples/truncated_wigner.jl:110-165), in FFT order (no fftshift).

    frames: N x B complex array (trajectories along axis 1).  Follows the
    example's steps: ft = fft(frames, axis 0) / N; nks = mean_b |ft|^2 - c/2;
    G2[m, n] = mean_b f(ft[m, b], ft[n, b], m == n, c) with
    f(x, y, d, c) = |x|^2 |y|^2 - (1 + d) c/2 (|x|^2 + |y|^2 - c/2),
    evaluated pair by pair as the example's G2 loop does (m, n over k), or,
    with pairwise=False, as the matrix product a a^T / B minus the same
    corrections (for large N).  Returns (nks, G2) in FP64."""
    u = np.asarray(frames, dtype=np.complex128)
    N, B = u.shape
    ft = np.fft.fft(u, axis=0) / N
    a = np.abs(ft) ** 2
    c = float(commutator)
    nks = a.mean(axis=1) - c / 2
    if not pairwise:
        nbar = a.mean(axis=1)
        G2 = (a @ a.T) / B - (1 + np.eye(N)) * c / 2 * (nbar[:, None] + nbar[None, :] - c / 2)
        return nks, G2
    G2 = np.empty((N, N))
    for n in range(N):
        for m in range(N):
            d = 1.0 if m == n else 0.0
            G2[m, n] = np.mean(a[m] * a[n] - (1 + d) * c / 2 * (a[m] + a[n] - c / 2))
    return nks, G2


def windowed_correlation(frames, w1, w2):
    """The `correlation` of test/windowed_ft.jl:31-49: per trajectory (axis 1)
    ft_j = ifftshift(fft(fftshift(frames .* w_j, axis 0), axis 0), axis 0) and
    g1[i, j] = sum_b conj(ft_2[j, b]) ft_1[i, b] / (N B), with w_j the windows'
    values at the grid indices."""
    u = np.asarray(frames, dtype=np.complex128)
    N, B = u.shape
    fts = []
    for w in (w1, w2):
        x = u * np.asarray(w, dtype=np.complex128)[:, None]
        fts.append(np.fft.ifftshift(np.fft.fft(np.fft.fftshift(x, axes=0), axis=0), axes=0))
    return fts[0] @ fts[1].conj().T / (N * B)


def analytic_commutation(L, N, w1, w2):
    """The vacuum expectation of test/windowed_ft.jl:51-59 for windows given
    as functions of r: sum_r exp(i (k - k') r) w1(r) conj(w2(r)) / (r_N - r_1) / 2
    on r = -L/2 + j L/N and k = -pi/L + 2 pi m / L."""
    rs = -L / 2 + np.arange(N) * (L / N)
    ks = -np.pi / L + 2 * np.pi / L * np.arange(N)
    w = np.asarray(w1(rs), dtype=np.complex128) * np.conj(np.asarray(w2(rs), dtype=np.complex128))
    ph = np.exp(1j * (ks[:, None, None] - ks[None, :, None]) * rs[None, None, :])
    return (ph * w[None, None, :]).sum(axis=2) / (rs[-1] - rs[0]) / 2


def density_g2_x(frames, commutator):
    """Normal-ordered <n(x) n(x + r)> along axis 0 of an ensemble frame,
    averaged over x and every other axis, with the ordering correction of
    examples/truncated_wigner.jl:140-157 (f(x, y, delta, c) with
    delta = [r == 0]), evaluated pointwise before averaging:
        G2[r] = mean f(psi(x), psi(x + r), r == 0, c),  n_phys = mean |psi|^2 - c/2.
    Returns (G2, n_phys)."""
    a = np.abs(np.asarray(frames, dtype=np.complex128)) ** 2
    c = float(commutator)
    N = a.shape[0]
    G2 = np.empty(N)
    for r in range(N):
        b = np.roll(a, -r, axis=0)  # b[x] = a[x + r]
        d = 1.0 if r == 0 else 0.0
        G2[r] = np.mean(a * b - (1 + d) * c / 2 * (a + b - c / 2))
    return G2, float(a.mean() - c / 2)
