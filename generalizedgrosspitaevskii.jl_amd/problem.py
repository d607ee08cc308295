"""Host-side mirror of GeneralizedGrossPitaevskii.jl's user API over libgpe_hip.

Mirrors, name for name:
    GrossPitaevskiiProblem(u0, lengths; dispersion, potential, nonlinearity,
        pump, position_noise_func, momentum_noise_func, noise_prototype, param)
                                                      src/problem.jl:97-123
    StrangSplitting()                                 src/strang_splitting.jl:7
    solve(prob, alg, tspan; dt, nsaves, save_start, rng, ...)
                                                      src/fixed_time_stepping.jl:79-81
    init(prob, alg, tspan; ...) / solve!(iter)        src/strang_splitting.jl:32-67,
                                                      src/fixed_time_stepping.jl:26-54
    direct_grid / reciprocal_grid                     src/problem.jl:129-139

Arrays keep Julia's index order: axis 0 is the first spatial dimension and
axes beyond the spatial ones are batch (trajectory) axes.  Host arrays are
numpy arrays (stored column-major, like Julia's); device arrays are torch
tensors on a HIP device with column-major strides (see ``device_empty``).

User closures keep the reference's signatures and are evaluated vectorised
over the grid: ``dispersion(k, param)``, ``potential(r, param)`` and
``pump(r, param, t)`` receive a tuple of arrays broadcastable to the spatial
grid; ``nonlinearity(u, param)`` and ``position_noise_func(u, r, param)``
receive the fields.  A closure returns a scalar-form value (number or
array), an ``SVector`` (a Python list, elementwise semantics) or an
``SMatrix`` (a list of lists, matrix semantics), as the reference's
StaticArrays forms.

Because a closure cannot cross the C ABI, ``init`` reduces each term to the
descriptor the library executes (the coverage limit of this path):
    dispersion   any closure -> D(k) table, exp(-i dt D) formed on the GPU
    potential    scalar/SVector closure -> V(r) table
    nonlinearity G_c(u) = c_c + sum_j g_cj |u_j|^2, fitted by probing the
                 closure and verified; anything else raises NotImplementedError
    pump         F(r,t) = profile(r) * envelope(t), verified on probe times
    noise        constant real eta per component
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Any, Optional

import numpy as np

from . import _lib as L


# ---------------------------------------------------------------- static arrays
def SVector(*xs):
    """Vector form (elementwise semantics), like StaticArrays' SVector."""
    if len(xs) == 1 and isinstance(xs[0], (list, tuple)):
        xs = tuple(xs[0])
    return list(xs)


def SMatrix(rows):
    """Matrix form (matrix semantics), like StaticArrays' SMatrix; rows is a
    list of rows.  ``SMatrix(x)`` of a non-list is the 1x1 matrix [[x]]."""
    if not isinstance(rows, (list, tuple)):
        return [[rows]]
    return [list(r) for r in rows]


def _form(x, shape):
    """(kind, complex array) of a closure result: 'S' shape, 'V' (M,)+shape,
    'M' (M,M)+shape."""
    if isinstance(x, (list, tuple)):
        if len(x) and all(isinstance(r, (list, tuple)) for r in x):
            m = len(x)
            out = np.empty((m, len(x[0])) + tuple(shape), dtype=np.complex128)
            for i, row in enumerate(x):
                if len(row) != len(x[0]):
                    raise ValueError("ragged SMatrix")
                for j, v in enumerate(row):
                    out[i, j] = np.broadcast_to(np.asarray(v, dtype=np.complex128), shape)
            return "M", out
        out = np.empty((len(x),) + tuple(shape), dtype=np.complex128)
        for i, v in enumerate(x):
            out[i] = np.broadcast_to(np.asarray(v, dtype=np.complex128), shape)
        return "V", out
    return "S", np.array(np.broadcast_to(np.asarray(x, dtype=np.complex128), shape))


# ------------------------------------------------------------------ arrays
def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _colmajor_strides(shape):
    st, acc = [], 1
    for n in shape:
        st.append(acc)
        acc *= n
    return tuple(st)


def device_empty(shape, dtype="complex64", device=0):
    """A torch tensor on HIP device ``device`` with Julia (column-major) layout."""
    import torch

    tdt = {"complex64": torch.complex64, "complex128": torch.complex128}[str(np.dtype(dtype))]
    return torch.empty_strided(tuple(shape), _colmajor_strides(shape), dtype=tdt, device=f"cuda:{device}")


def _as_complex(x):
    """complex.(u0) of src/problem.jl:114."""
    if _is_torch(x):
        import torch

        if not torch.is_complex(x):
            x = x.to(torch.complex64 if x.dtype in (torch.float32, torch.float16, torch.bfloat16) else torch.complex128)
        return x
    a = np.asarray(x)
    if not np.iscomplexobj(a):
        a = a.astype(np.complex64 if a.dtype == np.float32 else np.complex128)
    return a


def _dtype_code(x):
    name = str(x.dtype).replace("torch.", "")
    if name == "complex64":
        return L.GPE_C64
    if name == "complex128":
        return L.GPE_C128
    raise TypeError(f"unsupported field dtype {name}")


# ---------------------------------------------------------------- problem
class GrossPitaevskiiProblem:
    """src/problem.jl:97-123."""

    def __init__(self, u0, lengths, *, dispersion=None, potential=None, nonlinearity=None, pump=None,
                 position_noise_func=None, momentum_noise_func=None, noise_prototype=None, param=None):
        u0 = tuple(u0)
        lengths = tuple(lengths)
        assert all(len(tuple(x.shape)) >= len(lengths) for x in u0)  # problem.jl:112
        assert all(tuple(x.shape) == tuple(u0[0].shape) for x in u0)  # problem.jl:113
        self.u0 = tuple(_as_complex(x) for x in u0)
        self.lengths = tuple(float(L_) for L_ in lengths)  # promote(lengths...)
        self.dispersion = dispersion
        self.potential = potential
        self.nonlinearity = nonlinearity
        self.pump = pump
        self.position_noise_func = position_noise_func
        self.momentum_noise_func = momentum_noise_func
        self.noise_prototype = noise_prototype
        self.param = param

    @property
    def N(self):
        return len(self.lengths)

    @property
    def M(self):
        return len(self.u0)

    def __repr__(self):
        return f"{self.N}D GrossPitaevskiiProblem"


class StrangSplitting:
    """src/strang_splitting.jl:7."""


def direct_grid(prob):
    """x_j = j * (L/N): src/problem.jl:129-133."""
    shape = tuple(prob.u0[0].shape)
    return tuple(np.arange(n, dtype=np.float64) * (L_ / n) for L_, n in zip(prob.lengths, shape))


def reciprocal_grid(prob):
    """fftfreq(N, 2pi N/L) in FFT order: src/problem.jl:135-139."""
    shape = tuple(prob.u0[0].shape)
    out = []
    for L_, n in zip(prob.lengths, shape):
        mult = ((2.0 * math.pi) * n / L_) / n
        nneg = ((n - 1) >> 1) + 1
        m = np.arange(n)
        out.append(np.where(m < nneg, m, m - n).astype(np.float64) * mult)
    return tuple(out)


def _bcast_axes(vecs):
    nd = len(vecs)
    out = []
    for d, v in enumerate(vecs):
        sh = [1] * nd
        sh[d] = v.shape[0]
        out.append(v.reshape(sh))
    return tuple(out)


# -------------------------------------------------------- descriptor reduction
def _interleave(a):
    """complex128 array -> contiguous float64 view, column-major flattening."""
    flat = np.asfortranarray(a, dtype=np.complex128).ravel(order="F")
    return np.ascontiguousarray(flat).view(np.float64)


def _table_desc(kind, arr, M, spatial, what):
    if kind == "S":
        return L.GPE_FORM_SCALAR, _interleave(arr)
    if kind == "V":
        if arr.shape[0] != M:
            raise ValueError(f"{what}: SVector of length {arr.shape[0]} for {M} components")
        return L.GPE_FORM_DIAG, _interleave(arr)
    if arr.shape[0] != M or arr.shape[1] != M:
        raise ValueError(f"{what}: SMatrix of size {arr.shape[:2]} for {M} components")
    return L.GPE_FORM_FULL, _interleave(arr)


def _probe_fields(M, K, seed):
    rng = np.random.default_rng(seed)
    mags = np.exp(rng.uniform(-2.0, 1.5, size=(M, K)))
    ph = rng.uniform(0, 2 * math.pi, size=(M, K))
    return tuple(mags[c] * np.exp(1j * ph[c]) for c in range(M))


class MatrixNonlinearity(NotImplementedError):
    """nonlinearity(u, param) returns a 2x2 SMatrix with off-diagonal terms."""


def fit_nonlinearity(f, M, param):
    """Reduce nonlinearity(u, param) to G_c = c_c + sum_j g_cj |u_j|^2.

    Returns (g[M][M], c[M]) complex or raises NotImplementedError."""
    K = 64
    u = _probe_fields(M, K, 0x6E4C)
    kind, arr = _form(f(u, param), (K,))
    if kind == "S":
        G = np.broadcast_to(arr, (M, K))
    elif kind == "V":
        if arr.shape[0] != M:
            raise ValueError(f"nonlinearity: SVector of length {arr.shape[0]} for {M} components")
        G = arr
    else:
        if arr.shape[:2] != (M, M):
            raise ValueError("nonlinearity: SMatrix size mismatch")
        off = max((np.max(np.abs(arr[i, j])) for i in range(M) for j in range(M) if i != j), default=0.0)
        if off != 0.0:
            raise MatrixNonlinearity("nonlinearity: non-diagonal SMatrix G(u) (use fit_nonlinearity_matrix)")
        G = np.stack([arr[i, i] for i in range(M)])
    A = np.column_stack([np.ones(K)] + [np.abs(u[j]) ** 2 for j in range(M)])
    g = np.zeros((M, M), dtype=np.complex128)
    c = np.zeros(M, dtype=np.complex128)
    for i in range(M):
        coef, *_ = np.linalg.lstsq(A.astype(np.complex128), G[i], rcond=None)
        fit = A @ coef
        scale = 1.0 + np.max(np.abs(G[i]))
        if np.max(np.abs(fit - G[i])) > 1e-11 * scale:
            raise NotImplementedError("nonlinearity: G(u) is not of the form c + sum_j g_j |u_j|^2 "
                                      "supported by libgpe_hip")
        c[i] = coef[0]
        g[i] = coef[1:]
    # snap coefficients that are zero up to fitting noise
    tiny = 1e-14 * (1.0 + np.max(np.abs(g)) + np.max(np.abs(c)))
    g[np.abs(g) < tiny] = 0
    c[np.abs(c) < tiny] = 0
    # confirm on fresh samples
    u2 = _probe_fields(M, K, 0x6E4D)
    kind2, arr2 = _form(f(u2, param), (K,))
    G2 = np.broadcast_to(arr2, (M, K)) if kind2 == "S" else (
        arr2 if kind2 == "V" else np.stack([arr2[i, i] for i in range(M)]))
    pred = c[:, None] + g @ np.stack([np.abs(x) ** 2 for x in u2])
    if np.max(np.abs(pred - G2)) > 1e-11 * (1.0 + np.max(np.abs(G2))):
        raise NotImplementedError("nonlinearity: probe verification failed")
    return g, c


def fit_nonlinearity_matrix(f, M, param):
    """Reduce an SMatrix nonlinearity(u, param) (src/kernels.jl:44, cis of a
    matrix) to G_ab = c_ab + sum_j g_abj |u_j|^2 for two components.

    Returns (g[a][b][j], c[a][b]) complex or raises NotImplementedError."""
    if M != 2:
        raise NotImplementedError("nonlinearity: matrix G(u) is supported for two components")
    K = 64
    coefs = []
    for seed in (0x6E4E, 0x6E4F):
        u = _probe_fields(M, K, seed)
        kind, arr = _form(f(u, param), (K,))
        if kind != "M" or arr.shape[:2] != (2, 2):
            raise NotImplementedError("nonlinearity: expected a 2x2 SMatrix")
        A = np.column_stack([np.ones(K)] + [np.abs(u[j]) ** 2 for j in range(M)]).astype(np.complex128)
        g = np.zeros((2, 2, 2), dtype=np.complex128)
        c = np.zeros((2, 2), dtype=np.complex128)
        for a in range(2):
            for b in range(2):
                coef, *_ = np.linalg.lstsq(A, arr[a, b], rcond=None)
                if np.max(np.abs(A @ coef - arr[a, b])) > 1e-11 * (1.0 + np.max(np.abs(arr[a, b]))):
                    raise NotImplementedError("nonlinearity: matrix entry is not c + sum_j g_j |u_j|^2")
                c[a, b] = coef[0]
                g[a, b] = coef[1:]
        tiny = 1e-14 * (1.0 + np.max(np.abs(g)) + np.max(np.abs(c)))
        g[np.abs(g) < tiny] = 0
        c[np.abs(c) < tiny] = 0
        coefs.append((g, c))
    (g, c), (g2, c2) = coefs
    if np.max(np.abs(g - g2)) > 1e-10 * (1 + np.max(np.abs(g))) or np.max(np.abs(c - c2)) > 1e-10 * (1 + np.max(np.abs(c))):
        raise NotImplementedError("nonlinearity: probe verification failed")
    return g, c


def noise_eta(f, M, param, spatial_axes):
    """Reduce position_noise_func(u, r, param) to a constant real eta per component."""
    vals = []
    for seed in (0x7A1, 0x7A2):
        u = _probe_fields(M, 8, seed)
        rng = np.random.default_rng(seed)
        r = tuple(rng.uniform(0, 1, size=8) * (ax.max() + 1) for ax in spatial_axes)
        kind, arr = _form(f(u, r, param), (8,))
        eta = np.broadcast_to(arr, (M, 8)) if kind == "S" else arr
        if kind == "M" or eta.shape[0] != M:
            raise NotImplementedError("position_noise_func: matrix form not supported")
        vals.append(np.array(eta))
    a, b = vals
    if np.max(np.abs(a - a[:, :1])) != 0 or np.max(np.abs(a - b)) != 0:
        raise NotImplementedError("position_noise_func: only constant eta is supported by libgpe_hip")
    eta = a[:, 0]
    if np.any(eta.imag != 0):
        raise NotImplementedError("position_noise_func: complex eta is not supported by libgpe_hip")
    return eta.real.copy()


# Host memory a sampled pump may take (FP64 planes of every component).
MAX_PUMP_SAMPLE_BYTES = 16 << 30


def pump_samples(prob, r_axes, spatial, times, table=None):
    """F(r, t) on the grid at every pump time, for a pump that does not factor
    into profile(r) * envelope(t): what evaluate_pump! (src/misc.jl:34-42)
    computes before each half-step, done once per solve.

    Returns (GPE_PUMP_SAMPLES, float64 view of M blocks of len(times) planes
    (each column-major over the grid), None)."""
    M = prob.M
    nsp = int(np.prod(spatial))
    nbytes = M * len(times) * nsp * 16
    if nbytes > MAX_PUMP_SAMPLE_BYTES:
        raise NotImplementedError(
            f"pump: F(r,t) does not factor and its {len(times)} samples of {nsp} points need {nbytes} bytes "
            f"(limit {MAX_PUMP_SAMPLE_BYTES}); not supported by libgpe_hip")
    out = np.empty((M, len(times), nsp), dtype=np.complex128)
    for j, t in enumerate(times):
        F = table(t)
        for c in range(M):
            out[c, j] = np.asfortranarray(F[c]).ravel(order="F")
    return L.GPE_PUMP_SAMPLES, np.ascontiguousarray(out).reshape(-1).view(np.float64), None


def pump_descriptor(prob, r_axes, spatial, times, allow_samples=True):
    """Reduce pump(r, param, t) to profile(r) * envelope(t), or, when it does
    not factor (and allow_samples), to its samples at every pump time.

    Returns (kind, profile float64 array, envelope float64 array or None)."""
    M = prob.M
    f = prob.pump
    nt = len(times)
    probe_idx = sorted(set(np.linspace(0, nt - 1, min(nt, 17)).astype(int).tolist()))

    def table(t):
        kind, arr = _form(f(r_axes, prob.param, t), spatial)
        if kind == "S":
            return np.broadcast_to(arr, (M,) + spatial).copy()
        if kind == "M":
            if arr.shape[1] != 1:
                raise NotImplementedError("pump: SMatrix pump must be M x 1")
            arr = arr[:, 0]
        if arr.shape[0] != M:
            raise ValueError(f"pump: SVector of length {arr.shape[0]} for {M} components")
        return arr

    tabs = {i: table(times[i]) for i in probe_idx}
    ref = max(probe_idx, key=lambda i: np.linalg.norm(tabs[i]))
    P = tabs[ref]
    normP = np.linalg.norm(P)
    if normP == 0:
        return L.GPE_PUMP_NONE, None, None
    # the point where the profile is largest carries the envelope
    c_star, *pt = np.unravel_index(np.argmax(np.abs(P)), P.shape)
    r_star = tuple(float(r_axes[d].ravel()[pt[d]]) for d in range(len(spatial)))
    p_star = P[(c_star, *pt)]

    def env_point(t):
        kind, arr = _form(f(r_star, prob.param, t), ())
        if kind == "S":
            v = arr
        elif kind == "V":
            v = arr[c_star]
        else:
            v = arr[c_star, 0]
        return complex(v)

    env = None
    try:  # closures written with numpy broadcast over an array of times
        kind, arr = _form(f(r_star, prob.param, np.asarray(times)), (nt,))
        arr = arr if kind == "S" else (arr[c_star] if kind == "V" else arr[c_star, 0])
        env = np.array(arr, dtype=np.complex128) / p_star
        # spot-check the vectorised evaluation against scalar calls
        for i in (0, nt // 2, nt - 1):
            if env[i] != env_point(times[i]) / p_star:
                env = None
                break
    except Exception:
        env = None
    if env is None:
        env = np.array([env_point(t) for t in times], dtype=np.complex128) / p_star
    # verify separability on the probe times and a few more
    rng = np.random.default_rng(0xF00)
    check = set(probe_idx) | set(rng.integers(0, nt, size=min(nt, 8)).tolist())
    for i in check:
        Fi = tabs[i] if i in tabs else table(times[i])
        err = np.linalg.norm(Fi - env[i] * P)
        if err > 1e-12 * max(np.linalg.norm(Fi), normP * abs(env[i]), 1e-300):
            if not allow_samples:
                raise NotImplementedError("pump: F(r,t) is not profile(r)*envelope(t); not supported here")
            return pump_samples(prob, r_axes, spatial, times, table)
    static = bool(np.all(env == env[0]))
    if static:
        P = P * env[0]
        env = None
    # spatially separable profile (product of per-dimension factors)?
    kind = L.GPE_PUMP_TABLE
    prof = P
    if len(spatial) == 1:
        kind = L.GPE_PUMP_SEPARABLE
        prof = P
    elif len(spatial) == 2:
        factors = []
        ok = True
        for c in range(M):
            Pc = P[c]
            if not np.any(Pc):
                factors.append((np.zeros(spatial[0], complex), np.zeros(spatial[1], complex)))
                continue
            i0, j0 = np.unravel_index(np.argmax(np.abs(Pc)), Pc.shape)
            fx = Pc[:, j0].copy()
            fy = Pc[i0, :] / Pc[i0, j0]
            if np.max(np.abs(np.outer(fx, fy) - Pc)) > 1e-14 * np.max(np.abs(Pc)):
                ok = False
                break
            factors.append((fx, fy))
        if ok:
            kind = L.GPE_PUMP_SEPARABLE
            prof = np.concatenate([np.concatenate([fx, fy]) for fx, fy in factors])
    if kind == L.GPE_PUMP_TABLE:
        prof = np.concatenate([np.asfortranarray(P[c]).ravel(order="F") for c in range(M)])
    elif len(spatial) == 1:
        prof = np.concatenate([P[c] for c in range(M)])
    prof = np.ascontiguousarray(prof, dtype=np.complex128).view(np.float64)
    envf = None if env is None else np.ascontiguousarray(env).view(np.float64)
    return kind, prof, envf


# --------------------------------------------------------------- rng
class PhiloxRNG:
    """Counter-based stream for the Wiener increments: key `seed`, the next
    half-step draw uses counter `counter` (advanced by 2 per step)."""

    def __init__(self, seed=0, counter=0):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.counter = int(counter)


_default_rng = PhiloxRNG(0x5EED)


def _rng_of(rng):
    if rng is None:
        return _default_rng
    if isinstance(rng, PhiloxRNG):
        return rng
    if isinstance(rng, (int, np.integer)):
        return PhiloxRNG(int(rng))
    raise TypeError("rng must be None, an int seed or a PhiloxRNG")


# --------------------------------------------------------------- time stepping
def resolve_fixed_timestepping(dt, tspan, nsaves):
    """src/fixed_time_stepping.jl:14-24, through the library."""
    lib = L.load()
    out = ctypes.c_double()
    sps = ctypes.c_int64()
    L.check(lib.gpe_resolve_timestepping(float(dt), float(tspan[0]), float(tspan[1]), int(nsaves),
                                         ctypes.byref(out), ctypes.byref(sps)))
    return out.value, sps.value


def pump_times(tspan, dt, nsaves):
    lib = L.load()
    _, sps = resolve_fixed_timestepping(dt, tspan, nsaves)
    n = 1 + 2 * sps * nsaves
    times = np.empty(n, dtype=np.float64)
    L.check(lib.gpe_pump_times(float(tspan[0]), float(tspan[1]), float(dt), int(nsaves),
                               times.ctypes.data, n))
    return times


# --------------------------------------------------------------- iterator
class StrangSplittingIterator:
    """What init(prob, StrangSplitting(), tspan; ...) returns
    (src/strang_splitting.jl:9-30): the device handle with every table set."""

    def __init__(self, prob, tspan, dt, nsaves, save_start=True, rng=None):
        lib = L.load()
        if prob.noise_prototype is not None and prob.position_noise_func is None:
            # _mul(AdditiveIdentity, SVector) has no method in the reference (src/kernels.jl:9,42)
            raise TypeError("noise_prototype without position_noise_func (no method in the reference)")
        self.prob = prob
        self.tspan = (float(tspan[0]), float(tspan[1]))
        self.nsaves = int(nsaves)
        self.save_start = bool(save_start)
        self.dt_in = float(dt)
        self.dt, self.steps_per_save = resolve_fixed_timestepping(dt, tspan, nsaves)
        self.rng = _rng_of(rng)
        u0 = prob.u0
        shape = tuple(u0[0].shape)
        N, M = prob.N, prob.M
        if N > 3:
            raise NotImplementedError("libgpe_hip supports 1D, 2D and 3D grids")
        spatial = shape[:N]
        self.shape, self.spatial = shape, spatial
        batch = int(np.prod(shape[N:])) if len(shape) > N else 1
        self.on_device = _is_torch(u0[0]) and u0[0].is_cuda
        dev = u0[0].device.index if self.on_device else 0
        desc = L.ProblemDesc()
        desc.ndim = N
        desc.ncomp = M
        for d in range(3):
            desc.n[d] = spatial[d] if d < N else 1
            desc.L[d] = prob.lengths[d] if d < N else 1.0
        desc.batch = batch
        desc.dtype = _dtype_code(u0[0])
        desc.device = dev or 0
        h = ctypes.c_void_p()
        L.check(lib.gpe_create(ctypes.byref(desc), ctypes.byref(h)))
        self.h = h
        self._lib = lib
        try:
            self._set_terms()
        except BaseException:
            lib.gpe_destroy(h)
            self.h = None
            raise

    TERMS = ("dispersion", "potential", "nonlinearity", "pump", "noise")

    def set_param(self, param, terms=TERMS):
        """Replace prob.param and re-evaluate the listed terms (closures of
        param) on this handle.  Tables of an unchanged size keep their device
        memory, so a parameter scan over one handle allocates nothing per
        point.  No reference counterpart (the reference builds a new problem
        and iterator per parameter set)."""
        unknown = set(terms) - set(self.TERMS)
        if unknown:
            raise ValueError(f"set_param: unknown terms {sorted(unknown)}")
        old = self.prob
        new = GrossPitaevskiiProblem(old.u0, old.lengths, dispersion=old.dispersion, potential=old.potential,
                                     nonlinearity=old.nonlinearity, pump=old.pump,
                                     position_noise_func=old.position_noise_func,
                                     momentum_noise_func=old.momentum_noise_func,
                                     noise_prototype=old.noise_prototype, param=param)
        self.prob = new
        self._set_terms(terms)

    SCHEDULES = {"auto": L.GPE_SCHEDULE_AUTO, "launches": L.GPE_SCHEDULE_LAUNCHES, "graph": L.GPE_SCHEDULE_GRAPH}

    def set_schedule(self, schedule):
        """How solve issues the 2D step loop: "launches" (two kernel launches
        per step), "graph" (the steps of a save block built once into a HIP
        graph and replayed per block; no noise, static envelope) or "auto"
        (the graph from 8 save blocks on).  The results are identical bit for
        bit.  No reference counterpart (the reference steps on the host)."""
        if schedule not in self.SCHEDULES:
            raise ValueError(f"schedule must be one of {sorted(self.SCHEDULES)}")
        L.check(self._lib.gpe_set_schedule(self.h, self.SCHEDULES[schedule]))

    PLANNERS = {"estimate": L.GPE_PLAN_ESTIMATE, "measure": L.GPE_PLAN_MEASURE, "auto": L.GPE_PLAN_AUTO}

    def set_planner(self, planner):
        """Launch-shape planner of the 2D passes, the counterpart of the FFTW
        planner flags behind get_fft_plans (src/misc.jl:53-58): "estimate"
        (built-in shapes), "measure" (the next solve first times every valid
        shape pair on this handle and keeps the fastest) or "auto" (the
        default: a plan measured before in this process for the same shape
        and terms, else "measure" for runs of >= 1000 steps, else
        "estimate").  Results are identical under every plan."""
        if planner not in self.PLANNERS:
            raise ValueError(f"planner must be one of {sorted(self.PLANNERS)}")
        L.check(self._lib.gpe_set_planner(self.h, self.PLANNERS[planner]))

    def set_plan(self, x_rows=0, y_cols=0):
        """Fix the rows per workgroup of the row pass and the columns per
        workgroup of the column pass (0: built-in), e.g. a measured plan."""
        L.check(self._lib.gpe_set_plan(self.h, int(x_rows), int(y_cols)))

    @property
    def plan(self):
        """dict(x_rows, y_cols) in use, and the last measurement's mean us per
        step of every (x_rows, y_cols) pair timed ({} before any)."""
        xr, yc = ctypes.c_int32(), ctypes.c_int32()
        cand = (ctypes.c_int32 * 12)()
        us = np.zeros(6)
        L.check(self._lib.gpe_get_plan(self.h, ctypes.byref(xr), ctypes.byref(yc), cand,
                                       us.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        meas = {(int(cand[2 * i]), int(cand[2 * i + 1])): float(us[i]) for i in range(6) if cand[2 * i]}
        return dict(x_rows=xr.value, y_cols=yc.value, measured_us=meas)

    @property
    def last_schedule(self):
        """The schedule the last 2D solve ran with ("launches" or "graph")."""
        v = ctypes.c_int32()
        L.check(self._lib.gpe_get_schedule(self.h, ctypes.byref(v)))
        return {code: name for name, code in self.SCHEDULES.items()}[v.value]

    def _set_terms(self, which=TERMS):
        prob, lib, h = self.prob, self._lib, self.h
        M, spatial = prob.M, self.spatial
        kax = _bcast_axes(reciprocal_grid(prob))
        rax = _bcast_axes(direct_grid(prob))
        self._keep = []  # the library copies every table during the set call
        if "dispersion" not in which:
            pass
        elif prob.dispersion is not None:
            kind, arr = _form(prob.dispersion(kax, prob.param), spatial)
            form, tab = _table_desc(kind, arr, M, spatial, "dispersion")
            self._keep.append(tab)
            L.check(lib.gpe_set_dispersion(h, form, tab.ctypes.data))
        else:
            L.check(lib.gpe_set_dispersion(h, L.GPE_FORM_NONE, None))
        if "potential" in which and prob.potential is not None:
            kind, arr = _form(prob.potential(rax, prob.param), spatial)
            if kind == "M":
                raise NotImplementedError("potential: SMatrix V(r) is not supported by libgpe_hip")
            form, tab = _table_desc(kind, arr, M, spatial, "potential")
            self._keep.append(tab)
            L.check(lib.gpe_set_potential(h, form, tab.ctypes.data))
        if "nonlinearity" in which and prob.nonlinearity is not None:
            try:
                g, c = fit_nonlinearity(prob.nonlinearity, M, prob.param)
            except MatrixNonlinearity:
                gm, cm = fit_nonlinearity_matrix(prob.nonlinearity, M, prob.param)
                # layout: g[(a + 2b) + 4j], c[a + 2b]
                gbuf = np.ascontiguousarray([gm[a, b, j] for j in range(2) for b in range(2) for a in range(2)],
                                            dtype=np.complex128).view(np.float64)
                cbuf = np.ascontiguousarray([cm[a, b] for b in range(2) for a in range(2)],
                                            dtype=np.complex128).view(np.float64)
                self._keep += [gbuf, cbuf]
                L.check(lib.gpe_set_nonlinearity_matrix(h, gbuf.ctypes.data, cbuf.ctypes.data))
                g = None
            if g is not None:
                # column-major g[i + M j]
                gbuf = np.ascontiguousarray(g.T.ravel()).view(np.float64)
                cbuf = np.ascontiguousarray(c).view(np.float64)
                self._keep += [gbuf, cbuf]
                L.check(lib.gpe_set_nonlinearity(h, gbuf.ctypes.data, cbuf.ctypes.data))
        if "pump" in which and prob.pump is not None:
            times = pump_times(self.tspan, self.dt_in, self.nsaves)
            kind, prof, env = pump_descriptor(prob, rax, spatial, times)
            if kind != L.GPE_PUMP_NONE:
                self._keep += [prof, env]
                # samples: one plane per pump time; else envelope samples (0: static)
                n_env = len(times) if kind == L.GPE_PUMP_SAMPLES else (0 if env is None else len(env) // 2)
                L.check(lib.gpe_set_pump(h, kind, prof.ctypes.data, None if env is None else env.ctypes.data,
                                         n_env))
        if "noise" in which and prob.noise_prototype is not None:
            proto = prob.noise_prototype
            first = proto[0] if isinstance(proto, (tuple, list)) else proto
            cplx = np.iscomplexobj(np.empty(0, dtype=str(first.dtype).replace("torch.", "")))
            eta = noise_eta(prob.position_noise_func, M, prob.param, direct_grid(prob))
            self._keep.append(eta)
            L.check(lib.gpe_set_noise(h, L.GPE_NOISE_COMPLEX if cplx else L.GPE_NOISE_REAL,
                                      eta.ctypes.data, self.rng.seed, self.rng.counter))

    def close(self):
        if getattr(self, "h", None):
            self._lib.gpe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ solve!
    def solve(self, result=None, save_frames=True, ensemble_moments=False):
        """solve!(iter) (src/fixed_time_stepping.jl:26-54): returns (ts, result).

        Extensions: ``save_frames=False`` returns result = None and writes no
        frames; ``ensemble_moments=True`` also returns (n, n2), the batch means
        of |u_c|^2 and |u_c|^4 of every saved frame (FP64, reduced on the
        device), one array of shape spatial + (nframes,) per component."""
        prob, lib = self.prob, self._lib
        nfr = self.nsaves + int(self.save_start)
        rshape = self.shape + (nfr,)
        if self.on_device:
            u0 = [x if _contig_colmajor(x) else _to_colmajor(x) for x in prob.u0]
            if result is None and save_frames:
                result = tuple(device_empty(rshape, str(x.dtype).replace("torch.", ""), x.device.index) for x in u0)
            uptr = [x.data_ptr() for x in u0]
            rptr = None if result is None else [x.data_ptr() for x in result]
            _order_after_torch(u0[0])
        else:
            u0 = [np.asfortranarray(x) for x in prob.u0]
            if result is None and save_frames:
                result = tuple(np.empty(rshape, dtype=x.dtype, order="F") for x in u0)
            uptr = [x.ctypes.data for x in u0]
            rptr = None if result is None else [x.ctypes.data for x in result]
        moments = None
        if ensemble_moments:
            mshape = (prob.M,) + tuple(self.spatial) + (nfr,)
            n1 = np.empty((prob.M, int(np.prod(self.spatial)) * nfr), dtype=np.float64)
            n2 = np.empty_like(n1)
            L.check(lib.gpe_set_observables(self.h, n1.ctypes.data, n2.ctypes.data))
        else:
            L.check(lib.gpe_set_observables(self.h, None, None))
        ts = np.empty(nfr, dtype=np.float64)
        try:
            L.check(lib.gpe_solve(self.h, L.ptr_array(uptr), self.tspan[0], self.tspan[1], self.dt_in, self.nsaves,
                                  int(self.save_start), None if rptr is None else L.ptr_array(rptr),
                                  ts.ctypes.data))
        finally:
            if ensemble_moments:
                lib.gpe_set_observables(self.h, None, None)
        if prob.noise_prototype is not None:
            self.rng.counter += 2 * self.steps_per_save * self.nsaves
        self._u0_keep = u0
        out_res = None if result is None else tuple(result)
        if ensemble_moments:
            shape_f = tuple(self.spatial) + (nfr,)
            moments = (tuple(n1[c].reshape(shape_f, order="F") for c in range(prob.M)),
                       tuple(n2[c].reshape(shape_f, order="F") for c in range(prob.M)))
            return ts, out_res, moments
        return ts, out_res

    def last_ms(self):
        ms = ctypes.c_float()
        L.check(self._lib.gpe_last_solve_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def _ensemble_frame(self, frame, who):
        """(pointer, keep-alive) of one 1D ensemble frame in this handle's
        dtype and layout."""
        if self.prob.N != 1:
            raise NotImplementedError(f"{who}: 1D ensembles only")
        n = self.spatial[0]
        want = str(self.prob.u0[0].dtype).replace("torch.", "")
        if _is_torch(frame) and frame.is_cuda:
            f = frame if (_contig_colmajor(frame) and str(frame.dtype).replace("torch.", "") == want) \
                else _to_colmajor(frame, want)
            ptr = f.data_ptr()
            _order_after_torch(f)
        else:
            f = np.asfortranarray(np.asarray(frame).astype(want))
            ptr = f.ctypes.data
        if int(np.prod(tuple(f.shape))) != n * int(np.prod(self.shape[1:] or (1,))):
            raise ValueError(f"{who}: frame of shape {tuple(f.shape)} does not match {self.shape}")
        return ptr, f

    def kspace_correlations(self, frame, commutator):
        """n(k) and G2(k, k') of one 1D ensemble frame, computed on the
        device as examples/truncated_wigner.jl:110-165 does on the host:
        ft = fft(u_b)/N per trajectory, nks = mean |ft|^2 - c/2,
        G2[k, k'] = mean_b f(ft(k), ft(k'), k == k', c).  frame: (N, batch)
        field of this handle's dtype (numpy or device tensor, e.g.
        sol[0][..., f]); commutator: 1/L in the example.  Returns (nks, G2)
        in FFT order (apply np.fft.fftshift for the example's axes)."""
        ptr, keep = self._ensemble_frame(frame, "kspace_correlations")
        n = self.spatial[0]
        nk = np.empty(n)
        G2 = np.empty((n, n), order="F")
        L.check(self._lib.gpe_kspace_g2(self.h, ptr, float(commutator), nk.ctypes.data, G2.ctypes.data))
        del keep
        return nk, np.ascontiguousarray(G2)

    def windowed_correlation(self, frame, w1, w2):
        """The windowed first-order correlation of test/windowed_ft.jl:31-49
        on the device: ft_j = ifftshift(fft(fftshift(w_j .* u_b))) per
        trajectory and g1[i, j] = sum_b conj(ft_2[j, b]) ft_1[i, b] / (N batch).
        w1, w2: the windows' values at the grid indices (length N, real or
        complex).  Returns the N x N complex matrix in the reference's layout."""
        ptr, keep = self._ensemble_frame(frame, "windowed_correlation")
        n = self.spatial[0]
        wa = np.ascontiguousarray(np.broadcast_to(np.asarray(w1, dtype=np.complex128), (n,)))
        wb = np.ascontiguousarray(np.broadcast_to(np.asarray(w2, dtype=np.complex128), (n,)))
        g = np.empty((n, n), dtype=np.complex128, order="F")
        L.check(self._lib.gpe_windowed_g1(self.h, ptr, wa.ctypes.data, wb.ctypes.data, g.ctypes.data))
        del keep
        # the library returns FFT order; the shift pair maps index k to
        # (k + N/2) mod N with the sign (-1)^k of the half-length shift
        idx = (np.arange(n) + n // 2) % n
        sgn = np.where(idx % 2 == 0, 1.0, -1.0)
        return g[np.ix_(idx, idx)] * sgn[:, None] * sgn[None, :]

    def density_correlation_x(self, frame, commutator):
        """g2(r) along dimension 0 of one ensemble frame (1D or 2D, one
        component, e.g. sol[c][..., f]), reduced on the device: returns
        (G2, n_phys, g2) with G2[r] the normal-ordered <n(x) n(x + r)>
        averaged over x, the other dimensions and the trajectories, n_phys the
        mean physical density and g2 = G2 / n_phys^2 (the homogeneous-state
        normalisation).  commutator: per grid point (1/dA in 2D, 1/dx in 1D);
        the ordering correction is that of examples/truncated_wigner.jl:140-157."""
        want = str(self.prob.u0[0].dtype).replace("torch.", "")
        if _is_torch(frame) and frame.is_cuda:
            f = frame if (_contig_colmajor(frame) and str(frame.dtype).replace("torch.", "") == want) \
                else _to_colmajor(frame, want)
            ptr = f.data_ptr()
            _order_after_torch(f)
        else:
            f = np.asfortranarray(np.asarray(frame).astype(want))
            ptr = f.ctypes.data
        if int(np.prod(tuple(f.shape))) != int(np.prod(self.shape)):
            raise ValueError(f"density_correlation_x: frame of shape {tuple(f.shape)} does not match {self.shape}")
        n = self.spatial[0]
        G2 = np.empty(n)
        nphys = np.empty(1)
        L.check(self._lib.gpe_density_g2_x(self.h, ptr, float(commutator), G2.ctypes.data, nphys.ctypes.data))
        del f
        with np.errstate(divide="ignore", invalid="ignore"):
            g2 = G2 / nphys[0] ** 2
        return G2, float(nphys[0]), g2

    def get_field(self):
        """The handle's current real-space field (host numpy arrays)."""
        out = tuple(np.empty(self.shape, dtype=str(x.dtype).replace("torch.", ""), order="F") for x in self.prob.u0)
        L.check(self._lib.gpe_get_field(self.h, L.ptr_array([x.ctypes.data for x in out])))
        return out

    def steady_state(self, u0=None, tol=1e-8, check_every=50, max_steps=200000):
        """Step with this iterator's resolved dt from u0 (default prob.u0) until
        ||u_{n+1} - u_n|| / ||u_{n+1}|| <= tol (FP64, every check_every steps).
        Returns (u, info) with info = dict(steps, residual, history, ms)."""
        lib = self._lib
        src = self.prob.u0 if u0 is None else tuple(u0)
        if _is_torch(src[0]) and src[0].is_cuda:  # device start field (either kind of iterator)
            want = self.prob.u0[0].dtype
            want = str(want).replace("torch.", "")
            u = [x if (_contig_colmajor(x) and str(x.dtype).replace("torch.", "") == want)
                 else _to_colmajor(x, want) for x in src]
            ptrs = [x.data_ptr() for x in u]
            _order_after_torch(u[0])
        else:
            u = [np.asfortranarray(np.asarray(x).astype(str(self.prob.u0[0].dtype))) for x in src]
            ptrs = [x.ctypes.data for x in u]
        steps = ctypes.c_int64()
        res = ctypes.c_double()
        nh = ctypes.c_int64()
        ms = ctypes.c_float()
        maxh = max_steps // check_every + 1
        hist = np.zeros(maxh)
        L.check(lib.gpe_steady_state(self.h, L.ptr_array(ptrs), self.dt, float(tol), int(check_every), int(max_steps),
                                     ctypes.byref(steps), ctypes.byref(res), hist.ctypes.data, maxh, ctypes.byref(nh),
                                     ctypes.byref(ms)))
        return self.get_field(), dict(steps=steps.value, residual=res.value, history=hist[:min(nh.value, maxh)],
                                      ms=ms.value)


def wigner_density_g2(n, n2, commutator):
    """Physical density and equal-point g2 from truncated-Wigner moments.

    n, n2: ensemble means of |psi|^2 and |psi|^4 per point (as returned by
    solve(..., ensemble_moments=True)); commutator: [psi, psi^dagger] of the
    discretised field (1/dA on a grid of cell area dA).  Symmetric ordering
    is undone as in examples/truncated_wigner.jl:110-165 (there for psi(k)
    pairs with delta_{k,k'}), here for the same point (delta = 1):
        n_phys = <|psi|^2> - c/2
        G2     = <|psi|^4> - 2 c <|psi|^2> + c^2/2
        g2     = G2 / n_phys^2
    Returns (n_phys, g2)."""
    n = np.asarray(n, dtype=np.float64)
    n2 = np.asarray(n2, dtype=np.float64)
    c = float(commutator)
    n_phys = n - c / 2
    G2 = n2 - 2 * c * n + c * c / 2
    with np.errstate(divide="ignore", invalid="ignore"):
        g2 = G2 / (n_phys * n_phys)
    return n_phys, g2


def steady_state(prob, alg, *, dt, tol=1e-8, check_every=50, max_steps=200000, u0=None):
    """Run prob (static pump) from u0 (default prob.u0) until the relative
    step-to-step change ||u_{n+1} - u_n|| / ||u_{n+1}|| <= tol.  No reference
    counterpart: the reference integrates a fixed tspan.  Returns (u, info)."""
    it = init(prob, alg, (0.0, float(dt)), dt=dt, nsaves=1)
    try:
        return it.steady_state(u0=u0, tol=tol, check_every=check_every, max_steps=max_steps)
    finally:
        it.close()


def _contig_colmajor(x):
    return tuple(x.stride()) == _colmajor_strides(tuple(x.shape))


def _order_after_torch(x):
    """The library runs on its own non-blocking HIP stream: wait for work
    that torch queued on the tensor's device (producers of u0, the layout
    copy above) before the library reads the buffers."""
    import torch

    torch.cuda.current_stream(x.device).synchronize()


def _to_colmajor(x, dtype=None):
    y = device_empty(tuple(x.shape), dtype or str(x.dtype).replace("torch.", ""), x.device.index)
    y.copy_(x)
    return y


def init(prob, alg, tspan, *, dt, nsaves, show_progress=False, progress=None, save_start=True,
         workgroup_size=(), rng=None):
    """src/strang_splitting.jl:32-67 (progress and workgroup_size are accepted
    for signature parity; the library picks its own launch shapes)."""
    if not isinstance(alg, StrangSplitting):
        raise TypeError("only StrangSplitting() is implemented")
    return StrangSplittingIterator(prob, tspan, dt, nsaves, save_start, rng)


def solve(prob, alg, tspan, *, dt, nsaves, show_progress=False, progress=None, save_start=True,
          workgroup_size=(), rng=None, save_frames=True, ensemble_moments=False):
    """src/fixed_time_stepping.jl:79-81: solve!(init(prob, alg, tspan; kwargs...)).
    save_frames / ensemble_moments: see StrangSplittingIterator.solve."""
    it = init(prob, alg, tspan, dt=dt, nsaves=nsaves, save_start=save_start, rng=rng)
    try:
        return it.solve(save_frames=save_frames, ensemble_moments=ensemble_moments)
    finally:
        it.close()


# --------------------------------------------------------------- sweeps
class SweepIterator:
    """Many independent 1D single-component problems of uneven grid size and
    parameters advanced together (gpe_sweep_* of include/gpe_hip.h).  The
    reference runs such a sweep as one solve per GrossPitaevskiiProblem."""

    def __init__(self, problems, tspan, dt, nsaves, save_start=True, rng=None):
        lib = L.load()
        problems = list(problems)
        if not problems:
            raise ValueError("empty sweep")
        for p in problems:
            if p.N != 1 or p.M != 1 or len(p.u0[0].shape) != 1:
                raise NotImplementedError("sweeps take 1D single-component problems without batch dims")
        dtypes = {str(p.u0[0].dtype).replace("torch.", "") for p in problems}
        if len(dtypes) != 1:
            raise TypeError("all instances of a sweep need the same field precision")
        self.problems = problems
        self.tspan = (float(tspan[0]), float(tspan[1]))
        self.dt_in = float(dt)
        self.nsaves = int(nsaves)
        self.save_start = bool(save_start)
        self.dt, self.steps_per_save = resolve_fixed_timestepping(dt, tspan, nsaves)
        self.rng = _rng_of(rng)
        self.n = np.array([p.u0[0].shape[0] for p in problems], dtype=np.int64)
        self.off = np.concatenate([[0], np.cumsum(self.n)[:-1]]).astype(np.int64)
        self.total = int(self.n.sum())
        Ls = np.array([p.lengths[0] for p in problems], dtype=np.float64)
        self.dtype = dtypes.pop()
        h = ctypes.c_void_p()
        L.check(lib.gpe_sweep_create(len(problems), self.n.ctypes.data, Ls.ctypes.data,
                                     L.GPE_C64 if self.dtype == "complex64" else L.GPE_C128, 0, ctypes.byref(h)))
        self.h, self._lib = h, lib
        try:
            self._set_terms()
        except BaseException:
            lib.gpe_sweep_destroy(h)
            self.h = None
            raise

    def _set_terms(self):
        lib, h, probs = self._lib, self.h, self.problems
        D = np.zeros(self.total, dtype=np.complex128)
        g = np.zeros(len(probs), dtype=np.complex128)
        c = np.zeros(len(probs), dtype=np.complex128)
        for i, p in enumerate(probs):
            sl = slice(self.off[i], self.off[i] + self.n[i])
            if p.dispersion is not None:
                kind, arr = _form(p.dispersion(reciprocal_grid(p), p.param), (self.n[i],))
                arr = arr if kind == "S" else (arr[0] if kind == "V" else arr[0, 0])
                D[sl] = arr
            if p.potential is not None:
                raise NotImplementedError("sweeps: potentials are not supported")
            if p.nonlinearity is not None:
                gi, ci = fit_nonlinearity(p.nonlinearity, 1, p.param)
                g[i], c[i] = gi[0, 0], ci[0]
        self._keep = [D, g, c]
        L.check(lib.gpe_sweep_set_dispersion(h, D.ctypes.data))
        L.check(lib.gpe_sweep_set_nonlinearity(h, g.ctypes.data, c.ctypes.data))
        if any(p.pump is not None for p in probs):
            times = pump_times(self.tspan, self.dt_in, self.nsaves)
            prof = np.zeros(self.total, dtype=np.complex128)
            envs = []
            for i, p in enumerate(probs):
                if p.pump is None:
                    envs.append(None)
                    continue
                kind, pr, env = pump_descriptor(p, (direct_grid(p)[0],), (self.n[i],), times,
                                                allow_samples=False)
                if kind == L.GPE_PUMP_NONE:
                    envs.append(None)
                    continue
                prof[self.off[i]:self.off[i] + self.n[i]] = pr.view(np.complex128)
                envs.append(None if env is None else env.view(np.complex128))
            ref = next((e for e in envs if e is not None), None)
            if ref is not None:
                j = int(np.argmax(np.abs(ref)))
                for i, e in enumerate(envs):
                    if probs[i].pump is None:
                        continue
                    e = np.ones_like(ref) if e is None else e
                    a = e[j] / ref[j]
                    if np.max(np.abs(e - a * ref)) > 1e-12 * np.max(np.abs(e)):
                        raise NotImplementedError("sweeps share one pump envelope up to a factor per instance")
                    prof[self.off[i]:self.off[i] + self.n[i]] *= a
            envf = None if ref is None else np.ascontiguousarray(ref).view(np.float64)
            proff = np.ascontiguousarray(prof).view(np.float64)
            self._keep += [proff, envf]
            L.check(lib.gpe_sweep_set_pump(h, proff.ctypes.data, None if envf is None else envf.ctypes.data,
                                           0 if envf is None else len(envf) // 2))
        noisy = [p.noise_prototype is not None for p in probs]
        if any(noisy):
            if not all(noisy):
                raise NotImplementedError("sweeps: give every instance a noise_prototype, or none")
            eta = np.zeros(len(probs))
            for i, p in enumerate(probs):
                if p.position_noise_func is None:
                    raise TypeError("noise_prototype without position_noise_func (no method in the reference)")
                eta[i] = noise_eta(p.position_noise_func, 1, p.param, direct_grid(p))[0]
            first = probs[0].noise_prototype
            first = first[0] if isinstance(first, (tuple, list)) else first
            cplx = np.iscomplexobj(np.empty(0, dtype=str(first.dtype).replace("torch.", "")))
            self._keep.append(eta)
            L.check(lib.gpe_sweep_set_noise(h, L.GPE_NOISE_COMPLEX if cplx else L.GPE_NOISE_REAL, eta.ctypes.data,
                                            self.rng.seed, self.rng.counter))

    def solve(self, save_frames=True):
        """Returns (ts, results): results[i] has shape (n_i, nsaves + save_start)."""
        lib = self._lib
        nfr = self.nsaves + int(self.save_start)
        u0 = np.concatenate([np.asarray(p.u0[0]) for p in self.problems]).astype(self.dtype)
        res = np.empty(self.total * nfr, dtype=self.dtype) if save_frames else None
        ts = np.empty(nfr)
        L.check(lib.gpe_sweep_solve(self.h, u0.ctypes.data, self.tspan[0], self.tspan[1], self.dt_in, self.nsaves,
                                    int(self.save_start), None if res is None else res.ctypes.data, ts.ctypes.data))
        if any(p.noise_prototype is not None for p in self.problems):
            self.rng.counter += 2 * self.steps_per_save * self.nsaves
        if res is None:
            return ts, None
        out = [res[self.off[i] * nfr:(self.off[i] + self.n[i]) * nfr].reshape((self.n[i], nfr), order="F")
               for i in range(len(self.problems))]
        return ts, out

    def last_ms(self):
        ms = ctypes.c_float()
        L.check(self._lib.gpe_sweep_last_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            self._lib.gpe_sweep_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_sweep(problems, alg, tspan, *, dt, nsaves, save_start=True, rng=None, save_frames=True):
    """Solve many 1D problems of uneven size in one go: the equivalent of
    [solve(p, alg, tspan; dt, nsaves) for p in problems] (ts is shared)."""
    if not isinstance(alg, StrangSplitting):
        raise TypeError("only StrangSplitting() is implemented")
    it = SweepIterator(problems, tspan, dt, nsaves, save_start, rng)
    try:
        return it.solve(save_frames=save_frames)
    finally:
        it.close()
