"""GPU parity with three and four components (docs/src/multicomponent_systems.md:
`u0 = (psi_1, psi_2, ...)`; diagonal operators as SVectors).

The M components of a line sit in M wavefront groups of one workgroup, as for
two.  |u_j|^2 crosses components through LDS for the coupled nonlinearity
G_c = c_c + sum_j g_cj |u_j|^2.  A diagonal exp(-i dt D) needs no exchange; a
D(k) that couples components (an M x M SMatrix) is exponentiated per k in FP64
and all M components of a k point meet in LDS.  Bars as in test_gpu_parity.py;
the oracle's M > 2 path is pinned in test_oracle.py (decoupled components
equal single runs, per-component norm conservation, closed-form coupled free
propagation).
"""
import math

import numpy as np
import pytest

from test_gpu_parity import FP32_TOL, FP64_TOL, check, run_both

pytestmark = pytest.mark.gpu


def terms_m(M, L, pump=True, noise=False):
    """Per-component dispersions, complex cross-coupled G, an SVector
    potential and a Gaussian pump on the even components."""
    rng = np.random.default_rng(100 + M)
    g = rng.uniform(-0.5, 1.0, (M, M)) + 1j * rng.uniform(-0.05, 0.0, (M, M))
    c0 = rng.uniform(-0.2, 0.2, M) - 1j * rng.uniform(0.0, 0.05, M)
    a = rng.uniform(0.2, 0.6, M)
    nd = len(L)

    def dispersion(k, p):
        kk = sum(x ** 2 for x in k)
        return [a[c] * kk - 0.02j * c + 0.1 * c for c in range(M)]

    def nonlinearity(u, p):
        return [c0[c] + sum(g[c, j] * abs(u[j]) ** 2 for j in range(M)) for c in range(M)]

    def potential(r, p):
        return [0.01 * (c + 1) * (r[0] - L[0] / 2) ** 2 + 0 * r[-1] for c in range(M)]

    def pumpf(r, p, t):
        f = 0.3 * np.exp(-sum((x - Lx / 2) ** 2 for x, Lx in zip(r, L)) / 2.0)
        return [f if c % 2 == 0 else 0 * f for c in range(M)]

    out = dict(dispersion=dispersion, nonlinearity=nonlinearity, potential=potential)
    if pump:
        out["pump"] = pumpf
    if noise:
        out["position_noise_func"] = lambda u, r, p: [0.05 * (c + 1) for c in range(M)]
        out["noise_prototype"] = True
    return out


def rand_fields(M, shape, seed):
    rng = np.random.default_rng(seed)
    return tuple(0.4 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(M))


@pytest.mark.parametrize("M", [3, 4])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, FP64_TOL), (np.complex64, FP32_TOL)])
def test_2d_many_components(ggp, orc, M, dtype, tol):
    shape, L = (64, 32), (8.0, 5.0)
    out = run_both(ggp, orc, rand_fields(M, shape, M), L, (0, 0.4), 0.02, 4, dtype=dtype, **terms_m(M, L))
    err = check(*out, tol)
    print(f"M={M} 2D {np.dtype(dtype).name}: rel L2 {err:.3e}")


@pytest.mark.parametrize("M", [3, 4])
def test_1d_batched_many_components_with_noise(ggp, orc, M):
    N, B, L = 128, 3, (12.0,)
    out = run_both(ggp, orc, rand_fields(M, (N, B), 20 + M), L, (0, 0.6), 0.03, 3, seed=314,
                   **terms_m(M, L, noise=True))
    err = check(*out, FP64_TOL)
    print(f"M={M} 1D batched noise: rel L2 {err:.3e}")


def test_3d_three_components(ggp, orc):
    shape, L = (16, 32, 16), (4.0, 6.0, 5.0)
    out = run_both(ggp, orc, rand_fields(3, shape, 7), L, (0, 0.3), 0.03, 2, **terms_m(3, L))
    err = check(*out, FP64_TOL)
    print(f"M=3 3D: rel L2 {err:.3e}")


@pytest.mark.parametrize("M,N,dtype", [(3, 2048, np.complex64), (4, 2048, np.complex64), (4, 1024, np.complex128)])
def test_longest_lines_for_many_components(ggp, orc, M, N, dtype):
    """The longest line each precision takes at M = 3, 4 (a workgroup holds
    M * N/8 threads and two LDS buffers of M lines)."""
    L = (40.0,)
    tol = FP64_TOL if dtype == np.complex128 else FP32_TOL
    out = run_both(ggp, orc, rand_fields(M, (N, 2), 30 + M), L, (0, 0.2), 0.02, 2, dtype=dtype,
                   **terms_m(M, L, pump=False))
    err = check(*out, tol)
    print(f"M={M} N={N} {np.dtype(dtype).name}: rel L2 {err:.3e}")


def coupled_dispersion(M, a=None, w=0.3):
    """D(k) = diag(a_c |k|^2 + losses) + w C with a complex non-Hermitian
    coupling C (no symmetry in k or between components)."""
    rng = np.random.default_rng(200 + M)
    a = rng.uniform(0.2, 0.6, M) if a is None else a
    C = rng.uniform(-1, 1, (M, M)) + 0.3j * rng.uniform(-1, 1, (M, M))

    def disp(k, p):
        kk = sum(x ** 2 for x in k)
        z = 0 * kk
        return [[(a[r] * kk - 0.02j * r if r == c else z) + w * C[r, c] + 0.05 * (r - c) * k[0] for c in range(M)]
                for r in range(M)]

    return disp


@pytest.mark.parametrize("M", [3, 4])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, FP64_TOL), (np.complex64, FP32_TOL)])
def test_2d_coupled_dispersion(ggp, orc, M, dtype, tol):
    shape, L = (64, 32), (8.0, 5.0)
    terms = terms_m(M, L)
    terms["dispersion"] = coupled_dispersion(M)
    out = run_both(ggp, orc, rand_fields(M, shape, 50 + M), L, (0, 0.4), 0.02, 4, dtype=dtype, **terms)
    err = check(*out, tol)
    print(f"M={M} coupled D 2D {np.dtype(dtype).name}: rel L2 {err:.3e}")


@pytest.mark.parametrize("M", [3, 4])
def test_1d_batched_coupled_dispersion_with_noise(ggp, orc, M):
    N, B, L = 256, 3, (12.0,)
    terms = terms_m(M, L, noise=True)
    terms["dispersion"] = coupled_dispersion(M)
    out = run_both(ggp, orc, rand_fields(M, (N, B), 60 + M), L, (0, 0.6), 0.03, 3, seed=271, **terms)
    err = check(*out, FP64_TOL)
    print(f"M={M} coupled D 1D batched noise: rel L2 {err:.3e}")


def test_3d_coupled_dispersion(ggp, orc):
    shape, L = (16, 32, 16), (4.0, 6.0, 5.0)
    terms = terms_m(3, L)
    terms["dispersion"] = coupled_dispersion(3)
    out = run_both(ggp, orc, rand_fields(3, shape, 70), L, (0, 0.3), 0.03, 2, **terms)
    err = check(*out, FP64_TOL)
    print(f"M=3 coupled D 3D: rel L2 {err:.3e}")


def test_coupled_dispersion_free_propagation_exact(ggp):
    """D(k) = |k|^2/2 I + W P, P a non-symmetric real 4 x 4, nothing else: the
    GPU frames equal the closed form exp(-i t D(k)) u~(k, 0) through P's
    eigenvectors (FP64, 1e-12); P != P^T, so a transposed row fails."""
    M, N, L, W = 4, 64, 6.0, 0.7
    rng = np.random.default_rng(9)
    P = rng.uniform(-1, 1, (M, M))
    lam, V = np.linalg.eig(P)
    Vi = np.linalg.inv(V)
    u0 = tuple(np.asfortranarray(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) for _ in range(M))

    def disp(k, p):
        kk = (k[0] ** 2 + k[1] ** 2) / 2
        z = 0 * kk
        return [[kk * (a == b) + W * P[a, b] + z for b in range(M)] for a in range(M)]

    ts, res = ggp.solve(ggp.GrossPitaevskiiProblem(u0, (L, L), dispersion=disp), ggp.StrangSplitting(), (0.0, 1.0),
                        dt=0.05, nsaves=4)
    k = 2 * np.pi * np.fft.fftfreq(N, L / N)
    kk = (k[:, None] ** 2 + k[None, :] ** 2) / 2
    U0 = np.stack([np.fft.fft2(x) for x in u0])
    worst = 0.0
    for f in range(1, 5):
        t = float(ts[f])
        mix = V @ np.diag(np.exp(-1j * W * lam * t)) @ Vi
        want = np.exp(-1j * kk * t)[None] * np.einsum("ab,bxy->axy", mix, U0)
        for c in range(M):
            got = np.fft.fft2(np.asarray(res[c])[..., f])
            worst = max(worst, float(np.linalg.norm(got - want[c]) / np.linalg.norm(want[c])))
    assert worst <= 1e-12, worst
    print(f"coupled D M=4 free propagation: rel L2 {worst:.3e}")


def test_many_components_matrix_dispersion_rules(ggp, orc):
    """A diagonal SMatrix D(k) is taken as its diagonal (no exchange between
    components); lines too long for M components are refused at creation."""
    M, shape, L = 3, (32, 32), (5.0, 5.0)

    def diag_matrix(k, p):
        kk = k[0] ** 2 + k[1] ** 2
        z = 0 * kk
        return [[kk / 2, z, z], [z, kk / 3 - 0.01j, z], [z, z, kk / 4 + 0.2]]

    out = run_both(ggp, orc, rand_fields(M, shape, 3), L, (0, 0.2), 0.02, 2, dispersion=diag_matrix,
                   nonlinearity=lambda u, p: [abs(u[0]) ** 2, abs(u[1]) ** 2 + 0.5 * abs(u[2]) ** 2, 0 * u[0]])
    check(*out, FP64_TOL)

    long = tuple(np.zeros((2048, 2), np.complex128, order="F") for _ in range(3))
    with pytest.raises(ggp.GpeUnsupported):
        ggp.solve(ggp.GrossPitaevskiiProblem(long, (10.0,)), ggp.StrangSplitting(), (0, 0.1), dt=0.05, nsaves=1)


@pytest.mark.parametrize("M,shape,dtype,noise", [
    (1, (4096, 16), np.complex64, False),    # 2D rows of 4096 points, 1024-thread row pass
    (2, (4096, 2), np.complex64, True),      # 1D whole-step kernel, two 4096-point components
    (3, (2048, 2), np.complex64, False),     # 1D, three 2048-point components
    (2, (2048, 2), np.complex128, False),    # 1D FP64, two 2048-point components
])
def test_long_lines_with_pump_and_potential(ggp, orc, M, shape, dtype, noise):
    """Kernels of more than 512 threads per workgroup read the pump profile,
    exp(-i dt/2 V) and the FFT twiddles where they are used; with every term
    on, their frames match the oracle as the register-held forms do."""
    nd = 2 if shape[1] == 16 else 1
    L = (60.0, 12.0) if nd == 2 else (60.0,)
    tol = FP64_TOL if dtype == np.complex128 else FP32_TOL
    out = run_both(ggp, orc, rand_fields(M, shape, 80 + M), L, (0, 0.2), 0.02, 2, dtype=dtype,
                   seed=(123 if noise else None), **terms_m(M, L, pump=True, noise=noise))
    err = check(*out, tol)
    print(f"M={M} {shape} {np.dtype(dtype).name} pump+V: rel L2 {err:.3e}")
