"""GPU parity on 3D grids (docs/src/spatial_grid.md:17-18: `lengths = (Lx, Ly, Lz)`).

A 3D step is the x-pass of the 2D schedule followed by three column passes:
FFT_y, then FFT_z -> exp(-i dt D) -> IFFT_z, then IFFT_y.  Every case is
compared with the FP64 oracle (whose 3D path is pinned by the closed forms in
test_oracle.py) on non-cubic grids with distinct extents and box lengths, so
a swapped axis or a wrong k index fails.  Bars as in test_gpu_parity.py.
"""
import math

import numpy as np
import pytest

from test_gpu_parity import FP32_TOL, FP64_TOL, check, rel_l2, run_both

pytestmark = pytest.mark.gpu


def aniso(k, p):
    return 0.5 * k[0] ** 2 + 0.3 * k[1] ** 2 + 0.2 * k[2] ** 2


@pytest.mark.parametrize("shape", [(16, 32, 64), (64, 16, 32), (32, 64, 16)])
def test_3d_free_propagation_exact(ggp, shape):
    """Anisotropic D only: the GPU field equals the exact spectral propagator."""
    Ls = (3.0, 5.0, 7.0)
    rng = np.random.default_rng(1)
    u0 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    prob = ggp.GrossPitaevskiiProblem((np.asfortranarray(u0),), Ls, dispersion=aniso)
    ts, res = ggp.solve(prob, ggp.StrangSplitting(), (0, 0.5), dt=0.05, nsaves=2)
    ks = [2 * np.pi * np.fft.fftfreq(n, L / n) for n, L in zip(shape, Ls)]
    D = 0.5 * ks[0][:, None, None] ** 2 + 0.3 * ks[1][None, :, None] ** 2 + 0.2 * ks[2][None, None, :] ** 2
    U0 = np.fft.fftn(u0)
    worst = 0.0
    for f, t in enumerate(ts):
        exact = np.fft.ifftn(np.exp(-1j * D * t) * U0)
        worst = max(worst, rel_l2(res[0][..., f], exact))
    assert worst <= FP64_TOL, worst
    print(f"3D free propagation {shape}: rel L2 {worst:.3e}")


def terms_3d(L):
    """Two components: non-symmetric 2x2 D(k) with a k_z-odd coupling (no
    folding, no symmetric record), a cross-coupled nonlinearity with loss, an
    SVector potential and a Gaussian pump on component 0 (a table: 3D pumps
    are not separated into per-axis factors)."""
    def dispersion(k, p):
        kk = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
        return [[kk / 2 - 0.1j, 0.3 * k[2] + 0.2], [0.4 + 0.1 * k[0] - 0.05j, kk / 3 - 0.02j]]

    def nonlinearity(u, p):
        return [abs(u[0]) ** 2 + 0.5 * abs(u[1]) ** 2, 0.3 * abs(u[0]) ** 2 + abs(u[1]) ** 2 - 0.01j]

    def potential(r, p):
        return [0.02 * (r[0] - L[0] / 2) ** 2, 0.01 * (r[2] - L[2] / 2) ** 2 - 0.005j]

    def pump(r, p, t):
        f = 0.5 * np.exp(-sum((x - Lx / 2) ** 2 for x, Lx in zip(r, L)) / 4.0)
        return [f, 0 * f]

    return dict(dispersion=dispersion, nonlinearity=nonlinearity, potential=potential, pump=pump)


@pytest.mark.parametrize("dtype,tol", [(np.complex128, FP64_TOL), (np.complex64, FP32_TOL)])
def test_3d_two_component_all_terms(ggp, orc, dtype, tol):
    shape, L = (32, 16, 64), (6.0, 4.0, 9.0)
    rng = np.random.default_rng(8)
    u0 = tuple(0.3 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(2))
    out = run_both(ggp, orc, u0, L, (0, 0.4), 0.02, 4, dtype=dtype, **terms_3d(L))
    err = check(*out, tol)
    print(f"3D two-component {np.dtype(dtype).name}: rel L2 {err:.3e}")


def test_3d_batched_noise(ggp, orc):
    """Trajectories of a 3D grid with Philox noise: the draw layout (pairs x
    and x + N/2 of every x-line, lines flattened over y, z, trajectory) is the
    one the oracle generates."""
    shape, B, L = (16, 32, 16), 3, (4.0, 6.0, 5.0)
    gam = 0.1
    u0 = (np.full(shape + (B,), 0.4 + 0.1j),)
    out = run_both(ggp, orc, u0, L, (0, 0.6), 0.05, 3, seed=2024,
                   dispersion=lambda k, p: aniso(k, p) - 0.5 - 1j * gam / 2,
                   nonlinearity=lambda u, p: abs(u[0]) ** 2, pump=lambda r, p, t: 0.8 + 0 * r[0],
                   position_noise_func=lambda u, r, p: math.sqrt(gam / 2), noise_prototype=True)
    err = check(*out, FP64_TOL)
    print(f"3D batched noise: rel L2 {err:.3e}")


def test_3d_matrix_nonlinearity(ggp, orc):
    """2x2 SMatrix G (per-point matrix exponential between transform-only
    x-passes) on a 3D grid with a scalar potential."""
    shape, L = (16, 32, 16), (5.0, 7.0, 4.0)
    rng = np.random.default_rng(9)
    u0 = tuple(0.3 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(2))

    def nonlinearity(u, p):
        n0, n1 = abs(u[0]) ** 2, abs(u[1]) ** 2
        return [[n0 + 0.6 * n1 - 0.05j, 0.4 + 0 * n0], [0.3 + 0 * n0, n1 + 0.6 * n0]]

    out = run_both(ggp, orc, u0, L, (0, 0.3), 0.02, 3, dispersion=aniso, nonlinearity=nonlinearity,
                   potential=lambda r, p: 0.05 * (r[1] - 3.5) ** 2)
    err = check(*out, FP64_TOL)
    print(f"3D matrix G: rel L2 {err:.3e}")


def test_3d_steady_state_matches_oracle_fixed_point(ggp, orc):
    """Steady-state loop on a 3D grid (the residual check runs through the
    three column passes): a homogeneously pumped, damped Kerr field converges
    to the fixed point of the Strang map, which the oracle reaches by plain
    stepping over the same grid."""
    shape, L = (16, 16, 32), (4.0, 4.0, 8.0)
    terms = dict(dispersion=lambda k, p: aniso(k, p) - 0.2 - 0.5j,
                 nonlinearity=lambda u, p: abs(u[0]) ** 2, pump=lambda r, p, t: 0.5 + 0 * r[0])
    u0 = (np.zeros(shape, complex),)
    prob = ggp.GrossPitaevskiiProblem((np.asfortranarray(u0[0]),), L, **terms)
    u, info = ggp.steady_state(prob, ggp.StrangSplitting(), dt=0.05, tol=1e-13, check_every=20, max_steps=4000)
    assert info["residual"] <= 1e-13, info
    _, res, _ = orc.solve(orc.Problem(u0, L, **terms), (0, 100.0), 0.05, 1)
    want = res[0][..., -1]
    err = rel_l2(np.asarray(u[0]), want)
    assert err <= 1e-11, err
    print(f"3D steady state after {info['steps']} steps: rel L2 {err:.3e}")
