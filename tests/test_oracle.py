"""Pins of the FP64 oracle (oracle/gpe_oracle.py) that do not retype its
formulas: known answers, closed-form solutions, invariants, an independent
library (scipy), and the reference's own analytic tests.

Reference tests restated here (expected values computed at test time in the
reference too; it stores no golden vectors):
    test/windowed_ft.jl          vacuum commutator of windowed FTs, rtol 7e-2
    test/exciton_polariton_test.jl  homogeneous steady-state relations, 3e-2
Closed forms replace StructuredLight's free_propagation / kerr_propagation
(test/free_propagation.jl:16, test/kerr_propagation.jl:17), which are not in
this image.
"""
import math

import numpy as np
import pytest
import scipy.linalg

from oracle import gpe_oracle as orc


# ------------------------------------------------ src/fixed_time_stepping.jl
@pytest.mark.parametrize("dt,tspan,nsaves,expect_dt,expect_sps", [
    (0.01, (0, 0.4), 64, 0.00625, 1),                     # README Kerr example
    (0.05, (0, 3300), 512, 0.04996366279069767, 129),     # test/bistability_cycle.jl
    (0.1, (0, 100), 256, 0.09765625, 4),                  # test/exciton_polariton_test.jl
    (4, (0, 200), 1, 4.0, 50),                            # test/windowed_ft.jl
    (0.05, (0, 500), 100, 0.05, 100),                     # 512^2 polariton workload
])
def test_resolve_fixed_timestepping_known_cases(dt, tspan, nsaves, expect_dt, expect_sps):
    got_dt, sps, t0 = orc.resolve_fixed_timestepping(dt, tspan, nsaves)
    assert got_dt == expect_dt and sps == expect_sps and t0 == tspan[0]


def test_time_accumulation_is_repeated_addition():
    # t += dt (fixed_time_stepping.jl:44): 10^4 steps of 0.05 do not land on 500
    p = orc.Problem((np.zeros(16, complex),), (1.0,))
    ts, _, _ = orc.solve(p, (0, 500), 0.05, 100)
    assert ts[-1] == 500.0000000000794
    assert ts[1] == sum([0.05] * 100)


# ----------------------------------------------------------- src/problem.jl
def test_grids():
    x = orc.direct_grid_1d(8.0, 128)
    assert x[0] == 0.0 and x[-1] == 127 * 0.0625 and np.all(np.diff(x) == 0.0625)
    k = orc.reciprocal_grid_1d(8.0, 128)
    # FFT order [0..N/2-1, -N/2..-1] * 2pi/L, bit-exact for power-of-two N
    m = np.r_[np.arange(64), np.arange(-64, 0)]
    assert np.array_equal(k, m * (2 * math.pi / 8.0))
    np.testing.assert_allclose(k, 2 * math.pi * np.fft.fftfreq(128, 8.0 / 128), rtol=1e-15)


# ------------------------------------------------------------------ Philox
def test_philox_known_answers():
    """Random123 known-answer vectors of Philox4x32-10."""
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kat:
        got = tuple(int(x) for x in orc.philox4x32_10(*ctr, *key))
        assert got == want


def test_xi_statistics():
    """E xi = 0, E|xi|^2 = 1 with Re, Im ~ N(0, 1/2) (src/problem.jl:38-41);
    real xi ~ N(0, 1); draws of different calls/components uncorrelated."""
    shape = (256, 512)
    z = orc.philox_xi(shape, 0, 1234, 7)
    n = z.size
    tol = 5 / math.sqrt(n)
    assert abs(z.mean()) < tol
    assert abs(np.mean(abs(z) ** 2) - 1) < 5 * tol
    assert abs(np.var(z.real) - 0.5) < 5 * tol and abs(np.var(z.imag) - 0.5) < 5 * tol
    assert abs(np.mean(z.real * z.imag)) < tol
    z2 = orc.philox_xi(shape, 0, 1234, 8)
    z3 = orc.philox_xi(shape, 1, 1234, 7)
    assert abs(np.mean(z * z2.conj())) < tol and abs(np.mean(z * z3.conj())) < tol
    r = orc.philox_xi(shape, 0, 1234, 7, complex_noise=False)
    assert np.all(r.imag == 0) and abs(np.var(r.real) - 1) < 10 * tol
    # Gaussian tails: P(|Re xi| > 3 sqrt(1/2)) = 0.0027
    frac = np.mean(abs(z.real) > 3 * math.sqrt(0.5))
    assert abs(frac - 0.0027) < 0.0006


# ------------------------------------------------------- matrix exponential
def test_expm_2x2_against_scipy():
    rng = np.random.default_rng(0)
    for scale in (1e-6, 1e-3, 0.1, 1.0, 5.0):
        A = (rng.standard_normal((2, 2, 50)) + 1j * rng.standard_normal((2, 2, 50))) * scale
        E = orc.expm_2x2(A)
        for i in range(50):
            ref = scipy.linalg.expm(A[:, :, i])
            assert np.max(abs(E[:, :, i] - ref)) <= 1e-14 * max(1.0, np.max(abs(ref)))


def test_expm_2x2_degenerate_and_nilpotent():
    # s -> 0: equal diagonal, zero coupling; nilpotent [[0,1],[0,0]]: exp = I + A
    A = np.zeros((2, 2, 3), complex)
    A[:, :, 0] = [[0.3 + 0.1j, 0], [0, 0.3 + 0.1j]]
    A[:, :, 1] = [[0, 1], [0, 0]]
    A[:, :, 2] = [[1e-9, 2e-9j], [3e-9, -1e-9]]
    E = orc.expm_2x2(A)
    np.testing.assert_allclose(E[:, :, 0], np.exp(0.3 + 0.1j) * np.eye(2), rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(E[:, :, 1], [[1, 1], [0, 1]], atol=1e-15)
    np.testing.assert_allclose(E[:, :, 2], scipy.linalg.expm(A[:, :, 2]), atol=1e-17)


# ------------------------------------------- closed-form evolutions (Strang)
def _gauss2d(N, L, x0, y0, w, kx=0.0):
    r = np.arange(N) * (L / N)
    return np.exp(-((r[:, None] - x0) ** 2 + (r[None, :] - y0) ** 2) / w ** 2 + 1j * kx * r[:, None])


def test_free_propagation_exact():
    """D = |k|^2/2 only: Strang reduces to the exact spectral propagator,
    u(k, t) = exp(-i |k|^2 t / 2) u(k, 0) (replaces StructuredLight's
    free_propagation of test/free_propagation.jl:16)."""
    N, L = 64, 10.0
    u0 = _gauss2d(N, L, 5.0, 4.0, 1.3, kx=2.0)
    ts, res, _ = orc.solve(orc.Problem((u0,), (L, L), dispersion=lambda k, p: (k[0] ** 2 + k[1] ** 2) / 2),
                           (0, 2.0), 0.07, 20)
    k = orc.reciprocal_grid_1d(L, N)
    kk = k[:, None] ** 2 + k[None, :] ** 2
    U0 = np.fft.fft2(u0)
    for f, t in enumerate(ts):
        exact = np.fft.ifft2(np.exp(-0.5j * kk * t) * U0)
        assert np.linalg.norm(res[0][..., f] - exact) <= 1e-12 * np.linalg.norm(exact)


def test_homogeneous_kerr_exact():
    """Constant field, D(0) = 0, G = g|u|^2: u(t) = u0 exp(-i g |u0|^2 t)."""
    a, g = 0.7 - 0.2j, -1.7
    u0 = np.full((32, 32), a)
    ts, res, _ = orc.solve(orc.Problem((u0,), (5.0, 5.0), dispersion=lambda k, p: (k[0] ** 2 + k[1] ** 2) / 2,
                                       nonlinearity=lambda u, p: g * abs(u[0]) ** 2), (0, 3.0), 0.01, 6)
    for f, t in enumerate(ts):
        np.testing.assert_allclose(res[0][..., f], a * np.exp(-1j * g * abs(a) ** 2 * t), rtol=2e-13)


def test_pump_trapezoid_with_reference_times():
    """No D, no G: every half-step adds (dt/4)(F(tau_prev) + F(tau)) with the
    pump times t+dt/2, t+dt taken AFTER t += dt (strang_splitting.jl:87,89,
    fixed_time_stepping.jl:44-45) and F_now = F(t0) from init (:58)."""
    N, L = 16, 4.0
    env = lambda t: 1.0 + 0.3 * t + 0.05j * t * t
    prof = np.exp(-((np.arange(N) * (L / N) - 2.0) ** 2))
    pump = lambda r, p, t: np.exp(-((r[0] - 2.0) ** 2)) * env(t)
    dt, nsaves, tspan = 0.1, 3, (0.0, 1.2)
    ts, res, it = orc.solve(orc.Problem((np.zeros(N, complex),), (L,), pump=pump), tspan, dt, nsaves)
    rdt, sps, _ = orc.resolve_fixed_timestepping(dt, tspan, nsaves)
    taus = [0.0]
    t = 0.0
    expect = np.zeros(N, complex)
    frames = []
    for s in range(sps * nsaves):
        t += rdt
        for tau in (t + rdt / 2, t + rdt):
            expect = expect + (rdt / 4) * (env(taus[-1]) + env(tau)) * prof
            taus.append(tau)
        if (s + 1) % sps == 0:
            frames.append(expect.copy())
    for f in range(nsaves):
        np.testing.assert_allclose(res[0][..., f + 1], frames[f], rtol=1e-14, atol=1e-15)
    # the lag matters: evaluating at the start-of-step times gives a visibly different answer
    naive = np.zeros(N, complex)
    for s in range(sps * nsaves):
        t0 = s * rdt
        naive = naive + (rdt / 2) * (env(t0) + env(t0 + rdt)) * prof
    assert np.linalg.norm(naive - frames[-1]) > 1e-3 * np.linalg.norm(frames[-1])


def test_pump_trapezoid_non_factoring_pump():
    """A pump that does not factor (a spot moving across a 2D box with a
    time-dependent phase), no D, no G: every half-step adds
    (dt/4)(F(r, tau_prev) + F(r, tau)) at the reference's lagged pump times,
    pointwise, so the frames are sums of closed-form samples."""
    N0, N1, L = 16, 8, (4.0, 2.0)
    x = np.arange(N0) * (L[0] / N0)
    y = np.arange(N1) * (L[1] / N1)

    def F(X, Y, t):
        return np.exp(-((X - 1.0 - t) ** 2 + (Y - 1.0) ** 2)) * (1.0 + 0.2j * t)

    pump = lambda r, p, t: F(r[0], r[1], t)
    dt, nsaves, tspan = 0.1, 3, (0.0, 0.9)
    ts, res, it = orc.solve(orc.Problem((np.zeros((N0, N1), complex),), L, pump=pump), tspan, dt, nsaves)
    rdt, sps, _ = orc.resolve_fixed_timestepping(dt, tspan, nsaves)
    X, Y = x[:, None], y[None, :]
    tau_prev, t = 0.0, 0.0
    expect = np.zeros((N0, N1), complex)
    frames = []
    for s in range(sps * nsaves):
        t += rdt
        for tau in (t + rdt / 2, t + rdt):
            expect = expect + (rdt / 4) * (F(X, Y, tau_prev) + F(X, Y, tau))
            tau_prev = tau
        if (s + 1) % sps == 0:
            frames.append(expect.copy())
    for f in range(nsaves):
        np.testing.assert_allclose(res[0][..., f + 1], frames[f], rtol=1e-13, atol=1e-15)
    # one profile times an envelope cannot reproduce it
    P0 = F(X, Y, 0.0)
    e = np.vdot(P0, frames[-1]) / np.vdot(P0, P0)
    assert np.linalg.norm(frames[-1] - e * P0) > 1e-2 * np.linalg.norm(frames[-1])


def test_norm_conserved_for_hermitian_problem():
    """Real D(k), real G, no pump/noise: every factor is unitary."""
    rng = np.random.default_rng(1)
    u0 = rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))
    ts, res, _ = orc.solve(orc.Problem((u0,), (6.0, 6.0), dispersion=lambda k, p: (k[0] ** 2 + k[1] ** 2) / 2,
                                       nonlinearity=lambda u, p: 0.8 * abs(u[0]) ** 2, potential=lambda r, p:
                                       0.5 * (r[0] - 3) ** 2), (0, 2.0), 0.02, 4)
    n0 = np.linalg.norm(u0)
    for f in range(res[0].shape[-1]):
        assert abs(np.linalg.norm(res[0][..., f]) / n0 - 1) < 1e-13


def test_strang_is_second_order():
    """Self-convergence on a focusing Kerr problem: error(dt)/error(dt/2) ~ 4
    (replaces the comparison with StructuredLight's kerr_propagation)."""
    N, L = 64, 12.0
    u0 = _gauss2d(N, L, 6.0, 6.0, 1.5).astype(complex)
    prob = lambda: orc.Problem((u0,), (L, L), dispersion=lambda k, p: (k[0] ** 2 + k[1] ** 2) / 2,
                               nonlinearity=lambda u, p: -2.0 * abs(u[0]) ** 2)
    out = {}
    for dt in (0.04, 0.02, 0.0025):
        _, res, _ = orc.solve(prob(), (0, 0.8), dt, 1)
        out[dt] = res[0][..., -1]
    e1 = np.linalg.norm(out[0.04] - out[0.0025])
    e2 = np.linalg.norm(out[0.02] - out[0.0025])
    assert 3.3 < e1 / e2 < 4.7


def test_static_array_forms_agree():
    """Scalar, SVector{1} and SMatrix{1,1} forms give the same solution
    (test/bistability_cycle.jl:67-71, multicomponent_systems.md:59)."""
    N, L = 32, 8.0
    u0 = (_gauss2d(N, L, 4, 4, 1.0),)
    d = lambda k, p: (k[0] ** 2 + k[1] ** 2) / 2 - 0.1j
    g = lambda u, p: 0.5 * abs(u[0]) ** 2
    f = lambda r, p, t: 0.2 * np.exp(-((r[0] - 4) ** 2 + (r[1] - 4) ** 2)) * (1 + t)
    forms = [(d, g, f),
             (lambda k, p: [d(k, p)], lambda u, p: [g(u, p)], lambda r, p, t: [f(r, p, t)]),
             (lambda k, p: [[d(k, p)]], lambda u, p: [[g(u, p)]], lambda r, p, t: [f(r, p, t)])]
    sols = []
    for dd, gg, ff in forms:
        _, res, _ = orc.solve(orc.Problem(u0, (L, L), dispersion=dd, nonlinearity=gg, pump=ff), (0, 0.5), 0.05, 2)
        sols.append(res[0])
    for s in sols[1:]:
        np.testing.assert_allclose(s, sols[0], rtol=1e-14, atol=1e-15)


def test_batch_broadcast_equals_single_runs():
    rng = np.random.default_rng(2)
    B = 3
    u0 = rng.standard_normal((32, B)) + 1j * rng.standard_normal((32, B))
    kw = dict(dispersion=lambda k, p: k[0] ** 2 / 2 - 0.05j, nonlinearity=lambda u, p: abs(u[0]) ** 2,
              pump=lambda r, p, t: 0.3 + 0 * r[0])
    _, batched, _ = orc.solve(orc.Problem((u0,), (6.0,), **kw), (0, 1.0), 0.05, 2)
    for b in range(B):
        _, single, _ = orc.solve(orc.Problem((u0[:, b].copy(),), (6.0,), **kw), (0, 1.0), 0.05, 2)
        assert np.array_equal(batched[0][:, b, :], single[0])


# --------------------------------------------------------------- noise, TWA
def test_noise_stationary_vacuum_occupation():
    """d psi = -gamma/2 psi dt - i eta dW: stationary <|psi|^2> = eta^2/gamma."""
    N, B, L = 32, 512, 10.0
    gam, eta = 0.4, 0.3
    prob = orc.Problem((np.zeros((N, B), complex),), (L,), dispersion=lambda k, p: -0.5j * gam + 0 * k[0],
                       position_noise_func=lambda u, r, p: eta, noise_prototype="complex", noise_seed=11)
    ts, res, _ = orc.solve(prob, (0, 60.0), 0.1, 1, save_start=False)
    occ = np.mean(abs(res[0][..., -1]) ** 2)
    expect = eta ** 2 / gam
    # discrete-time correction of the exponential integrator is O(gamma dt)
    assert abs(occ / expect - 1) < 0.03


def test_reference_windowed_ft_commutator():
    """test/windowed_ft.jl: vacuum fluctuations of a damped 1D field with
    eta = sqrt(gamma / 2 dx); the windowed-FT correlation must match the
    analytic commutator within rtol 7e-2 (same workload, Philox noise)."""
    hb = 0.6582
    gam = 0.1 / hb
    m = hb ** 2 / 2.5
    d0 = 0.49 / hb
    L, N = 20.0, 64
    dL = L / N
    param = dict(d0=d0, m=m, gam=gam, hb=hb, dL=dL)

    def dispersion(ks, p):
        return -1j * p["gam"] / 2 + p["hb"] * ks[0] ** 2 / (2 * p["m"]) - p["d0"]

    prob = orc.Problem((np.zeros((N, 10 ** 4), complex),), (L,), dispersion=dispersion, param=param,
                       position_noise_func=lambda psi, r, p: math.sqrt(p["gam"] / 2 / p["dL"]),
                       noise_prototype="complex", noise_seed=1234)
    _, res, _ = orc.solve(prob, (0, 200), 4, 1, save_start=False)
    sol = res[0][..., 0]
    rs = -L / 2 + np.arange(N) * (L / N)
    ks = -math.pi / L + np.arange(N) * (2 * math.pi / L)
    window = lambda x, x0, w: np.exp(-(x - x0) ** 2 / w ** 2)

    def correlation(w1, w2):
        s1 = sol * w1[:, None]
        s2 = sol * w2[:, None]
        f1 = np.fft.ifftshift(np.fft.fft(np.fft.fftshift(s1, 0), axis=0), 0)
        f2 = np.fft.ifftshift(np.fft.fft(np.fft.fftshift(s2, 0), axis=0), 0)
        # g1[i, j] = dot(f2[j, :], f1[i, :]) / length(sol), dot conjugating its first argument
        return f1 @ f2.conj().T / sol.size

    def analytic(w1, w2):
        out = np.empty((N, N), complex)
        for i, k in enumerate(ks):
            for j, kp in enumerate(ks):
                out[i, j] = np.sum(np.exp(1j * (k - kp) * rs) * w1 * np.conj(w2))
        return out / (rs[-1] - rs[0]) / 2

    for x0 in (-1, 0, 1):
        for w in (3, 4, 5):
            w1, w2 = window(rs, x0, w), window(rs, -x0, w)
            corr, an = correlation(w1, w2), analytic(w1, w2)
            assert np.linalg.norm(corr - an) <= 7e-2 * max(np.linalg.norm(corr), np.linalg.norm(an))


def test_reference_exciton_polariton_steady_state():
    """test/exciton_polariton_test.jl: 2x2 SMatrix dispersion, SVector
    nonlinearity and Gaussian photon pump; homogeneous steady-state relations
    at the centre within 3e-2 (in the loss convention of the dispersion used,
    examples/exciton_polariton.jl:117,121)."""
    hb = 0.654
    Wr = 5.07 / 2 / hb
    gx = 0.0015 / hb
    gc = 0.07 / 0.6571 / hb
    wx, wc = 1484.44 / hb, 1482.76 / hb
    m = hb ** 2 / (2 * 2e-1)
    dx, dc = wc - wx, 0.0
    A, w, g = 2.0, 100.0, 1e-2 / hb
    L, N = 256.0, 128

    def dispersion(k, p):
        Dcc = hb * (k[0] ** 2 + k[1] ** 2) / (2 * m) - dc - 1j * gc
        return [[Dcc, Wr], [Wr, -dx - 1j * gx]]

    prob = orc.Problem((np.zeros((N, N), complex),) * 2, (L, L), dispersion=dispersion,
                       nonlinearity=lambda u, p: [0 * u[0], g * abs(u[1]) ** 2],
                       pump=lambda r, p, t: [A * np.exp(-((r[0] - L / 2) ** 2 + (r[1] - L / 2) ** 2) / w ** 2),
                                             0 * r[0] * r[1]])
    _, res, _ = orc.solve(prob, (0, 100), 0.1, 256)
    nx = abs(res[1][N // 2, N // 2, -1]) ** 2
    nc = abs(res[0][N // 2, N // 2, -1]) ** 2
    r1 = abs(Wr - (dx + 1j * gx - g * nx) * (dc + 1j * gc) / Wr) ** 2 * nx / abs(A) ** 2
    r2 = abs(dx + 1j * gx - g * nx) ** 2 * nx / Wr ** 2 / nc
    assert abs(r1 - 1) < 3e-2 and abs(r2 - 1) < 3e-2


# ------------------------------------------- momentum-space pair correlations
def test_kspace_g2_coherent_plane_wave_closed_form(orc):
    """Every trajectory the plane wave alpha exp(2 pi i k0 x / N): |ft(k)|^2 =
    |alpha|^2 delta_{k,k0}, so n(k) and G2(k, k') have closed forms."""
    N, B, k0, alpha, c = 16, 5, 3, 0.7 - 0.4j, 1 / 8.0
    x = np.arange(N)
    frames = np.repeat((alpha * np.exp(2j * np.pi * k0 * x / N))[:, None], B, axis=1)
    nks, G2 = orc.kspace_g2(frames, c)
    A = abs(alpha) ** 2
    want_n = np.full(N, -c / 2)
    want_n[k0] += A
    assert np.allclose(nks, want_n, rtol=0, atol=1e-15)
    want = np.full((N, N), c * c / 4)
    np.fill_diagonal(want, c * c / 2)
    want[k0, :] = want[:, k0] = -(c / 2) * (A - c / 2)
    want[k0, k0] = A * A - 2 * c * A + c * c / 2
    assert np.allclose(G2, want, rtol=0, atol=1e-14)


def test_kspace_g2_vacuum_is_uncorrelated(orc):
    """Wigner vacuum: ft(k) complex Gaussian with <|ft|^2> = c/2 independently
    per k.  Normal-ordered moments vanish: n(k) -> 0 and G2(k, k') -> 0
    (diagonal: <|z|^4> = 2 <|z|^2>^2 for a complex Gaussian)."""
    N, B, c = 16, 40000, 0.25
    rng = np.random.default_rng(7)
    ft = (rng.standard_normal((N, B)) + 1j * rng.standard_normal((N, B))) * np.sqrt(c / 4)
    frames = np.fft.ifft(ft, axis=0) * N
    nks, G2 = orc.kspace_g2(frames, c)
    assert np.max(abs(nks)) < 5 * (c / 2) / np.sqrt(B)
    assert np.max(abs(G2)) < 8 * (c / 2) ** 2 / np.sqrt(B) * 2


def test_kspace_g2_matrix_form_equals_pairwise(orc):
    rng = np.random.default_rng(3)
    frames = rng.standard_normal((32, 7)) + 1j * rng.standard_normal((32, 7))
    n1, g1 = orc.kspace_g2(frames, 0.3)
    n2, g2 = orc.kspace_g2(frames, 0.3, pairwise=False)
    assert np.array_equal(n1, n2)
    assert np.allclose(g1, g2, rtol=1e-13, atol=1e-13 * np.max(abs(g1)))
    assert np.allclose(g1, g1.T, rtol=0, atol=0)


# ------------------------------------------------ windowed correlation
def test_windowed_correlation_plane_wave_layout(orc):
    """One trajectory exp(2 pi i k0 n / N) with windows 1 and i: after the
    fftshift / ifftshift pair the only entry sits at i0 = (k0 + N/2) mod N and
    equals conj(i N) N / N = -i N, which fixes the output shift and which
    factor is conjugated (the vacuum check below is blind to both)."""
    N, k0 = 16, 5
    u = np.exp(2j * np.pi * k0 * np.arange(N) / N)[:, None]
    g = orc.windowed_correlation(u, np.ones(N), 1j * np.ones(N))
    i0 = (k0 + N // 2) % N
    want = np.zeros((N, N), complex)
    want[i0, i0] = -1j * N
    assert np.allclose(g, want, rtol=0, atol=1e-12)


def test_windowed_correlation_vacuum_matches_analytic(orc):
    """test/windowed_ft.jl's check with vacuum samples drawn directly:
    [psi(r), psi^dagger(r')] = delta N/L, so <|psi|^2>_W = N/(2L); the
    windowed correlation matches the analytic commutator within the
    reference's rtol 7e-2 for its nine window pairs."""
    L, N, B = 20.0, 64, 10000
    rng = np.random.default_rng(1)
    u = (rng.standard_normal((N, B)) + 1j * rng.standard_normal((N, B))) * np.sqrt(N / L / 4)
    rs = -L / 2 + np.arange(N) * (L / N)
    for x0 in (-1, 0, 1):
        for w in (3, 4, 5):
            f1 = lambda r, x0=x0, w=w: np.exp(-(r - x0) ** 2 / w ** 2)
            f2 = lambda r, x0=x0, w=w: np.exp(-(r + x0) ** 2 / w ** 2)
            corr = orc.windowed_correlation(u, f1(rs), f2(rs))
            an = orc.analytic_commutation(L, N, f1, f2)
            assert np.linalg.norm(corr - an) <= 7e-2 * max(np.linalg.norm(corr), np.linalg.norm(an))


# ------------------------------------------------ density correlation g2(r)
def test_density_g2_x_closed_forms(orc):
    """A homogeneous coherent field and a static density wave
    n(x) = 1 + eps cos(2 pi x / N) (the same in every trajectory) have closed
    forms: <n(x) n(x + r)>_x = |a|^4, resp. 1 + eps^2/2 cos(2 pi r / N)."""
    c = 0.2
    alpha = 0.8 + 0.3j
    G2, nphys = orc.density_g2_x(np.full((16, 4, 3), alpha), c)
    A = abs(alpha) ** 2
    want = A * A - (1 + (np.arange(16) == 0)) * c / 2 * (2 * A - c / 2)
    assert np.allclose(G2, want, rtol=0, atol=1e-14) and abs(nphys - (A - c / 2)) < 1e-15
    N, eps = 32, 0.4
    x = np.arange(N)
    u = np.sqrt(1 + eps * np.cos(2 * np.pi * x / N))[:, None, None] * np.ones((1, 8, 2))
    G2, nphys = orc.density_g2_x(u, c)
    r = np.arange(N)
    want = 1 + eps ** 2 / 2 * np.cos(2 * np.pi * r / N) - (1 + (r == 0)) * c / 2 * (2 - c / 2)
    assert np.allclose(G2, want, rtol=0, atol=1e-13) and abs(nphys - (1 - c / 2)) < 1e-14


def test_density_g2_x_vacuum_is_uncorrelated(orc):
    N, B, c = 16, 20000, 4.0
    rng = np.random.default_rng(9)
    u = (rng.standard_normal((N, B)) + 1j * rng.standard_normal((N, B))) * np.sqrt(c / 4)
    G2, nphys = orc.density_g2_x(u, c)
    scale = (c / 2) ** 2 / np.sqrt(N * B)
    assert abs(nphys) < 6 * (c / 2) / np.sqrt(N * B)
    assert np.max(abs(G2)) < 30 * scale


def test_matrix_nonlinearity_constant_rabi_closed_form(orc):
    """A constant 2x2 SMatrix G (Rabi coupling, complex detuning) and no
    dispersion: every half-step is u <- exp(-i dt/2 G) u, so after time t the
    field is exp(-i t G) u0 exactly (scipy's expm)."""
    G = np.array([[0.3 - 0.05j, 0.7], [0.7, -0.2]])
    rng = np.random.default_rng(4)
    u0 = tuple(rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8)) for _ in range(2))
    prob = orc.Problem(u0, (4.0, 4.0), nonlinearity=lambda u, p: [[G[0, 0] + 0 * u[0], G[0, 1] + 0 * u[0]],
                                                                   [G[1, 0] + 0 * u[0], G[1, 1] + 0 * u[0]]])
    ts, res, _ = orc.solve(prob, (0, 1.0), 0.05, 2)
    E = scipy.linalg.expm(-1j * ts[-1] * G)
    want = np.einsum("ab,bxy->axy", E, np.stack(u0))
    got = np.stack([res[c][..., -1] for c in range(2)])
    assert np.max(abs(got - want)) < 1e-12 * np.max(abs(want))


@pytest.mark.parametrize("form", ["scalar", "svector"])
def test_matrix_nonlinearity_with_constant_potential_closed_form(orc, form):
    """A constant non-symmetric 2x2 SMatrix G with a constant potential and no
    dispersion.  Julia's product rules (src/kernels.jl:9-15) fix the half-step
    factor: a scalar V gives the matrix exp(-i dt/2 G) exp(-i dt/2 V), so the
    field after time t is exp(-i t G) exp(-i t V) u0; an SVector V gives the
    vector e = exp(-i dt/2 G) exp(-i dt/2 V) (a matrix-vector product) that
    multiplies each component, so u_c(t) = e_c^(2 n) u_c(0) after n steps.
    G01 != G10 makes a transposed product fail."""
    G = np.array([[0.3 - 0.05j, 0.7], [0.2 + 0.1j, -0.2]])
    rng = np.random.default_rng(5)
    u0 = tuple(rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8)) for _ in range(2))
    V = 0.4 if form == "scalar" else [0.4, -0.3 - 0.02j]
    prob = orc.Problem(u0, (4.0, 4.0), nonlinearity=lambda u, p: [[G[0, 0] + 0 * u[0], G[0, 1] + 0 * u[0]],
                                                                   [G[1, 0] + 0 * u[0], G[1, 1] + 0 * u[0]]],
                       potential=(lambda r, p: V + 0 * r[0]) if form == "scalar"
                       else (lambda r, p: [V[0] + 0 * r[0], V[1] + 0 * r[0]]))
    dt, T = 0.05, 1.0
    ts, res, _ = orc.solve(prob, (0, T), dt, 2)
    got = np.stack([res[c][..., -1] for c in range(2)])
    if form == "scalar":
        E = scipy.linalg.expm(-1j * ts[-1] * G) * np.exp(-1j * ts[-1] * V)
        want = np.einsum("ab,bxy->axy", E, np.stack(u0))
    else:
        n = int(round(ts[-1] / dt))
        e = scipy.linalg.expm(-0.5j * dt * G) @ np.exp(-0.5j * dt * np.asarray(V))
        want = (e ** (2 * n))[:, None, None] * np.stack(u0)
    assert np.max(abs(got - want)) < 1e-11 * np.max(abs(want))


# ------------------------------------------------------------------ 3D grids
def test_free_propagation_3d_exact():
    """3D grid, anisotropic D = a k_x^2 + b k_y^2 + c k_z^2 on a non-cubic box
    (docs/src/spatial_grid.md:17-18 lists 3D domains): Strang is the exact
    spectral propagator, checked with scipy's FFT.  Distinct extents, lengths
    and coefficients per axis make a swapped axis fail."""
    import scipy.fft

    n, Ls, co = (16, 8, 32), (3.0, 5.0, 7.0), (0.5, 0.3, 0.2)
    rng = np.random.default_rng(21)
    u0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    disp = lambda k, p: co[0] * k[0] ** 2 + co[1] * k[1] ** 2 + co[2] * k[2] ** 2
    ts, res, _ = orc.solve(orc.Problem((u0,), Ls, dispersion=disp), (0, 0.5), 0.05, 2)
    ks = [orc.reciprocal_grid_1d(L, m) for L, m in zip(Ls, n)]
    D = co[0] * ks[0][:, None, None] ** 2 + co[1] * ks[1][None, :, None] ** 2 + co[2] * ks[2][None, None, :] ** 2
    U0 = scipy.fft.fftn(u0)
    for f, t in enumerate(ts):
        exact = scipy.fft.ifftn(np.exp(-1j * D * t) * U0)
        assert np.linalg.norm(res[0][..., f] - exact) <= 1e-12 * np.linalg.norm(exact)


def test_two_component_plane_wave_3d_exact():
    """A single 3D plane wave exp(i q.r) in both components under a 2x2
    non-symmetric D(k) and no real-space terms: the mode evolves as
    expm(-i t D(q)) (scipy), and every other mode stays empty."""
    n, Ls = (8, 16, 8), (2.0, 4.0, 3.0)
    m = (1, 3, 6)  # mode indices (FFT order; 6 of 8 is the negative -2)
    ks = [orc.reciprocal_grid_1d(L, N) for L, N in zip(Ls, n)]
    q = np.array([ks[d][m[d]] for d in range(3)])
    r = np.meshgrid(*[np.arange(N) * (L / N) for L, N in zip(Ls, n)], indexing="ij")
    wave = np.exp(1j * (q[0] * r[0] + q[1] * r[1] + q[2] * r[2]))
    a0 = np.array([0.6 - 0.1j, 0.2 + 0.5j])
    u0 = (a0[0] * wave, a0[1] * wave)

    def disp(k, p):
        kk = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
        return [[kk / 2 - 0.1j, 0.3 * k[2] + 0.2], [0.4 - 0.05j * k[0], kk / 3]]

    ts, res, _ = orc.solve(orc.Problem(u0, Ls, dispersion=disp), (0, 1.0), 0.1, 2)
    kk = float(q @ q)
    Dq = np.array([[kk / 2 - 0.1j, 0.3 * q[2] + 0.2], [0.4 - 0.05j * q[0], kk / 3]])
    for f, t in enumerate(ts):
        a = scipy.linalg.expm(-1j * t * Dq) @ a0
        for c in range(2):
            np.testing.assert_allclose(res[c][..., f], a[c] * wave, atol=1e-13)


# ------------------------------------------------------- M > 2 components
def test_three_components_decoupled_equal_single_runs():
    """M = 3 with a diagonal (SVector) D, a potential and G_c depending on
    |u_c|^2 alone: each component evolves as its own one-component problem,
    which the scalar pins above fix (docs/src/multicomponent_systems.md: the
    tuple u0 = (psi_1, psi_2, ...) orders the components)."""
    N, L = 32, 6.0
    rng = np.random.default_rng(12)
    u0 = tuple(0.5 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) for _ in range(3))
    ds = [lambda k: (k[0] ** 2 + k[1] ** 2) / 2, lambda k: 0.3 * k[0] ** 2 - 0.1j, lambda k: 0.2 * k[1] ** 2 + 0.4]
    gs = [1.0, -0.5 + 0.05j, 0.3]
    vs = [0.1, 0.0, -0.2]
    prob = orc.Problem(u0, (L, L), dispersion=lambda k, p: [d(k) for d in ds],
                       nonlinearity=lambda u, p: [gs[c] * abs(u[c]) ** 2 for c in range(3)],
                       potential=lambda r, p: [vs[c] * r[0] + 0 * r[1] for c in range(3)])
    ts, res, _ = orc.solve(prob, (0, 0.5), 0.05, 2)
    for c in range(3):
        single = orc.Problem((u0[c],), (L, L), dispersion=lambda k, p, c=c: ds[c](k),
                             nonlinearity=lambda u, p, c=c: gs[c] * abs(u[0]) ** 2,
                             potential=lambda r, p, c=c: vs[c] * r[0] + 0 * r[1])
        _, r1, _ = orc.solve(single, (0, 0.5), 0.05, 2)
        assert np.max(abs(res[c] - r1[0])) <= 1e-13 * np.max(abs(r1[0]))


def test_four_components_coupled_norm_conserved():
    """M = 4 with real diagonal D, real cross-coupled G (sum_j g_cj |u_j|^2)
    and a real potential: every factor of the split step is unitary, so the
    norm of each component is conserved to rounding."""
    N, L = 64, 8.0
    rng = np.random.default_rng(13)
    u0 = tuple(rng.standard_normal(N) + 1j * rng.standard_normal(N) for _ in range(4))
    g = rng.uniform(-1, 1, (4, 4))
    prob = orc.Problem(u0, (L,), dispersion=lambda k, p: [(c + 1) * k[0] ** 2 / 4 for c in range(4)],
                       nonlinearity=lambda u, p: [sum(g[c, j] * abs(u[j]) ** 2 for j in range(4)) for c in range(4)],
                       potential=lambda r, p: [np.cos(r[0] * (c + 1)) for c in range(4)])
    ts, res, _ = orc.solve(prob, (0, 2.0), 0.01, 4)
    for c in range(4):
        n0 = np.sum(abs(u0[c]) ** 2)
        for f in range(len(ts)):
            assert abs(np.sum(abs(res[c][..., f]) ** 2) - n0) <= 1e-12 * n0


def test_three_components_coupled_dispersion_exact():
    """D(k) = |k|^2/2 I + W P with a non-symmetric 3 x 3 coupling P and
    nothing else: every Strang step is exp(-i dt D(k)) on the transforms, so
    u(t) is closed-form through the eigenvectors of P (independent of the
    oracle's matrix exponential; P != P^T, so a transposed operand fails)."""
    N, L, W = 32, 6.0, 0.7
    P = np.array([[0.0, 1.0, 0.3], [0.2, 0.5, -0.4], [0.9, -0.1, -0.2]])
    lam, V = np.linalg.eig(P)
    Vi = np.linalg.inv(V)
    rng = np.random.default_rng(5)
    u0 = tuple(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)) for _ in range(3))

    def disp(k, p):
        kk = (k[0] ** 2 + k[1] ** 2) / 2
        z = 0 * kk
        return [[kk * (a == b) + W * P[a, b] + z for b in range(3)] for a in range(3)]

    tspan, dt, nsaves = (0.0, 0.6), 0.05, 3
    ts, res, _ = orc.solve(orc.Problem(u0, (L, L), dispersion=disp), tspan, dt, nsaves)
    k = 2 * np.pi * np.fft.fftfreq(N, L / N)
    kk = (k[:, None] ** 2 + k[None, :] ** 2) / 2
    U0 = np.stack([np.fft.fft2(x) for x in u0])
    for f in range(1, nsaves + 1):
        t = ts[f]
        mix = V @ np.diag(np.exp(-1j * W * lam * t)) @ Vi
        want = np.exp(-1j * kk * t)[None] * np.einsum("ab,bxy->axy", mix, U0)
        for c in range(3):
            got = np.fft.fft2(res[c][..., f])
            assert np.linalg.norm(got - want[c]) <= 1e-12 * np.linalg.norm(want[c])
