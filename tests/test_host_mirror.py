"""Host-side mirror of the reference API: closure -> descriptor reduction
and the reference's argument checks (CPU only)."""
import math

import numpy as np
import pytest


def test_problem_checks_like_reference(ggp):
    # src/problem.jl:112-113
    with pytest.raises(AssertionError):
        ggp.GrossPitaevskiiProblem((np.zeros(8),), (1.0, 1.0))
    with pytest.raises(AssertionError):
        ggp.GrossPitaevskiiProblem((np.zeros((8, 8)), np.zeros((8, 4))), (1.0, 1.0))
    p = ggp.GrossPitaevskiiProblem((np.zeros((8, 8)),), (2, 3))
    assert p.u0[0].dtype == np.complex128 and p.lengths == (2.0, 3.0) and repr(p) == "2D GrossPitaevskiiProblem"
    p32 = ggp.GrossPitaevskiiProblem((np.zeros(8, np.float32),), (1,))
    assert p32.u0[0].dtype == np.complex64


def test_grids_match_reference_definitions(ggp):
    p = ggp.GrossPitaevskiiProblem((np.zeros((16, 8)),), (4.0, 2.0))
    x, y = ggp.direct_grid(p)
    assert np.array_equal(x, np.arange(16) * 0.25) and np.array_equal(y, np.arange(8) * 0.25)
    kx, ky = ggp.reciprocal_grid(p)
    assert np.array_equal(kx, np.r_[np.arange(8), np.arange(-8, 0)] * (2 * math.pi / 4.0))


def test_nonlinearity_fit(ggp):
    from gpe_hip_amd import problem as P

    g, c = P.fit_nonlinearity(lambda u, p: [0 * u[0], p * abs(u[1]) ** 2], 2, 0.015)
    assert np.allclose(g, [[0, 0], [0, 0.015]], atol=1e-17) and np.all(c == 0)
    g, c = P.fit_nonlinearity(lambda u, p: [[abs(u[0]) ** 2 - 0.5j, 0], [0, 2 * abs(u[0]) ** 2 + abs(u[1]) ** 2]],
                              2, None)
    assert np.allclose(g, [[1, 0], [2, 1]]) and np.allclose(c, [-0.5j, 0])
    g, c = P.fit_nonlinearity(lambda u, p: -6 * abs(u[0]) ** 2 / 2, 1, None)
    assert np.allclose(g, [[-3]]) and np.allclose(c, [0])
    with pytest.raises(NotImplementedError):
        P.fit_nonlinearity(lambda u, p: abs(u[0]) ** 4, 1, None)
    with pytest.raises(NotImplementedError):
        P.fit_nonlinearity(lambda u, p: [[0 * u[0], 1 + 0 * u[0]], [1 + 0 * u[0], 0 * u[0]]], 2, None)


def test_fit_nonlinearity_matrix(ggp):
    """An off-diagonal SMatrix G(u) (Rabi coupling plus cross-phase terms)
    reduces to G_ab = c_ab + sum_j g_abj |u_j|^2."""
    from gpe_hip_amd import problem as P

    def G(u, p):
        n0, n1 = abs(u[0]) ** 2, abs(u[1]) ** 2
        return [[p["g"] * n0 + p["g12"] * n1 - 0.1j, p["W"] + 0 * n0], [p["W"] + 0.2 * n1, p["g"] * n1 + p["g12"] * n0]]

    g, c = P.fit_nonlinearity_matrix(G, 2, dict(g=1.5, g12=0.5, W=0.3))
    want_c = np.array([[-0.1j, 0.3], [0.3, 0]])
    want_g = np.zeros((2, 2, 2), complex)
    want_g[0, 0] = [1.5, 0.5]
    want_g[1, 1] = [0.5, 1.5]
    want_g[1, 0] = [0, 0.2]
    assert np.allclose(c, want_c, atol=1e-13) and np.allclose(g, want_g, atol=1e-13)
    with pytest.raises(NotImplementedError):
        P.fit_nonlinearity_matrix(lambda u, p: [[abs(u[0]) ** 4, 0 * u[0]], [1 + 0 * u[0], 0 * u[0]]], 2, None)


def test_noise_eta(ggp):
    from gpe_hip_amd import problem as P

    axes = (np.arange(8) * 1.0,)
    assert np.allclose(P.noise_eta(lambda u, r, p: 0.25, 1, None, axes), [0.25])
    assert np.allclose(P.noise_eta(lambda u, r, p: [0.1, 0.2], 2, None, axes), [0.1, 0.2])
    with pytest.raises(NotImplementedError):
        P.noise_eta(lambda u, r, p: 0.1 * abs(u[0]), 1, None, axes)
    with pytest.raises(NotImplementedError):
        P.noise_eta(lambda u, r, p: 0.1j, 1, None, axes)


def test_pump_descriptor_separable_and_static(ggp):
    from gpe_hip_amd import _lib as Lb
    from gpe_hip_amd import problem as P

    N, L = 32, 8.0
    p = ggp.GrossPitaevskiiProblem((np.zeros((N, N)),) * 2, (L, L),
                                   pump=lambda r, q, t: [np.exp(-((r[0] - 4) ** 2 + (r[1] - 4) ** 2)), 0 * r[0]])
    rax = P._bcast_axes(ggp.direct_grid(p))
    kind, prof, env = P.pump_descriptor(p, rax, (N, N), np.linspace(0, 1, 21))
    assert kind == Lb.GPE_PUMP_SEPARABLE and env is None and prof.size == 2 * 2 * (N + N)
    # a non-separable profile falls back to the full table
    p2 = ggp.GrossPitaevskiiProblem((np.zeros((N, N)),), (L, L),
                                    pump=lambda r, q, t: np.exp(-((r[0] - 4) ** 2) * (1 + (r[1] - 4) ** 2)))
    kind, prof, env = P.pump_descriptor(p2, rax, (N, N), np.linspace(0, 1, 21))
    assert kind == Lb.GPE_PUMP_TABLE and prof.size == 2 * N * N
    # time dependence F = P(r) e(t) gives one envelope sample per pump time
    times = np.linspace(0, 2, 41)
    p3 = ggp.GrossPitaevskiiProblem((np.zeros(N),), (L,), pump=lambda r, q, t: np.exp(-r[0] ** 2) * math.sqrt(t))
    kind, prof, env = P.pump_descriptor(p3, P._bcast_axes(ggp.direct_grid(p3)), (N,), times)
    e = env.view(np.complex128)
    # profile taken at the probe time of largest norm (t = 2): F = P(r) e(t), e(2) = 1
    assert e.shape == (41,) and np.allclose(e, np.sqrt(times / 2.0))
    # not separable in time: sampled on the grid at every pump time, planes
    # in time order (sweeps keep the factored forms only)
    p4 = ggp.GrossPitaevskiiProblem((np.zeros(N),), (L,), pump=lambda r, q, t: np.exp(-(r[0] - t) ** 2))
    ax = P._bcast_axes(ggp.direct_grid(p4))
    kind, prof, env = P.pump_descriptor(p4, ax, (N,), times)
    assert kind == Lb.GPE_PUMP_SAMPLES and env is None
    x = ggp.direct_grid(p4)[0]
    want = np.exp(-(x[None, :] - times[:, None]) ** 2)
    assert np.array_equal(prof.view(np.complex128).reshape(len(times), N), want.astype(complex))
    with pytest.raises(NotImplementedError):
        P.pump_descriptor(p4, ax, (N,), times, allow_samples=False)


def test_sampled_pump_planes_are_column_major_per_component(ggp):
    """Two components, 2D: block c holds plane j of component c, x fastest."""
    from gpe_hip_amd import _lib as Lb
    from gpe_hip_amd import problem as P

    N0, N1, L = 8, 4, (4.0, 2.0)

    def pump(r, q, t):
        f = (r[0] + 10 * r[1]) * (1 + t) + 1j * t * r[0] ** 2
        return [f, 0 * f + 3 * t]

    p = ggp.GrossPitaevskiiProblem((np.zeros((N0, N1)),) * 2, L, pump=pump)
    times = np.array([0.0, 0.5, 1.0])
    kind, prof, env = P.pump_descriptor(p, P._bcast_axes(ggp.direct_grid(p)), (N0, N1), times)
    assert kind == Lb.GPE_PUMP_SAMPLES
    blk = prof.view(np.complex128).reshape(2, len(times), N0 * N1)
    x, y = ggp.direct_grid(p)
    for j, t in enumerate(times):
        f0 = (x[:, None] + 10 * y[None, :]) * (1 + t) + 1j * t * x[:, None] ** 2
        assert np.array_equal(blk[0, j], f0.ravel(order="F"))
        assert np.array_equal(blk[1, j], np.full(N0 * N1, 3 * t, complex))


def test_noise_prototype_without_noise_func_is_an_error(ggp):
    """_mul(AdditiveIdentity, SVector) has no method in the reference (Q9)."""
    prob = ggp.GrossPitaevskiiProblem((np.zeros(16, complex),), (1.0,), noise_prototype=(np.zeros(16, complex),))
    with pytest.raises(TypeError):
        ggp.init(prob, ggp.StrangSplitting(), (0, 1), dt=0.1, nsaves=1)
    with pytest.raises(TypeError):
        ggp.init(prob, object(), (0, 1), dt=0.1, nsaves=1)


def test_wigner_density_g2_closed_forms(ggp):
    """Symmetric-ordering correction of the truncated-Wigner moments on
    synthetic Wigner samples with known answers: vacuum (n = 0, no pairs),
    coherent state (g2 = 1), thermal state (g2 = 2)."""
    rng = np.random.default_rng(11)
    c = 1 / 0.0625  # commutator 1/dA
    S = 400000

    def cgauss(var):
        return np.sqrt(var / 2) * (rng.standard_normal(S) + 1j * rng.standard_normal(S))

    def moments(psi):
        a = np.abs(psi) ** 2
        return a.mean(), (a * a).mean()

    n, g2 = ggp.wigner_density_g2(*moments(cgauss(c / 2)), c)
    assert abs(n) < 0.01 * c
    alpha = 3.0 * np.sqrt(c)
    n, g2 = ggp.wigner_density_g2(*moments(alpha + cgauss(c / 2)), c)
    assert abs(n / alpha ** 2 - 1) < 0.01 and abs(g2 - 1) < 0.01
    nbar = 2.0 * c
    n, g2 = ggp.wigner_density_g2(*moments(cgauss(nbar + c / 2)), c)
    assert abs(n / nbar - 1) < 0.02 and abs(g2 - 2) < 0.05


def test_pump_descriptor_3d_is_a_column_major_table(ggp):
    """3D pumps go to the library as a full table (the separable form has
    per-axis factors for 1D and 2D only), flattened column-major like the
    field: entry (i, j, k) of component c at c * prod(n) + i + n0 (j + n1 k)."""
    from gpe_hip_amd import _lib as Lb
    from gpe_hip_amd import problem as P

    n, L = (16, 8, 4), (4.0, 2.0, 1.0)
    f = lambda r, q, t: [np.exp(-((r[0] - 2) ** 2 + 3 * (r[1] - 1) ** 2 + 5 * (r[2] - 0.5) ** 2)), 0 * r[0]]
    p = ggp.GrossPitaevskiiProblem((np.zeros(n),) * 2, L, pump=f)
    rax = P._bcast_axes(ggp.direct_grid(p))
    kind, prof, env = P.pump_descriptor(p, rax, n, np.linspace(0, 1, 9))
    assert kind == Lb.GPE_PUMP_TABLE and env is None
    tab = prof.view(np.complex128)
    assert tab.size == 2 * int(np.prod(n))
    want = f(tuple(np.meshgrid(*[np.arange(m) * (l / m) for m, l in zip(n, L)], indexing="ij")), None, 0)[0]
    i, j, k = 5, 3, 2
    assert tab[i + n[0] * (j + n[1] * k)] == want[i, j, k]
    assert not np.any(tab[int(np.prod(n)):])
