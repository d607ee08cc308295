"""Uneven-size instance sweeps (gpe_sweep_*): every instance against the FP64
oracle run on its own (the reference's way: one solve per problem), and the
bistability hysteresis of a reduced uneven-size sweep."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def ramp_problem(ggp_or_orc, N, L, delta, gam, g, amp, is_oracle, noise=None, dtype=np.complex128, T=200.0):
    """1D homogeneous-pump polariton, D = k^2/2 - delta - i gamma, G = g|u|^2,
    pump amplitude amp * 3 (1 - |2t/T - 1|)."""
    def dispersion(k, p):
        return k[0] ** 2 / 2 - delta - 1j * gam

    def pump(r, p, t):
        return amp * 3.0 * (1 - np.abs(2 * np.asarray(t) / T - 1)) + 0 * r[0]

    kw = dict(dispersion=dispersion, nonlinearity=lambda u, p: g * abs(u[0]) ** 2, pump=pump)
    u0 = (np.zeros(N, dtype=dtype),)
    if is_oracle:
        if noise is not None:
            kw.update(position_noise_func=lambda u, r, p: noise[0], noise_prototype="complex", noise_seed=noise[1],
                      noise_pair_base=noise[2])
        return ggp_or_orc.Problem((np.zeros(N, complex),), (L,), **kw)
    if noise is not None:
        kw.update(position_noise_func=lambda u, r, p: noise[0], noise_prototype=(np.empty(N, dtype),))
    return ggp_or_orc.GrossPitaevskiiProblem(u0, (L,), **kw)


def test_sweep_instances_match_oracle(ggp, orc):
    sizes = [64, 16, 256, 1024, 64, 4096, 32, 512]
    rng = np.random.default_rng(44)
    params = [(N, 20.0 + 10 * i, rng.uniform(0.2, 1.5), 0.1 + 0.02 * i, 1.0 - 0.05 * i, 0.3 + 0.1 * i, 0.05 * (i % 3))
              for i, N in enumerate(sizes)]
    seed = 321
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    probs = [ramp_problem(ggp, N, L, d, gm, g, a, False, noise=(eta, seed, 0) if True else None)
             for (N, L, d, gm, g, a, eta) in params]
    tspan, dt, nsaves = (0, 6.0), 0.05, 6
    ts, res = ggp.solve_sweep(probs, ggp.StrangSplitting(), tspan, dt=dt, nsaves=nsaves, rng=ggp.PhiloxRNG(seed))
    worst = 0.0
    for i, (N, L, d, gm, g, a, eta) in enumerate(params):
        op = ramp_problem(orc, N, L, d, gm, g, a, True, noise=(eta, seed, int(off[i]) // 2))
        ts_o, res_o, _ = orc.solve(op, tspan, dt, nsaves)
        assert np.array_equal(ts, ts_o)
        assert res[i].shape == res_o[0].shape
        err = np.linalg.norm(res[i] - res_o[0]) / np.linalg.norm(res_o[0])
        worst = max(worst, err)
    assert worst <= 1e-12, worst


def test_sweep_without_frames_and_deterministic(ggp):
    probs = [ramp_problem(ggp, N, 50.0, 0.5, 0.1, 1.0, 0.5, False, dtype=np.complex64) for N in (256, 1024, 256)]
    ts, res = ggp.solve_sweep(probs, ggp.StrangSplitting(), (0, 2.0), dt=0.05, nsaves=2, save_frames=False)
    assert res is None and ts.shape == (3,)
    ts, res = ggp.solve_sweep(probs, ggp.StrangSplitting(), (0, 2.0), dt=0.05, nsaves=2)
    # identical instances of different offsets without noise evolve identically
    np.testing.assert_array_equal(res[0], res[2])


def test_sweep_bistability_hysteresis(ggp):
    """Uneven sizes drawn from {256..2048}, detuning delta_i in [0.4, 1.5],
    gamma = 0.1, g = 1, homogeneous pump 0 -> 3 -> 0 over 4000 steps.
    Inside an instance's bistable window n(gamma^2 + (g n - delta)^2) = A^2 the
    down-ramp density exceeds the up-ramp density; at the top of the ramp the
    field sits near the steady-state curve; by the end (pump back to 0, the
    upper branch surviving down to its lower turning point) it has fallen well
    below the top density."""
    rng = np.random.default_rng(0x5EED0004)
    n_inst = 24
    sizes = rng.choice([256, 512, 1024, 2048], n_inst)
    deltas = rng.uniform(0.4, 1.5, n_inst)
    gam, g, T = 0.1, 1.0, 200.0
    probs = [ramp_problem(ggp, int(N), 200.0, d, gam, g, 1.0, False, dtype=np.complex64, T=T)
             for N, d in zip(sizes, deltas)]
    ts, res = ggp.solve_sweep(probs, ggp.StrangSplitting(), (0, T), dt=0.05, nsaves=200)
    A = 3.0 * (1 - np.abs(2 * ts / T - 1))
    for i, d in enumerate(deltas):
        n = np.mean(np.abs(res[i]) ** 2, axis=0)
        assert np.all(np.isfinite(n))
        # homogeneous pump keeps the field homogeneous
        assert np.max(np.abs(np.abs(res[i][:, -50]) ** 2 - n[-50])) < 1e-3 * max(n[-50], 1e-6)
        # bistable window in n: dI/dn = 0 at n = (2 d +- sqrt(d^2 - 3 gam^2)) / (3 g)
        disc = math.sqrt(d * d - 3 * gam * gam)
        n_lo, n_hi = (2 * d - disc) / (3 * g), (2 * d + disc) / (3 * g)
        I = lambda x: x * (gam ** 2 + (g * x - d) ** 2)
        a_mid = math.sqrt(math.sqrt(I(n_lo) * I(n_hi)))
        up = np.argmin(np.abs(A[: len(A) // 2] - a_mid))
        down = len(A) // 2 + np.argmin(np.abs(A[len(A) // 2:] - a_mid))
        assert n[down] > 1.5 * n[up], (i, d, n[up], n[down])
        top = len(A) // 2
        assert abs(I(n[top]) / A[top] ** 2 - 1) < 0.15, (i, I(n[top]), A[top] ** 2)
        assert n[-1] < 0.5 * n[top]
