"""Steady-state residual (gpe_steady_state): the residual the library forms
against the one computed from the oracle's consecutive frames, and a
converged run whose final field equals a plain solve of the same length."""
import numpy as np
import pytest

from tests.test_gpu_parity import polariton_terms

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ndim", [1, 2])
def test_residual_matches_oracle(ggp, orc, ndim):
    N = 64
    L = 25.0
    shape = (N,) * ndim
    u0 = (np.zeros(shape, complex), np.zeros(shape, complex))
    terms = polariton_terms(L, A=1.0, w=6.0)
    if ndim == 1:
        t2 = dict(terms)
        t2["pump"] = lambda r, p, t: [p["A"] * np.exp(-((r[0] - p["L"] / 2) ** 2) / p["w"] ** 2), 0 * r[0]]
        terms = t2
    dt, k = 0.05, 40
    prob = ggp.GrossPitaevskiiProblem(u0, (L,) * ndim, **terms)
    it = ggp.init(prob, ggp.StrangSplitting(), (0.0, dt), dt=dt, nsaves=1)
    u, info = it.steady_state(tol=1e-300, check_every=k, max_steps=2 * k)
    it.close()
    assert info["steps"] == 2 * k and len(info["history"]) == 2
    _, res_o, _ = orc.solve(orc.Problem(u0, (L,) * ndim, **terms), (0, 2 * k * dt), dt, 2 * k)
    for j, s in enumerate((k, 2 * k)):
        a = np.concatenate([res_o[c][..., s].ravel() for c in range(2)])
        b = np.concatenate([res_o[c][..., s - 1].ravel() for c in range(2)])
        r = np.linalg.norm(a - b) / np.linalg.norm(a)
        assert abs(info["history"][j] / r - 1) < 1e-9
    for c in range(2):
        assert np.linalg.norm(u[c] - res_o[c][..., -1]) <= 1e-12 * np.linalg.norm(res_o[c][..., -1])


def test_steady_state_converges(ggp):
    N, L = 64, 25.0
    u0 = (np.zeros((N, N), complex), np.zeros((N, N), complex))
    prob = ggp.GrossPitaevskiiProblem(u0, (L, L), **polariton_terms(L, A=1.0, w=6.0))
    u, info = ggp.steady_state(prob, ggp.StrangSplitting(), dt=0.05, tol=1e-8, check_every=100, max_steps=400000)
    assert info["residual"] <= 1e-8 and info["steps"] < 400000
    h = info["history"]
    assert h[-1] < 1e-3 * h[0]
    # near a fixed point: a residual r per step leaves a distance of about
    # r / (gamma_min dt) = 1e-8 / (0.005 * 0.05) = 4e-5 along the slowest
    # (exciton-like) modes, so 100 more steps move the field by less than that
    ts, res = ggp.solve(ggp.GrossPitaevskiiProblem(u, (L, L), **polariton_terms(L, A=1.0, w=6.0)),
                        ggp.StrangSplitting(), (0, 5.0), dt=0.05, nsaves=1)
    for c in range(2):
        assert np.linalg.norm(res[c][..., -1] - u[c]) <= 2e-4 * np.linalg.norm(u[c])


def test_steady_state_graph_replay_is_bit_identical(ggp):
    """Without noise the steps of a check block before the two compared
    fields replay as one prebuilt graph; residual history, step count and
    final field equal the launch-per-step run bit for bit."""
    N, L = 64, 25.0
    u0 = (np.zeros((N, N), complex), np.zeros((N, N), complex))
    prob = ggp.GrossPitaevskiiProblem(u0, (L, L), **polariton_terms(L, A=1.0, w=6.0))
    out = {}
    for sched in ("launches", "auto"):
        it = ggp.init(prob, ggp.StrangSplitting(), (0.0, 0.05), dt=0.05, nsaves=1)
        it.set_schedule(sched)
        u, info = it.steady_state(tol=1e-300, check_every=25, max_steps=200)
        out[sched] = (u, info, it.last_schedule)
        it.close()
    (ua, ia, sa), (ub, ib, sb) = out["launches"], out["auto"]
    assert (sa, sb) == ("launches", "graph")
    assert ia["steps"] == ib["steps"] == 200
    assert np.array_equal(ia["history"], ib["history"])
    for c in range(2):
        assert np.array_equal(ua[c], ub[c])
