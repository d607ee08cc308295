"""Warm-start host pieces on the device: parameter reuse of one handle,
concurrent steady-state generation, and the warm solve with its cold
fallback (the network itself is checked in test_gpu_fno.py)."""
import math

import numpy as np
import pytest

from tests.test_gpu_parity import polariton_terms

pytestmark = pytest.mark.gpu

N, L = 64, 25.0


def _make(ggp, dtype=np.complex128):
    def make(p):
        t = polariton_terms(L, A=p["A"], w=p["w"])
        t["param"] = dict(t["param"], dc=p["d"], dx=p["d"])
        u0 = (np.zeros((N, N), dtype), np.zeros((N, N), dtype))
        return ggp.GrossPitaevskiiProblem(u0, (L, L), **t)
    return make


def _params(n, seed=5):
    rng = np.random.default_rng(seed)
    return [dict(A=rng.uniform(0.5, 1.5), w=rng.uniform(4, 8), d=rng.uniform(0, 1)) for _ in range(n)]


def _full(ggp, p):
    """Parameters as the problem closures read them."""
    t = polariton_terms(L, A=p["A"], w=p["w"])
    return dict(t["param"], dc=p["d"], dx=p["d"])


def test_set_param_matches_a_fresh_handle(ggp):
    make = _make(ggp)
    p1, p2 = _params(2)
    it = ggp.init(make(p1), ggp.StrangSplitting(), (0.0, 2.5), dt=0.05, nsaves=5)
    it.solve()
    it.set_param(_full(ggp, p2), ("dispersion", "pump"))
    ts_a, res_a = it.solve()
    it.close()
    ts_b, res_b = ggp.solve(make(p2), ggp.StrangSplitting(), (0.0, 2.5), dt=0.05, nsaves=5)
    assert np.array_equal(ts_a, ts_b)
    for c in range(2):
        assert np.array_equal(res_a[c], res_b[c])


def test_concurrent_generation_matches_sequential(ggp):
    from gpe_hip_amd import warmstart as W

    make = _make(ggp)
    ps = [_full(ggp, p) for p in _params(6, seed=9)]
    mk = lambda q: make(dict(A=q["A"], w=q["w"], d=q["dc"]))
    fields, info = W.generate_steady_states(mk, ps, dt=0.05, tol=1e-6, check_every=100, max_steps=40000, workers=3)
    for i, q in enumerate(ps):
        u, inf = ggp.steady_state(mk(q), ggp.StrangSplitting(), dt=0.05, tol=1e-6, check_every=100,
                                  max_steps=40000)
        assert inf["steps"] == info[i]["steps"]
        for c in range(2):
            assert np.array_equal(fields[i, c], u[c])


def test_solve_warm_converges_to_the_cold_state_and_falls_back(ggp):
    torch = pytest.importorskip("torch")
    from gpe_hip_amd import warmstart as W

    make = _make(ggp)
    ps = [_full(ggp, p) for p in _params(17, seed=13)]
    mk = lambda q: make(dict(A=q["A"], w=q["w"], d=q["dc"]))
    fields, _ = W.generate_steady_states(mk, ps[:16], dt=0.05, tol=1e-6, check_every=100, max_steps=40000,
                                         workers=4)
    x = np.arange(N) * (L / N)
    r2 = (x[:, None] - L / 2) ** 2 + (x[None, :] - L / 2) ** 2

    def feats(qs):
        X = np.empty((len(qs), 4, N, N), np.float32)
        for i, q in enumerate(qs):
            X[i, 0] = q["A"] * np.exp(-r2 / q["w"] ** 2)
            X[i, 1] = q["dc"]
        X[:, 2] = (x / L)[:, None]
        X[:, 3] = (x / L)[None, :]
        return torch.tensor(X, device="cuda:0")

    ws = W.WarmStart(N, 4, 4, width=8, modes=8, layers=2, seed=0)
    losses = ws.fit(feats(ps[:16]), torch.tensor(W.fields_to_channels(fields), device="cuda:0"), epochs=3, batch=4,
                    log_every=0)
    assert all(math.isfinite(v) for v in losses)
    it = ggp.init(mk(ps[16]), ggp.StrangSplitting(), (0.0, 0.05), dt=0.05, nsaves=1)
    uc, ic = it.steady_state(tol=1e-9, check_every=50, max_steps=400000)
    uw, iw = W.solve_warm(it, ws, feats(ps[16:]), tol=1e-9, check_every=50, max_steps=400000)
    assert iw["start"] == "warm" and iw["residual"] <= 1e-9
    nrm = math.sqrt(sum(np.linalg.norm(v) ** 2 for v in uc))
    assert math.sqrt(sum(np.linalg.norm(a - b) ** 2 for a, b in zip(uw, uc))) <= 1e-5 * nrm
    # a budget too small for the warm run: the cold rerun is taken (and also stops at the budget)
    uf, inf = W.solve_warm(it, ws, feats(ps[16:]), tol=1e-9, check_every=10, max_steps=20)
    assert inf["start"] == "cold-fallback" and inf["steps"] == 20
    it.close()
