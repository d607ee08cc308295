"""GPU parity: libgpe_hip (through the C ABI, via the host mirror) against the
FP64 oracle on identical inputs.

Bars: FP64 fields -> relative L2 <= 1e-12 on every saved frame; FP32 fields
-> relative L2 against the FP64 oracle <= 1e-3 (the measured value of each
case is recorded in gpurun_out/parity_errors.jsonl and stated in the docs); ts
bit-identical.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-3


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def run_both(ggp, orc, u0, lengths, tspan, dt, nsaves, dtype=np.complex128, save_start=True, seed=None, planner=None,
             **terms):
    ts_o, res_o, _ = orc.solve(orc.Problem(u0, lengths, noise_seed=seed or 0, **_oracle_terms(terms)), tspan, dt,
                               nsaves, save_start)
    gu0 = tuple(np.asfortranarray(x.astype(dtype)) for x in u0)
    gterms = dict(terms)
    if "noise_prototype" in gterms:
        gterms["noise_prototype"] = tuple(np.empty(x.shape, gterms.pop("noise_dtype", np.complex128)) for x in u0)
    else:
        gterms.pop("noise_dtype", None)
    rng = ggp.PhiloxRNG(seed or 0) if seed is not None else None
    it = ggp.init(ggp.GrossPitaevskiiProblem(gu0, lengths, **gterms), ggp.StrangSplitting(), tspan, dt=dt,
                  nsaves=nsaves, save_start=save_start, rng=rng)
    if planner is not None:
        it.set_planner(planner)
    ts, res = it.solve()
    it.close()
    return ts_o, res_o, ts, res


def _oracle_terms(terms):
    t = dict(terms)
    if "noise_prototype" in t:
        t["noise_prototype"] = "real" if np.dtype(t.pop("noise_dtype", np.complex128)).kind == "f" else "complex"
    else:
        t.pop("noise_dtype", None)
    return t


def check(ts_o, res_o, ts, res, tol):
    assert np.array_equal(np.asarray(ts), np.asarray(ts_o))
    worst = 0.0
    for c in range(len(res_o)):
        assert res[c].shape == res_o[c].shape
        for f in range(res_o[c].shape[-1]):
            worst = max(worst, rel_l2(res[c][..., f].astype(np.complex128), res_o[c][..., f]))
    _record(worst, res[0].dtype)
    assert worst <= tol, f"rel L2 {worst:.3e} > {tol:.1e}"
    return worst


def _record(err, dtype):
    import json
    import os

    os.makedirs("gpurun_out", exist_ok=True)
    name = os.environ.get("PYTEST_CURRENT_TEST", "?").split(" ")[0]
    with open("gpurun_out/parity_errors.jsonl", "a") as f:
        f.write(json.dumps({"test": name, "dtype": str(dtype), "rel_l2": err}) + "\n")


# ------------------------------------------------------------ C1: README case
def c1_inputs(N=128):
    L = 8.0
    rs = np.arange(N) * (L / N)
    u0 = (np.exp(-(rs[:, None] - 4) ** 2 - (rs[None, :] - 4) ** 2).astype(complex),)
    return u0, (L, L)


def disp_kin(k, p):
    return sum(x ** 2 for x in k) / 2


@pytest.mark.parametrize("dtype,tol", [(np.complex128, FP64_TOL), (np.complex64, FP32_TOL)])
def test_c1_readme_kerr(ggp, orc, dtype, tol):
    """128^2, L=8, D=|k|^2/2, G=-6|u|^2, dt 0.01 -> 0.00625, 64 frames."""
    u0, lengths = c1_inputs()
    out = run_both(ggp, orc, u0, lengths, (0, 0.4), 0.01, 64, dtype=dtype, dispersion=disp_kin,
                   nonlinearity=lambda u, p: p["g"] * abs(u[0]) ** 2, param={"g": -6.0})
    err = check(*out, tol)
    print(f"C1 {np.dtype(dtype).name}: rel L2 {err:.3e}")


# --------------------------------------------- two components, 2x2 dispersion
def polariton_terms(L, A=1.0, w=None):
    w = w or L / 4

    def dispersion(k, p):
        kk = sum(x ** 2 for x in k)
        return [[p["hb"] * kk / (2 * p["m"]) - p["dc"] - 1j * p["gc"], p["W"]],
                [p["W"], -p["dx"] - 1j * p["gx"]]]

    def nonlinearity(u, p):
        return [0 * u[0], p["g"] * abs(u[1]) ** 2]

    def pump(r, p, t):
        f = p["A"] * np.exp(-sum((x - p["L"] / 2) ** 2 for x in r) / p["w"] ** 2)
        return [f, 0 * f]

    param = dict(hb=0.6582, m=0.18194, dc=0.0, dx=0.0, gc=0.05, gx=0.005, W=2.5, g=0.015, A=A, w=w, L=L)
    return dict(dispersion=dispersion, nonlinearity=nonlinearity, pump=pump, param=param)


@pytest.mark.parametrize("N", [64, 128])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, FP64_TOL), (np.complex64, FP32_TOL)])
def test_two_component_pumped(ggp, orc, N, dtype, tol):
    L = 100.0 * N / 512
    u0 = (np.zeros((N, N), complex), np.zeros((N, N), complex))
    out = run_both(ggp, orc, u0, (L, L), (0, 10.0), 0.05, 10, dtype=dtype, **polariton_terms(L))
    err = check(*out, tol)
    print(f"two-component {N}^2 {np.dtype(dtype).name}: rel L2 {err:.3e}")


# --------------------------------------------------- 1D, time-dependent pump
def test_1d_ramped_pump(ggp, orc):
    """Bistability-style 1D problem with the pump ramp (pump times follow the
    reference's t+dt/2, t+dt after t += dt)."""
    N, L = 256, 256.0
    p = dict(g=0.01, d=0.3, gam=0.1, w0=1483.0, kz=27.0, Imax=0.6, width=50.0, tmax=60.0, L=L)

    def disp(k, p):
        return -1j * p["gam"] / 2 + p["w0"] * (1 + k[0] ** 2 / (2 * p["kz"] ** 2)) - (p["w0"] + p["d"])

    def I(t):
        v = -p["Imax"] * t * (t - p["tmax"]) * 4 / p["tmax"] ** 2
        return np.maximum(v, 0.0)

    def pump(x, p, t):
        return np.exp(-((x[0] - p["L"] / 2) ** 2) / p["width"] ** 2) * np.sqrt(I(t))

    u0 = (np.zeros(N, complex),)
    out = run_both(ggp, orc, u0, (L,), (0, 60), 0.05, 12, dispersion=disp,
                   nonlinearity=lambda u, p: p["g"] * abs(u[0]) ** 2, pump=pump, param=p)
    check(*out, FP64_TOL)


# ------------------------------------------------------- every FFT length
@pytest.mark.parametrize("N", [16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_1d_sizes_batched(ggp, orc, N):
    L = float(N) / 8
    rng = np.random.default_rng(N)
    B = 3
    u0 = (rng.standard_normal((N, B)) + 1j * rng.standard_normal((N, B)),)
    out = run_both(ggp, orc, u0, (L,), (0, 0.5), 0.05, 2, dispersion=disp_kin,
                   nonlinearity=lambda u, p: 0.3 * abs(u[0]) ** 2)
    check(*out, FP64_TOL)


@pytest.mark.parametrize("shape", [(16, 16), (32, 64), (64, 32), (256, 16), (16, 512), (1024, 32)])
def test_2d_shapes(ggp, orc, shape):
    rng = np.random.default_rng(sum(shape))
    u0 = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape),)
    lengths = (shape[0] / 4.0, shape[1] / 2.0)
    out = run_both(ggp, orc, u0, lengths, (0, 0.2), 0.02, 2, dispersion=disp_kin,
                   nonlinearity=lambda u, p: 0.1 * abs(u[0]) ** 2)
    check(*out, FP64_TOL)


def test_2d_batched_two_component(ggp, orc):
    rng = np.random.default_rng(7)
    shape = (32, 32, 3)
    u0 = tuple(rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(2))
    out = run_both(ggp, orc, u0, (8.0, 8.0), (0, 0.4), 0.05, 2, **polariton_terms(8.0))
    check(*out, FP64_TOL)


def test_2d_batched_wide_column_shape(ggp, orc):
    """Enough trajectories that the column pass takes its wide workgroup shape
    (16 columns of 128 for 16 x 128 grids: 2048 workgroups), with the 2x2
    dispersion exchange between components."""
    rng = np.random.default_rng(21)
    shape = (16, 128, 2048)
    u0 = tuple((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 0.1 for _ in range(2))
    out = run_both(ggp, orc, u0, (4.0, 32.0), (0, 0.1), 0.05, 1, planner="estimate", **polariton_terms(32.0))
    check(*out, FP64_TOL)


# -------------------------------------------------------------- noise (TWA)
@pytest.mark.parametrize("noise_dtype", [np.complex128, np.float64])
def test_noise_1d_philox(ggp, orc, noise_dtype):
    N, B, L = 64, 8, 20.0
    gam = 0.1
    u0 = (np.zeros((N, B), complex),)
    out = run_both(ggp, orc, u0, (L,), (0, 4.0), 0.1, 4, seed=1234,
                   dispersion=lambda k, p: -1j * gam / 2 + k[0] ** 2 / 2,
                   position_noise_func=lambda u, r, p: math.sqrt(gam / 2 / (L / N)),
                   noise_prototype=True, noise_dtype=noise_dtype)
    check(*out, FP64_TOL)


def test_noise_2d_batched_kerr(ggp, orc):
    N, B, L = 32, 4, 8.0
    u0 = (np.full((N, N, B), 0.5 + 0j),)
    out = run_both(ggp, orc, u0, (L, L), (0, 1.0), 0.05, 2, seed=99,
                   dispersion=lambda k, p: (k[0] ** 2 + k[1] ** 2) / 2 - 0.5 - 0.1j,
                   nonlinearity=lambda u, p: abs(u[0]) ** 2, pump=lambda r, p, t: 1.2 + 0 * r[0],
                   position_noise_func=lambda u, r, p: 0.3, noise_prototype=True)
    check(*out, FP64_TOL)


# ---------------------------------------------------------------- potential
@pytest.mark.parametrize("ncomp", [1, 2])
def test_potential(ggp, orc, ncomp):
    N, L = 64, 10.0
    rng = np.random.default_rng(3)
    u0 = tuple(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)) for _ in range(ncomp))

    def pot(r, p):
        v = 0.5 * ((r[0] - 5) ** 2 + (r[1] - 5) ** 2)
        return v if ncomp == 1 else [v, 2 * v - 0.01j]

    out = run_both(ggp, orc, u0, (L, L), (0, 0.3), 0.01, 3, dispersion=disp_kin, potential=pot)
    check(*out, FP64_TOL)


def test_save_start_false_and_frames(ggp, orc):
    u0, lengths = c1_inputs(64)
    out = run_both(ggp, orc, u0, lengths, (0, 0.2), 0.01, 5, save_start=False, dispersion=disp_kin)
    check(*out, FP64_TOL)


# ------------------------------------------ general two-component forms
def test_two_component_nonsymmetric_coupled(ggp, orc):
    """Full non-symmetric 2x2 D(k) with a k-linear (odd) coupling, so neither
    the symmetric nor the folded table applies; cross-phase nonlinearity
    couples |u_1|^2 and |u_2|^2; noise on both components."""
    N, B, L = 32, 2, 8.0
    rng = np.random.default_rng(5)
    u0 = tuple(rng.standard_normal((N, N, B)) * 0.3 + 1j * rng.standard_normal((N, N, B)) * 0.3 for _ in range(2))

    def disp(k, p):
        kk = k[0] ** 2 + k[1] ** 2
        return [[kk / 2 - 0.1j, 0.4 * k[0] + 0.2], [0.4 * k[0] - 0.3j, kk / 3 - 0.05j]]

    out = run_both(ggp, orc, u0, (L, L), (0, 0.6), 0.05, 3, seed=77, dispersion=disp,
                   nonlinearity=lambda u, p: [abs(u[0]) ** 2 + 0.5 * abs(u[1]) ** 2,
                                              0.3 * abs(u[0]) ** 2 + abs(u[1]) ** 2 - 0.02j],
                   position_noise_func=lambda u, r, p: [0.1, 0.2], noise_prototype=True)
    check(*out, FP64_TOL)


def test_two_component_1d_full_dispersion(ggp, orc):
    N, B, L = 128, 4, 30.0
    rng = np.random.default_rng(6)
    u0 = tuple(rng.standard_normal((N, B)) + 1j * rng.standard_normal((N, B)) for _ in range(2))
    out = run_both(ggp, orc, u0, (L,), (0, 1.0), 0.02, 4, **polariton_terms(L))
    check(*out, FP64_TOL)


# ------------------------------------------------ ensemble observables
def test_ensemble_moments_match_oracle(ggp, orc):
    N, B, L = 32, 16, 8.0
    u0 = (np.full((N, N, B), 0.3 + 0.1j),)
    terms = dict(dispersion=lambda k, p: (k[0] ** 2 + k[1] ** 2) / 2 - 0.5 - 0.05j,
                 nonlinearity=lambda u, p: abs(u[0]) ** 2, pump=lambda r, p, t: 1.2 + 0 * r[0],
                 position_noise_func=lambda u, r, p: 0.3)
    ts_o, res_o, _ = orc.solve(orc.Problem(u0, (L, L), noise_prototype="complex", noise_seed=5, **terms),
                               (0, 2.0), 0.05, 4)
    prob = ggp.GrossPitaevskiiProblem(u0, (L, L), noise_prototype=(np.empty((N, N, B), complex),), **terms)
    for frames in (True, False):
        ts, res, (n1, n2) = ggp.solve(prob, ggp.StrangSplitting(), (0, 2.0), dt=0.05, nsaves=4,
                                      rng=ggp.PhiloxRNG(5), save_frames=frames, ensemble_moments=True)
        assert (res is None) != frames
        a2 = abs(res_o[0]) ** 2
        want1 = a2.mean(axis=2)
        want2 = (a2 ** 2).mean(axis=2)
        assert n1[0].shape == want1.shape
        assert np.max(abs(n1[0] - want1)) <= 1e-12 * np.max(want1)
        assert np.max(abs(n2[0] - want2)) <= 1e-12 * np.max(want2)
        if frames:
            assert rel_l2(res[0], res_o[0]) <= FP64_TOL


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_twa_stationary_vacuum_occupation(ggp, dtype):
    """Truncated Wigner vacuum: D = |k|^2/2 - delta - i gamma/2 with
    eta = sqrt(gamma / (2 dA)) keeps <|psi|^2> = 1/(2 dA) (the pairing of
    test/windowed_ft.jl:24,28), starting from the Wigner vacuum itself."""
    N, B, L = 64, 256, 16.0
    dA = (L / N) ** 2
    gam, delta = 0.1, 0.5
    rng = np.random.default_rng(9)
    u0 = ((rng.standard_normal((N, N, B)) + 1j * rng.standard_normal((N, N, B))) / math.sqrt(2) / math.sqrt(2 * dA),)
    prob = ggp.GrossPitaevskiiProblem(tuple(x.astype(dtype) for x in u0), (L, L),
                                      dispersion=lambda k, p: (k[0] ** 2 + k[1] ** 2) / 2 - delta - 0.5j * gam,
                                      position_noise_func=lambda u, r, p: math.sqrt(gam / (2 * dA)),
                                      noise_prototype=(np.empty((N, N, B), dtype),))
    ts, res, (n1, n2) = ggp.solve(prob, ggp.StrangSplitting(), (0, 40.0), dt=0.05, nsaves=4,
                                  rng=ggp.PhiloxRNG(2024), save_frames=False, ensemble_moments=True)
    target = 1 / (2 * dA)
    for f in range(n1[0].shape[-1]):
        assert abs(n1[0][..., f].mean() / target - 1) < 0.02
        # complex Gaussian field: <|psi|^4> = 2 <|psi|^2>^2
        assert abs(n2[0][..., f].mean() / (2 * n1[0][..., f].mean() ** 2) - 1) < 0.03
        # symmetric ordering undone (examples/truncated_wigner.jl:110-165): no particles, no pairs
        n_phys, _ = ggp.wigner_density_g2(n1[0][..., f].mean(), n2[0][..., f].mean(), 1 / dA)
        assert abs(n_phys) < 0.02 * target


# ------------------------------------------------------------ step graph
def _solve_with(ggp, prob, tspan, dt, nsaves, schedule, save_start=True, moments=False):
    it = ggp.init(prob, ggp.StrangSplitting(), tspan, dt=dt, nsaves=nsaves, save_start=save_start)
    try:
        it.set_schedule(schedule)
        out = it.solve(ensemble_moments=moments)
        return out, it.last_schedule
    finally:
        it.close()


def _graph_case(name):
    if name == "kerr64_f64":
        u0, lengths = c1_inputs(64)
        terms = dict(dispersion=disp_kin, nonlinearity=lambda u, p: -6.0 * abs(u[0]) ** 2)
        return u0, lengths, terms, np.complex128, (0, 0.5), 0.01, 10, False
    if name == "polariton128_f32":
        L = 100.0 * 128 / 512
        u0 = (np.zeros((128, 128), complex), np.zeros((128, 128), complex))
        return u0, (L, L), polariton_terms(L), np.complex64, (0, 5.0), 0.05, 10, False
    # batched 32x32 x 4, potential, ensemble moments of every frame
    N, B, L = 32, 4, 8.0
    rs = np.arange(N) * (L / N)
    base = np.exp(-(rs[:, None] - 4) ** 2 - (rs[None, :] - 3) ** 2).astype(complex)
    u0 = (np.stack([base * (1 + 0.1 * b) for b in range(B)], axis=2),)
    terms = dict(dispersion=lambda k, p: (k[0] ** 2 + k[1] ** 2) / 2 - 0.3j,
                 potential=lambda r, p: 0.5 * ((r[0] - 4) ** 2 + (r[1] - 4) ** 2),
                 nonlinearity=lambda u, p: abs(u[0]) ** 2)
    return u0, (L, L), terms, np.complex128, (0, 1.0), 0.02, 10, True


@pytest.mark.parametrize("save_start", [True, False])
@pytest.mark.parametrize("case", ["kerr64_f64", "polariton128_f32", "batched32_f64_potential_moments"])
def test_step_graph_matches_launches_bit_for_bit(ggp, orc, case, save_start):
    """The replayed step graph issues the per-step launches with the same
    arguments, so frames, ts and ensemble moments agree bit for bit with the
    launch-per-step schedule (every frame is one replay), and the FP64 frames
    meet the oracle bar."""
    u0, lengths, terms, dtype, tspan, dt, nsaves, moments = _graph_case(case)
    prob = ggp.GrossPitaevskiiProblem(tuple(np.asfortranarray(x.astype(dtype)) for x in u0), lengths, **terms)
    a, sched_a = _solve_with(ggp, prob, tspan, dt, nsaves, "launches", save_start, moments)
    b, sched_b = _solve_with(ggp, prob, tspan, dt, nsaves, "graph", save_start, moments)
    assert (sched_a, sched_b) == ("launches", "graph")
    assert np.array_equal(a[0], b[0])
    for x, y in zip(a[1], b[1]):
        assert np.array_equal(x, y)
    if moments:
        for pa, pb in zip(a[2], b[2]):
            for x, y in zip(pa, pb):
                assert np.array_equal(x, y)
    if dtype == np.complex128:
        ts_o, res_o, _ = orc.solve(orc.Problem(u0, lengths, **terms), tspan, dt, nsaves, save_start)
        check(ts_o, res_o, b[0], b[1], FP64_TOL)


def test_step_graph_schedule_rules(ggp):
    """AUTO replays a graph from 8 save blocks of >= 2 steps on; an explicit
    graph schedule refuses noise (each step draws new Philox counters)."""
    u0, lengths = c1_inputs(32)
    prob = ggp.GrossPitaevskiiProblem(u0, lengths, dispersion=disp_kin,
                                      nonlinearity=lambda u, p: -1.0 * abs(u[0]) ** 2)
    for nsaves, want in ((8, "graph"), (4, "launches")):
        it = ggp.init(prob, ggp.StrangSplitting(), (0, 0.16), dt=0.01, nsaves=nsaves)
        it.solve()
        assert it.last_schedule == want, (nsaves, it.last_schedule)
        it.close()
    it = ggp.init(prob, ggp.StrangSplitting(), (0, 0.08), dt=0.01, nsaves=8)  # one step per save
    it.solve()
    assert it.last_schedule == "launches"
    it.close()
    noisy = ggp.GrossPitaevskiiProblem(u0, lengths, dispersion=disp_kin, position_noise_func=lambda u, r, p: 0.1,
                                       noise_prototype=(np.empty((32, 32), complex),))
    it = ggp.init(noisy, ggp.StrangSplitting(), (0, 0.16), dt=0.01, nsaves=8, rng=ggp.PhiloxRNG(1))
    it.set_schedule("graph")
    with pytest.raises(ggp.GpeUnsupported):
        it.solve()
    it.set_schedule("auto")
    it.solve()
    assert it.last_schedule == "launches"
    it.close()


# ------------------------------------------------------------ measured plans
@pytest.mark.parametrize("case", ["kerr64_f64", "polariton128_f32", "batched32_f64_potential_moments"])
def test_measured_plan_matches_estimate_bit_for_bit(ggp, case):
    """GPE_PLAN_MEASURE times every valid launch shape on the handle and keeps
    the fastest; the shapes change no arithmetic, so every plan, measured or
    imposed, gives the built-in plan's frames bit for bit."""
    u0, lengths, terms, dtype, tspan, dt, nsaves, _ = _graph_case(case)
    prob = ggp.GrossPitaevskiiProblem(tuple(np.asfortranarray(x.astype(dtype)) for x in u0), lengths, **terms)
    it = ggp.init(prob, ggp.StrangSplitting(), tspan, dt=dt, nsaves=nsaves)
    it.set_planner("estimate")
    ref = it.solve()
    est = it.plan
    assert est["measured_us"] == {}
    it.set_planner("measure")
    out = it.solve()
    plan = it.plan
    meas = plan["measured_us"]
    assert (plan["x_rows"], plan["y_cols"]) == min(meas, key=meas.get)
    assert all(v > 0 for v in meas.values())
    for x, y in zip(ref[1], out[1]):
        assert np.array_equal(x, y)
    shapes = list(meas)
    assert len(shapes) >= 2
    for w, c in shapes:
        it.set_plan(w, c)
        got = it.solve()
        assert it.plan["x_rows"] == w and it.plan["y_cols"] == c
        for x, y in zip(ref[1], got[1]):
            assert np.array_equal(x, y), (w, c)
    with pytest.raises(ggp.GpeError):
        it.set_plan(3, 0)
    it.close()


def test_step_graphs_on_concurrent_handles(ggp):
    """Distinct handles on several host threads, each building and replaying
    step graphs (solve and steady state) while the main thread keeps the
    default stream busy with synchronous copies: every run succeeds and equals
    its sequential twin bit for bit (the graphs are built node by node, so no
    stream is ever in capture mode)."""
    import threading

    import torch

    u0, lengths = c1_inputs(32)
    probs = [ggp.GrossPitaevskiiProblem(u0, lengths, dispersion=disp_kin,
                                        nonlinearity=lambda u, p, g=g: g * abs(u[0]) ** 2) for g in (-1.0, -2.0, 0.5, 1.5)]

    def run(prob):
        it = ggp.init(prob, ggp.StrangSplitting(), (0, 0.4), dt=0.005, nsaves=10)
        it.set_schedule("graph")
        ts, res = it.solve()
        u, info = it.steady_state(tol=1e-300, check_every=8, max_steps=64)
        it.close()
        return res[0], u[0], info["history"]

    seq = [run(p) for p in probs]
    out = [None] * len(probs)
    errs = []

    def worker(i):
        try:
            for _ in range(3):
                out[i] = run(probs[i])
        except Exception as e:  # reported below
            errs.append(e)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(probs))]
    for t in threads:
        t.start()
    x = torch.zeros(1 << 20, device="cuda")
    while any(t.is_alive() for t in threads):
        x.cpu()  # synchronous device-to-host copies on the default stream
    for t in threads:
        t.join()
    assert not errs, errs
    for (a1, b1, h1), (a2, b2, h2) in zip(seq, out):
        assert np.array_equal(a1, a2) and np.array_equal(b1, b2) and np.array_equal(h1, h2)


def test_auto_planner_measures_long_runs_once_per_process(ggp):
    """AUTO (the default) keeps the built-in shapes for a short run, measures
    for a run of >= 1000 steps, and a second handle of the same shape and terms
    reuses that plan without measuring; the frames agree bit for bit."""
    rng = np.random.default_rng(5)
    u0 = (np.asfortranarray(rng.standard_normal((16, 32, 3)) + 1j * rng.standard_normal((16, 32, 3))) * 0.2,)
    prob = ggp.GrossPitaevskiiProblem(u0, (4.0, 8.0), dispersion=disp_kin,
                                      nonlinearity=lambda u, p: 0.7 * abs(u[0]) ** 2)
    it = ggp.init(prob, ggp.StrangSplitting(), (0, 0.5), dt=0.01, nsaves=5)  # 50 steps
    short = it.solve()
    assert it.plan["measured_us"] == {}
    it.close()
    it = ggp.init(prob, ggp.StrangSplitting(), (0, 10.0), dt=0.01, nsaves=10)  # 1000 steps
    it.solve()
    first = it.plan
    assert first["measured_us"] and (first["x_rows"], first["y_cols"]) == min(first["measured_us"],
                                                                                key=first["measured_us"].get)
    it.close()
    it = ggp.init(prob, ggp.StrangSplitting(), (0, 10.0), dt=0.01, nsaves=10)
    it.solve()
    again = it.plan
    it.close()
    assert again["measured_us"] == {} and (again["x_rows"], again["y_cols"]) == (first["x_rows"], first["y_cols"])
    it = ggp.init(prob, ggp.StrangSplitting(), (0, 0.5), dt=0.01, nsaves=5)
    it.solve()  # a short run now reuses the kept plan
    assert (it.plan["x_rows"], it.plan["y_cols"]) == (first["x_rows"], first["y_cols"])
    res = it.solve()
    it.close()
    assert np.array_equal(short[1][0], res[1][0])


# ------------------------------------------------ matrix nonlinearity (SMatrix G)
def rabi_terms(L, W=0.4, gamma=0.05):
    """Two coupled condensates: G = [[g n0 + g12 n1 - i gamma, W], [W, g n1 + g12 n0]]
    (an SMatrix nonlinearity, src/kernels.jl:44), a Gaussian pump on
    component 0 and the free dispersion."""
    def nonlinearity(u, p):
        n0, n1 = abs(u[0]) ** 2, abs(u[1]) ** 2
        return [[p["g"] * n0 + p["g12"] * n1 - 1j * p["gam"], p["W"] + 0 * n0],
                [p["W"] + 0 * n0, p["g"] * n1 + p["g12"] * n0]]

    def pump(r, p, t):
        f = 0.8 * np.exp(-sum((x - p["L"] / 2) ** 2 for x in r) / (p["L"] / 5) ** 2)
        return [f, 0 * f]

    return dict(dispersion=disp_kin, nonlinearity=nonlinearity, pump=pump,
                param=dict(g=1.0, g12=0.6, W=W, gam=gamma, L=L))


@pytest.mark.parametrize("dtype,tol", [(np.complex128, FP64_TOL), (np.complex64, FP32_TOL)])
@pytest.mark.parametrize("save_start", [True, False])
def test_matrix_nonlinearity_matches_oracle(ggp, orc, dtype, tol, save_start):
    N, L = 64, 12.0
    rng = np.random.default_rng(17)
    u0 = tuple(0.3 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) for _ in range(2))
    out = run_both(ggp, orc, u0, (L, L), (0, 1.0), 0.02, 5, dtype=dtype, save_start=save_start, **rabi_terms(L))
    err = check(*out, tol)
    print(f"matrix G {np.dtype(dtype).name}: rel L2 {err:.3e}")


@pytest.mark.parametrize("dtype,tol", [(np.complex128, FP64_TOL), (np.complex64, FP32_TOL)])
def test_matrix_nonlinearity_1d_batched_matches_oracle(ggp, orc, dtype, tol):
    N, B, L = 128, 3, 12.0
    rng = np.random.default_rng(23)
    u0 = tuple(0.3 * (rng.standard_normal((N, B)) + 1j * rng.standard_normal((N, B))) for _ in range(2))
    out = run_both(ggp, orc, u0, (L,), (0, 1.0), 0.02, 5, dtype=dtype, **rabi_terms(L))
    err = check(*out, tol)
    print(f"matrix G 1D {np.dtype(dtype).name}: rel L2 {err:.3e}")


@pytest.mark.parametrize("noise_dtype,shape", [(np.complex128, (64, 64)), (np.float64, (128, 4))])
def test_matrix_nonlinearity_with_noise_matches_oracle(ggp, orc, noise_dtype, shape):
    """Philox noise with a matrix G, drawn with the fused kernels' layout, in
    2D (complex noise) and 1D batched (real noise)."""
    nd = 2 if shape == (64, 64) else 1
    L = 12.0
    rng = np.random.default_rng(29)
    u0 = tuple(0.3 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(2))
    out = run_both(ggp, orc, u0, (L,) * nd, (0, 0.4), 0.02, 4, seed=77, noise_prototype=True,
                   noise_dtype=noise_dtype, position_noise_func=lambda u, r, p: [0.2, 0.1], **rabi_terms(L))
    err = check(*out, FP64_TOL)
    print(f"matrix G + noise {nd}D: rel L2 {err:.3e}")


@pytest.mark.parametrize("form", ["scalar", "svector"])
def test_matrix_nonlinearity_with_potential_matches_oracle(ggp, orc, form):
    """A potential with a matrix G follows the reference's product rules: a
    scalar V scales exp(-i dt/2 G); an SVector V turns it into a vector that
    multiplies the fields elementwise."""
    N, L = 64, 12.0
    rng = np.random.default_rng(31)
    u0 = tuple(0.3 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) for _ in range(2))
    if form == "scalar":
        pot = lambda r, p: 0.05 * ((r[0] - 6) ** 2 + (r[1] - 6) ** 2)
    else:
        pot = lambda r, p: [0.05 * (r[0] - 6) ** 2, 0.03 * (r[1] - 6) ** 2 - 0.01j]
    out = run_both(ggp, orc, u0, (L, L), (0, 0.4), 0.02, 4, potential=pot, **rabi_terms(L))
    err = check(*out, FP64_TOL)
    print(f"matrix G + {form} V: rel L2 {err:.3e}")
