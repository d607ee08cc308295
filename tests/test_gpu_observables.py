"""Momentum-space occupation n(k) and pair correlation G2(k, k') of 1D
ensemble frames (gpe_kspace_g2) against the oracle's restatement of
examples/truncated_wigner.jl:110-165 on identical frames.

Bars: FP64 frames -> relative L2 (over all entries) <= 1e-12; FP32 frames
(transform in FP32, sums in FP64) -> <= 1e-5."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _handle(ggp, N, B, dtype, L=10.0):
    u0 = (np.zeros((N, B), dtype),)
    prob = ggp.GrossPitaevskiiProblem(u0, (L,), dispersion=lambda k, p: k[0] ** 2 / 2)
    return ggp.init(prob, ggp.StrangSplitting(), (0.0, 0.1), dt=0.05, nsaves=1)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("N,B,dtype,tol", [(16, 37, np.complex128, 1e-12), (64, 300, np.complex128, 1e-12),
                                           (128, 33, np.complex64, 1e-5), (256, 20, np.complex128, 1e-12),
                                           (2048, 19, np.complex128, 1e-12), (1024, 24, np.complex64, 1e-5)])
def test_kspace_g2_matches_oracle(ggp, orc, N, B, dtype, tol):
    rng = np.random.default_rng(N + B)
    frames = ((rng.standard_normal((N, B)) + 1j * rng.standard_normal((N, B))) * 0.3 + 0.5).astype(dtype)
    c = 1 / 10.0
    it = _handle(ggp, N, B, dtype)
    nk, G2 = it.kspace_correlations(frames, c)
    it.close()
    pairwise = N <= 128
    nk_o, G2_o = orc.kspace_g2(frames.astype(np.complex128), c, pairwise=pairwise)
    assert nk.shape == (N,) and G2.shape == (N, N)
    e_n, e_g = _rel(nk, nk_o), _rel(G2, G2_o)
    print(f"N={N} B={B} {np.dtype(dtype).name}: n(k) {e_n:.2e}, G2 {e_g:.2e}")
    assert e_n <= tol and e_g <= tol


def test_kspace_g2_device_frame_and_twa_vacuum(ggp, orc):
    """A device frame from a truncated-Wigner run gives the host frame's
    result bit for bit; for the undriven, undamped vacuum ensemble the
    normal-ordered moments stay zero within their statistical error."""
    import torch

    N, B, L = 64, 4096, 16.0
    dx = L / N
    c = 1 / dx  # [psi(x), psi^dagger(y)] = delta_xy / dx on the grid
    rng = np.random.default_rng(11)
    u = (rng.standard_normal((N, B)) + 1j * rng.standard_normal((N, B))) * np.sqrt(c / 4)
    prob = ggp.GrossPitaevskiiProblem((u,), (L,), dispersion=lambda k, p: k[0] ** 2 / 2)
    it = ggp.init(prob, ggp.StrangSplitting(), (0.0, 1.0), dt=0.05, nsaves=2)
    ts, res = it.solve()
    frame = np.asfortranarray(res[0][..., -1])
    ck = c / N  # commutator of ft = fft(u)/N: delta_kk' / (N dx) = 1/L
    nk_h, G2_h = it.kspace_correlations(frame, ck)
    dev = ggp.device_empty((N, B), "complex128", 0)
    dev.copy_(torch.from_numpy(frame))
    nk_d, G2_d = it.kspace_correlations(dev, ck)
    it.close()
    assert np.array_equal(nk_h, nk_d) and np.array_equal(G2_h, G2_d)
    nk_o, G2_o = orc.kspace_g2(frame, ck, pairwise=False)
    assert _rel(G2_h, G2_o) <= 1e-12 and _rel(nk_h, nk_o) <= 1e-12
    # free evolution keeps the vacuum: n(k) ~ 0, G2 ~ 0 (scale (c/2)^2 / sqrt(B))
    assert np.max(abs(nk_h)) < 6 * (ck / 2) / np.sqrt(B)
    assert np.max(abs(G2_h)) < 20 * (ck / 2) ** 2 / np.sqrt(B)


@pytest.mark.parametrize("N,B,dtype,tol", [(16, 21, np.complex128, 1e-12), (64, 300, np.complex128, 1e-12),
                                           (512, 17, np.complex64, 1e-5), (2048, 9, np.complex128, 1e-12)])
def test_windowed_correlation_matches_oracle(ggp, orc, N, B, dtype, tol):
    rng = np.random.default_rng(2 * N + B)
    frames = ((rng.standard_normal((N, B)) + 1j * rng.standard_normal((N, B))) * 0.3).astype(dtype)
    x = np.arange(N) / N
    w1 = np.exp(-((x - 0.4) / 0.2) ** 2)
    w2 = np.exp(-((x - 0.6) / 0.3) ** 2) * np.exp(0.7j * x)
    it = _handle(ggp, N, B, dtype)
    g = it.windowed_correlation(frames, w1, w2)
    it.close()
    g_o = orc.windowed_correlation(frames.astype(np.complex128), w1, w2)
    e = _rel(g, g_o)
    print(f"windowed N={N} B={B} {np.dtype(dtype).name}: {e:.2e}")
    assert g.shape == (N, N) and e <= tol


def test_reference_windowed_ft_vacuum(ggp, orc):
    """test/windowed_ft.jl end to end: the damped, noise-driven polariton
    field relaxes to the vacuum, and the windowed correlation of 10^4
    trajectories (computed on the device) matches the analytic commutator
    within the reference's rtol 7e-2 for its nine window pairs."""
    L, N, B = 20.0, 64, 10 ** 4
    dL = L / N
    hb = 0.6582
    gam, m, d0 = 0.1 / hb, hb ** 2 / 2.5, 0.49 / hb
    param = dict(hb=hb, gam=gam, m=m, d0=d0, dL=dL)
    u0 = (np.zeros((N, B), complex),)
    prob = ggp.GrossPitaevskiiProblem(
        u0, (L,), param=param,
        dispersion=lambda k, p: -1j * p["gam"] / 2 + p["hb"] * k[0] ** 2 / (2 * p["m"]) - p["d0"],
        position_noise_func=lambda u, r, p: np.sqrt(p["gam"] / 2 / p["dL"]),
        noise_prototype=(np.empty((N, B), complex),))
    it = ggp.init(prob, ggp.StrangSplitting(), (0.0, 200.0), dt=4.0, nsaves=1, save_start=False,
                  rng=ggp.PhiloxRNG(1234))
    ts, res = it.solve()
    frame = res[0][..., 0]
    rs = -L / 2 + np.arange(N) * dL
    worst = 0.0
    for x0 in (-1, 0, 1):
        for w in (3, 4, 5):
            f1 = lambda r, x0=x0, w=w: np.exp(-(r - x0) ** 2 / w ** 2)
            f2 = lambda r, x0=x0, w=w: np.exp(-(r + x0) ** 2 / w ** 2)
            corr = it.windowed_correlation(frame, f1(rs), f2(rs))
            an = orc.analytic_commutation(L, N, f1, f2)
            err = np.linalg.norm(corr - an) / max(np.linalg.norm(corr), np.linalg.norm(an))
            worst = max(worst, err)
            assert err <= 7e-2, (x0, w, err)
    it.close()
    print(f"windowed_ft vacuum: worst relative deviation {worst:.3f} (bar 7e-2)")


@pytest.mark.parametrize("shape,dtype,tol", [((64, 32, 3), np.complex128, 1e-12), ((256, 40), np.complex64, 1e-5),
                                             ((1024, 7), np.complex128, 1e-12), ((16, 16, 50), np.complex128, 1e-12)])
def test_density_g2_x_matches_oracle(ggp, orc, shape, dtype, tol):
    rng = np.random.default_rng(sum(shape))
    frames = ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 0.5 + 0.3).astype(dtype)
    nd = 2 if len(shape) == 3 else 1
    u0 = (np.zeros(shape, dtype),)
    prob = ggp.GrossPitaevskiiProblem(u0, (10.0,) * nd, dispersion=lambda k, p: sum(x ** 2 for x in k) / 2)
    it = ggp.init(prob, ggp.StrangSplitting(), (0.0, 0.1), dt=0.05, nsaves=1)
    c = 0.37
    G2, nphys, g2 = it.density_correlation_x(frames, c)
    it.close()
    G2_o, n_o = orc.density_g2_x(frames.astype(np.complex128), c)
    e = _rel(G2, G2_o)
    print(f"density g2 {shape} {np.dtype(dtype).name}: {e:.2e}")
    assert e <= tol and abs(nphys - n_o) <= tol * abs(n_o)
    assert np.allclose(g2, G2 / nphys ** 2)


def test_density_g2_x_twa_vacuum(ggp):
    """A free 2D vacuum ensemble through solve: the normal-ordered density
    correlation stays zero within its statistical error at every separation."""
    N, B, L = 32, 2048, 8.0
    dA = (L / N) ** 2
    c = 1 / dA
    rng = np.random.default_rng(13)
    u = (rng.standard_normal((N, N, B)) + 1j * rng.standard_normal((N, N, B))) * np.sqrt(c / 4)
    prob = ggp.GrossPitaevskiiProblem((u,), (L, L), dispersion=lambda k, p: (k[0] ** 2 + k[1] ** 2) / 2)
    it = ggp.init(prob, ggp.StrangSplitting(), (0.0, 1.0), dt=0.05, nsaves=1)
    ts, res = it.solve()
    G2, nphys, g2 = it.density_correlation_x(res[0][..., -1], c)
    it.close()
    scale = (c / 2) ** 2 / np.sqrt(N * N * B)
    assert abs(nphys) < 6 * (c / 2) / np.sqrt(N * N * B)
    assert np.max(abs(G2)) < 30 * scale
