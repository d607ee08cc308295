"""GPU parity for pumps F(r, t) that do not factor into profile(r) * envelope(t).

The reference evaluates the pump on the grid before every half-step
(evaluate_pump!, src/misc.jl:34-42) for any closure.  The mirror sends such a
pump as its samples at every pump time (GPE_PUMP_SAMPLES) and the half-steps
read F_now and F_next from them.  Each case moves a Gaussian spot across the
box with a time-dependent phase, so no single profile fits two times, and
compares every saved frame with the FP64 oracle (bars as in
test_gpu_parity.py).  The cases cover the fused 2D row pass (single and
batched grids), the 1D whole-step kernel with Philox noise, the 3D schedule
and the per-point matrix-nonlinearity kernel.
"""
import numpy as np
import pytest

from test_gpu_parity import FP32_TOL, FP64_TOL, check, disp_kin, polariton_terms, rabi_terms, run_both

pytestmark = pytest.mark.gpu


def moving_spot(L, nd, comp=0, M=1, v=1.5, w=None, amp=0.8):
    """F_comp(r, t) = amp exp(-|r - c(t)|^2 / w^2 + i t), c(t) moving along x."""
    w = w or L / 6

    def pump(r, p, t):
        d2 = (r[0] - L / 4 - v * t) ** 2
        for x in r[1:nd]:
            d2 = d2 + (x - L / 2) ** 2
        f = amp * np.exp(-d2 / w ** 2 + 1j * t)
        out = [0 * f] * M
        out[comp] = f
        return out if M > 1 else f

    return pump


@pytest.mark.parametrize("dtype,tol", [(np.complex128, FP64_TOL), (np.complex64, FP32_TOL)])
def test_2d_two_component_moving_pump(ggp, orc, dtype, tol):
    N = 64
    L = 100.0 * N / 512
    terms = polariton_terms(L)
    terms["pump"] = moving_spot(L, 2, comp=0, M=2, v=0.8, amp=1.0)
    u0 = (np.zeros((N, N), complex), np.zeros((N, N), complex))
    out = run_both(ggp, orc, u0, (L, L), (0, 4.0), 0.05, 8, dtype=dtype, **terms)
    err = check(*out, tol)
    print(f"moving pump 2D {np.dtype(dtype).name}: rel L2 {err:.3e}")


def test_2d_batched_moving_pump_kerr(ggp, orc):
    """Trajectories share the sampled planes: the plane offset is the line's
    spatial index, not its index over the batch."""
    N, B, L = 64, 3, 10.0
    rng = np.random.default_rng(31)
    u0 = (0.2 * (rng.standard_normal((N, N, B)) + 1j * rng.standard_normal((N, N, B))),)
    out = run_both(ggp, orc, u0, (L, L), (0, 0.6), 0.02, 3, dispersion=disp_kin,
                   nonlinearity=lambda u, p: 1.0 * abs(u[0]) ** 2 - 0.1j, pump=moving_spot(L, 2), param={})
    err = check(*out, FP64_TOL)
    print(f"moving pump 2D batched: rel L2 {err:.3e}")


@pytest.mark.parametrize("dtype,tol", [(np.complex128, FP64_TOL), (np.complex64, FP32_TOL)])
def test_1d_batched_moving_pump_with_noise(ggp, orc, dtype, tol):
    """1D whole-step kernel (lines in registers for a save block) with
    Philox noise: sample indices advance inside the kernel."""
    N, B, L = 256, 4, 40.0
    rng = np.random.default_rng(37)
    u0 = (0.1 * (rng.standard_normal((N, B)) + 1j * rng.standard_normal((N, B))),)
    out = run_both(ggp, orc, u0, (L,), (0, 2.0), 0.02, 5, dtype=dtype, seed=91, noise_prototype=True,
                   position_noise_func=lambda u, r, p: 0.05,
                   dispersion=lambda k, p: k[0] ** 2 / 2 - 0.5 - 0.1j,
                   nonlinearity=lambda u, p: abs(u[0]) ** 2, pump=moving_spot(L, 1, v=4.0), param={})
    err = check(*out, tol)
    print(f"moving pump 1D + noise {np.dtype(dtype).name}: rel L2 {err:.3e}")


def test_3d_moving_pump(ggp, orc):
    shape, L = (32, 16, 16), (6.0, 4.0, 5.0)

    def pump(r, p, t):
        d2 = (r[0] - 1.5 - 2.0 * t) ** 2 + (r[1] - 2.0) ** 2 + (r[2] - 2.5 + t) ** 2
        return 0.7 * np.exp(-d2 + 0.5j * t)

    rng = np.random.default_rng(41)
    u0 = (0.2 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)),)
    out = run_both(ggp, orc, u0, L, (0, 0.4), 0.02, 4, dispersion=lambda k, p: 0.5 * sum(x ** 2 for x in k) - 0.05j,
                   nonlinearity=lambda u, p: 0.5 * abs(u[0]) ** 2, pump=pump, param={})
    err = check(*out, FP64_TOL)
    print(f"moving pump 3D: rel L2 {err:.3e}")


@pytest.mark.parametrize("nd", [1, 2])
def test_matrix_nonlinearity_moving_pump(ggp, orc, nd):
    N, L = (128, 12.0) if nd == 1 else (64, 12.0)
    shape = (N, 3) if nd == 1 else (N, N)
    rng = np.random.default_rng(43)
    u0 = tuple(0.3 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(2))
    terms = rabi_terms(L)
    terms["pump"] = moving_spot(L, nd, comp=0, M=2, v=3.0)
    out = run_both(ggp, orc, u0, (L,) * nd, (0, 1.0), 0.02, 5, **terms)
    err = check(*out, FP64_TOL)
    print(f"matrix G, moving pump {nd}D: rel L2 {err:.3e}")


def test_sampled_pump_error_paths(ggp):
    """A sampled pump covers one solve: the steady-state loop refuses it, and a
    table whose sample count does not match the solve is rejected."""
    import ctypes

    N, L = 32, 8.0
    prob = ggp.GrossPitaevskiiProblem((np.zeros((N, N), np.complex128),), (L, L), dispersion=disp_kin,
                                      pump=moving_spot(L, 2), param={})
    with pytest.raises(Exception, match="static"):
        ggp.steady_state(prob, ggp.StrangSplitting(), dt=0.05, max_steps=200, check_every=50)
    it = ggp.init(prob, ggp.StrangSplitting(), (0.0, 1.0), dt=0.1, nsaves=5)
    lib = ggp.problem.L.load()
    one = np.zeros(2 * N * N)
    rc = lib.gpe_set_pump(it.h, ggp.problem.L.GPE_PUMP_SAMPLES, one.ctypes.data, None, 1)
    assert rc == 0
    u = np.zeros((N, N), np.complex128, order="F")
    out = np.zeros((N, N), np.complex128, order="F")
    ptrs = (ctypes.c_void_p * 1)(u.ctypes.data)
    res = (ctypes.c_void_p * 1)(out.ctypes.data)
    ts = np.zeros(2)
    rc = lib.gpe_solve(it.h, ptrs, ctypes.c_double(0.0), ctypes.c_double(0.1), ctypes.c_double(0.1), 1, 0, res,
                       ts.ctypes.data)
    assert rc != 0 and "sampled pump has 1 samples, this solve needs 3" in lib.gpe_last_error().decode()
    rc = lib.gpe_set_pump(it.h, ggp.problem.L.GPE_PUMP_SAMPLES, one.ctypes.data, one.ctypes.data, 1)
    assert rc != 0 and "no envelope" in lib.gpe_last_error().decode()
    it.close()
