"""Throughput of the Strang-splitting step on one MI355X.

Workload (default): the two-component exciton-polariton problem on a 512x512
grid, L = 100 um, 2x2 dispersion [hbar k^2/2m - dc - i gc, W; W, -dx - i gx]
(hbar = 0.6582 meV ps, m = 3.2e-5 m_e = 0.181940 meV ps^2 um^-2, dc = dx = 0,
gc = 0.05, gx = 0.005, W = 2.5), exciton nonlinearity g|u2|^2 with g = 0.015,
Gaussian pump of waist 25 um and amplitude 1 on the photon, u0 = 0, dt = 0.05,
one saved frame per 100 steps (the 10^4-step run saves 100 frames).  Fields
are complex FP32 (bulk stepping); the FP64 run of the same workload is
reported beside it.

One bench "step" is one full Strang step of the whole 512x512x2 field; the
timed region is `--steps` steps of solve(prob, StrangSplitting(), ...) with
the initial field and the result frames resident in HBM.

Printed: one JSON line with metric "grid-point-steps/s", a `roofline` object
for the dominant kernel (HIP-event timed durations, algorithmic bytes per
launch) and a `cpu_baseline` object (the FP64 numpy oracle on a bounded
sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
SAVE_EVERY = 100
N_GRID = 512
L_BOX = 100.0


def polariton_param():
    me = 5685.63  # electron mass in meV ps^2 um^-2
    return dict(hb=0.6582, m=3.2e-5 * me, dc=0.0, dx=0.0, gc=0.05, gx=0.005, W=2.5, g=0.015, A=1.0, w=25.0,
                L=L_BOX)


def dispersion(k, p):
    kk = k[0] ** 2 + k[1] ** 2
    return [[p["hb"] * kk / (2 * p["m"]) - p["dc"] - 1j * p["gc"], p["W"]], [p["W"], -p["dx"] - 1j * p["gx"]]]


def nonlinearity(u, p):
    return [0 * u[0], p["g"] * abs(u[1]) ** 2]


def pump(r, p, t):
    f = p["A"] * np.exp(-((r[0] - p["L"] / 2) ** 2 + (r[1] - p["L"] / 2) ** 2) / p["w"] ** 2)
    return [f, 0 * f]


def terms():
    return dict(dispersion=dispersion, nonlinearity=nonlinearity, pump=pump, param=polariton_param())


def read_traffic(dtype, kernel):
    """HBM bytes per launch of `kernel` ("x_pass" / "y_pass") from the
    committed rocprofv3 PMC summary (None when absent)."""
    path = os.path.join(ROOT, "profiles", "x_pass_y_pass_512sq_2comp_pmc.md")
    if not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_launch[{dtype},{kernel}]"
    for line in open(path):
        if line.strip().startswith(key):
            return float(line.split("=")[1].split()[0])
    return None


def bracket_overhead(raw_x, raw_y, step_ms):
    """Per-launch latency added by bracketing every launch with HIP events.
    A step without events is one x-pass plus one y-pass; with events each
    launch carries the same extra o, so raw_x + raw_y = step + 2 o."""
    o = max(0.0, (raw_x + raw_y - step_ms) / 2)
    return o, o


def gpu_run(g, torch, dtype, steps, warmup, dt=0.05, plan=None):
    """Returns dict with wall/device timing and per-kernel durations.  plan =
    (x_rows, y_cols) fixes the launch shapes; None lets the first handle
    measure them and reuses its choice."""
    shape = (N_GRID, N_GRID)
    u0 = []
    for _ in range(2):
        t = g.device_empty(shape, dtype, 0)
        t.zero_()
        u0.append(t)
    prob = g.GrossPitaevskiiProblem(tuple(u0), (L_BOX, L_BOX), **terms())

    def make_iter(nsteps):
        nsaves = nsteps // SAVE_EVERY if nsteps % SAVE_EVERY == 0 else 1
        return g.init(prob, g.StrangSplitting(), (0.0, nsteps * dt), dt=dt, nsaves=nsaves)

    it_w = make_iter(warmup)
    if plan is not None:
        it_w.set_plan(*plan)
    else:
        it_w.set_planner("measure")
    it_w.solve()  # measures the launch plan unless one is given
    plan = it_w.plan
    it_w.close()
    it = make_iter(steps)
    it.set_plan(plan["x_rows"], plan["y_cols"])
    res = tuple(g.device_empty(shape + (it.nsaves + 1,), dtype, 0) for _ in range(2))
    it.solve(result=res)  # untimed: first touch of the frame buffers
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it.solve(result=res)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = it.last_ms()
    sched = it.last_schedule
    # per-kernel durations: HIP events around every launch of 200 more steps
    import ctypes

    msx, msy, mse = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    g._lib.check(g.load_library().gpe_profile_kernels(it.h, 200, ctypes.byref(msx), ctypes.byref(msy),
                                                      ctypes.byref(mse)))
    finite = all(bool(torch.isfinite(r).all()) for r in res)
    it.close()
    ox, oy = bracket_overhead(msx.value, msy.value, dev_ms / steps)
    return dict(wall=wall, dev_ms=dev_ms, ms_x=msx.value - ox, ms_y=msy.value - oy, plan=plan, schedule=sched,
                ms_x_raw=msx.value, ms_y_raw=msy.value, ms_event=mse.value, ms_overhead=ox, finite=finite)


def twa_run(g, torch, steps, batch=1024, n=256, dtype="complex64"):
    """Batched truncated-Wigner ensemble: `batch` trajectories of the driven-
    dissipative 1-component polariton on an n x n grid, L = 64 (dA = 1/16),
    D = |k|^2/2 - delta - i gamma (delta = 0.5, gamma = 0.1), g = 1,
    homogeneous pump F = 1.2, eta = sqrt(gamma / (2 dA)), Philox noise,
    u0 = Wigner vacuum <|u0|^2> = 1/(2 dA), dt = 0.05, ensemble moments every
    100 steps (no frames).  Returns timing and the last frame's mean density."""
    L = 64.0
    dA = (L / n) ** 2
    gam, delta = 0.1, 0.5
    shape = (n, n, batch)
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(0x5EED0003)
    u = g.device_empty(shape, dtype, 0)
    rdt = torch.float32 if dtype == "complex64" else torch.float64
    re = torch.randn(tuple(reversed(shape)), generator=gen, device="cuda:0", dtype=rdt)
    im = torch.randn(tuple(reversed(shape)), generator=gen, device="cuda:0", dtype=rdt)
    u.copy_(torch.complex(re, im).permute(2, 1, 0) / math.sqrt(2) / math.sqrt(2 * dA))
    del re, im
    proto = (u,)
    prob = g.GrossPitaevskiiProblem((u,), (L, L),
                                    dispersion=lambda k, p: (k[0] ** 2 + k[1] ** 2) / 2 - delta - 1j * gam,
                                    nonlinearity=lambda v, p: abs(v[0]) ** 2, pump=lambda r, p, t: 1.2 + 0 * r[0],
                                    position_noise_func=lambda v, r, p: math.sqrt(gam / (2 * dA)),
                                    noise_prototype=proto)
    dt = 0.05
    nsaves = max(1, steps // SAVE_EVERY)
    it = g.init(prob, g.StrangSplitting(), (0.0, steps * dt), dt=dt, nsaves=nsaves, rng=g.PhiloxRNG(0x5EED0003))
    it_w = g.init(prob, g.StrangSplitting(), (0.0, 20 * dt), dt=dt, nsaves=1, rng=g.PhiloxRNG(1))
    it_w.set_planner("measure")
    it_w.solve(save_frames=False, ensemble_moments=True)  # also measures the launch plan
    plan = it_w.plan
    it.set_plan(plan["x_rows"], plan["y_cols"])  # reuse it (FFTW-wisdom style)
    it_w.close()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ts, _, (n1, n2) = it.solve(save_frames=False, ensemble_moments=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    import ctypes

    # g2(r) along x of the final ensemble field, reduced on the device (untimed above)
    fld = g.device_empty((n, n, batch), dtype, 0)
    g._lib.check(g.load_library().gpe_get_field(it.h, g._lib.ptr_array([fld.data_ptr()])))
    torch.cuda.synchronize()
    t_g2 = time.perf_counter()
    G2, nphys, g2r = it.density_correlation_x(fld, 1 / dA)
    t_g2 = time.perf_counter() - t_g2
    del fld
    msx, msy, mse = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    g._lib.check(g.load_library().gpe_profile_kernels(it.h, 20, ctypes.byref(msx), ctypes.byref(msy),
                                                      ctypes.byref(mse)))
    it.close()
    ox, _ = bracket_overhead(msx.value, msy.value, wall * 1e3 / steps)
    msx.value -= ox
    msy.value -= ox
    npts = n * n * batch
    csz = 8 if dtype == "complex64" else 16
    launch_bytes = 2 * csz * npts
    dom = max(msx.value, msy.value)
    return dict(workload=f"truncated-Wigner ensemble, {batch} x {n}x{n}, M=1, Philox noise, moments every 100 steps",
                value=npts * steps / wall, unit="grid-point-steps/s", ms_per_step=wall * 1e3 / steps, steps=steps,
                dtype="c64" if dtype == "complex64" else "c128",
                hbm_equiv_frac=npts * steps / wall * 2 * 2 * csz / 1e9 / HBM_PEAK_GBS,
                x_pass_us=msx.value * 1e3, y_pass_us=msy.value * 1e3,
                dominant_kernel_GBs=launch_bytes / (dom * 1e-3) / 1e9,
                dominant_kernel_frac=launch_bytes / (dom * 1e-3) / 1e9 / HBM_PEAK_GBS,
                mean_density_last_frame=float(n1[0][..., -1].mean()), vacuum_density=1 / (2 * dA),
                g2_r_along_x=dict(r=[0, 1, 2, 4, 8, 16], g2=[float(g2r[r]) for r in (0, 1, 2, 4, 8, 16)],
                                  n_phys=nphys, device_ms=t_g2 * 1e3,
                                  note="the workload pairs amplitude damping gamma with noise sqrt(gamma/(2 dA)), "
                                       "which holds <|psi|^2>_W near 1/(4 dA), below the Wigner vacuum "
                                       "1/(2 dA) (the vacuum pairing is -i gamma/2 with that noise, "
                                       "test/windowed_ft.jl:24,28); the ordering-corrected density is then "
                                       "negative and g2 is reported as computed, not as a physical "
                                       "quantity"),
                launch_plan=dict(x_rows=plan["x_rows"], y_cols=plan["y_cols"]))


def sweep_run(g, torch, ninst=4096, steps=4000, dtype=np.complex64):
    """Uneven-size bistability sweep: `ninst` 1D instances with N_i drawn from
    {256, 512, 1024, 2048} and detuning delta_i from U[0.2, 1.5] (seed
    0x5EED0004), L = 200, gamma = 0.1, g = 1, D = k^2/2 - delta_i - i gamma,
    homogeneous pump amplitude 0 -> 3 -> 0 over the run, dt = 0.05, one frame
    per 20 steps (instance-major frames in host memory, copied once)."""
    rng = np.random.default_rng(0x5EED0004)
    sizes = rng.choice([256, 512, 1024, 2048], ninst)
    deltas = rng.uniform(0.2, 1.5, ninst)
    dt = 0.05
    T = steps * dt
    gam = 0.1
    probs = []
    for N, d in zip(sizes, deltas):
        probs.append(g.GrossPitaevskiiProblem(
            (np.zeros(int(N), dtype=dtype),), (200.0,),
            dispersion=(lambda dd: (lambda k, p: k[0] ** 2 / 2 - dd - 1j * gam))(d),
            nonlinearity=lambda u, p: abs(u[0]) ** 2,
            pump=lambda r, p, t: 3.0 * (1 - np.abs(2 * np.asarray(t) / T - 1)) + 0 * r[0]))
    t_setup = time.perf_counter()
    it = g.SweepIterator(probs, (0.0, T), dt, steps // 20)
    t_setup = time.perf_counter() - t_setup
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ts, res = it.solve(save_frames=False)
    wall = time.perf_counter() - t0
    dev_ms = it.last_ms()
    it.close()
    pts = int(sizes.sum())
    return dict(workload=f"uneven-size bistability sweep, {ninst} x 1D N in {{256..2048}}, {steps} steps, "
                         "one launch per length and 20-step block", value=pts * steps / wall,
                unit="grid-point-steps/s", ms_total=wall * 1e3, device_ms=dev_ms, steps=steps, points=pts,
                dtype="c64" if dtype == np.complex64 else "c128", host_setup_s=t_setup,
                bound="VALU (each line stays in registers for 20 steps; HBM traffic 16 B per point per 20 steps)")


C5_SEED = 0x5EED0005
BF16_DENSE_PEAK_TFS = 2500.0


def c5_param(A, w, delta):
    """The polariton model above with pump amplitude A, waist w and
    detuning delta = dc = dx."""
    p = polariton_param()
    p.update(A=float(A), w=float(w), dc=float(delta), dx=float(delta))
    return p


def c5_draw(n, rng):
    return [c5_param(rng.uniform(0.5, 2.5), rng.uniform(10, 40), rng.uniform(0, 1.5)) for _ in range(n)]


def c5_features(params, n_grid=256, L=L_BOX):
    """Input maps of the surrogate: pump profile A exp(-|r - c|^2 / w^2),
    detuning, x / L, y / L (float32, (n, 4, N, N), axis 0 of a plane = x)."""
    x = np.arange(n_grid) * (L / n_grid)
    X = np.empty((len(params), 4, n_grid, n_grid), dtype=np.float32)
    r2 = (x[:, None] - L / 2) ** 2 + (x[None, :] - L / 2) ** 2
    for i, p in enumerate(params):
        X[i, 0] = p["A"] * np.exp(-r2 / p["w"] ** 2)
        X[i, 1] = p["dc"]
    X[:, 2] = (x / L)[:, None]
    X[:, 3] = (x / L)[None, :]
    return X


def band_residual(u, m):
    """||u - P u|| / ||u|| for the complex fields u (tuple of (N, N)), P the
    retained-mode low-pass applied to Re and Im planes (kx in {0..m-1,
    N-m..N-1}, ky in {0..m-1}, Hermitian completion): how much of a steady
    state lies outside the surrogate's band."""
    num = den = 0.0
    for f in u:
        for plane in (f.real, f.imag):
            N = plane.shape[0]
            F = np.fft.fft2(plane)
            kx = np.r_[np.arange(m), np.arange(N - m, N)]
            G = np.zeros_like(F)
            G[kx, :m] = F[kx, :m] * np.where(np.arange(m) == 0, 1.0, 2.0)
            num += np.linalg.norm(plane - np.fft.ifft2(G).real) ** 2
            den += np.linalg.norm(plane) ** 2
    return math.sqrt(num / den)


def band_limited(u, m):
    """The retained-mode low-pass of band_residual, applied to complex fields."""
    out = []
    for f in u:
        parts = []
        for plane in (f.real, f.imag):
            N = plane.shape[0]
            F = np.fft.fft2(plane)
            kx = np.r_[np.arange(m), np.arange(N - m, N)]
            G = np.zeros_like(F)
            G[kx, :m] = F[kx, :m] * np.where(np.arange(m) == 0, 1.0, 2.0)
            parts.append(np.fft.ifft2(G).real)
        out.append(parts[0] + 1j * parts[1])
    return tuple(out)


def warmstart_diagnosis(it, ws, Xte, params, torch):
    """Steps to residual <= 1e-8 from controlled starts around the converged
    state u*: the network's prediction with and without the band limit, u*
    band-limited, u* (1 + 1e-2), and u* plus broadband noise of relative
    size 1e-3.  Separates the error the refinement pays for."""
    rows = []
    for i, p in enumerate(params):
        it.set_param(p, ("dispersion", "pump"))
        us, _ = it.steady_state(tol=1e-10, check_every=50, max_steps=400000)
        nrm = math.sqrt(sum(np.linalg.norm(x) ** 2 for x in us))
        r = np.random.default_rng(i)
        noise = [r.standard_normal(x.shape) + 1j * r.standard_normal(x.shape) for x in us]
        s = 1e-3 * nrm / math.sqrt(sum(np.linalg.norm(x) ** 2 for x in noise))
        starts = {
            "cold": None,
            "fno_band_limited": tuple(x[0] for x in ws.predict_fields(Xte[i:i + 1], dtype=torch.complex128)),
            "fno_raw": tuple(x[0] for x in ws.predict_fields(Xte[i:i + 1], dtype=torch.complex128,
                                                             band_limit=False)),
            "exact_band_limited": band_limited(us, ws.net.m),
            "exact_x_1.01": tuple(x * 1.01 for x in us),
            "exact_plus_1e-3_broadband": tuple(x + s * n for x, n in zip(us, noise)),
        }
        row = dict(A=p["A"], w=p["w"], delta=p["dc"], out_of_band=band_residual(us, ws.net.m))
        for k, u0 in starts.items():
            _, inf = it.steady_state(u0=u0, tol=1e-8, check_every=50, max_steps=400000)
            row[k] = inf["steps"]
        rows.append(row)
        print(f"  controlled starts {row}", flush=True)
    return rows


def c5_problem(g, dtype, n_grid=256, L=L_BOX):
    def make(p):
        u0 = (np.zeros((n_grid, n_grid), dtype), np.zeros((n_grid, n_grid), dtype))
        return g.GrossPitaevskiiProblem(u0, (L, L), dispersion=dispersion, nonlinearity=nonlinearity, pump=pump,
                                        param=p)
    return make


def warmstart_run(g, torch, args):
    """Learned warm start on the polariton model at 256x256 (L = 100): steady
    states for n_train drawn (A ~ U[0.5, 2.5], w ~ U[10, 40], delta ~ U[0, 1.5];
    seed 0x5EED0005) from u0 = 0 in FP32 (tol 3e-6, 8 handles in flight),
    FNO (width 32, 16 modes, 4 layers) trained on them, then n_test unseen
    cases run to residual <= 1e-8 in FP64 twice: cold from u0 = 0 and warm
    from the FNO prediction.  Time = device time of the refinement (+ the
    FNO inference for warm), both checked every 50 steps."""
    from gpe_hip_amd import warmstart as W

    rng = np.random.default_rng(C5_SEED)
    n_tr, n_te = args.ws_train, args.ws_test
    p_train = c5_draw(n_tr, rng)
    p_test = c5_draw(n_te, rng)
    N = 256
    rec = {}
    if args.ws_load:
        ws = W.WarmStart.load(args.ws_load)
        rec["train"] = dict(loaded=args.ws_load)
    else:
        print(f"[warmstart] generating {n_tr} FP32 steady states", flush=True)
        t0 = time.perf_counter()
        fields, info = W.generate_steady_states(c5_problem(g, np.complex64), p_train, dt=0.05, tol=3e-6,
                                                check_every=200, max_steps=40000, workers=8,
                                                progress=max(1, n_tr // 8))
        t_gen = time.perf_counter() - t0
        conv = np.array([i["residual"] <= 3e-6 for i in info])
        rec["data"] = dict(n=n_tr, seconds=t_gen, ms_per_case=1e3 * t_gen / n_tr, tol=3e-6, dtype="c64",
                           converged=int(conv.sum()), steps_mean=float(np.mean([i["steps"] for i in info])),
                           residual_max=float(max(i["residual"] for i in info)))
        print(f"[warmstart] data: {rec['data']}", flush=True)
        keep = np.nonzero(np.array([i["residual"] <= 1e-4 for i in info]))[0]
        X = torch.tensor(c5_features([p_train[i] for i in keep], N), device="cuda:0")
        Y = torch.tensor(W.fields_to_channels(fields[keep]), device="cuda:0")
        del fields
        ws = W.WarmStart(N, 4, 4, width=32, modes=16, layers=4, seed=1)
        print(f"[warmstart] training on {len(keep)} cases", flush=True)
        t0 = time.perf_counter()
        losses = ws.fit(X, Y, epochs=args.ws_epochs, batch=16, lr=args.ws_lr, time_budget=args.ws_time)
        t_train = time.perf_counter() - t0
        rec["train"] = dict(n=int(len(keep)), epochs_run=len(losses), epochs=args.ws_epochs, seconds=t_train,
                            loss_first=losses[0], loss_last=losses[-1])
        del X, Y
        if args.ws_save:
            ws.save(args.ws_save)
    # inference roofline: one case at a time, as in the warm start
    Xte = torch.tensor(c5_features(p_test, N), device="cuda:0")
    ops = ws.net.ops
    ws.predict(Xte[:1])
    torch.cuda.synchronize()
    f0, b0 = ops.flops, ops.bytes
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record()
    for _ in range(reps):
        ws.predict(Xte[:1])
    e1.record()
    torch.cuda.synchronize()
    inf_ms = e0.elapsed_time(e1) / reps
    flops = (ops.flops - f0) / reps
    nbytes = (ops.bytes - b0) / reps
    rec["fno_inference"] = dict(ms=inf_ms, flops=flops, operand_bytes=nbytes,
                                achieved_TFs=flops / (inf_ms * 1e-3) / 1e12,
                                achieved_GBs=nbytes / (inf_ms * 1e-3) / 1e9)
    print(f"[warmstart] inference {rec['fno_inference']}", flush=True)
    # refinement: cold vs warm, FP64, one handle
    it = g.init(c5_problem(g, np.complex128)(p_test[0]), g.StrangSplitting(), (0.0, 0.05), dt=0.05, nsaves=1)
    cold_ms, warm_ms, cold_steps, warm_steps, err, same, outband = [], [], [], [], [], [], []
    fallbacks = 0
    t0 = time.perf_counter()
    for i, p in enumerate(p_test):
        it.set_param(p, ("dispersion", "pump"))
        uc, ic = it.steady_state(tol=1e-8, check_every=50, max_steps=400000)
        uw, iw = W.solve_warm(it, ws, Xte[i:i + 1], tol=1e-8, check_every=50, max_steps=400000)
        fallbacks += iw["start"] != "warm"
        pn = [x[0].cpu().numpy() for x in ws.predict_fields(Xte[i:i + 1], dtype=torch.complex128)]
        nrm = math.sqrt(sum(np.linalg.norm(uc[c]) ** 2 for c in range(2)))
        err.append(math.sqrt(sum(np.linalg.norm(pn[c] - uc[c]) ** 2 for c in range(2))) / nrm)
        same.append(math.sqrt(sum(np.linalg.norm(uw[c] - uc[c]) ** 2 for c in range(2))) / nrm)
        outband.append(band_residual(uc, ws.net.m))
        cold_ms.append(ic["ms"])
        warm_ms.append(iw["ms"])
        cold_steps.append(ic["steps"])
        warm_steps.append(iw["steps"])
        if (i + 1) % max(1, n_te // 8) == 0:
            print(f"  refined {i + 1}/{n_te}  cold {np.mean(cold_ms):.1f} ms  warm {np.mean(warm_ms):.1f} ms  "
                  f"fno err median {np.median(err):.2e}  ({time.perf_counter() - t0:.0f} s)", flush=True)
    diag = warmstart_diagnosis(it, ws, Xte, p_test[:args.ws_diag], torch) if args.ws_diag else None
    it.close()
    cold_ms, warm_ms = np.array(cold_ms), np.array(warm_ms)
    err, same = np.array(err), np.array(same)
    out = {
        "metric": "time to steady state (residual <= 1e-8), cold / FNO-warm",
        "value": float(cold_ms.sum() / warm_ms.sum()),
        "unit": "x (ratio of summed device times)",
        "higher_is_better": True,
        "n_gpus": 1,
        "dtype": "refine c128; FNO bf16 x bf16 -> f32",
        "data": "synthetic: steady states generated by this solver",
        "config": {"workload": "polariton model 256x256, L=100, A~U[0.5,2.5], w~U[10,40], delta~U[0,1.5] "
                               "(seed 0x5EED0005); FNO width 32, 16 modes, 4 layers",
                   "n_train": n_tr, "n_test": n_te, "check_every": 50},
        "cold_ms_mean": float(cold_ms.mean()), "warm_ms_mean": float(warm_ms.mean()),
        "cold_steps_mean": float(np.mean(cold_steps)), "warm_steps_mean": float(np.mean(warm_steps)),
        "per_case_speedup_median": float(np.median(cold_ms / warm_ms)),
        "fno_rel_err": {"median": float(np.median(err)), "p90": float(np.quantile(err, 0.9)),
                        "max": float(err.max())},
        "warm_vs_cold_final_rel_diff_max": float(same.max()),
        "steady_state_out_of_band_max": float(np.max(outband)),
        "cases_other_branch": int((same > 1e-4).sum()),
        "cold_fallbacks": int(fallbacks),
        "controlled_starts": diag,
        "roofline": {"bound": "hbm", "kernel": "FNO inference, one case (gemm_kernel<bf16>, bias/GELU, "
                                               "band limit)",
                     "achieved": rec["fno_inference"]["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": rec["fno_inference"]["achieved_GBs"] / HBM_PEAK_GBS, "traffic": None,
                     "note": "about 90 launches for one case: launch latency, not bandwidth, sets this; "
                             "bench.py --fno-bench gives the batched throughput"},
        **rec,
    }
    print(json.dumps(out))


def size_sweep(g, torch, sizes=(256, 512, 1024, 2048, 4096), target_s=0.5):
    """The headline workload (two-component polariton, FP32 fields, device
    frames every 100 steps) on N x N grids with the grid spacing kept
    (L = 100 N / 512): where the path leaves the latency regime.  Each size
    runs as many steps (a multiple of 100) as fill about target_s."""
    rows = []
    for n in sizes:
        L = L_BOX * n / N_GRID
        shape = (n, n)
        u0 = []
        for _ in range(2):
            t = g.device_empty(shape, "complex64", 0)
            t.zero_()
            u0.append(t)
        p = polariton_param()
        p.update(L=L)
        prob = g.GrossPitaevskiiProblem(tuple(u0), (L, L), dispersion=dispersion, nonlinearity=nonlinearity,
                                        pump=pump, param=p)
        it = g.init(prob, g.StrangSplitting(), (0.0, 100 * 0.05), dt=0.05, nsaves=1)
        it.solve(save_frames=False)  # warm-up and per-step estimate
        torch.cuda.synchronize()
        est = max(it.last_ms() / 100, 1e-3)
        it.close()
        steps = max(100, int(target_s * 1e3 / est) // 100 * 100)
        it = g.init(prob, g.StrangSplitting(), (0.0, steps * 0.05), dt=0.05, nsaves=steps // 100)
        res = tuple(g.device_empty(shape + (it.nsaves + 1,), "complex64", 0) for _ in range(2))
        it.solve(result=res)  # untimed: first touch of the frames; measures the launch plan
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it.solve(result=res)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        plan, sched = it.plan, it.last_schedule
        it.close()
        del res
        gps = n * n * steps / wall
        rows.append(dict(n=n, steps=steps, us_per_step=wall * 1e6 / steps, gps=gps,
                         hbm_equiv_frac=gps * 64 / 1e9 / HBM_PEAK_GBS, state_mib=2 * n * n * 8 / 2 ** 20,
                         plan=dict(x_rows=plan["x_rows"], y_cols=plan["y_cols"],
                                   measured_us_per_step={f"{k[0]},{k[1]}": v
                                                         for k, v in plan["measured_us"].items()}),
                         step_loop=sched))
        print(f"  {n}x{n}: {rows[-1]['us_per_step']:.1f} us/step, {gps:.3e} gps/s, "
              f"{100 * rows[-1]['hbm_equiv_frac']:.1f} % of HBM at 64 B/gps; plan (x_rows, y_cols) = "
              f"({plan['x_rows']}, {plan['y_cols']}), measured us per step "
              f"{ {k: round(v, 2) for k, v in plan['measured_us'].items()} }, {sched}",
              flush=True)
    return {"metric": "grid-point-steps/s by grid size (two-component polariton, FP32)", "rows": rows}


def grid3d_run(g, torch, n=128, target_s=0.5):
    """The polariton terms on an n^3 grid (two components, FP32 fields, box
    L = 100 n / 512 per axis, frames every 100 steps into device memory).  A 3D
    step is four passes (x-pass, FFT_y, the z-pass with exp(-i dt D), IFFT_y):
    128 B per grid-point-step at FP32."""
    L = L_BOX * n / N_GRID
    shape = (n, n, n)
    u0 = []
    for _ in range(2):
        t = g.device_empty(shape, "complex64", 0)
        t.zero_()
        u0.append(t)
    p = polariton_param()
    p.update(L=L)
    prob = g.GrossPitaevskiiProblem(tuple(u0), (L, L, L), dispersion=dispersion, nonlinearity=nonlinearity,
                                    pump=pump, param=p)
    it = g.init(prob, g.StrangSplitting(), (0.0, 100 * 0.05), dt=0.05, nsaves=1)
    it.solve(save_frames=False)
    torch.cuda.synchronize()
    est = max(it.last_ms() / 100, 1e-3)
    it.close()
    steps = max(100, int(target_s * 1e3 / est) // 100 * 100)
    it = g.init(prob, g.StrangSplitting(), (0.0, steps * 0.05), dt=0.05, nsaves=steps // 100)
    res = tuple(g.device_empty(shape + (it.nsaves + 1,), "complex64", 0) for _ in range(2))
    it.solve(result=res)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it.solve(result=res)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    finite = bool(torch.isfinite(torch.view_as_real(res[0][..., -1])).all().item())
    it.close()
    gps = n ** 3 * steps / wall
    return {"workload": f"two-component polariton on {n}^3, FP32, frames every 100 steps (4 passes per step)",
            "value": gps, "unit": "grid-point-steps/s", "ms_per_step": wall * 1e3 / steps, "steps": steps,
            "dtype": "c64", "bytes_per_gps": 128, "hbm_equiv_frac": gps * 128 / 1e9 / HBM_PEAK_GBS,
            "state_mib": 2 * n ** 3 * 8 / 2 ** 20, "finite": finite}


def fno_bench(torch, batch=16, reps=10):
    """Throughput of the warm-start network's kernels alone: the 256x256 FNO
    (width 32, 16 modes, 4 layers, random weights) forward on `batch` cases
    and one training step (forward, backward, Adam) on the same batch.
    FLOPs are counted from the products issued (2 m n k per product)."""
    from gpe_hip_amd.fno import FNO

    net = FNO(256, in_ch=4, out_ch=4, width=32, modes=16, layers=4, seed=0)
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(0x5EED0005)
    x = torch.randn(batch, 4, 256, 256, device="cuda:0", generator=gen)
    t = torch.randn(batch, 4, 256, 256, device="cuda:0", generator=gen)
    out = {}

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        f0, b0 = net.ops.flops, net.ops.bytes
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / n
        fl = (net.ops.flops - f0) / n
        by = (net.ops.bytes - b0) / n
        return dict(ms=ms, gflop=fl / 1e9, tflops=fl / ms / 1e9, operand_gb=by / 1e9, gbs=by / ms / 1e6,
                    flop_per_byte=fl / by)

    out["forward_one_case"] = timed(lambda: net.forward(x[:1]), reps)
    out["forward_batch"] = timed(lambda: net.forward(x), reps)

    def train():
        net.zero_grad()
        net.loss_and_grad(x, t)
        net.adam_step(1e-4)

    out["train_step_batch"] = timed(train, max(2, reps // 2))
    fb = out["forward_batch"]
    return {
        "metric": "FNO forward throughput (warm-start network, 256x256)",
        "value": fb["tflops"],
        "unit": "TFLOP/s",
        "note": "about 8 flop per operand byte, far below the bf16 ridge (~312): the forward is bound by memory; "
                "the roofline object uses operand bytes (each operand read once, outputs written, accumulated "
                "outputs also read)",
        "higher_is_better": True,
        "n_gpus": 1,
        "dtype": "bf16 x bf16 -> f32",
        "data": "synthetic (random weights and inputs)",
        "config": {"workload": f"FNO width 32, 16 modes, 4 layers, {batch} cases of 256x256x4"},
        "roofline": {"bound": "hbm", "kernel": "gemm_kernel<bf16> and bias/GELU (whole forward)",
                     "achieved": fb["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": fb["gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "mfma_frac": fb["tflops"] / BF16_DENSE_PEAK_TFS},
        **out,
    }


def cpu_baseline(seconds=12.0):
    """FP64 numpy oracle (one host core) on a bounded sample of the same
    workload: as many steps as fit in about `seconds`, from the state after a
    short warm-up."""
    sys.path.insert(0, ROOT)
    from oracle import gpe_oracle as orc

    dt = 0.05
    u0 = (np.zeros((N_GRID, N_GRID), complex),) * 2
    prob = orc.Problem(u0, (L_BOX, L_BOX), **terms())
    it = orc.Iterator(prob, (0.0, 1e9), dt, 1)
    it.dt = dt
    t = 0.0
    for _ in range(2):  # warm-up
        t += dt
        it.step(t, dt)
    n = 0
    t0 = time.perf_counter()
    while True:
        t += dt
        it.step(t, dt)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    gps = N_GRID * N_GRID * n / el
    return dict(value=gps, unit="grid-point-steps/s", cores=1, kind="port",
                sample=f"{n} Strang steps of the 512x512 two-component FP64 workload with the numpy oracle "
                       f"({el:.1f} s, single host thread)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-twa", action="store_true", help="skip the batched truncated-Wigner line")
    ap.add_argument("--twa-steps", type=int, default=2000)
    ap.add_argument("--no-sweep", action="store_true", help="skip the uneven-size sweep line")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D grid line")
    ap.add_argument("--plan", type=lambda v: tuple(int(x) for x in v.split(",")), default=None,
                    help="x_rows,y_cols: fix the headline's launch shapes instead of measuring them (profiling "
                         "runs use it: a profiler's per-dispatch cost skews the planner's timings)")
    ap.add_argument("--dtype", default="complex64", choices=["complex64", "complex128"],
                    help="field precision of the headline run")
    ap.add_argument("--warmstart", action="store_true",
                    help="run only the learned warm-start experiment (data, training, cold vs warm refinement)")
    ap.add_argument("--ws-train", type=int, default=4096)
    ap.add_argument("--ws-test", type=int, default=512)
    ap.add_argument("--ws-epochs", type=int, default=40)
    ap.add_argument("--ws-lr", type=float, default=1e-3)
    ap.add_argument("--ws-time", type=float, default=420.0, help="training time budget in seconds")
    ap.add_argument("--ws-save", default=None, help="write the trained surrogate (.npz)")
    ap.add_argument("--ws-load", default=None, help="skip data and training, load a surrogate (.npz)")
    ap.add_argument("--ws-diag", type=int, default=0, help="controlled-start comparison on this many test cases")
    ap.add_argument("--fno-bench", action="store_true", help="run only the warm-start network's throughput")
    ap.add_argument("--size-sweep", action="store_true", help="run only the headline workload at 256..4096 squared")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain device copy of 256 MiB (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the multi-GPU path is design-only)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    if args.calibrate_copy:
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": x.numel() * 4}))
        return
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge

    g = ge.load_package()
    if args.warmstart:
        warmstart_run(g, torch, args)
        return
    if args.fno_bench:
        print(json.dumps(fno_bench(torch)))
        return
    if args.size_sweep:
        print(json.dumps(size_sweep(g, torch)))
        return
    runs = {}
    for dtype in ([args.dtype] + [d for d in ("complex64", "complex128") if d != args.dtype]):
        runs[dtype] = gpu_run(g, torch, dtype, args.steps, args.warmup, plan=args.plan)
    head = runs[args.dtype]
    npts = N_GRID * N_GRID
    gps = npts * args.steps / head["wall"]
    csz = 8 if args.dtype == "complex64" else 16
    b_alg_gps = 2 * 2 * 2 * csz  # passes x (read+write) x components x bytes per complex value
    launch_bytes = 2 * 2 * csz * npts  # one pass: read + write of both components
    dominant = "y" if head["ms_y"] >= head["ms_x"] else "x"
    dur_ms = max(head["ms_x"], head["ms_y"])
    achieved = launch_bytes / (dur_ms * 1e-3) / 1e9
    out = {
        "metric": "grid-point-steps/s",
        "value": gps,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64" if args.dtype == "complex64" else "c128",
        "data": "synthetic (u0 = 0, the polariton parameters above)",
        "config": {"workload": "two-component exciton-polariton Strang step, 512x512, M=2, dt=0.05, "
                               "frame every 100 steps", "grid": [N_GRID, N_GRID], "components": 2,
                   "global_batch": 1, "parallelism": "single GPU",
                   "launch_plan": {"x_rows": head["plan"]["x_rows"], "y_cols": head["plan"]["y_cols"],
                                   "planner": "measure" if args.plan is None else "fixed"},
                   "step_loop": head["schedule"]},
        "hbm_equiv": {"bytes_per_gps": b_alg_gps, "achieved_GBs": gps * b_alg_gps / 1e9,
                      "frac_of_peak": gps * b_alg_gps / 1e9 / HBM_PEAK_GBS},
        "device_ms_per_step": head["dev_ms"] / args.steps,
        "roofline": {
            "bound": "hbm",
            "kernel": "y_pass_kernel (FFT_y, exp(-i dt D), IFFT_y)" if dominant == "y" else
            "x_pass_kernel (IFFT_x, half-steps, FFT_x)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": read_traffic(args.dtype, "y_pass" if dominant == "y" else "x_pass"),
            "avg_launch_us": dur_ms * 1e3,
            "algorithmic_bytes_per_launch": launch_bytes,
            "x_pass_us": head["ms_x"] * 1e3,
            "y_pass_us": head["ms_y"] * 1e3,
            "timing": "HIP events around every launch on the library's stream, each minus the bracketing "
                      "latency o = (x_raw + y_raw - step)/2 = %.2f us, step timed with events around the whole "
                      "loop only (empty event pair: %.2f us)" % (head["ms_overhead"] * 1e3, head["ms_event"] * 1e3),
        },
        "other_precision": {
            d: {"value": npts * args.steps / r["wall"], "ms_per_step": r["wall"] * 1e3 / args.steps,
                "x_pass_us": r["ms_x"] * 1e3, "y_pass_us": r["ms_y"] * 1e3,
                "hbm_equiv_frac": npts * args.steps / r["wall"] * (64 if d == "complex64" else 128) / 1e9
                / HBM_PEAK_GBS}
            for d, r in runs.items() if d != args.dtype
        },
        "finite": all(r["finite"] for r in runs.values()),
    }
    other = {}
    if not args.no_twa:
        other["twa_1024x256sq"] = twa_run(g, torch, args.twa_steps)
    if not args.no_sweep:
        other["sweep_4096x1d"] = sweep_run(g, torch)
    if not args.no_3d:
        other["polariton_128cube"] = grid3d_run(g, torch)
    if other:
        out["other_workloads"] = other
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
