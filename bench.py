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


def gpu_run(g, torch, dtype, steps, warmup, dt=0.05):
    """Returns dict with wall/device timing and per-kernel durations."""
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
    it_w.solve()
    it_w.close()
    it = make_iter(steps)
    res = tuple(g.device_empty(shape + (it.nsaves + 1,), dtype, 0) for _ in range(2))
    it.solve(result=res)  # untimed: first touch of the frame buffers
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it.solve(result=res)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = it.last_ms()
    # per-kernel durations: HIP events around every launch of 200 more steps
    import ctypes

    msx, msy, mse = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    g._lib.check(g.load_library().gpe_profile_kernels(it.h, 200, ctypes.byref(msx), ctypes.byref(msy),
                                                      ctypes.byref(mse)))
    finite = all(bool(torch.isfinite(r).all()) for r in res)
    it.close()
    ox, oy = bracket_overhead(msx.value, msy.value, dev_ms / steps)
    return dict(wall=wall, dev_ms=dev_ms, ms_x=msx.value - ox, ms_y=msy.value - oy,
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
    it_w.solve(save_frames=False, ensemble_moments=True)
    it_w.close()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ts, _, (n1, n2) = it.solve(save_frames=False, ensemble_moments=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    import ctypes

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
                mean_density_last_frame=float(n1[0][..., -1].mean()), vacuum_density=1 / (2 * dA))


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
                         "one launch per size and 20-step block", value=pts * steps / wall,
                unit="grid-point-steps/s", ms_total=wall * 1e3, device_ms=dev_ms, steps=steps, points=pts,
                dtype="c64" if dtype == np.complex64 else "c128", host_setup_s=t_setup,
                bound="VALU (each line stays in registers for 20 steps; HBM traffic 16 B per point per 20 steps)")


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
    ap.add_argument("--dtype", default="complex64", choices=["complex64", "complex128"],
                    help="field precision of the headline run")
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
    runs = {}
    for dtype in ([args.dtype] + [d for d in ("complex64", "complex128") if d != args.dtype]):
        runs[dtype] = gpu_run(g, torch, dtype, args.steps, args.warmup)
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
                   "global_batch": 1, "parallelism": "single GPU"},
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
    if other:
        out["other_workloads"] = other
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
