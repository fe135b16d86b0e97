#!/usr/bin/env python3
"""Benchmark of the k-space neutrino hot path on one MI355X.

Step = one PM-step pass of the hot path over a 1024^3 r2c FP64 grid
(1024 x 1024 x 513 complex, 8.6 GB, resident in HBM): (a) k-shell binning into
1024 bins, then (c) the per-mode delta_nu/delta_cdm rescale.  value = 48 B per
mode (16 B read by the bin sweep + 32 B read/write by the rescale sweep) times
the mode count, divided by the time of the two sweeps (HIP events on the null
stream both sweeps run on).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "kspace-neutrinos_amd"), os.path.join(ROOT, "tests")]

HBM_PEAK_GBS = 8000.0
DP = C.POINTER(C.c_double)


def bbks_amplitude(N, box, rms=1e-2, ns=0.96, gamma=0.2 / 1000.0):
    """A such that sum over modes (with multiplicity) of A k^ns T^2 = rms^2 N^6."""
    from scipy.signal import fftconvolve
    kv = np.arange(N)
    kv = np.where(kv <= N // 2, kv, kv - N)
    k2xy = (kv[:, None] ** 2 + kv[None, :] ** 2).ravel()
    hxy = np.bincount(k2xy).astype(float)
    kz = np.arange(N // 2 + 1)
    hz = np.zeros(kz[-1] ** 2 + 1)
    hz[kz ** 2] = np.where((kz == 0) | (kz == N // 2), 1.0, 2.0)
    n = np.rint(fftconvolve(hxy, hz)).astype(np.int64)
    k2 = np.nonzero(n)[0]
    k2 = k2[k2 > 0]
    k = np.sqrt(k2) * 2 * math.pi / box
    q = k / gamma
    t = np.log(1 + 2.34 * q) / (2.34 * q) * (1 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25
    shape = np.sum(n[k2] * k ** ns * t * t)
    return rms ** 2 * float(N) ** 6 / shape


def cpu_baseline(N, nrbins, box, d_grid, torch, nslab):
    """The oracle (C restatement, 1 core) on nslab slabs of the same grid."""
    import oracle_py as O
    F = d_grid[:nslab].cpu().numpy()
    modes = F.shape[0] * F.shape[1] * F.shape[2]
    t0 = time.perf_counter()
    n, p, c, k = O.total_powerspectrum(F, nrbins, 0, nslab)
    t1 = time.perf_counter()
    logkk = np.log(k[:n] * 2 * math.pi / box)
    O.rescale(F, box, logkk, np.full(n, 0.3), 0.05)
    t2 = time.perf_counter()
    return {"value": round(48.0 * modes / (t2 - t0) / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": "%d of %d slabs of the %d^3 grid (%d modes): oracle total_powerspectrum %.2fs + rescale %.2fs"
                      % (nslab, N, N, modes, t1 - t0, t2 - t1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--nrbins", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-slabs", type=int, default=192)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the hot path here is the single-GPU kernel set; "
              "multi-GPU sharding is described in DESIGN.md and not run.", file=sys.stderr)
        sys.exit(2)
    import torch
    import conftest
    conftest.load_testhost()
    import knhip
    kn = knhip.lib()

    N, nrbins, box = args.n, args.nrbins, 512000.0
    modes = N * N * (N // 2 + 1)
    grid = torch.empty((N, N, N // 2 + 1), dtype=torch.complex128, device="cuda")
    amp = bbks_amplitude(N, box)
    assert kn.kn_synth_grid(grid.data_ptr(), N, 0, N, box, amp, 0.96, 0.2 / 1000.0, 3) == 0
    d_power = torch.empty(nrbins + 1, dtype=torch.float64, device="cuda")
    # ratio table on the binned k's (a fixed toy table: the rescale sweep's cost does not depend on it)
    power, keffs = np.zeros(nrbins), np.zeros(nrbins)
    count = np.zeros(nrbins, dtype=np.int64)
    n = kn.total_powerspectrum(N, grid.data_ptr(), nrbins, 0, N, power.ctypes.data_as(DP),
                               count.ctypes.data_as(C.POINTER(C.c_longlong)), keffs.ctypes.data_as(DP), 0)
    logkk = np.log(keffs[:n] * 2 * math.pi / box)
    ratio = 0.3 * np.exp(-np.exp(logkk) * 1e3)
    tab = torch.from_numpy(np.concatenate([logkk, ratio])).cuda()
    norm = 1e-6  # keeps the grid values stable over many steps

    def step(ev=None):
        if ev:
            ev[0].record()
        kn.kn_bin_power_device(grid.data_ptr(), N, nrbins, 0, N, d_power.data_ptr(), d_power.data_ptr() + 8 * nrbins,
                               None)
        if ev:
            ev[1].record()
        kn.kn_rescale_device(grid.data_ptr(), N, 0, N, box, tab.data_ptr(), tab.data_ptr() + 8 * n, n, norm, None)
        if ev:
            ev[2].record()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(evs[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_bin = np.array([e[0].elapsed_time(e[1]) for e in evs]) * 1e-3
    t_res = np.array([e[1].elapsed_time(e[2]) for e in evs]) * 1e-3
    t_step = t_bin + t_res
    value = 48.0 * modes / t_step.mean() / 1e9
    res_gbs = 32.0 * modes / t_res.mean() / 1e9
    bin_gbs = 16.0 * modes / t_bin.mean() / 1e9
    out = {
        "metric": "P(k) bin+rescale GB/s at 1024^3",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step.mean() * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "bin+rescale of a %d^3 r2c FP64 Gaussian field (BBKS, 512 Mpc/h), %d log k-bins"
                               % (N, nrbins), "grid_modes": modes, "nrbins": nrbins},
        "roofline": {"bound": "hbm", "achieved": round(res_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(res_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_rescale (32 B/mode algorithmic)"},
        "bin_sweep": {"ms": round(t_bin.mean() * 1e3, 4), "GB/s": round(bin_gbs, 1),
                      "frac": round(bin_gbs / HBM_PEAK_GBS, 4)},
        "rescale_sweep": {"ms": round(t_res.mean() * 1e3, 4), "GB/s": round(res_gbs, 1)},
        "wall_s": round(wall, 3),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N, nrbins, box, grid, torch, min(args.cpu_slabs, N))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
