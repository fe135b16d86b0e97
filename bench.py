#!/usr/bin/env python3
"""Benchmark of the k-space neutrino hot path on one MI355X.

Step = one PM-step pass of the hot path over a 1024^3 r2c FP64 grid
(1024 x 1024 x 513 complex, 8.6 GB, resident in HBM): (a) k-shell binning into
1024 bins, then (c) the per-mode delta_nu/delta_cdm rescale.  value = 48 B per
mode (16 B read by the bin sweep + 32 B read/write by the rescale sweep) times
the mode count, divided by the time of the two sweeps (HIP events on the null
stream both sweeps run on).

The same line carries the stream ceilings this part sustains for the two
sweeps' access patterns (plain kernels of the grid's size,
libkn_stream_ceiling.so) and the two-sweep bound they imply, and the CPU
baseline: the oracle (the C restatement of the reference's path) built with
the reference's compile flags (-O2 -ffast-math), timed on one core and on all
usable cores of the host (at most 16).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024] [--no-cpu] [--no-ceiling]
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
# HBM bytes per k_rescale dispatch at 1024^3, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch) scaled by what the same counters read for a 2 GiB
# device copy (FETCH_SIZE x2.0000, WRITE_SIZE x1.0000); commands and tables in
# profiles/bin_rescale_1024_hbm_counters.txt.
RESCALE_1024_FETCH_KIB, RESCALE_1024_WRITE_KIB = 4202615.6, 8405006.9
FETCH_SCALE, WRITE_SCALE = 2.0000, 1.0000


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


def cpu_baseline(N, nrbins, box, d_grid, nslab, one_core_slabs=32):
    """The oracle built with the reference's flags on nslab slabs of the same
    grid (bin sums + rescale): one core on the first one_core_slabs slabs, then
    all usable cores (at most 16) on all nslab slabs, contiguous slab ranges
    per thread as the reference's MPI ranks split the slabs."""
    import cpu_timing as CT
    info = CT.host_info()
    F = d_grid[:nslab].cpu().numpy()
    logkk = np.log(np.geomspace(2 * math.pi / box, math.sqrt(3) * math.pi * N / box, 256))
    r = CT.grid_sweeps(F, nrbins, box, logkk, np.full(256, 0.3), 0.05, info["workers"], min(one_core_slabs, nslab))
    del F
    return {"value": round(r["all_cores_GBs"], 4), "unit": "GB/s", "cores": info["workers"], "kind": "port",
            "sample": "%d of %d slabs of the %d^3 grid, bin sums + rescale, oracle built -O2 -ffast-math, %d threads "
                      "over slab ranges, %.2fs" % (nslab, N, N, info["workers"], r["all_cores_s"]),
            "one_core": {"value": round(r["one_core_GBs"], 4), "cores": 1,
                         "sample": "%d slabs, %.2fs" % (r["one_core_slabs"], r["one_core_s"])},
            "host": {"nproc": info["nproc"], "usable": info["usable"], "cpu_model": info["model"]}}


def stream_ceiling(grid, modes, reps=10):
    """GB/s of a plain read-only stream (16 B/mode) and an in-place read+write
    stream (32 B/mode) over the grid's bytes, with and without the
    non-temporal hint (libkn_stream_ceiling.so, null stream, HIP events),
    and the bound they put on two dependent sweeps of 48 B/mode."""
    import torch
    L = C.CDLL(os.path.join(ROOT, "kspace-neutrinos_amd", "libkn_stream_ceiling.so"))
    L.kn_stream_ceiling.restype = C.c_int
    L.kn_stream_ceiling.argtypes = [C.c_void_p, C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_void_p]
    sink = torch.zeros(1, dtype=torch.float64, device="cuda")
    out = {}
    for kind, name, nbytes, slots in ((0, "read", 16, 2), (1, "read_write", 32, 8)):
        for nt in (0, 1):
            assert L.kn_stream_ceiling(grid.data_ptr(), modes, kind, nt, slots, sink.data_ptr()) == 0
            torch.cuda.synchronize()
            ts = []
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                L.kn_stream_ceiling(grid.data_ptr(), modes, kind, nt, slots, sink.data_ptr())
                e1.record()
                e1.synchronize()
                ts.append(e0.elapsed_time(e1) * 1e-3)
            out["%s%s_GBs" % (name, "_nt" if nt else "")] = round(nbytes * modes / float(np.median(ts)) / 1e9, 1)
    t2 = 16.0 * modes / (out["read_nt_GBs"] * 1e9) + 32.0 * modes / (out["read_write_nt_GBs"] * 1e9)
    out["two_sweep_bound_GBs"] = round(48.0 * modes / t2 / 1e9, 1)
    out["two_sweep_bound_frac_of_peak"] = round(out["two_sweep_bound_GBs"] / HBM_PEAK_GBS, 4)
    out["note"] = ("plain kernels of the grid's size: read-only and in-place read+write streams, 512 threads per "
                   "block, contiguous ranges per block; the bound is 48 B/mode over 16 B at the read ceiling plus "
                   "32 B at the read+write ceiling, both non-temporal")
    return out


def dnu_bench(kn, nk, na=2000, calls=3, cpu_k=16, no_cpu=False):
    """get_delta_nu_combined (the per-step free-streaming integral) on a hybrid
    history of the 1024^3 step: nk bins, na stored steps over a in [0.01, 1]
    with delta_tot = a sqrt(P(k)),
    masses 0.06/0.06/0.3 eV with particles from a = 0.3333 (vcrit 500 km/s), so
    two distinct masses: 2 nk QAG integrals per call.  Plain path: one C-ABI call,
    inputs on the host as the reference's caller holds them."""
    from dnu_common import Pair, synthetic_history
    P = Pair([0.06, 0.06, 0.3], 0.3, nk, 0.01 + na / 100.0 + 0.02, hybrid=(500.0, 0.3333))
    logs, rows, k, init = synthetic_history(na, nk, 1.0, noise=0.0)  # delta_tot = a sqrt(P(k))
    P.set_history(logs, rows, init)
    out = np.zeros(nk)
    kk = np.ascontiguousarray(k)
    kn.get_delta_nu_combined(C.byref(P.d), 1.0, kk.ctypes.data_as(DP), out.ctypes.data_as(DP))  # warm-up
    t = []
    for _ in range(calls):
        t0 = time.perf_counter()
        kn.get_delta_nu_combined(C.byref(P.d), 1.0, kk.ctypes.data_as(DP), out.ctypes.data_as(DP))
        t.append(time.perf_counter() - t0)
    tg = float(np.median(t))
    # the QAG kernel alone, HIP events around its launch on its stream
    kn.kn_dnu_qag_timing(1)
    tq = []
    for _ in range(calls):
        kn.get_delta_nu_combined(C.byref(P.d), 1.0, kk.ctypes.data_as(DP), out.ctypes.data_as(DP))
        tq.append(kn.kn_dnu_last_qag_ms())
    kn.kn_dnu_qag_timing(0)
    items = C.c_int()
    total = kn.kn_dnu_last_intervals(C.byref(items))
    evals = 61 * (2 * total - items.value)          # QAG: the 61-point rule 2n - 1 times per item
    flops = 570.0 * evals                           # hybrid integrand, flops per evaluation (see DESIGN)
    tqm = float(np.median(tq)) * 1e-3
    ksteps = 2.0 * nk * na
    res = {"workload": "get_delta_nu_combined, hybrid 0.06/0.06/0.3 eV, %d k x %d steps, a=1" % (nk, na),
           "ms_per_call": round(tg * 1e3, 3), "kstep_per_s": round(ksteps / tg, 1),
           "note": "k.step = one (k bin, stored step) pair per distinct mass; plain C-ABI call, host in/out",
           "roofline": {"bound": "fp64", "kernel": "k_dnu_qag_mix", "achieved": round(flops / tqm / 1e12, 3),
                        "peak": 78.6, "unit": "TFLOP/s", "frac": round(flops / tqm / 1e12 / 78.6, 4),
                        "kernel_ms": round(tqm * 1e3, 4), "qag_items": items.value,
                        "intervals": int(total), "evaluations": int(evals),
                        "evaluations_per_kstep": round(evals / ksteps, 3),
                        "note": "algorithmic FP64 work: integrand evaluations (61 (2n - 1) per item ending "
                                "with n subintervals) x 570 flops (hybrid integrand, add/mul 1, fma 2, div "
                                "and sqrt 10, transcendentals 20), over the QAG kernel's HIP-event time; "
                                "the kernel is bound by its serial bisection chain, not FP64 issue"}}
    if not no_cpu:
        import cpu_timing as CT
        sel = np.linspace(0, nk - 1, cpu_k).astype(int)
        Q = Pair([0.06, 0.06, 0.3], 0.3, cpu_k, 0.01 + na / 100.0 + 0.02, hybrid=(500.0, 0.3333))
        Q.set_history(logs, rows[sel], init[sel])
        rc, o = Q.T.get_delta_nu_combined(1.0, k[sel])       # the strict oracle, for the error
        res["max_rel_err_vs_oracle"] = float(np.max(np.abs(out[sel] / o - 1)))
        Q.free()
        amp = rows[sel, -1]                                   # delta_tot = amp a, a_max = 1
        r = CT.dnu_one_core([([0.06, 0.06, 0.3], 0.3, (500.0, 0.3333), na, amp, k[sel], 1.0, None)])
        res["cpu_baseline"] = {"kstep_per_s": round(r["kstep_per_s"], 1), "cores": 1, "kind": "port",
                               "sample": "%d of the %d k bins, oracle built -O2 -ffast-math, get_delta_nu_combined "
                                         "%.2fs (the reference computes this on every rank, one core each)"
                                         % (cpu_k, nk, r["s"])}
    P.free()
    return res


class _HCtx(C.Structure):
    _fields_ = [("omnu", C.c_void_p), ("Omega_nonu", C.c_double), ("OmegaLambda", C.c_double),
                ("Hubble", C.c_double)]


def build_histories(kn, nhist, rng, draw):
    """nhist tables with their own Omega_nu and H(a): draw(rng) -> (na, nk, m, Omega_m).
    delta_tot = D(a) sqrt(P(k)) (1 + 0.1 N(0,1)) with D(a) = a and the noise drawn
    per k (smooth along a), log-spaced a in [0.01, 1] and k in
    [2 pi/512, sqrt(3) pi 1024/512] h/Mpc; returns (keep, hists, data) with the
    kn_batch_history of each table."""
    import conftest
    host = conftest.load_testhost()
    import knhip
    host.kn_test_hctx_init.argtypes = [C.POINTER(_HCtx), C.c_void_p, C.c_double, C.c_double]
    hfn = C.cast(host.kn_test_hubble_ctx, C.c_void_p).value
    keep, hists, data = [], [], []
    for h in range(nhist):
        na, nk, m, om0 = draw(rng)
        om = knhip.OmegaNu()
        kn.init_omega_nu(C.byref(om), (C.c_double * 3)(m, 0.0, 0.0), 0.01, 0.7, 2.7255)
        d = knhip.DeltaTotTable()
        kn.allocate_delta_tot_table(C.byref(d), nk, 0.01, 0.01 + na / 100.0 + 0.02, om0, C.byref(om),
                                    conftest.UNIT_TIME, conftest.UNIT_LENGTH, 0)
        sf, rows, init = knhip.table_views(d)
        logs = np.linspace(math.log(0.01), 0.0, na)
        k = np.exp(np.linspace(math.log(2 * math.pi / 512), math.log(math.sqrt(3) * math.pi * 1024 / 512), nk)) / 1e3
        q = k * 1000 / 0.2
        t = np.log(1 + 2.34 * q) / (2.34 * q) * (1 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25
        amp = np.sqrt(1e9 * (k * 1000) ** 0.96 * t * t) * (1 + 0.1 * rng.standard_normal(nk))
        sf[:na] = logs
        rows[:nk, :na] = np.outer(amp, np.exp(logs))
        init[:] = 0.4 * rows[:nk, 0]
        d.ia = na
        ctx = _HCtx()
        host.kn_test_hctx_init(C.byref(ctx), C.addressof(om), om0, conftest.UNIT_TIME)
        hh = knhip.BatchHistory()
        assert kn.kn_batch_history_from_table(C.byref(d), 1.0, m, C.byref(hh)) == 0
        hh.hubble, hh.hubble_ctx = hfn, C.addressof(ctx)
        keep.append((om, d, ctx))
        hists.append(hh)
        data.append(dict(na=na, nk=nk, m=m, om0=om0, logs=logs, rows=np.array(rows[:nk, :na]), k=k,
                         init=np.array(init), amp=amp))
    return keep, hists, data


def plain_vs_group(kn, torch, keep, hists, data, reps=3, seed=5, background=False):
    """The same histories through the per-history path (get_delta_nu, then
    kn_get_delta_nu_adjoint, per history; host tables) and through one kn_group
    (solve + adjoint, eager and as a replayed graph; inputs resident in HBM)."""
    import conftest
    host = conftest.load_testhost()
    import knhip
    nhist = len(hists)
    g_all = np.random.default_rng(seed).uniform(-1, 1, sum(x["nk"] for x in data))
    plain_out, t_fwd, t_adj, off = [], 0.0, 0.0, 0
    for (om, d, ctx), x in zip(keep, data):
        host.kn_test_init_hubble(C.byref(om), x["om0"], conftest.UNIT_TIME)
        out = np.zeros(x["nk"])
        kk = np.ascontiguousarray(x["k"])
        t0 = time.perf_counter()
        kn.get_delta_nu(C.byref(d), 1.0, kk.ctypes.data_as(DP), out.ctypes.data_as(DP), x["m"])
        t_fwd += time.perf_counter() - t0
        gb = np.ascontiguousarray(g_all[off:off + x["nk"]])
        db, ib, mb = np.zeros(x["nk"] * x["na"]), np.zeros(x["nk"]), C.c_double()
        t0 = time.perf_counter()
        rc = kn.kn_get_delta_nu_adjoint(C.byref(d), 1.0, kk.ctypes.data_as(DP), x["m"], gb.ctypes.data_as(DP),
                                        None, db.ctypes.data_as(DP), ib.ctypes.data_as(DP), C.byref(mb))
        t_adj += time.perf_counter() - t0
        assert rc == 0
        plain_out.append(out)
        off += x["nk"]
    t0 = time.perf_counter()
    B = knhip.Batch(hists, keep=keep, api="kn_group")
    t_create = time.perf_counter() - t0
    for i, x in enumerate(data):
        B.load(i, x["logs"], x["rows"], x["k"], x["init"])
    B._put("gbar", 0, g_all)
    modes = {}
    for ns in (0, 4):  # merged launches, and per-history chains on 4 streams
        assert kn.kn_group_set_streams(B.h, ns) == 0
        B.solve()
        B.adjoint()
        t0 = time.perf_counter()
        for _ in range(2):
            assert kn.kn_group_solve(B.h, None) == 0
            assert kn.kn_group_adjoint(B.h, None) == 0
        modes["merged" if ns == 0 else "streams%d" % ns] = round((time.perf_counter() - t0) / 2, 4)
    assert kn.kn_group_autotune(B.h, None) == 0
    nstream = kn.kn_group_streams(B.h)
    B.solve()
    B.adjoint()
    torch.cuda.synchronize()
    ts_f, ts_fa = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        assert kn.kn_group_solve(B.h, None) == 0
        t1 = time.perf_counter()
        assert kn.kn_group_adjoint(B.h, None) == 0
        t2 = time.perf_counter()
        ts_f.append(t1 - t0)
        ts_fa.append(t2 - t0)
    s = torch.cuda.Stream()
    assert kn.kn_group_graph_capture(B.h, 1, None, None, C.c_void_p(s.cuda_stream)) == 0
    assert kn.kn_group_graph_launch(B.h, C.c_void_p(s.cuda_stream)) == 0
    s.synchronize()
    tg = []
    for _ in range(reps):
        t0 = time.perf_counter()
        kn.kn_group_graph_launch(B.h, C.c_void_p(s.cuda_stream))
        s.synchronize()
        tg.append(time.perf_counter() - t0)
    assert kn.kn_group_status(B.h) == 0
    dev = B.delta_nu()
    diff = max(float(np.max(np.abs(dev[B.off_k[i]:B.off_k[i + 1]] / plain_out[i] - 1))) for i in range(nhist))
    B.close()
    bgres = None
    if background:  # the same histories with H from device backgrounds (Omega_nu and H tabulated on the GPU)
        t0 = time.perf_counter()
        bgs, hb = [], []
        for hh, x in zip(hists, data):
            bg = knhip.Background(x["om0"], 0.7, 3.24077929e-18 * conftest.UNIT_TIME, [x["m"], 0.0, 0.0], a0=0.01)
            h2 = knhip.BatchHistory.from_buffer_copy(hh)
            h2.hubble, h2.hubble_ctx = bg.hubble_fn, bg.ctx
            bgs.append(bg)
            hb.append(h2)
        t1 = time.perf_counter()
        Bb = knhip.Batch(hb, keep=bgs, api="kn_group")
        t2 = time.perf_counter()
        for i, x in enumerate(data):
            Bb.load(i, x["logs"], x["rows"], x["k"], x["init"])
        Bb.solve()
        db = Bb.delta_nu()
        bgres = {"create_s": round(t2 - t0, 4), "backgrounds_s": round(t1 - t0, 4),
                 "max_rel_diff_vs_host_h": float(np.max(np.abs(db / dev - 1))),
                 "note": "one kn_background per history (its density table by 200 device QAGs) and the group's "
                         "H tables built on the device in two launches, against create with the host's H "
                         "sampled per history"}
        Bb.close()
        for bg in bgs:
            bg.close()
    steps = pm_steps(kn, torch, keep, hists, data, plain_out, g_all)
    for om, d, ctx in keep:
        kn.free_delta_tot_table(C.byref(d))
    ksteps = sum(x["nk"] * x["na"] for x in data)
    tf, tfa, tgr = float(np.median(ts_f)), float(np.median(ts_fa)), float(np.median(tg))
    return {"ksteps": int(ksteps),
            "plain": {"forward_s": round(t_fwd, 4), "forward_adjoint_s": round(t_fwd + t_adj, 4),
                      "forward_kstep_per_s": round(ksteps / t_fwd, 1),
                      "note": "one get_delta_nu + one kn_get_delta_nu_adjoint C-ABI call per history, host in/out, "
                              "H(a) table rebuilt per history (one hubble_function per process)"},
            "group": {"forward_s": round(tf, 4), "forward_adjoint_s": round(tfa, 4),
                      "graph_forward_adjoint_s": round(tgr, 4), "create_s": round(t_create, 3),
                      "streams": nstream, "forward_adjoint_s_by_mode": modes,
                      "forward_kstep_per_s": round(ksteps / tf, 1),
                      "note": "kn_group: histories resident in HBM; the autotuner's launch mode (streams 0: merged "
                              "launches, one grid per phase for all histories; n > 0: per-history launch chains on "
                              "n HIP streams); create (one H table per history) once, outside the step"},
            "speedup_forward_adjoint": round((t_fwd + t_adj) / min(tfa, tgr), 2),
            "speedup_forward": round(t_fwd / tf, 2),
            "max_rel_diff_group_vs_plain": diff,
            "create_on_device_background": bgres,
            "pm_step": steps}


def pm_steps(kn, torch, keep, hists, data, plain_out, g_all, nsteps=8):
    """The per-PM-step update of the same histories: each created with its
    last nsteps stored steps missing (na_max = the full count), then stepped
    nsteps times, each step appending the next stored column and solving +
    differentiating at its a.  Eager: kn_group_append (host checks) +
    kn_group_solve + kn_group_adjoint per step.  Graph: the group on its
    capacity plan, one captured graph of append + solve + adjoint replayed per
    step (the step's log a and column copied device to device into the
    captured arrays first).  Both end at a = 1 with every stored step, where
    they are compared with the per-history path."""
    import knhip
    nhist = len(hists)
    short = []
    for hh, x in zip(hists, data):
        h2 = knhip.BatchHistory.from_buffer_copy(hh)
        h2.na = x["na"] - nsteps
        h2.na_max = x["na"]
        h2.a = math.exp(x["logs"][x["na"] - nsteps - 1])
        h2.a_max = 1.0
        short.append(h2)
    res, outs = {}, {}
    for mode in ("eager", "graph"):
        B = knhip.Batch(short, keep=keep, api="kn_group")
        for i, x in enumerate(data):
            B.load(i, x["logs"][:x["na"] - nsteps], x["rows"][:, :x["na"] - nsteps], x["k"], x["init"])
        B._put("gbar", 0, g_all)
        off = np.array(B.off_k[:nhist + 1])
        la = torch.tensor(np.array([[x["logs"][x["na"] - nsteps + t] for x in data] for t in range(nsteps)]),
                          dtype=torch.float64, device="cuda")
        cols = np.zeros((nsteps, int(off[-1])))
        for i, x in enumerate(data):
            for t in range(nsteps):
                cols[t, off[i]:off[i + 1]] = x["rows"][:, x["na"] - nsteps + t]
        cols = torch.tensor(cols, dtype=torch.float64, device="cuda")
        s = torch.cuda.Stream()
        if mode == "graph":
            B.set_capacity_plan(True)
            B.stage_step(np.zeros(nhist), np.zeros(int(off[-1])))
            B.capture(with_adjoint=True, with_append=True, stream=C.c_void_p(s.cuda_stream))
        torch.cuda.synchronize()
        ts = []
        for t in range(nsteps):
            t0 = time.perf_counter()
            if mode == "eager":
                assert kn.kn_group_append(B.h, C.c_void_p(la[t].data_ptr()), C.c_void_p(cols[t].data_ptr()),
                                          None) == 0
                assert kn.kn_group_solve(B.h, None) == 0
                assert kn.kn_group_adjoint(B.h, None) == 0
            else:
                with torch.cuda.stream(s):
                    B.step_buf[0].copy_(la[t])
                    B.step_buf[1].copy_(cols[t])
                B.graph_launch(C.c_void_p(s.cuda_stream))
                s.synchronize()
            ts.append(time.perf_counter() - t0)
        assert kn.kn_group_status(B.h) == 0
        outs[mode] = (B.delta_nu(), B.mnu_bar())
        res[mode + "_s_per_step"] = round(float(np.median(ts)), 5)
        B.close()
    dev = outs["graph"][0]
    res["graph_equals_eager"] = bool(np.array_equal(outs["graph"][0], outs["eager"][0]) and
                                     np.array_equal(outs["graph"][1], outs["eager"][1]))
    res["max_rel_diff_vs_plain"] = max(float(np.max(np.abs(dev[B.off_k[i]:B.off_k[i + 1]] / plain_out[i] - 1)))
                                       for i in range(nhist))
    res["steps"] = nsteps
    res["note"] = ("per PM step: append one stored step to every history, solve and differentiate; eager = "
                   "kn_group_append + solve + adjoint; graph = one replay of the captured append + solve + adjoint "
                   "(capacity plan: launches sized for na_max)")
    return res


def c4_cpu_baseline(data, nsample=4, nall=64):
    """The oracle built with the reference's flags on whole histories of the
    same workload (every k bin, each history its own cosmology): one core on
    nsample histories spread over the set, all usable cores (at most 16) on
    nall histories, one history per job."""
    import cpu_timing as CT
    info = CT.host_info()

    def job(h):
        dd = data[h]
        return ([dd["m"], 0.0, 0.0], dd["om0"], None, dd["na"], dd["amp"], dd["k"], 1.0, dd["m"])
    one = np.linspace(0, len(data) - 1, nsample).astype(int)
    allh = np.linspace(0, len(data) - 1, min(nall, len(data))).astype(int)
    r1 = CT.dnu_one_core([job(h) for h in one])
    rn = CT.dnu_all_cores([job(h) for h in allh], info["workers"])
    return {"kstep_per_s": round(rn["kstep_per_s"], 1), "cores": info["workers"], "kind": "port",
            "sample": "%d whole histories of the %d (every %dth), oracle built -O2 -ffast-math get_delta_nu, one "
                      "history per process, %.2fs" % (len(allh), len(data), max(1, len(data) // len(allh)), rn["s"]),
            "one_core": {"kstep_per_s": round(r1["kstep_per_s"], 1), "cores": 1,
                         "sample": "%d whole histories (%s), %.2fs"
                                   % (nsample, ", ".join("%d: nk %d Na %d" % (h, data[h]["nk"], data[h]["na"])
                                                         for h in one), r1["s"])},
            "host": {"nproc": info["nproc"], "usable": info["usable"], "cpu_model": info["model"]}}


def c4_batching(kn, torch, nhist=256, seed=1000, reps=3, no_cpu=False):
    """256 independent histories: Na ~ U[100, 3000], nk ~ U[64, 512], one
    species m ~ U[0.06, 0.6] eV, Omega_m ~ U[0.25, 0.35], each with its own
    H(a) (noise drawn per k: drawn per (k, a) entry it makes the reference's own
    QAG fail, GSL error -> terminate 2001, on long histories)."""
    draw = lambda r: (int(r.integers(100, 3001)), int(r.integers(64, 513)), float(r.uniform(0.06, 0.6)),
                      float(r.uniform(0.25, 0.35)))
    keep, hists, data = build_histories(kn, nhist, np.random.default_rng(seed), draw)
    res = plain_vs_group(kn, torch, keep, hists, data, reps, background=True)
    if not no_cpu:
        res["cpu_baseline"] = c4_cpu_baseline(data)
    res["workload"] = ("%d histories, Na~U[100,3000], nk~U[64,512], m~U[0.06,0.6] eV, Omega_m~U[0.25,0.35]; "
                       "delta_tot = D(a) sqrt(P(k)) (1 + 0.1 N(0,1)) with D(a) = a and the noise drawn per k "
                       "(drawn per (k, a) entry it makes the reference's own QAG fail, terminate 2001)" % nhist)
    return res


def c5_group(kn, torch, nhist=64, seed=2000, reps=3):
    """The differentiation workload's batch: 64 histories of 1024 k x 2000
    steps, m ~ U[0.06, 0.6] eV (Omega_m = 0.3), forward + adjoint."""
    draw = lambda r: (2000, 1024, float(r.uniform(0.06, 0.6)), 0.3)
    keep, hists, data = build_histories(kn, nhist, np.random.default_rng(seed), draw)
    res = plain_vs_group(kn, torch, keep, hists, data, reps)
    res["workload"] = "%d histories of 1024 k x 2000 steps, m~U[0.06,0.6] eV, Omega_m=0.3" % nhist
    return res


def c5_adjoints(kn, torch):
    """Adjoint timings at the differentiation workload's shapes: get_delta_nu's
    adjoint for one 1024 k x 2000 step history (m = 0.3 eV), and the binning
    adjoint on a 256^3 grid (one forward pass of 16 B/mode + one 32 B/mode sweep)."""
    from dnu_common import Pair, synthetic_history
    import knhip
    nk, na = 1024, 2000
    P = Pair([0.3, 0.0, 0.0], 0.3, nk, 0.01 + na / 100.0 + 0.02)
    logs, rows, k, init = synthetic_history(na, nk, 1.0, noise=0.0)
    P.set_history(logs, rows, init)
    g = np.random.default_rng(3).uniform(-1, 1, nk)
    db, ib, dnu, mb = np.zeros(nk * na), np.zeros(nk), np.zeros(nk), C.c_double()
    kk = np.ascontiguousarray(k)
    args = (C.byref(P.d), 1.0, kk.ctypes.data_as(DP), 0.3, g.ctypes.data_as(DP), dnu.ctypes.data_as(DP),
            db.ctypes.data_as(DP), ib.ctypes.data_as(DP), C.byref(mb))
    assert kn.kn_get_delta_nu_adjoint(*args) == 0
    t = []
    for _ in range(3):
        t0 = time.perf_counter()
        kn.kn_get_delta_nu_adjoint(*args)
        t.append(time.perf_counter() - t0)
    ta = float(np.median(t))
    P.free()
    N, nrbins = 256, 256
    d = torch.empty((N, N, N // 2 + 1), dtype=torch.complex128, device="cuda")
    assert kn.kn_synth_grid(d.data_ptr(), N, 0, N, 512000.0, 1e3, 0.96, 0.2 / 1000.0, 5) == 0
    power, keffs = np.zeros(nrbins), np.zeros(nrbins)
    count = np.zeros(nrbins, dtype=np.int64)
    n = kn.total_powerspectrum(N, d.data_ptr(), nrbins, 0, N, power.ctypes.data_as(DP),
                               count.ctypes.data_as(C.POINTER(C.c_longlong)), keffs.ctypes.data_as(DP), 0)
    gb = np.ascontiguousarray(np.random.default_rng(4).uniform(-1, 1, n))
    fb = torch.empty_like(d)
    kn.kn_bin_power_adjoint(d.data_ptr(), N, nrbins, 0, N, gb.ctypes.data_as(DP), n, fb.data_ptr(), None)
    t = []
    for _ in range(5):
        t0 = time.perf_counter()
        kn.kn_bin_power_adjoint(d.data_ptr(), N, nrbins, 0, N, gb.ctypes.data_as(DP), n, fb.data_ptr(), None)
        t.append(time.perf_counter() - t0)
    tb = float(np.median(t))
    modes = N * N * (N // 2 + 1)
    return {"delta_nu_adjoint": {"workload": "kn_get_delta_nu_adjoint, 1 species 0.3 eV, %d k x %d steps" % (nk, na),
                                 "ms_per_call": round(ta * 1e3, 3), "kstep_per_s": round(nk * na / ta, 1),
                                 "note": "includes the forward with recorded subdivisions; host in/out"},
            "bin_adjoint": {"workload": "kn_bin_power_adjoint, 256^3, 256 bins",
                            "ms_per_call": round(tb * 1e3, 3), "GB_per_s": round(48.0 * modes / tb / 1e9, 1),
                            "note": "forward bin pass (16 B/mode) + adjoint sweep (32 B/mode), host sync included"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--nrbins", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ceiling", action="store_true", help="skip the stream-ceiling measurement")
    ap.add_argument("--cpu-slabs", type=int, default=256)
    ap.add_argument("--no-dnu", action="store_true", help="skip the free-streaming integral timings")
    ap.add_argument("--only-dnu", action="store_true",
                    help="only the get_delta_nu_combined timing (757 k x 2000 steps), for profiling")
    ap.add_argument("--only-batch", action="store_true",
                    help="only the 256-history plain vs grouped delta_nu + adjoint timing")
    ap.add_argument("--batch-hists", type=int, default=256, help="histories for --only-batch")
    ap.add_argument("--only-adjoint-batch", action="store_true",
                    help="only the 64-history 1024 k x 2000 step plain vs grouped forward + adjoint timing")
    ap.add_argument("--only-ceiling", action="store_true",
                    help="only the stream-ceiling kernels on a 1024^3 grid's bytes, for profiling")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 2 GiB (reference for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the hot path here is the single-GPU kernel set; "
              "this build has no multi-GPU run (the slab decomposition is a design note).", file=sys.stderr)
        sys.exit(2)
    import torch
    if args.copy_calibration:
        nbytes = 2 << 30
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}))
        return
    if args.only_ceiling:
        N = args.n
        modes = N * N * (N // 2 + 1)
        g = torch.zeros((N, N, N // 2 + 1), dtype=torch.complex128, device="cuda")
        print(json.dumps({"stream_ceiling": stream_ceiling(g, modes)}))
        return
    import conftest
    conftest.load_testhost()
    import knhip
    kn = knhip.lib()
    if args.only_dnu:
        print(json.dumps({"delta_nu": dnu_bench(kn, 757, no_cpu=True)}))
        return
    if args.only_batch:
        print(json.dumps({"delta_nu_batch": c4_batching(kn, torch, nhist=args.batch_hists)}))
        return
    if args.only_adjoint_batch:
        print(json.dumps({"group_64x1024x2000": c5_group(kn, torch)}))
        return

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
        "metric": "P(k) bin+rescale GB/s at 1024^3; delta_nu k*step/s batched vs plain @1e-9",
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
                     "frac": round(res_gbs / HBM_PEAK_GBS, 4),
                     "traffic": (round(1024 * (RESCALE_1024_FETCH_KIB * FETCH_SCALE + RESCALE_1024_WRITE_KIB * WRITE_SCALE))
                                 if N == 1024 else None),
                     "kernel": "k_rescale_rows_pf (32 B/mode algorithmic, %d modes per launch)" % modes},
        "bin_sweep": {"ms": round(t_bin.mean() * 1e3, 4), "GB/s": round(bin_gbs, 1),
                      "frac": round(bin_gbs / HBM_PEAK_GBS, 4)},
        "rescale_sweep": {"ms": round(t_res.mean() * 1e3, 4), "GB/s": round(res_gbs, 1)},
        "wall_s": round(wall, 3),
    }
    # the rescale fused with the PM Green's-function multiply that follows it
    # in a force pass, against the two separate sweeps (rescale, then Green's
    # function alone), on the same grid
    def timed(fn, reps=5):
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        return float(np.median(ts))
    asmth = 1.25 * box / N
    t_res = timed(lambda: kn.kn_rescale_device(grid.data_ptr(), N, 0, N, box, tab.data_ptr(), tab.data_ptr() + 8 * n,
                                               n, norm, None))
    t_green = timed(lambda: kn.kn_rescale_green_device(grid.data_ptr(), N, 0, N, box, None, None, 0, 0.0, asmth, None))
    t_fused = timed(lambda: kn.kn_rescale_green_device(grid.data_ptr(), N, 0, N, box, tab.data_ptr(),
                                                       tab.data_ptr() + 8 * n, n, norm, asmth, None))
    out["rescale_green_fused"] = {
        "workload": "rescale + PM Green's function (CiC deconvolution, r_s = 1.25 cells), %d^3" % N,
        "fused_ms": round(t_fused * 1e3, 4), "two_sweeps_ms": round((t_res + t_green) * 1e3, 4),
        "rescale_ms": round(t_res * 1e3, 4), "green_ms": round(t_green * 1e3, 4),
        "fused_GBs": round(32.0 * modes / t_fused / 1e9, 1),
        "note": "one read+write sweep (32 B/mode) in place of two (64 B/mode)"}
    # the Green's factors shrink the field (high-k modes towards denormals): regenerate it
    assert kn.kn_synth_grid(grid.data_ptr(), N, 0, N, box, amp, 0.96, 0.2 / 1000.0, 3) == 0
    if not args.no_ceiling:
        sc = stream_ceiling(grid, modes)
        sc["frac_of_bound"] = round(value / sc["two_sweep_bound_GBs"], 4)
        out["stream_ceiling"] = sc
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N, nrbins, box, grid, min(args.cpu_slabs, N))
    if not args.no_dnu:
        del grid
        torch.cuda.empty_cache()
        out["delta_nu"] = dnu_bench(kn, n, no_cpu=args.no_cpu)
        out["delta_nu_batch"] = c4_batching(kn, torch, no_cpu=args.no_cpu)
        out["adjoints"] = c5_adjoints(kn, torch)
        out["adjoints"]["group_64x1024x2000"] = c5_group(kn, torch)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
