"""CPU-baseline timings: the oracle (the C restatement of the reference's
path) built with the reference's own compile flags (-O2 -ffast-math,
reference Makefile:2; oracle/_build/libkn_oracle_fast.so), timed on one core
and on all the cores this process may use:

* binning + rescale: the slabs split over threads, as the reference splits
  them over MPI ranks (ctypes releases the GIL around each C call);
* the free-streaming integral: k bins or whole histories split over worker
  processes (the oracle keeps its H(a) in process globals), started with
  "spawn" so that no worker inherits the GPU state of the caller.

Imported only by bench.py's cpu_baseline legs; nothing here runs on the GPU
and nothing here is the thing measured there.
"""
import concurrent.futures as cf
import ctypes as C
import math
import multiprocessing
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
UNIT_LENGTH = 3.085678e21          # kpc/h in cm (delta_tot_table_test.c:90)
UNIT_TIME = UNIT_LENGTH / 1e5


def host_info(max_workers=16):
    """Cores usable by this process (capped at max_workers) and the CPU model."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable": avail, "workers": max(1, min(avail, max_workers)), "model": model}


# ------------------------------------------------------------ grid sweeps

def grid_sweeps(F, nrbins, box, logkk, ratio, norm, workers, one_core_slabs):
    """Binning sums + rescale (in place) of the host slabs F[nslab][N][N/2+1]
    with the fast oracle: one core on the first one_core_slabs slabs, then all
    workers on all slabs (contiguous slab ranges, one per thread)."""
    sys.path.insert(0, HERE)
    import oracle_py as O
    L = O.lib_fast()
    nslab, N = F.shape[0], F.shape[1]
    per_slab = N * (N // 2 + 1)

    def chunk(s0, s1):
        view = F[s0:s1]
        O.bin_sums(view, nrbins, startslab=s0, L=L)
        O.rescale(view, box, logkk, ratio, norm, startslab=s0, L=L, in_place=True)

    t0 = time.perf_counter()
    chunk(0, one_core_slabs)
    t1 = time.perf_counter()
    one = 48.0 * one_core_slabs * per_slab / (t1 - t0) / 1e9
    bounds = np.linspace(0, nslab, workers + 1).astype(int)
    with cf.ThreadPoolExecutor(workers) as ex:
        t2 = time.perf_counter()
        list(ex.map(lambda i: chunk(bounds[i], bounds[i + 1]), range(workers)))
        t3 = time.perf_counter()
    allc = 48.0 * nslab * per_slab / (t3 - t2) / 1e9
    return {"one_core_GBs": one, "one_core_s": t1 - t0, "one_core_slabs": one_core_slabs,
            "all_cores_GBs": allc, "all_cores_s": t3 - t2, "all_cores_slabs": int(nslab)}


# --------------------------------------------- free-streaming integrals

def _worker_init():
    os.environ["KN_ORACLE_BUILD"] = "fast"
    sys.path.insert(0, HERE)


def _history_table(O, masses, omega0, hybrid, logs, rows, init):
    oom = O.omega_nu(masses, a0=0.01)
    if hybrid:
        vcrit, tcrit = hybrid
        O.lib().orc_init_hybrid_nu(C.byref(oom.hybnu), (C.c_double * 3)(*masses), vcrit, 2.99792458e10 / 1e5, tcrit,
                                   oom.kBtnu)
    O.lib().orc_init_hubble(C.byref(oom), omega0, UNIT_TIME)
    na, nk = len(logs), rows.shape[0]
    T = O.Table(oom, nk, 0.01, 0.01 + na / 100.0 + 0.02, omega0, UNIT_TIME, UNIT_LENGTH)
    T.scalefact[:na] = logs
    T.delta_tot[:, :na] = rows
    T.delta_nu_init[:] = init
    T.d.ia = na
    return oom, T


def _dnu_job(job):
    """One history: stored steps log a on [log 0.01, log a_max] (na points),
    delta_tot = amp[k] a, delta_nu_init = 0.4 delta_tot at the first step;
    get_delta_nu_combined at a = a_max (single_mass None) or get_delta_nu for
    that mass.  Returns (seconds, k.steps, rc)."""
    import oracle_py as O
    masses, omega0, hybrid, na, amp, k, a_max, single_mass = job
    logs, rows = history_rows(na, len(amp), amp, a_max)
    oom, T = _history_table(O, masses, omega0, hybrid, logs, rows, 0.4 * rows[:, 0])
    t0 = time.perf_counter()
    if single_mass is None:
        rc, _ = T.get_delta_nu_combined(a_max, k)
        nmass = len({m for m in masses if m > 0})
    else:
        rc, _ = T.get_delta_nu(a_max, k, single_mass)
        nmass = 1
    dt = time.perf_counter() - t0
    return dt, nmass * len(amp) * na, rc


def _noop(_):
    return os.getpid()


def _pool(workers):
    return cf.ProcessPoolExecutor(workers, mp_context=multiprocessing.get_context("spawn"), initializer=_worker_init)


def dnu_one_core(jobs):
    """The jobs one after another in one worker process: k.step/s of the
    integrator time (table set-up excluded)."""
    with _pool(1) as ex:
        res = [ex.submit(_dnu_job, j).result() for j in jobs]
    assert all(r[2] == 0 for r in res)
    t = sum(r[0] for r in res)
    return {"kstep_per_s": sum(r[1] for r in res) / t, "s": t, "ksteps": int(sum(r[1] for r in res))}


def dnu_all_cores(jobs, workers):
    """The jobs spread over `workers` processes (each worker started before
    the clock): k.step/s of the wall time."""
    with _pool(workers) as ex:
        list(ex.map(_noop, range(workers)))
        t0 = time.perf_counter()
        res = list(ex.map(_dnu_job, jobs))
        t1 = time.perf_counter()
    assert all(r[2] == 0 for r in res)
    return {"kstep_per_s": sum(r[1] for r in res) / (t1 - t0), "s": t1 - t0, "ksteps": int(sum(r[1] for r in res))}


def history_rows(na, nk, amp, a_max=1.0):
    """log a on [log 0.01, log a_max] (na points) and delta_tot = amp[k] a."""
    logs = np.linspace(math.log(0.01), math.log(a_max), na)
    return logs, np.outer(amp, np.exp(logs))
