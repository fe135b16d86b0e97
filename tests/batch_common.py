"""Shared helpers of the batched-solver tests (GPU and CPU-emulation builds):
matched product/oracle cosmologies with a per-history H(a) context, mixed
history sets, and comparison of adjoints against the oracle."""
import ctypes as C
import math

import numpy as np

import knhip
import oracle_py as O
from conftest import UNIT_LENGTH, UNIT_TIME, load_testhost
from dnu_common import Pair, synthetic_history

RTOL = 1e-9


def rel(a, b):
    return float(np.max(np.abs(a / b - 1)))


class HCtx(C.Structure):
    _fields_ = [("omnu", C.c_void_p), ("Omega_nonu", C.c_double), ("OmegaLambda", C.c_double),
                ("Hubble", C.c_double)]


class Cosmo:
    """One cosmology on both sides: product omega_nu + per-history H context, oracle omega_nu.
    With background=L (a library build), the histories' H is a kn_background of L
    (Omega_nu and the H panels tabulated on the device) instead of the host callback."""

    def __init__(self, masses, hybrid=None, omega0=0.3, background=None):
        self.P = Pair(masses, omega0, 4, 2.0, hybrid=hybrid)
        self.masses = masses
        self.omega0 = omega0
        host = load_testhost()
        self.ctx = HCtx()
        host.kn_test_hctx_init.argtypes = [C.POINTER(HCtx), C.c_void_p, C.c_double, C.c_double]
        host.kn_test_hctx_init(C.byref(self.ctx), C.addressof(self.P.om), omega0, UNIT_TIME)
        self.hfn = C.cast(host.kn_test_hubble_ctx, C.c_void_p).value
        self.hctx = C.addressof(self.ctx)
        self.bg = None
        if background is not None:
            self.bg = knhip.Background(omega0, 0.7, 3.24077929e-18 * UNIT_TIME, masses, a0=0.01, L=background)
            self.hfn, self.hctx = self.bg.hubble_fn, self.bg.ctx

    def history(self, nk, na, na_max, a, a_max, mnu):
        """BatchHistory as get_delta_nu derives it (kn_batch_history_from_table on the product table)."""
        d = self.P.d
        d.ia = na
        hh = knhip.BatchHistory()
        assert knhip.lib().kn_batch_history_from_table(C.byref(d), a, mnu, C.byref(hh)) == 0
        hh.nk, hh.na, hh.na_max, hh.a_max = nk, na, na_max, a_max
        hh.hubble = self.hfn
        hh.hubble_ctx = self.hctx
        return hh

    def oracle(self, nk, logs, rows, init):
        O.lib().orc_init_hubble(C.byref(self.P.oom), self.omega0, UNIT_TIME)
        na = len(logs)
        T = O.Table(self.P.oom, nk, 0.01, max(2.0, 0.01 + na / 100.0 + 0.02), self.omega0, UNIT_TIME, UNIT_LENGTH)
        T.scalefact[:na] = logs
        T.delta_tot[:, :na] = rows
        T.delta_nu_init[:] = init
        T.d.ia = na
        return T

    def free(self):
        if self.bg is not None:
            self.bg.close()
        self.P.free()


def adjoint_case(P, a, k, m, seed=3, L=None, fn="kn_get_delta_nu_adjoint"):
    nk, na = P.nk, P.d.ia
    g = np.random.default_rng(seed).uniform(-1, 1, nk)
    dnu, db, ib = np.zeros(nk), np.zeros(nk * na), np.zeros(nk)
    mb = C.c_double()
    kk = np.ascontiguousarray(k, dtype=np.float64)
    dp = C.POINTER(C.c_double)
    rc = getattr(L or knhip.lib(), fn)(C.byref(P.d), a, kk.ctypes.data_as(dp), m, g.ctypes.data_as(dp),
                                             dnu.ctypes.data_as(dp), db.ctypes.data_as(dp), ib.ctypes.data_as(dp),
                                             C.byref(mb))
    assert rc == 0
    orc = P.T.get_delta_nu_adjoint(a, k, m, g)
    assert orc[0] == 0
    return (dnu, db.reshape(nk, na), ib, mb.value), orc


def compare_adjoint(name, gpu, orc):
    dnu, db, ib, mb = gpu
    _, odnu, odb, oib, omb = orc
    e_f = rel(dnu, odnu)
    scale = np.maximum(np.max(np.abs(odb), axis=1, keepdims=True), 1e-300)
    e_y = float(np.max(np.abs(db - odb) / scale))
    e_i = float(np.max(np.abs(ib - oib) / np.max(np.abs(oib))))
    e_m = abs(mb / omb - 1) if omb != 0 else abs(mb)
    print("  %s: delta_nu %.2e  d/d delta_tot %.2e  d/d init %.2e  d/d mnu %.2e" % (name, e_f, e_y, e_i, e_m))
    assert e_f < RTOL and e_y < RTOL and e_i < RTOL and e_m < RTOL


def mixed_batch(ncosmo_hists=16, seed=0, extra=0, L=None, memcpy=None, api="kn_group", streams=None,
                background=False):
    """Histories over three cosmologies with mixed lengths and k counts, in one kn_group
    (background: their H from kn_background objects of the same library)."""
    rng = np.random.default_rng(seed)
    bgl = (L or knhip.lib()) if background else None
    cosmos = [Cosmo([0.15, 0, 0], background=bgl), Cosmo([0.1, 0.1, 0.1], background=bgl),
              Cosmo([0.3, 0.06, 0.06], hybrid=(500.0, 0.3333), background=bgl)]
    specs = []
    for i in range(ncosmo_hists * len(cosmos)):
        c = i % len(cosmos)
        na = int(rng.choice([2, 3, 7, 25, 60, 110, 150]))
        nk = int(rng.choice([4, 8, 16, 24]))
        amax = float(np.exp(math.log(0.01) + (na / 150.0) * (math.log(0.95) - math.log(0.01)))) if na > 2 else 0.03
        logs, rows, k, init = synthetic_history(na, nk, amax, seed=i, noise=0.05)
        m = cosmos[c].masses[0] if c != 1 else 0.1
        specs.append(dict(c=c, na=na, nk=nk, amax=amax, logs=logs, rows=rows, k=k, init=init, m=m))
    hists = [cosmos[s["c"]].history(s["nk"], s["na"], s["na"] + extra, s["amax"], 1.0 if extra else s["amax"],
                                    s["m"]) for s in specs]
    B = knhip.Batch(hists, keep=cosmos, L=L, memcpy=memcpy, api=api)
    if streams is not None:
        B.set_streams(streams)
    for i, s in enumerate(specs):
        B.load(i, s["logs"], s["rows"], s["k"], s["init"])
    return cosmos, specs, B




def _sync(stream):
    if stream is not None:
        import torch
        torch.cuda.synchronize()


def append_flow(L=None, memcpy=None, nk=6, steps=6, mode="eager", stream=None, background=False):
    """Histories whose derived state changes while steps are appended, each
    step compared with the oracle's get_delta_nu at the new a (mode "eager":
    kn_group_append then solve; "async": on the capacity plan,
    kn_group_append_async then solve; "graph": one captured graph of append +
    solve + adjoint replayed per step, the step staged before each replay):
      0: created with one stored step (initial-condition piece only), then integrates;
      1: hybrid (vcrit 500 km/s, particles from a = 0.3333): qc and the tolerance switch on;
      2: hybrid with vcrit 1e5 km/s: past a = 0.3333 particles are nearly all
         the neutrinos and only the initial-condition piece remains;
      3: massless, then given a mass by set_mass before the last step.
    background: the histories' H from kn_background objects of the same library.
    Returns the worst relative error over all steps."""
    bgl = (L or knhip.lib()) if background else None
    cosmos = [Cosmo([0.2, 0, 0], background=bgl), Cosmo([0.3, 0, 0], hybrid=(500.0, 0.3333), background=bgl),
              Cosmo([0.3, 0, 0], hybrid=(1e5, 0.3333), background=bgl)]
    rng = np.random.default_rng(17)
    a_start = [0.0101, 0.25, 0.25, 0.25]
    specs = []
    for i, c in enumerate([0, 1, 2, 0]):
        m = cosmos[c].masses[0] if i != 3 else 0.0
        if i == 0:
            logs = np.array([math.log(0.01)])
            k = np.geomspace(2e-4, 5e-2, nk)
            rows = np.outer(np.linspace(1.0, 2.0, nk), [1.0])
            init = 0.4 * rows[:, 0]
        else:
            logs, rows, k, init = synthetic_history(20, nk, a_start[i], seed=40 + i, noise=0.02)
        specs.append(dict(c=c, m=m, logs=logs, rows=rows, k=k, init=init, a=a_start[i]))
    hists = [cosmos[s["c"]].history(nk, len(s["logs"]), len(s["logs"]) + steps + 1, s["a"], 0.6, s["m"]) for s in specs]
    B = knhip.Batch(hists, keep=cosmos, L=L, memcpy=memcpy)
    for i, s in enumerate(specs):
        B.load(i, s["logs"], s["rows"], s["k"], s["init"])
    if mode != "eager":
        B.set_capacity_plan(True)
    if mode == "graph":
        B._put("gbar", 0, np.zeros(B.off_k[B.n]))
        B.capture(with_adjoint=True, with_append=True, stream=stream)
    worst = 0.0
    for step in range(steps + 1):
        if step:
            loga, rows = np.zeros(B.n), np.zeros(B.off_k[B.n])
            for i, s in enumerate(specs):
                s["a"] = 0.25 + 0.05 * step if i else math.exp(s["logs"][-1]) * 1.5
                r = s["rows"][:, -1] * (1.05 + 0.01 * rng.standard_normal(nk))
                s["logs"] = np.append(s["logs"], math.log(s["a"]))
                s["rows"] = np.concatenate([s["rows"], r[:, None]], axis=1)
                loga[i] = math.log(s["a"])
                rows[B.off_k[i]:B.off_k[i] + nk] = r
            if mode == "eager":
                B.append(loga, rows)
            elif mode == "async":
                B.append_async(loga, rows)
            else:
                B.stage_step(loga, rows)
        if step == steps:
            specs[3]["m"] = 0.15
            B.set_mass(3, 0.15)
        if mode == "graph" and step:
            B.graph_launch(stream)
            _sync(stream)
            assert B.status() == 0
        else:
            B.solve()
        for i, s in enumerate(specs):
            T = cosmos[s["c"]].oracle(nk, s["logs"], s["rows"], s["init"])
            rc, o = T.get_delta_nu(s["a"], s["k"], s["m"])
            assert rc == 0
            e = rel(B.delta_nu(i), o)
            print("  step %d hist %d (Na=%d a=%.4f m=%g): rel err %.2e" % (step, i, len(s["logs"]), s["a"], s["m"], e))
            worst = max(worst, e)
    # a step that does not follow the last stored one is rejected before anything changes
    sf = B._get("scalefact", 0, B.off_s[B.n])
    bad = np.array([s["logs"][-1] for s in specs])
    if mode == "eager":
        with pytest_raises_code(3002):
            B.append(bad, np.zeros(B.off_k[B.n]))
    elif mode == "async":
        B.append_async(bad, np.zeros(B.off_k[B.n]))
        assert B.status() == 3002 and B.status() == 0  # reported once
    else:
        B.stage_step(bad, np.zeros(B.off_k[B.n]))
        B.graph_launch(stream)
        _sync(stream)
        assert B.status() == 3002 and B.status() == 0
    assert np.array_equal(B._get("scalefact", 0, B.off_s[B.n]), sf)
    B.solve()
    for i, s in enumerate(specs):
        T = cosmos[s["c"]].oracle(nk, s["logs"], s["rows"], s["init"])
        rc, o = T.get_delta_nu(s["a"], s["k"], s["m"])
        worst = max(worst, rel(B.delta_nu(i), o))
    B.close()
    for c in cosmos:
        c.free()
    return worst


def pytest_raises_code(code):
    import pytest
    return pytest.raises(RuntimeError, match="code %d" % code)


def compare_subdivisions(B, cosmos, specs):
    """Each history's final QAG subdivisions (interval lists in QUADPACK's
    workspace order) against the oracle's recorded ones: the same number of
    intervals and the same end points bit for bit, i.e. the same sequence of
    bisection decisions.  Returns (items compared, total intervals)."""
    items = total = 0
    for i, s in enumerate(specs):
        T = cosmos[s["c"]].oracle(s["nk"], s["logs"], s["rows"], s["init"])
        rc, _, oivl, onint = T.get_delta_nu_record(s["amax"], s["k"], s["m"])
        assert rc == 0
        ivl, nint = B.subdivision(i)
        np.testing.assert_array_equal(nint, onint, err_msg="history %d: interval counts" % i)
        for k in range(s["nk"]):
            n = int(nint[k])
            assert np.array_equal(ivl[k, :n], oivl[k, :n]), "history %d k %d: interval lists differ" % (i, k)
            items += n > 0
            total += n
    return items, total


def checkpoint_flow(tmp_path, L=None, memcpy=None):
    """(f)2 on a group: history states out to host tables (kn_group_export),
    through the reference's state file (save_all_nu_state / read_all_nu_state,
    7 significant digits) and back into another group (kn_group_import).
    Returns (worst error of the file round trip against the oracle on the
    reloaded values, whether the direct export -> import round trip solved to
    the same delta_nu bit for bit)."""
    L = L or knhip.lib()
    kn = knhip.lib()
    cosmos, specs, B = mixed_batch(2, seed=4, extra=3, L=L, memcpy=memcpy)
    B.solve()
    # a second group on the same histories, loaded with other rows
    hists = [cosmos[s["c"]].history(s["nk"], s["na"], s["na"] + 3, s["amax"], 1.0, s["m"]) for s in specs]
    B2 = knhip.Batch(hists, keep=cosmos, L=L, memcpy=memcpy)
    for i, s in enumerate(specs):
        B2.load(i, s["logs"], s["rows"] * 1.5, s["k"], s["init"])
    worst, same = 0.0, True
    for i, s in enumerate(specs):
        c = cosmos[s["c"]]
        d = knhip.DeltaTotTable()
        kn.allocate_delta_tot_table(C.byref(d), s["nk"], 0.01, 2.0, c.omega0, C.byref(c.P.om), UNIT_TIME,
                                    UNIT_LENGTH, 0)
        assert L.kn_group_export(B.h, i, C.byref(d)) == 0
        sf, dt, init = knhip.table_views(d)
        na = len(s["logs"])
        assert d.ia == na and d.nk == s["nk"]
        assert np.array_equal(sf[:na], s["logs"]) and np.array_equal(dt[:s["nk"], :na], s["rows"])
        assert np.array_equal(init, s["init"])
        assert np.array_equal(np.ctypeslib.as_array(d.wavenum, shape=(s["nk"],)), s["k"])
        assert np.array_equal(np.ctypeslib.as_array(d.delta_nu_last, shape=(s["nk"],)), B.delta_nu(i))
        # straight back into the other group: the same solve
        assert L.kn_group_import(B2.h, i, C.byref(d), s["amax"]) == 0
        # through the state file into a fresh table, then into the other group
        path = str(tmp_path / ("state%d.txt" % i))
        kn.save_all_nu_state(C.byref(d), path.encode())
        d2 = knhip.DeltaTotTable()
        kn.allocate_delta_tot_table(C.byref(d2), s["nk"], 0.01, 2.0, c.omega0, C.byref(c.P.om), UNIT_TIME,
                                    UNIT_LENGTH, 0)
        kn.read_all_nu_state(C.byref(d2), path.encode())
        assert d2.ia == na
        sf2, dt2, init2 = knhip.table_views(d2)
        init2[:] = s["init"]
        np.ctypeslib.as_array(d2.wavenum, shape=(s["nk"],))[:] = s["k"]
        s["file"] = (sf2[:na].copy(), dt2[:s["nk"], :na].copy())
        s["d2"] = d2
        kn.free_delta_tot_table(C.byref(d))
    B2.solve()
    for i, s in enumerate(specs):
        same = same and np.array_equal(B2.delta_nu(i), B.delta_nu(i))
    for i, s in enumerate(specs):
        # the file's last log a (7 digits) may round past log amax: integrate to the later of the two
        a2 = math.exp(s["file"][0][-1])
        while math.log(a2) < s["file"][0][-1]:  # exp and log round: a2 no earlier than the stored step
            a2 = float(np.nextafter(a2, 2.0))
        s["a2"] = max(s["amax"], a2)
        assert L.kn_group_import(B2.h, i, C.byref(s["d2"]), s["a2"]) == 0
    B2.solve()
    for i, s in enumerate(specs):
        logs2, rows2 = s["file"]
        T = cosmos[s["c"]].oracle(s["nk"], logs2, rows2, s["init"])
        rc, o = T.get_delta_nu(s["a2"], s["k"], s["m"])
        assert rc == 0
        worst = max(worst, rel(B2.delta_nu(i), o))
    # a before the last stored step, and a table too small for the history
    d = specs[0]["d2"]
    assert L.kn_group_import(B2.h, 0, C.byref(d), math.exp(specs[0]["file"][0][-1]) * 0.99) == 3004
    small = knhip.DeltaTotTable()
    kn.allocate_delta_tot_table(C.byref(small), 2, 0.01, 2.0, 0.3, C.byref(cosmos[0].P.om), UNIT_TIME,
                                UNIT_LENGTH, 0)
    assert L.kn_group_export(B2.h, 0, C.byref(small)) == 3002
    kn.free_delta_tot_table(C.byref(small))
    for s in specs:
        kn.free_delta_tot_table(C.byref(s["d2"]))
    B.close()
    B2.close()
    for c in cosmos:
        c.free()
    return worst, same
