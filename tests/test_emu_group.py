"""(b) grouped get_delta_nu (kn_group_*) in the CPU SIMT emulation build of
csrc/delta_nu.hip (tests/emu), against the oracle.

A group solves many histories with the per-call kernels, one HIP stream per
history slot; in the emulation every kernel and copy runs when issued, so
these tests check the per-history launch sequences, offsets, packing of
histories with spare capacity, the initial-condition kernels, the adjoint
bookkeeping and the host-side checks.  Bars as on the device: 1e-9 relative.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import knhip
from batch_common import RTOL, Cosmo, compare_adjoint, mixed_batch, rel
from conftest import ROOT
from dnu_common import synthetic_history

DNU_EMU_LIB = os.path.join(ROOT, "tests", "emu", "_build", "libkn_dnu_emu.so")


def host_memcpy(dst, src, n, kind):
    C.memmove(dst, src, n)


@pytest.fixture(scope="module")
def emu(kn):
    if not os.path.exists(DNU_EMU_LIB):
        raise RuntimeError("emulation build missing: make -C tests/emu")
    return knhip.bind(C.CDLL(DNU_EMU_LIB))


def _forward_errors(B, cosmos, specs):
    worst = 0.0
    for i, s in enumerate(specs):
        T = cosmos[s["c"]].oracle(s["nk"], s["logs"], s["rows"], s["init"])
        rc, o = T.get_delta_nu(s["amax"], s["k"], s["m"])
        assert rc == 0
        worst = max(worst, rel(B.delta_nu(i), o))
    return worst


@pytest.mark.parametrize("streams", [0, 4], ids=["merged", "streams4"])
def test_emu_group_mixed_forward_and_adjoint(emu, streams):
    """Twelve histories over three cosmologies (one hybrid), Na from 2 to 150,
    some with spare capacity (rows read and written at stride na_max),
    forward and adjoint, in merged launches and on four streams."""
    cosmos, specs, B = mixed_batch(4, seed=2, extra=3, L=emu, memcpy=host_memcpy, api="kn_group", streams=streams)
    assert B.streams() == streams
    B.solve()
    worst = _forward_errors(B, cosmos, specs)
    print("  %d histories: worst forward rel err %.3e" % (B.n, worst))
    assert worst < RTOL
    g = np.random.default_rng(7).uniform(-1, 1, B.off_k[B.n])
    B.adjoint(g)
    mb = B.mnu_bar()
    for i, s in enumerate(specs):
        T = cosmos[s["c"]].oracle(s["nk"], s["logs"], s["rows"], s["init"])
        gi = g[B.off_k[i]:B.off_k[i] + s["nk"]]
        orc = T.get_delta_nu_adjoint(s["amax"], s["k"], s["m"], gi)
        compare_adjoint("hist %d (Na=%d nk=%d)" % (i, s["na"], s["nk"]),
                        (B.delta_nu(i), B.dtot_bar(i, s["na"]), B.init_bar(i), mb[i]), orc)
        # entries beyond the stored steps are zero
        hs = B.hists[i]
        full = B._get("dtot_bar", B.off_y[i], hs.nk * hs.na_max).reshape(hs.nk, hs.na_max)
        assert np.all(full[:, s["na"]:] == 0)
    B.close()
    for c in cosmos:
        c.free()


def test_emu_group_append_steps_and_capacity(emu):
    """Two appended steps per history, then a third that overflows na_max:
    3004 before anything is written, and the group's status stays 3004."""
    cosmos, specs, B = mixed_batch(2, seed=4, extra=2, L=emu, memcpy=host_memcpy, api="kn_group")
    rng = np.random.default_rng(5)
    for step in range(2):
        loga, rows = np.zeros(B.n), np.zeros(B.off_k[B.n])
        for i, s in enumerate(specs):
            new_a = math.exp(s["logs"][-1]) * 1.02
            s["logs"] = np.append(s["logs"], math.log(new_a))
            r = s["rows"][:, -1] * 1.02 * (1 + 0.01 * rng.standard_normal(s["nk"]))
            s["rows"] = np.concatenate([s["rows"], r[:, None]], axis=1)
            s["amax"] = new_a
            loga[i] = math.log(new_a)
            rows[B.off_k[i]:B.off_k[i] + s["nk"]] = r
        B.append(loga, rows)
        B.solve()
        worst = _forward_errors(B, cosmos, specs)
        print("  append step %d: worst rel err %.3e" % (step, worst))
        assert worst < RTOL
    sf_before = B._get("scalefact", 0, B.off_s[B.n])
    with pytest.raises(RuntimeError, match="code 3004"):
        B.append(np.zeros(B.n) + math.log(0.99), np.zeros(B.off_k[B.n]))
    assert np.array_equal(B._get("scalefact", 0, B.off_s[B.n]), sf_before)
    assert B.status() == 3004
    B.close()
    for c in cosmos:
        c.free()


def test_emu_group_set_mass(emu):
    """A new mass for a plain and a hybrid history (qc follows the mass)."""
    cosmos, specs, B = mixed_batch(1, seed=6, L=emu, memcpy=host_memcpy, api="kn_group")
    B.solve()
    for i, s in enumerate(specs):
        if s["na"] > 2:
            s["m"] = s["m"] * 1.3
            B.set_mass(i, s["m"])
    B.solve()
    worst = _forward_errors(B, cosmos, specs)
    print("  after set_mass: worst rel err %.3e" % worst)
    assert worst < RTOL
    B.close()
    for c in cosmos:
        c.free()


def test_emu_group_out_of_range_a_reports_2001(emu):
    """An upper limit beyond the stored steps makes the integrand leave the
    delta_tot table: GSL_EDOM -> 2001, with the history named."""
    c = Cosmo([0.15, 0, 0])
    logs, rows, k, init = synthetic_history(10, 4, 0.2)
    hh = c.history(4, 10, 10, 0.25, 0.25, 0.15)
    B = knhip.Batch([hh], keep=(c,), L=emu, memcpy=host_memcpy, api="kn_group")
    B.load(0, logs, rows, k, init)
    with pytest.raises(RuntimeError, match="code 2001.*history 0"):
        B.solve()
    B.close()
    c.free()


def test_emu_group_argument_checks(emu):
    c = Cosmo([0.15, 0, 0])
    hh = c.history(4, 10, 10, 0.25, 0.25, 0.15)
    g = C.c_void_p()
    big = knhip.BatchHistory.from_buffer_copy(hh)
    big.na_max = 20000  # long histories: the adjoint splits each row in parts
    assert emu.kn_group_create(C.byref(g), 1, C.byref(big)) == 0
    emu.kn_group_destroy(g)
    bad = knhip.BatchHistory.from_buffer_copy(hh)
    bad.na_max = 3  # fewer stored steps than na
    assert emu.kn_group_create(C.byref(g), 1, C.byref(bad)) == 3002
    bad = knhip.BatchHistory.from_buffer_copy(hh)
    bad.a = bad.time_transfer
    assert emu.kn_group_create(C.byref(g), 1, C.byref(bad)) == 3002
    assert emu.kn_group_create(C.byref(g), 1, C.byref(hh)) == 0
    # a step needs both d_loga and d_rows; a null stream cannot be captured; no graph yet
    assert emu.kn_group_graph_capture(g, 0, C.c_void_p(8), None, C.c_void_p(1)) == 3002
    assert emu.kn_group_graph_capture(g, 0, None, None, None) == 3002
    assert emu.kn_group_graph_launch(g, None) == 3002
    assert emu.kn_group_set_mass(g, 1, 0.2) == 3002 and emu.kn_group_set_mass(g, 0, 0.0) == 3002
    emu.kn_group_destroy(g)
    c.free()


def test_emu_fslength_two_passes_equal_one_qag_per_node(emu):
    """The fslength nodes' first pass (two nodes per wavefront, the first
    Gauss-Kronrod estimate accepted by QAG's own test) plus the second pass
    (full QAG for the rest) give the values of one full QAG per node, bit for
    bit: for a smooth H(a) (no node needs the second pass) and for an H(a)
    with a sharp step near a = 0.1 (the nodes whose range crosses it do)."""
    lib = C.CDLL(DNU_EMU_LIB)
    HUB = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
    lib.emu_fs_two_pass_check.restype = C.c_int
    lib.emu_fs_two_pass_check.argtypes = [HUB, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_double,
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]
    smooth = HUB(lambda a, ctx: 0.1 * math.sqrt(0.3 / a ** 3 + 0.7))
    step = HUB(lambda a, ctx: 0.1 * math.sqrt(0.3 / a ** 3 + 0.7) * (1.3 + 0.3 * math.tanh((math.log(a) - math.log(0.1)) / 0.002)))
    nfs = 48
    for name, fn, want_listed in (("smooth", smooth, False), ("step", step, True)):
        two, full = (C.c_double * nfs)(), (C.c_double * nfs)()
        listed = lib.emu_fs_two_pass_check(fn, None, math.log(0.01), 0.0, nfs, 1.0, two, full)
        a, b = np.array(two[:]), np.array(full[:])
        print("  %s H: %d of %d nodes to the second pass" % (name, listed, nfs))
        assert np.array_equal(a, b)
        assert np.all(a[:-1] > 0) and a[-1] == 0.0
        assert (listed > 0) == want_listed


def test_emu_fslength_across_the_density_switch_matches_the_oracle():
    """The host's neutrino density switches from its spline table to the
    non-relativistic series at a = 100 kT/m (omega_nu_single.c:171-183), where
    H jumps by ~1e-8 relative.  For m = 0.379 eV that point, ln a = -3.115,
    lies inside the integration range: the H panel that holds it is split
    there, so the fslength nodes (two passes over the panel table) match the
    oracle's per-node QAG on the exact H to rounding.  Without the split the
    nodes whose Gauss-Kronrod points fall in that panel were 2e-10 off, and
    the adjoint's d/d delta_tot near the switch 1e-9 off."""
    import oracle_py as O
    lib = C.CDLL(DNU_EMU_LIB)
    lib.emu_fs_two_pass_check.restype = C.c_int
    lib.emu_fs_two_pass_check.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_double,
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]
    m, om0 = 0.379, 0.3
    c = Cosmo([m, 0.0, 0.0], omega0=om0)
    x_switch = math.log(100 * c.P.om.kBtnu / m)
    x0, nfs = math.log(0.01), 2400
    assert x0 < x_switch < 0.0
    two, full = (C.c_double * nfs)(), (C.c_double * nfs)()
    lib.emu_fs_two_pass_check(c.hfn, C.addressof(c.ctx), x0, 0.0, nfs, 1.0, two, full)
    O.lib().orc_init_hubble(C.byref(c.P.oom), om0, 3.085678e21 / 1e5)
    xs = x0 + np.arange(nfs) * (0.0 - x0) / (nfs - 1.0)
    orc = np.array([O.lib().orc_fslength(float(v), 0.0, 1.0) for v in xs[:-1]])
    err = float(np.max(np.abs(np.array(two[:-1]) / orc - 1)))
    print("  fslength at %d nodes across ln a = %.4f: max rel err %.2e" % (nfs, x_switch, err))
    assert err < 1e-13
    assert np.array_equal(np.array(two[:]), np.array(full[:]))
    c.free()


def test_emu_group_capacity_plan_across_the_adjoint_part_boundary(emu):
    """A history stepped from 1022 to 1026 stored steps on the capacity plan:
    the adjoint's knot parts go from one of 1022 knots to two of 513 (the
    part length is not monotone in the step count), and the launch sized for
    na_max must still hold the longest part; every step's forward and
    gradients against the stored-step plan (bit for bit) and the oracle."""
    c = Cosmo([0.2, 0.0, 0.0])
    na0, nk, steps = 1022, 3, 4
    logs, rows, k, init = synthetic_history(na0 + steps, nk, 1.0, seed=31, noise=0.0)
    outs = []
    for cap in (False, True):
        hh = c.history(nk, na0, na0 + steps, math.exp(logs[na0 - 1]), 1.0, 0.2)
        B = knhip.Batch([hh], keep=(c,), L=emu, memcpy=host_memcpy, api="kn_group")
        B.load(0, logs[:na0], rows[:, :na0], k, init)
        if cap:
            B.set_capacity_plan(True)
        g = np.array([0.3, -1.0, 0.7])
        res = []
        for t in range(steps):
            if cap:
                B.append_async([logs[na0 + t]], rows[:, na0 + t])
            else:
                B.append([logs[na0 + t]], rows[:, na0 + t])
            B.solve()
            B.adjoint(g)
            na = na0 + t + 1
            res.append((B.delta_nu(0), B.dtot_bar(0, na), B.init_bar(0), B.mnu_bar()))
            if cap:
                T = c.oracle(nk, logs[:na], rows[:, :na], init)
                orc = T.get_delta_nu_adjoint(math.exp(logs[na - 1]), k, 0.2, g)
                compare_adjoint("step %d (Na=%d)" % (t, na), res[-1][:3] + (res[-1][3][0],), orc)
        assert B.status() == 0
        outs.append(res)
        B.close()
    for ra, rb in zip(*outs):
        for a, b in zip(ra, rb):
            assert np.array_equal(a, b)
    c.free()


def _group_arrays(B):
    n_y = B.off_y[B.n]
    return (B.delta_nu(), B._get("dtot_bar", 0, n_y), B._get("init_bar", 0, B.off_k[B.n]), B.mnu_bar())


def test_emu_group_merged_equals_streams_bit_for_bit(emu):
    """The merged launches run each history's blocks with the per-history
    launches' device code and inputs: identical results, forward and adjoint."""
    out = []
    for streams in (0, 2):
        cosmos, specs, B = mixed_batch(3, seed=9, extra=2, L=emu, memcpy=host_memcpy, api="kn_group", streams=streams)
        B.solve()
        B.adjoint(np.random.default_rng(3).uniform(-1, 1, B.off_k[B.n]))
        out.append(_group_arrays(B))
        B.close()
        for c in cosmos:
            c.free()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_emu_group_merged_with_ic_only_histories(emu):
    """Massless and one-step histories (initial-condition piece only) between
    integrated ones, in merged launches: forward against the oracle, adjoint
    rows of dL/d delta_tot zero for them, init_bar and mnu_bar as the oracle."""
    c = Cosmo([0.15, 0, 0])
    cases = [(20, 8, 0.2, 0.15), (20, 8, 0.2, 0.0), (1, 6, 0.011, 0.15), (30, 12, 0.3, 0.15), (12, 70, 0.15, 0.0)]
    hists, data = [], []
    for i, (na, nk, amax, m) in enumerate(cases):
        logs, rows, k, init = synthetic_history(na, nk, amax, seed=20 + i, noise=0.02) if na > 1 else \
            (np.array([math.log(0.01)]), np.ones((nk, 1)), np.geomspace(1e-3, 1.0, nk), np.ones(nk))
        hists.append(c.history(nk, na, na + 3, amax, 1.0, m))
        data.append((na, nk, amax, m, logs, rows, k, init))
    B = knhip.Batch(hists, keep=(c,), L=emu, memcpy=host_memcpy, api="kn_group")
    assert B.streams() == 0
    for i, (na, nk, amax, m, logs, rows, k, init) in enumerate(data):
        B.load(i, logs, rows, k, init)
    B.solve()
    g = np.random.default_rng(11).uniform(-1, 1, B.off_k[B.n])
    B.adjoint(g)
    mb = B.mnu_bar()
    for i, (na, nk, amax, m, logs, rows, k, init) in enumerate(data):
        T = c.oracle(nk, logs, rows, init)
        rc, o = T.get_delta_nu(amax, k, m)
        assert rc == 0
        assert rel(B.delta_nu(i), o) < RTOL
        gi = g[B.off_k[i]:B.off_k[i] + nk]
        hs = B.hists[i]
        full = B._get("dtot_bar", B.off_y[i], hs.nk * hs.na_max).reshape(hs.nk, hs.na_max)
        assert np.all(full[:, na:] == 0)
        if m == 0.0 or na < 2:
            assert np.all(full == 0)
        orc = T.get_delta_nu_adjoint(amax, k, m, gi)
        compare_adjoint("hist %d (Na=%d nk=%d m=%g)" % (i, na, nk, m),
                        (B.delta_nu(i), B.dtot_bar(i, na), B.init_bar(i), mb[i]), orc)
    B.close()
    c.free()


def test_emu_group_set_streams_checks(emu):
    c = Cosmo([0.15, 0, 0])
    hh = c.history(4, 10, 10, 0.25, 0.25, 0.15)
    g = C.c_void_p()
    assert emu.kn_group_create(C.byref(g), 1, C.byref(hh)) == 0
    assert emu.kn_group_streams(g) == 0
    assert emu.kn_group_set_streams(g, 9) == 3002 and emu.kn_group_set_streams(g, -1) == 3002
    assert emu.kn_group_set_streams(g, 8) == 0 and emu.kn_group_streams(g) == 8
    emu.kn_group_destroy(g)
    c.free()


@pytest.mark.parametrize("limit", [200, 31, 8])
def test_emu_wavefront_error_order_equals_quadpack(limit):
    """QAG's error-ordered interval list kept by all 64 lanes (searches by
    ballot, parallel shifts) against QUADPACK's serial dqpsrt (qag_sort), over
    199 bisections whose new error estimates decrease with noise and repeat
    exactly (ties), so both scans stop at equal keys; every list and every
    choice of the next interval identical.  limit 200 reaches the branch that
    maintains fewer than all entries (last >= limit / 2 + 2)."""
    lib = C.CDLL(DNU_EMU_LIB)
    lib.emu_qag_sort_check.restype = None
    lib.emu_qag_sort_check.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(C.c_int)]
    rng = np.random.default_rng(limit)
    nsteps = 199
    levels = np.array([1.0, 0.5, 0.5, 0.25, 0.125, 0.0625])
    errs = rng.choice(levels, 2 * nsteps) * 0.97 ** np.repeat(np.arange(nsteps), 2)
    errs[::7] = errs[1::7][:len(errs[::7])]  # more exact repeats
    e = (C.c_double * len(errs))(*errs)
    bad = (C.c_int * nsteps)()
    lib.emu_qag_sort_check(e, nsteps, limit, bad)
    assert sum(bad[:min(nsteps, limit - 1)]) == 0


@pytest.mark.parametrize("mode", ["eager", "async"])
def test_emu_group_append_rederives_hybrid_and_ic_state(emu, mode):
    """Appends that change what get_delta_nu derives (delta_tot_table.c:513-555):
    a group created at one stored step, hybrid histories crossing nu_crit_time
    (one of them into initial-condition-only), a massless history given a mass,
    and a non-increasing step rejected with nothing changed.  "async": the
    group on its capacity plan (launches for na_max, live sizes read from
    device memory) and the step appended by the device alone."""
    from batch_common import append_flow
    worst = append_flow(L=emu, memcpy=host_memcpy, mode=mode)
    print("  worst rel err %.3e" % worst)
    assert worst < RTOL


def test_emu_group_capacity_plan_equals_the_stored_step_plan(emu):
    """The same histories solved and differentiated with launches for the
    stored steps and with launches for na_max (capacity plan): every output
    bit for bit, before and after a device append."""
    outs = []
    for cap in (False, True):
        cosmos, specs, B = mixed_batch(3, seed=21, extra=3, L=emu, memcpy=host_memcpy, api="kn_group")
        if cap:
            B.set_capacity_plan(True)
        rng = np.random.default_rng(22)
        g = rng.uniform(-1, 1, B.off_k[B.n])
        res = []
        for step in range(2):
            if step:
                loga, rows = np.zeros(B.n), np.zeros(B.off_k[B.n])
                for i, s in enumerate(specs):
                    loga[i] = s["logs"][-1] + 0.02
                    rows[B.off_k[i]:B.off_k[i] + s["nk"]] = s["rows"][:, -1] * 1.03
                if cap:
                    B.append_async(loga, rows)
                else:
                    B.append(loga, rows)
            B.solve()
            B.adjoint(g)
            res += _group_arrays(B)
        assert B.status() == 0
        outs.append(res)
        B.close()
        for c in cosmos:
            c.free()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_emu_group_subdivisions_equal_the_oracle(emu):
    """The device QAG (run by the emulation) makes the oracle's bisection
    decisions: per item the same interval count and the same interval list,
    bit for bit, over three cosmologies (one hybrid) and Na 2..150."""
    from batch_common import compare_subdivisions
    cosmos, specs, B = mixed_batch(4, seed=3, L=emu, memcpy=host_memcpy)
    B.solve()
    items, total = compare_subdivisions(B, cosmos, specs)
    print("  %d items, %d intervals identical" % (items, total))
    assert items > 0
    B.close()
    for c in cosmos:
        c.free()


def test_emu_group_checkpoint_through_host_tables(emu, tmp_path):
    """kn_group_export / kn_group_import with the reference's state file in
    between: the direct round trip solves bit for bit as before; the file's
    values (7 significant digits) solve to the oracle's delta_nu on them."""
    from batch_common import checkpoint_flow
    worst, same = checkpoint_flow(tmp_path, L=emu, memcpy=host_memcpy)
    print("  state file round trip: worst rel err %.2e vs the oracle; direct round trip identical: %s" % (worst, same))
    assert same and worst < RTOL
