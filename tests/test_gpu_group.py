"""(b) grouped get_delta_nu on the GPU (kn_group_*): many histories, each with
its own H(a), mass, stored steps and k bins, solved with the per-call kernels'
device code in merged launches (one grid per phase for all histories) and on
several HIP streams; forward and adjoint against the oracle, merged against
streams bit for bit, appended steps, new masses, graph replays and the
autotuner.

Bars: delta_nu and its gradients within 1e-9 relative of the oracle (the
d/d delta_tot rows relative to their largest entry); graph replays and
re-solves bit-identical to the eager solve they repeat.
"""
import ctypes as C
import math

import numpy as np
import pytest

import knhip
from batch_common import RTOL, Cosmo, compare_adjoint, mixed_batch, rel
from dnu_common import synthetic_history

pytestmark = pytest.mark.gpu


def _forward_errors(B, cosmos, specs):
    worst = 0.0
    for i, s in enumerate(specs):
        T = cosmos[s["c"]].oracle(s["nk"], s["logs"], s["rows"], s["init"])
        rc, o = T.get_delta_nu(s["amax"], s["k"], s["m"])
        assert rc == 0
        worst = max(worst, rel(B.delta_nu(i), o))
    return worst


def _free(cosmos):
    for c in cosmos:
        c.free()


def _group_arrays(B):
    return (B.delta_nu(), B._get("dtot_bar", 0, B.off_y[B.n]), B._get("init_bar", 0, B.off_k[B.n]), B.mnu_bar())


@pytest.mark.parametrize("streams", [0, 4], ids=["merged", "streams4"])
@pytest.mark.parametrize("extra", [0, 3])
def test_group_mixed_forward_and_adjoint(kn, torch_cuda, extra, streams):
    """48 histories over three cosmologies (one hybrid), Na 2..150, nk 4..24;
    extra > 0 leaves spare capacity (rows read and written at stride na_max)."""
    print("group mixed, extra capacity %d, streams %d" % (extra, streams), flush=True)
    cosmos, specs, B = mixed_batch(16, seed=2, extra=extra, api="kn_group", streams=streams)
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
    B.close()
    _free(cosmos)


def test_group_append_steps_and_capacity(kn, torch_cuda):
    print("group append", flush=True)
    cosmos, specs, B = mixed_batch(4, seed=4, extra=2, api="kn_group")
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
    with pytest.raises(RuntimeError, match="code 3004"):
        B.append(np.zeros(B.n) + math.log(0.99), np.zeros(B.off_k[B.n]))
    assert B.status() == 3004
    B.close()
    _free(cosmos)


def test_group_set_mass_and_out_of_range(kn, torch_cuda):
    print("group set_mass", flush=True)
    cosmos, specs, B = mixed_batch(2, seed=6, api="kn_group")
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
    _free(cosmos)
    c = Cosmo([0.15, 0, 0])
    logs, rows, k, init = synthetic_history(10, 4, 0.2)
    hh = c.history(4, 10, 10, 0.25, 0.25, 0.15)
    B = knhip.Batch([hh], keep=(c,), api="kn_group")
    B.load(0, logs, rows, k, init)
    print("group out-of-range a", flush=True)
    with pytest.raises(RuntimeError, match="code 2001.*history 0"):
        B.solve()
    B.close()
    c.free()


def test_group_graph_replays_and_autotune(kn, torch_cuda):
    """solve + adjoint captured once on a torch stream and replayed three
    times with eager solves in between (each replay bit-identical to the eager
    result), then a new mass read at replay; the autotuner's choice gives the
    same results bit for bit."""
    torch = torch_cuda
    cosmos, specs, B = mixed_batch(4, seed=8, api="kn_group")
    g = np.random.default_rng(9).uniform(-1, 1, B.off_k[B.n])
    B.solve()
    B.adjoint(g)
    ref = (B.delta_nu(), B._get("dtot_bar", 0, B.off_y[B.n]), B._get("init_bar", 0, B.off_k[B.n]), B.mnu_bar())
    s = torch.cuda.Stream()
    print("group graph capture", flush=True)
    assert kn.kn_group_graph_capture(B.h, 1, None, None, C.c_void_p(s.cuda_stream)) == 0
    for rep in range(3):
        B._put("delta_nu", 0, np.zeros(B.off_k[B.n]))
        B._put("mnu_bar", 0, np.zeros(B.n))
        print("  replay %d" % rep, flush=True)
        assert kn.kn_group_graph_launch(B.h, C.c_void_p(s.cuda_stream)) == 0
        s.synchronize()
        assert kn.kn_group_status(B.h) == 0
        got = (B.delta_nu(), B._get("dtot_bar", 0, B.off_y[B.n]), B._get("init_bar", 0, B.off_k[B.n]), B.mnu_bar())
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
        B.solve()  # eager work between replays
    i = next(j for j, sp in enumerate(specs) if sp["na"] > 2)
    specs[i]["m"] *= 1.2
    B.set_mass(i, specs[i]["m"])
    assert kn.kn_group_graph_launch(B.h, C.c_void_p(s.cuda_stream)) == 0
    s.synchronize()
    replay = B.delta_nu()
    B.solve()
    assert np.array_equal(replay, B.delta_nu())
    worst = _forward_errors(B, cosmos, specs)
    print("  new mass through the graph: worst rel err %.3e" % worst)
    assert worst < RTOL
    print("group autotune", flush=True)
    assert kn.kn_group_autotune(B.h, None) == 0
    ns = kn.kn_group_streams(B.h)
    assert ns in (0, 1, 2, 4, 8)
    before = B.delta_nu()
    B.solve()
    assert np.array_equal(before, B.delta_nu())
    print("  autotuned streams: %d (0: merged launches)" % ns)
    # the graph is stale after autotune (new launch plan)
    assert kn.kn_group_graph_launch(B.h, C.c_void_p(s.cuda_stream)) == 3002
    B.close()
    _free(cosmos)


def test_group_at_batching_workload_shape(kn, torch_cuda):
    """24 histories drawn like the 256-history batching workload (Na ~
    U[100, 3000], nk ~ U[64, 512], m ~ U[0.06, 0.6] eV, Omega_m ~ U[0.25, 0.35],
    per-k noise), forward and adjoint on the GPU; 3 sampled k per history
    against the oracle."""
    rng = np.random.default_rng(1000)
    cosmos, specs, hists = [], [], []
    for h in range(24):
        na, nk = int(rng.integers(100, 3001)), int(rng.integers(64, 513))
        m, om0 = float(rng.uniform(0.06, 0.6)), float(rng.uniform(0.25, 0.35))
        c = Cosmo([m, 0.0, 0.0], omega0=om0)
        logs = np.linspace(math.log(0.01), 0.0, na)
        k = np.exp(np.linspace(math.log(2 * math.pi / 512), math.log(math.sqrt(3) * math.pi * 1024 / 512), nk)) / 1e3
        q = k * 1000 / 0.2
        t = np.log(1 + 2.34 * q) / (2.34 * q) * (1 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25
        amp = np.sqrt(1e9 * (k * 1000) ** 0.96 * t * t) * (1 + 0.1 * rng.standard_normal(nk))
        rows = np.outer(amp, np.exp(logs))
        cosmos.append(c)
        specs.append(dict(na=na, nk=nk, m=m, logs=logs, rows=rows, k=k, init=0.4 * rows[:, 0]))
        hists.append(c.history(nk, na, na, 1.0, 1.0, m))
    B = knhip.Batch(hists, keep=cosmos, api="kn_group")
    for i, s in enumerate(specs):
        B.load(i, s["logs"], s["rows"], s["k"], s["init"])
    print("group at the batching workload shape: %d histories" % B.n, flush=True)
    B.solve()
    # dL/d delta_nu on 3 sampled k per history and zero elsewhere: d/d mnu is
    # then the sum over those k alone, which the oracle gives from their rows
    idxs = [np.array(sorted(rng.choice(s["nk"], 3, replace=False))) for s in specs]
    g = np.zeros(B.off_k[B.n])
    for i, idx in enumerate(idxs):
        g[B.off_k[i] + idx] = rng.uniform(-1, 1, len(idx))
    B.adjoint(g)
    mb = B.mnu_bar()
    worst_f, worst_b, worst_m = 0.0, 0.0, 0.0
    for i, s in enumerate(specs):
        idx = idxs[i]
        T = cosmos[i].oracle(len(idx), s["logs"], s["rows"][idx], s["init"][idx])
        orc = T.get_delta_nu_adjoint(1.0, s["k"][idx], s["m"], g[B.off_k[i] + idx])
        assert orc[0] == 0
        dnu = B.delta_nu(i)[idx]
        worst_f = max(worst_f, rel(dnu, orc[1]))
        db = B.dtot_bar(i, s["na"])[idx]
        scale = np.maximum(np.max(np.abs(orc[2]), axis=1, keepdims=True), 1e-300)
        worst_b = max(worst_b, float(np.max(np.abs(db - orc[2]) / scale)),
                      float(np.max(np.abs(B.init_bar(i)[idx] - orc[3]) / np.max(np.abs(orc[3])))))
        worst_m = max(worst_m, abs(mb[i] / orc[4] - 1))
    print("  sampled k: forward %.3e, d/d delta_tot and d/d init %.3e, d/d mnu %.3e" % (worst_f, worst_b, worst_m))
    assert worst_f < RTOL and worst_b < RTOL and worst_m < RTOL
    merged = _group_arrays(B)
    B.set_streams(4)
    print("  the same on 4 streams", flush=True)
    B.solve()
    B.adjoint(g)
    for a, b in zip(merged, _group_arrays(B)):
        assert np.array_equal(a, b)
    B.close()
    _free(cosmos)


def test_group_with_a_stepped_hubble_matches_the_emulation(kn, torch_cuda):
    """A history whose H(a) has a sharp step near a = 0.1: the fslength nodes
    whose range crosses it need the second (full QAG) pass.  The device result
    is compared with the same group run by the CPU emulation build of the same
    kernels (tests/emu), which the oracle-backed emulation tests pin."""
    import os
    from conftest import ROOT
    emu_path = os.path.join(ROOT, "tests", "emu", "_build", "libkn_dnu_emu.so")
    emu = knhip.bind(C.CDLL(emu_path))
    HUB = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
    hub = HUB(lambda a, ctx: 0.1 * math.sqrt(0.3 / a ** 3 + 0.7) *
              (1.3 + 0.3 * math.tanh((math.log(a) - math.log(0.1)) / 0.002)))
    logs, rows, k, init = synthetic_history(40, 8, 0.5)
    hh = knhip.BatchHistory()
    hh.nk, hh.na, hh.na_max, hh.a, hh.a_max = 8, 40, 40, 0.5, 0.5
    hh.time_transfer, hh.mnu, hh.kBtnu, hh.light = 0.01, 0.15, 1.68e-4, 1.0e5
    hh.delta_nu_prefac, hh.qc, hh.relerr, hh.nufrac_low, hh.vcrit, hh.ic_only = 1e-9, 0.0, 1e-6, 0.0, 0.0, 0
    hh.hubble = C.cast(hub, C.c_void_p).value
    out = []
    for L, mc in ((None, None), (emu, lambda d, s_, n, kind: C.memmove(d, s_, n))):
        B = knhip.Batch([hh], keep=(hub,), L=L, memcpy=mc, api="kn_group")
        B.load(0, logs, rows, k, init)
        print("stepped H: %s" % ("device" if L is None else "emulation"), flush=True)
        B.solve()
        out.append(B.delta_nu(0))
        B.close()
    e = rel(out[0], out[1])
    print("  device vs emulation: %.2e" % e)
    assert e < 1e-12


def test_group_merged_with_ic_only_histories(kn, torch_cuda):
    """Massless and one-step histories (initial-condition piece only) between
    integrated ones in merged launches: forward and adjoint against the
    oracle, their dL/d delta_tot rows zero."""
    c = Cosmo([0.15, 0, 0])
    cases = [(20, 8, 0.2, 0.15), (20, 8, 0.2, 0.0), (1, 6, 0.011, 0.15), (30, 12, 0.3, 0.15), (12, 70, 0.15, 0.0)]
    hists, data = [], []
    for i, (na, nk, amax, m) in enumerate(cases):
        logs, rows, k, init = synthetic_history(na, nk, amax, seed=20 + i, noise=0.02) if na > 1 else \
            (np.array([math.log(0.01)]), np.ones((nk, 1)), np.geomspace(1e-3, 1.0, nk), np.ones(nk))
        hists.append(c.history(nk, na, na + 3, amax, 1.0, m))
        data.append((na, nk, amax, m, logs, rows, k, init))
    B = knhip.Batch(hists, keep=(c,), api="kn_group")
    assert B.streams() == 0
    for i, (na, nk, amax, m, logs, rows, k, init) in enumerate(data):
        B.load(i, logs, rows, k, init)
    print("group merged with IC-only histories", flush=True)
    B.solve()
    g = np.random.default_rng(11).uniform(-1, 1, B.off_k[B.n])
    B.adjoint(g)
    mb = B.mnu_bar()
    for i, (na, nk, amax, m, logs, rows, k, init) in enumerate(data):
        T = c.oracle(nk, logs, rows, init)
        hs = B.hists[i]
        full = B._get("dtot_bar", B.off_y[i], hs.nk * hs.na_max).reshape(hs.nk, hs.na_max)
        assert np.all(full[:, na:] == 0)
        if m == 0.0 or na < 2:
            assert np.all(full == 0)
        orc = T.get_delta_nu_adjoint(amax, k, m, g[B.off_k[i]:B.off_k[i] + nk])
        compare_adjoint("hist %d (Na=%d nk=%d m=%g)" % (i, na, nk, m),
                        (B.delta_nu(i), B.dtot_bar(i, na), B.init_bar(i), mb[i]), orc)
    B.close()
    c.free()


def test_group_merged_equals_streams_bit_for_bit(kn, torch_cuda):
    """Merged launches and per-history launches on two streams run the same
    device code on the same inputs: identical forward and adjoint results."""
    out = []
    for streams in (0, 2):
        print("group mixed, streams %d" % streams, flush=True)
        cosmos, specs, B = mixed_batch(6, seed=9, extra=2, api="kn_group", streams=streams)
        B.solve()
        B.adjoint(np.random.default_rng(3).uniform(-1, 1, B.off_k[B.n]))
        out.append(_group_arrays(B))
        B.close()
        _free(cosmos)
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", ["eager", "async", "graph"])
def test_group_append_rederives_hybrid_and_ic_state(kn, torch_cuda, mode):
    """On the device: a group created at one stored step, hybrid histories
    crossing nu_crit_time (one into initial-condition-only), a massless
    history given a mass, a non-increasing step rejected with nothing changed;
    every step against the oracle (delta_tot_table.c:513-555).  "async": the
    capacity plan and the device-only append; "graph": one captured graph of
    append + solve + adjoint replayed for every step."""
    from batch_common import append_flow
    print("group append flow (%s)" % mode, flush=True)
    s = torch_cuda.cuda.Stream() if mode == "graph" else None
    worst = append_flow(mode=mode, stream=C.c_void_p(s.cuda_stream) if s is not None else None)
    print("  worst rel err %.3e" % worst)
    assert worst < RTOL


def test_group_subdivisions_equal_the_oracle(kn, torch_cuda):
    """The device QAG makes the oracle's bisection decisions: per item the
    same interval count and interval list, bit for bit (48 histories, three
    cosmologies, one hybrid, Na 2..150)."""
    from batch_common import compare_subdivisions
    print("group subdivisions vs oracle", flush=True)
    cosmos, specs, B = mixed_batch(16, seed=3)
    B.solve()
    items, total = compare_subdivisions(B, cosmos, specs)
    print("  %d items, %d intervals identical" % (items, total))
    assert items > 0
    B.close()
    _free(cosmos)


def test_group_checkpoint_through_host_tables(kn, torch_cuda, tmp_path):
    """kn_group_export / kn_group_import on the device, with the reference's
    state file in between (save_all_nu_state / read_all_nu_state): the
    direct round trip solves bit for bit as before, the file's values solve to
    the oracle's delta_nu on them."""
    from batch_common import checkpoint_flow
    print("group checkpoint through host tables", flush=True)
    worst, same = checkpoint_flow(tmp_path)
    print("  state file round trip: worst rel err %.2e vs the oracle; direct round trip identical: %s" % (worst, same))
    assert same and worst < RTOL
