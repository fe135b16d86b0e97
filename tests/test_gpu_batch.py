"""(b) batched free-streaming integrals and the adjoints on the GPU, through
the C ABI, against the oracle.

Bars: delta_nu within 1e-9 relative of the oracle's QAG forward; gradients
within 1e-9 of the oracle's adjoint (relative to the largest entry of each
row for d/d delta_tot), which the oracle's own tests pin by central
differences.  torch.autograd.gradcheck runs on the C-ABI adjoint with
central differences of relative step 1e-6.
"""
import ctypes as C
import math

import numpy as np
import pytest

import knhip
from batch_common import RTOL, Cosmo, adjoint_case, compare_adjoint, mixed_batch, rel
from dnu_common import Pair, synthetic_history

# The persistent batched solver (one launch with device work queues) did not
# complete a single-history solve + adjoint on the device (no fault, no
# progress; cause not established) and was retired from the library: these
# device tests stay skipped and are kept as the record of that path.  The
# batched workloads run through kn_group (tests/test_gpu_group.py), and the
# adjoint through kn_get_delta_nu_adjoint (tests/test_gpu_delta_nu.py).
pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="persistent batched solver retired: its device run did not complete, "
                                      "cause not established")]


@pytest.mark.parametrize("case", [
    dict(name="cspline", masses=[0.15, 0, 0], Na=60, nk=16, amax=0.5, hybrid=None),
    dict(name="linear Na=2", masses=[0.2, 0, 0], Na=2, nk=8, amax=0.03, hybrid=None),
    dict(name="hybrid", masses=[0.3, 0.06, 0.06], Na=50, nk=12, amax=0.9, hybrid=(500.0, 0.3333)),
    dict(name="noisy", masses=[0.3, 0, 0], Na=120, nk=12, amax=1.0, hybrid=None, noise=0.1),
])
def test_get_delta_nu_adjoint_matches_oracle(kn, torch_cuda, case):
    P = Pair(case["masses"], 0.3, case["nk"], 0.01 + case["Na"] / 100.0 + 0.02, hybrid=case["hybrid"])
    logs, rows, k, init = synthetic_history(case["Na"], case["nk"], case["amax"], noise=case.get("noise", 0.0))
    P.set_history(logs, rows, init)
    print("adjoint", case["name"], flush=True)
    gpu, orc = adjoint_case(P, case["amax"], k, case["masses"][0])
    compare_adjoint(case["name"], gpu, orc)
    P.free()


def test_adjoint_ic_only_massless(kn, torch_cuda):
    """mnu = 0: IC piece only; d/d delta_tot vanishes and d/d init is specialJ * (1 + ...)."""
    P = Pair([0.15, 0, 0], 0.3, 8, 0.5)
    logs, rows, k, init = synthetic_history(20, 8, 0.2)
    P.set_history(logs, rows, init)
    gpu, orc = adjoint_case(P, 0.2, k, 0.0)
    dnu, db, ib, mb = gpu
    assert np.all(db == 0) and mb == 0.0
    assert rel(dnu, orc[1]) < 1e-14 and rel(ib, orc[3]) < 1e-14
    P.free()


def test_batch_matches_oracle_mixed(kn, torch_cuda):
    cosmos, specs, B = mixed_batch(16)
    print("batch of %d histories, %d items" % (B.n, B.off_k[B.n]), flush=True)
    B.solve()
    worst = 0.0
    for i, s in enumerate(specs):
        T = cosmos[s["c"]].oracle(s["nk"], s["logs"], s["rows"], s["init"])
        rc, o = T.get_delta_nu(s["amax"], s["k"], s["m"])
        assert rc == 0
        worst = max(worst, rel(B.delta_nu(i), o))
    print("  batched forward: worst rel err vs oracle %.3e" % worst)
    assert worst < RTOL
    # adjoint of every history for random seeds, against the oracle's adjoint on a subset
    g = np.random.default_rng(7).uniform(-1, 1, B.off_k[B.n])
    B.adjoint(g)
    mb = B.mnu_bar()
    for i in range(0, B.n, 5):
        s = specs[i]
        T = cosmos[s["c"]].oracle(s["nk"], s["logs"], s["rows"], s["init"])
        gi = g[B.off_k[i]:B.off_k[i] + s["nk"]]
        orc = T.get_delta_nu_adjoint(s["amax"], s["k"], s["m"], gi)
        gpu = (B.delta_nu(i), B.dtot_bar(i, s["na"]), B.init_bar(i), mb[i])
        compare_adjoint("hist %d (Na=%d nk=%d)" % (i, s["na"], s["nk"]), gpu, orc)
    B.close()
    for c in cosmos:
        c.free()


def test_batch_graph_append_replays(kn, torch_cuda):
    """Capture append + forward + adjoint once; replay three steps with new rows
    written into the captured input buffers between replays; each replay
    matches the oracle on the history extended by the same rows."""
    import torch
    cosmos, specs, B = mixed_batch(4, seed=1, extra=3)
    nitems = B.off_k[B.n]
    d_loga = torch.zeros(B.n, dtype=torch.float64, device="cuda")
    d_rows = torch.zeros(nitems, dtype=torch.float64, device="cuda")
    g = np.random.default_rng(9).uniform(-1, 1, nitems)
    B._put("gbar", 0, g)
    side = torch.cuda.Stream()  # default priority; the null stream cannot be captured
    stream = side.cuda_stream
    assert kn.kn_batch_graph_capture(B.h, 1, None, None, None) == 3002
    assert kn.kn_batch_graph_capture(B.h, 1, C.c_void_p(d_loga.data_ptr()), C.c_void_p(d_rows.data_ptr()),
                                     C.c_void_p(stream)) == 0
    rng = np.random.default_rng(5)
    for step in range(3):
        loga, rows = np.zeros(B.n), np.zeros(nitems)
        for i, s in enumerate(specs):
            new_a = math.exp(s["logs"][-1]) * 1.02
            s["logs"] = np.append(s["logs"], math.log(new_a))
            r = s["rows"][:, -1] * 1.02 * (1 + 0.01 * rng.standard_normal(s["nk"]))
            s["rows"] = np.concatenate([s["rows"], r[:, None]], axis=1)
            s["amax"] = new_a
            loga[i] = math.log(new_a)
            rows[B.off_k[i]:B.off_k[i] + s["nk"]] = r
        d_loga.copy_(torch.from_numpy(loga))
        d_rows.copy_(torch.from_numpy(rows))
        torch.cuda.synchronize()
        print("graph replay", step, flush=True)
        assert kn.kn_batch_graph_launch(B.h, C.c_void_p(stream)) == 0
        torch.cuda.synchronize()
        assert kn.kn_batch_status(B.h) == 0
        worst, worst_m = 0.0, 0.0
        mb = B.mnu_bar()
        for i, s in enumerate(specs):
            T = cosmos[s["c"]].oracle(s["nk"], s["logs"], s["rows"], s["init"])
            rc, o = T.get_delta_nu(s["amax"], s["k"], s["m"])
            worst = max(worst, rel(B.delta_nu(i), o))
            if i % 3 == 0:
                gi = g[B.off_k[i]:B.off_k[i] + s["nk"]]
                orc = T.get_delta_nu_adjoint(s["amax"], s["k"], s["m"], gi)
                worst_m = max(worst_m, abs(mb[i] / orc[4] - 1))
        print("  replay %d: forward %.3e, d/d mnu %.3e" % (step, worst, worst_m))
        assert worst < RTOL and worst_m < RTOL
    # capacity: a fourth append overflows na_max and is reported, not written
    assert kn.kn_batch_graph_launch(B.h, C.c_void_p(stream)) == 0
    torch.cuda.synchronize()
    assert kn.kn_batch_status(B.h) == 3004
    B.close()
    for c in cosmos:
        c.free()


def test_batch_reports_out_of_range_a(kn, torch_cuda):
    """a beyond the last stored log a: the reference's interpolator fails with
    GSL_EDOM -> terminate(2001); the batch reports 2001."""
    c = Cosmo([0.15, 0, 0])
    logs, rows, k, init = synthetic_history(10, 4, 0.2)
    hh = c.history(4, 10, 10, 0.25, 0.25, 0.15)
    B = knhip.Batch([hh], keep=(c,))
    B.load(0, logs, rows, k, init)
    with pytest.raises(RuntimeError, match="code 2001"):
        B.solve()
    B.close()
    c.free()


def test_autograd_gradcheck_and_oracle_gradient(kn, torch_cuda):
    import torch
    P = Pair([0.15, 0, 0], 0.3, 6, 0.5)
    logs, rows, k, init = synthetic_history(12, 6, 0.3, noise=0.05)
    P.set_history(logs, rows, init)
    state = knhip.DeltaNuState(P.d, 0.3, k, 0.15)
    F = knhip.delta_nu_function()
    dt = torch.tensor(rows, dtype=torch.float64, device="cuda", requires_grad=True)
    di = torch.tensor(init, dtype=torch.float64, device="cuda", requires_grad=True)
    m = torch.tensor(0.15, dtype=torch.float64, device="cuda", requires_grad=True)
    print("gradcheck", flush=True)
    ok = torch.autograd.gradcheck(lambda a, b, c: F.apply(a, b, c, state), (dt, di, m), eps=1e-6, atol=1e-10,
                                  rtol=1e-6, nondet_tol=1e-12)
    assert ok
    # against the FP64 CPU reference gradient (oracle adjoint)
    g = torch.linspace(-1, 1, 6, dtype=torch.float64, device="cuda")
    out = F.apply(dt, di, m, state)
    (out * g).sum().backward()
    orc = P.T.get_delta_nu_adjoint(0.3, k, 0.15, g.cpu().numpy())
    e = max(float(np.max(np.abs(dt.grad.cpu().numpy() - orc[2]) / np.max(np.abs(orc[2])))),
            rel(di.grad.cpu().numpy(), orc[3]), abs(float(m.grad) / orc[4] - 1))
    print("  autograd vs oracle gradient: %.3e" % e)
    assert e < RTOL
    state.batch.close()
    P.free()
